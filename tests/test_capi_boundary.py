"""CPU checks of the C ABI boundary: libqse_hip.so loads, exports every function
include/qse_hip.h declares, and reports errors as status codes (no compute)."""
import re
import subprocess

import numpy as np

from quantumsimevo_amd import capi
from oracle_lib import ROOT


def header_functions():
    text = (ROOT / "include" / "qse_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(qse_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (qse_\w+)", out))
    for name in header_functions():
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_op_layout():
    assert capi.OP_DTYPE.itemsize == 104
    assert capi.OP_DTYPE.fields["controls"][1] == 16
    assert capi.OP_DTYPE.fields["m"][1] == 40


def test_status_strings_and_errors_without_device():
    lib = capi.load()
    assert lib.qse_abi_version() == 1
    assert lib.qse_status_string(0) == b"ok"
    assert lib.qse_status_string(1) == b"invalid argument"
    if capi.device_count() == 0:
        import ctypes
        h = ctypes.c_void_p()
        st = lib.qse_ctx_create(0, ctypes.byref(h))
        assert st == 2  # QSE_ERR_NO_DEVICE
        assert b"no HIP device" in lib.qse_last_error()
    # invalid arguments are rejected before any device work
    assert lib.qse_ctx_create(0, None) == 1


def test_amp_update_counts():
    from quantumsimevo_amd import circuits as C
    n = 10
    N = 1 << n
    g = C.GateList(n).hadamard(0).pauli_z(1).cnot(0, 1).cphase(0.1, 0, 1).mcphase(0.2, [1, 2], 3).swap(1, 2)
    assert C.amp_updates(g.ops(), n) == N + N / 2 + N / 2 + N / 4 + N / 8 + N / 2
    # QFT-30 figure of the worked example: 1.5703e11 amplitude updates
    assert abs(C.amp_updates(C.qft(30).ops(), 30) / 1.5703e11 - 1) < 1e-3
    ops = C.qaoa_maxcut(24, C.random_regular_graph(24, 3, seed=1), 8).ops()
    assert len(ops) == 24 + 8 * (36 * 3 + 24)
    assert np.sum(ops["param"] >= 0) == 8 * (36 + 24)


def test_hardware_check_without_device():
    """HardwareCheck::report with no visible device (HardwareCheck.cpp:35-54's
    'not detected' branch): unavailable, no sizes, host threads counted."""
    from quantumsimevo_amd import _quantumsimevo as Q
    if capi.device_count() > 0:
        return
    r = Q.device_report(0)
    assert not r["available"] and r["max_qubits_c128"] == 0 and r["host_threads"] > 0

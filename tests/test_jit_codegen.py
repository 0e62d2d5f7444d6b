"""CPU checks of the run-time specialised passes (quantumsimevo_amd/csrc/qse_jit_gen.cpp).

The generated HIP source of a planned program is compiled with g++ behind a host
shim (tests/jit_shim: one std::thread per workgroup thread, a barrier for
__syncthreads) and executed pass by pass; the result must reproduce the oracle
(complex128: 1e-12 max abs; complex64: 2e-5). The same source must compile with
hiprtc for gfx950 (compilation needs no GPU)."""
import ctypes as C
import math
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import circuits as G
from test_fused_plan import emu

SHIM = O.ROOT / "tests" / "jit_shim"


def _prepare(n, k, L, ops, params, relabel, prec, table=0):
    lib = emu()
    lib.emu_jit_prepare2.restype = C.c_long
    lib.emu_jit_prepare2.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                                      C.c_char_p, C.c_long]
    pp = None if params is None else np.ascontiguousarray(params, np.float64).ctypes.data
    need = lib.emu_jit_prepare2(n, k, L, ops.ctypes.data, len(ops), pp, relabel, prec, table, None, 0)
    assert need > 0, f"planner status {-need}"
    buf = C.create_string_buffer(need + 1)
    assert lib.emu_jit_prepare2(n, k, L, ops.ctypes.data, len(ops), pp, relabel, prec, table, buf, need + 1) == need
    return buf.value.decode()


def _program():
    """Serialized program of the prepared plan (the device program buffer), 8-byte aligned."""
    lib = emu()
    lib.emu_jit_program.restype = C.c_long
    lib.emu_jit_program.argtypes = [C.c_void_p]
    size = lib.emu_jit_program(None)
    buf = np.zeros((size + 7) // 8, np.float64)
    lib.emu_jit_program(buf.ctypes.data)
    return buf


def _passes():
    lib = emu()
    k, oop, ns = C.c_int(), C.c_int(), C.c_int()
    out = []
    for p in range(lib.emu_jit_pass_info(-1, None, None, None)):
        lib.emu_jit_pass_info(p, C.byref(k), C.byref(oop), C.byref(ns))
        out.append((k.value, oop.value, ns.value))
    return out


def _compile_host(src, npasses, tmp, table_form=0):
    name = "qse_jit_tpass" if table_form else "qse_jit_pass"
    table = "shim_kernel shim_kernels[] = {" + ", ".join(f"{name}{p}" for p in range(npasses)) + "};\n"
    cpp = Path(tmp) / "jit_host.cpp"
    cpp.write_text(f'#include "{SHIM / "prelude.hpp"}"\n' + src + f'#include "{SHIM / "epilogue.hpp"}"\n' + table)
    so = Path(tmp) / "libjit_host.so"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-pthread", "-Wno-attributes",
                        "-ffp-contract=off", "-o", str(so), str(cpp)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    lib = C.CDLL(str(so))
    lib.shim_launch.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_ulonglong, C.c_void_p]
    return lib


def run_jit(n, k, L, ops, psi, params=None, relabel=-1, prec=0, table=0):
    """Generated program executed on the host; returns the final state (complex128).
    table=1: the structure-keyed form, coefficients read from the serialized program."""
    pp = None if params is None else np.ascontiguousarray(params, np.float64)
    assert emu().emu_check_plan(n, k, L, ops.ctypes.data, len(ops), None if pp is None else pp.ctypes.data,
                                relabel) == 0
    src = _prepare(n, k, L, ops, params, relabel, prec, table)
    prog = _program()
    passes = _passes()
    dt = np.complex128 if prec == 0 else np.complex64
    with tempfile.TemporaryDirectory() as tmp:
        lib = _compile_host(src, len(passes), tmp, table)
        cur = np.ascontiguousarray(psi, dtype=dt).copy()
        other = np.zeros_like(cur)
        for p, (kp, oop, ns) in enumerate(passes):
            ntiles = 1 << (n - kp)
            table = np.zeros((max(ntiles * ns, 1), 4))
            if ns:
                emu().emu_jit_slot_table(p, table.ctypes.data_as(C.c_void_p))
            dst = other if oop else cur
            assert lib.shim_launch(p, 1 << (kp - 4), cur.ctypes.data, dst.ctypes.data, table.ctypes.data,
                                   ntiles, prog.ctypes.data) == 0
            if oop:
                cur, other = other, cur
    return cur.astype(np.complex128), src, passes


def check(n, k, L, ops, params=None, prec=0, seed=3, relabels=(-1, 0), table=0):
    psi = G.random_state(n, seed)
    want = O.run_ops(psi, n, ops, params)
    tol = 1e-12 if prec == 0 else 2e-5
    for rl in relabels:
        got, _, passes = run_jit(n, k, L, ops, psi, params, rl, prec, table)
        err = float(np.max(np.abs(got - want)))
        assert err < tol, (n, k, L, rl, prec, err, passes)


def test_qft_and_iqft():
    n = 12
    check(n, 9, 3, G.qft(n).ops())
    check(n, 9, 3, G.iqft(n).ops(), relabels=(1, 3))


def test_random_layered_shapes():
    for (n, k, L) in ((13, 10, 2), (14, 9, 5), (12, 8, 0)):
        check(n, k, L, G.random_layered(n, 6, seed=n + k).ops(), relabels=(-1,))


def test_every_gate_kind():
    n = 11
    ops = G.random_mixed(n, 300, 5).ops()
    check(n, 8, 4, ops, seed=6)
    check(n, 9, 0, ops, seed=6, relabels=(2, 3))


def test_qaoa_slots_and_terms():
    # k = 8 leaves 4 external qubits: ZZ terms on them go through per-tile slots
    n = 12
    edges = G.random_regular_graph(n, 3, seed=1)
    ops = G.qaoa_maxcut(n, edges, 3).ops()
    params = np.random.default_rng(2).random(6) * math.pi
    check(n, 8, 2, ops, params, relabels=(-1,))


def test_complex64():
    n = 12
    check(n, 9, 3, G.random_layered(n, 5, seed=9).ops(), prec=1, relabels=(-1,))
    check(n, 9, 3, G.qft(n).ops(), prec=1, relabels=(-1,))


def test_default_shape_qft14():
    # complex128 default shape: 12 tile qubits, 6 contiguous low qubits
    check(14, 12, 6, G.qft(14).ops(), relabels=(-1,))


def test_hiprtc_compiles_for_gfx950():
    rtc_path = "/opt/rocm/lib/libhiprtc.so"
    if not os.path.exists(rtc_path):
        pytest.skip("hiprtc not installed")
    rtc = C.CDLL(rtc_path)
    for prec, table in ((0, 0), (1, 0), (0, 1), (1, 1)):
        src = _prepare(14, 12, 6, G.qft(14).ops(), None, -1, prec, table)
        prog = C.c_void_p()
        assert rtc.hiprtcCreateProgram(C.byref(prog), src.encode(), b"t.hip", 0, None, None) == 0
        opts = [b"--offload-arch=gfx950", b"-O3", b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
        arr = (C.c_char_p * len(opts))(*opts)
        rc = rtc.hiprtcCompileProgram(prog, len(opts), arr)
        if rc != 0:
            n = C.c_size_t()
            rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
            log = C.create_string_buffer(n.value + 1)
            rtc.hiprtcGetProgramLog(prog, log)
            pytest.fail(log.value.decode()[-4000:])
        size = C.c_size_t()
        assert rtc.hiprtcGetCodeSize(prog, C.byref(size)) == 0 and size.value > 0
        rtc.hiprtcDestroyProgram(C.byref(prog))


def test_structure_keyed_form():
    """Coefficients read from the program buffer (the kernels the adjoint's
    internal runs use): every gate kind, slots and terms, both precisions."""
    n = 12
    ops = G.random_mixed(n, 300, 5).ops()
    check(n, 8, 4, ops, seed=6, relabels=(-1,), table=1)
    edges = G.random_regular_graph(n, 3, seed=1)
    q = G.qaoa_maxcut(n, edges, 3).ops()
    params = np.random.default_rng(2).random(6) * math.pi
    check(n, 8, 2, q, params, relabels=(-1,), table=1)
    check(n, 9, 3, G.random_layered(n, 5, seed=9).ops(), prec=1, relabels=(-1,), table=1)


def test_structure_key_ignores_values():
    lib = emu()
    lib.emu_same_structure.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    n = 12
    edges = G.random_regular_graph(n, 3, seed=1)
    q = G.qaoa_maxcut(n, edges, 2).ops()
    p1 = np.random.default_rng(3).random(4) * math.pi
    p2 = np.random.default_rng(4).random(4) * math.pi
    assert lib.emu_same_structure(n, 9, 3, q.ctypes.data, p1.ctypes.data, q.ctypes.data, p2.ctypes.data, len(q)) == 1
    # another graph (same gate count) is another structure
    other = G.qaoa_maxcut(n, G.random_regular_graph(n, 3, seed=2), 2).ops()
    assert len(other) == len(q)
    assert lib.emu_same_structure(n, 9, 3, q.ctypes.data, p1.ctypes.data, other.ctypes.data, p1.ctypes.data,
                                  len(q)) == 0


@pytest.fixture
def paired():
    """Generate complex128 passes with paired tiles (two tiles per workgroup
    iteration, loads of both issued together)."""
    lib = emu()
    lib.emu_jit_set_pairing.argtypes = [C.c_int]
    lib.emu_jit_set_pairing(1)
    yield
    lib.emu_jit_set_pairing(0)


def test_paired_tiles(paired):
    """Paired-tile passes: QFT with and without the out-of-place relabelling
    pass, a layered circuit, QAOA with per-tile slots and terms, every gate
    kind in the structure-keyed form; complex64 passes stay unpaired."""
    n = 12
    src = _prepare(n, 8, 3, G.qft(n).ops(), None, -1, 0)
    assert "ntiles / 2" in src and "lb15" in src
    check(n, 8, 3, G.qft(n).ops())
    check(n, 9, 2, G.random_layered(n, 6, seed=4).ops(), relabels=(-1,))
    edges = G.random_regular_graph(n, 3, seed=1)
    q = G.qaoa_maxcut(n, edges, 3).ops()
    params = np.random.default_rng(2).random(6) * math.pi
    check(n, 8, 2, q, params, relabels=(-1,))
    check(n, 8, 4, G.random_mixed(n, 300, 5).ops(), seed=6, relabels=(-1,), table=1)
    assert "ntiles / 2" not in _prepare(n, 8, 3, G.qft(n).ops(), None, -1, 1)


@pytest.mark.parametrize("pairing", [0, 1])
def test_direct_load_and_store(pairing):
    """Passes of 4+ stages take the first stage's amplitudes straight from the
    state and store the last stage's straight back where the lane mapping stays
    coalesced: plain, paired (complex128) and prefetching (complex64) kernels,
    literal and structure-keyed forms, against the oracle."""
    import re
    lib = emu()
    lib.emu_jit_set_pairing.argtypes = [C.c_int]
    lib.emu_jit_set_pairing(pairing)
    try:
        n, k, L = 14, 10, 3
        ops = G.random_layered(n, 12, seed=3).ops()
        for prec in (0, 1):
            src = _prepare(n, k, L, ops, None, -1, prec)
            marks = re.findall(r"direct load (\w+), direct store (\w+)", src)
            assert any(a == "yes" for a, _ in marks) and any(b == "yes" for _, b in marks), marks
            check(n, k, L, ops, prec=prec, relabels=(-1,))
        check(n, k, L, ops, relabels=(-1,), table=1)
        check(16, 12, 3, G.random_layered(16, 16, seed=5).ops(), relabels=(-1, 0))
    finally:
        lib.emu_jit_set_pairing(0)

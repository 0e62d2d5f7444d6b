"""CPU checks of the run-time generated joint passes of the adjoint reverse
sweep (jit_joint_source, quantumsimevo_amd/csrc/qse_jit_gen.cpp).

The forward state psi and lambda = O psi come from the oracle; the reverse
sweep's chunks (build_reverse_plan) are emitted as HIP source, compiled with g++
behind the host shim (tests/jit_shim, one std::thread per workgroup thread, one
workgroup, lanes exchanged through a buffer) and run pass by pass on psi and
lambda together. Their probe sums, weighted by the plan's probe references, must
give the oracle's adjoint gradient (orc_adjoint_expect, itself pinned to
parameter shift) to 1e-10, and the swept psi must keep its norm. The same source
must compile for gfx950."""
import ctypes as C
import math
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import circuits as G
from test_fused_plan import emu

SHIM = O.ROOT / "tests" / "jit_shim"


def _lib():
    lib = emu()
    lib.emu_joint_prepare.restype = C.c_long
    lib.emu_joint_prepare.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_char_p, C.c_long]
    lib.emu_joint_program.restype = C.c_long
    lib.emu_joint_program.argtypes = [C.c_void_p]
    lib.emu_joint_slot_table.argtypes = [C.c_int, C.c_void_p]
    return lib


def _compile(src, npasses, tmp):
    table = ("shim_joint_kernel shim_joint_kernels[] = {" + ", ".join(f"qse_jit_jpass{p}" for p in range(npasses)) +
             "};\n")
    cpp = Path(tmp) / "joint_host.cpp"
    cpp.write_text(f'#include "{SHIM / "prelude.hpp"}"\n' + src + f'#include "{SHIM / "epilogue_joint.hpp"}"\n' + table)
    so = Path(tmp) / "libjoint_host.so"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-pthread", "-Wno-attributes",
                        "-ffp-contract=off", "-o", str(so), str(cpp)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    lib = C.CDLL(str(so))
    lib.shim_joint_launch.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_ulonglong,
                                      C.c_void_p, C.c_void_p]
    return lib


def apply_obs(psi, n, masks, weights):
    idx = np.arange(1 << n, dtype=np.uint64)
    lam = np.zeros_like(psi)
    for m, w in zip(masks, weights):
        sign = 1.0 - 2.0 * (np.bitwise_count(idx & np.uint64(m)) & 1)
        lam += w * sign * psi
    return lam


def joint_gradient(n, k, L, ops, params, masks, weights, sources=None):
    lib = _lib()
    pp = np.ascontiguousarray(params, np.float64)
    psi = O.run_ops(O.basis(n), n, ops, pp)
    lam = apply_obs(psi, n, masks, weights)
    psi = np.ascontiguousarray(psi, np.complex128).copy()
    lam = np.ascontiguousarray(lam, np.complex128).copy()
    grad = np.zeros(len(pp))
    nch = lib.emu_joint_prepare(n, k, L, ops.ctypes.data, len(ops), pp.ctypes.data, -1, None, 0)
    assert nch >= 1, f"reverse plan status {-nch}"
    for ch in range(nch):
        need = lib.emu_joint_prepare(n, k, L, ops.ctypes.data, len(ops), pp.ctypes.data, ch, None, 0)
        buf = C.create_string_buffer(need + 1)
        assert lib.emu_joint_prepare(n, k, L, ops.ctypes.data, len(ops), pp.ctypes.data, ch, buf, need + 1) == need
        src = buf.value.decode()
        if sources is not None:
            sources.append(src)
        size = lib.emu_joint_program(None)
        prog = np.zeros((size + 7) // 8, np.float64)
        lib.emu_joint_program(prog.ctypes.data)
        npass = lib.emu_joint_pass_info(-1, None, None, None, None)
        with tempfile.TemporaryDirectory() as tmp:
            sh = _compile(src, npass, tmp)
            for p in range(npass):
                kp, ns, npr, pb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
                lib.emu_joint_pass_info(p, C.byref(kp), C.byref(ns), C.byref(npr), C.byref(pb))
                ntiles = 1 << (n - kp.value)
                table = np.zeros((max(ntiles * ns.value, 1), 4))
                if ns.value:
                    lib.emu_joint_slot_table(p, table.ctypes.data)
                part = np.zeros(max(npr.value, 1))
                assert sh.shim_joint_launch(p, 1 << (kp.value - 4), psi.ctypes.data, lam.ctypes.data, table.ctypes.data,
                                            ntiles, prog.ctypes.data, part.ctypes.data) == 0
                for j in range(npr.value):
                    par, coeff = C.c_int(), C.c_double()
                    lib.emu_joint_probe_ref(pb.value + j, C.byref(par), C.byref(coeff))
                    grad[par.value] += coeff.value * part[j]
    return grad, psi


def check(n, k, L, ops, params, masks, weights):
    _, want = O.adjoint_expect(n, ops, params, masks, weights)
    got, psi_end = joint_gradient(n, k, L, ops, params, masks, weights)
    assert np.max(np.abs(got - want)) < 1e-10, (got, want)
    # the sweep is unitary on psi
    assert abs(float(np.vdot(psi_end, psi_end).real) - 1.0) < 1e-12


def test_qaoa_maxcut():
    n = 12
    edges = G.random_regular_graph(n, 3, seed=1)
    ops = G.qaoa_maxcut(n, edges, 2).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(2).random(4) * math.pi
    check(n, 9, 3, ops, params, masks, w)


@pytest.mark.parametrize("k", [10, 11])
def test_full_waves(k):
    # 64 and 128 threads: the wave sums take the DPP path on the device
    n = 12
    edges = G.random_regular_graph(n, 3, seed=4)
    ops = G.qaoa_maxcut(n, edges, 2).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(9).random(4) * math.pi
    check(n, k, 3, ops, params, masks, w)


def test_ansatz_every_parameterised_kind():
    # RX, RY, RZ, PHASE and CPHASE parameters, shared parameters, CNOT ladders;
    # 8-qubit tiles leave external qubits for controls and Z-strings
    n = 11
    rng = np.random.default_rng(5)
    g = G.GateList(n)
    npar = 0
    for layer in range(3):
        for q in range(n):
            kind = (q + layer) % 4
            if kind == 0:
                g.rotate_x(0.0, q, param=npar, scale=1.0)
            elif kind == 1:
                g.rotate_y(0.0, q, param=npar, scale=-0.5)
            elif kind == 2:
                g.rotate_z(0.1, q, param=npar, scale=2.0)
            else:
                g.phase(0.0, q, param=npar, scale=1.0)
            npar += 1
        for q in range(layer % 2, n - 1, 2):
            g.cnot(q, q + 1)
        g.cphase(0.0, layer, n - 1 - layer, param=npar, scale=1.0)
        npar += 1
        g.hadamard(layer)
    ops = g.ops()
    params = rng.random(npar) * 2 * math.pi
    masks = [1 << q for q in range(n)] + [(1 << 2) | (1 << 9)]
    w = [1.0] * n + [0.5]
    check(n, 8, 2, ops, params, masks, w)


def test_sources_compile_for_gfx950():
    hiprtc = pytest.importorskip("quantumsimevo_amd.capi")
    n = 12
    edges = G.random_regular_graph(n, 3, seed=1)
    ops = G.qaoa_maxcut(n, edges, 1).ops()
    masks, w = G.maxcut_observable(edges)
    srcs = []
    joint_gradient(n, 11, 3, ops, np.array([0.4, 1.1]), masks, w, srcs)
    with tempfile.TemporaryDirectory() as tmp:
        for i, src in enumerate(srcs):
            f = Path(tmp) / f"joint{i}.hip"
            f.write_text("#include <hip/hip_runtime.h>\n" + src)
            r = subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O3", "-Xclang",
                                "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", "-o",
                                str(Path(tmp) / f"joint{i}.o"), str(f)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-4000:]
    assert hiprtc is not None

"""CPU checks of the fusion planner: the fused-pass program it emits, executed by
a host emulation of the kernel (tests/fused_emu.cpp), must reproduce the oracle."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import _build
from quantumsimevo_amd import circuits as G

_emu = None


def emu():
    global _emu
    if _emu is None:
        path = O.ROOT / "tests" / "_build" / "libfused_emu.so"
        if not path.exists():
            _build.build_test_emulator()
        lib = C.CDLL(str(path))
        lib.emu_run_fused.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                                       C.c_int]
        lib.emu_check_plan.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        _emu = lib
    return _emu


def run(n, k, L, ops, psi, params=None, relabel=None):
    """Emulated fused run; with relabel=None both planner modes must agree."""
    outs = []
    for rl in ((0, 1) if relabel is None else (relabel,)):
        s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        npass = C.c_int(0)
        pp = None if p is None else p.ctypes.data
        assert emu().emu_check_plan(n, k, L, ops.ctypes.data, len(ops), pp, rl) == 0
        rc = emu().emu_run_fused(n, k, L, ops.ctypes.data, len(ops), pp, s.ctypes.data, C.byref(npass), rl)
        assert rc == 0
        outs.append((s, npass.value))
    if len(outs) == 2:
        assert np.max(np.abs(outs[0][0] - outs[1][0])) < 1e-13
    return outs[-1]


SHAPES = [(10, 8, 4), (11, 8, 0), (12, 9, 3), (13, 10, 2), (12, 8, 1), (14, 9, 5)]  # last: 2L > k


@pytest.mark.parametrize("n,k,L", SHAPES)
def test_random_layered(n, k, L):
    ops = G.random_layered(n, 6, seed=n + k).ops()
    psi = G.random_state(n, 3)
    got, npass = run(n, k, L, ops, psi)
    assert np.max(np.abs(got - O.run_ops(psi, n, ops))) < 1e-12


@pytest.mark.parametrize("n,k,L", SHAPES)
def test_qft_and_iqft(n, k, L):
    psi = G.random_state(n, 4)
    ops = G.qft(n).ops()
    got, npass = run(n, k, L, ops, psi)
    assert np.max(np.abs(got - O.run_ops(psi, n, ops))) < 1e-12
    ops = G.iqft(n).ops()
    got, _ = run(n, k, L, ops, psi)
    assert np.max(np.abs(got - O.run_ops(psi, n, ops))) < 1e-12


def test_every_gate_kind_mixed():
    n = 11
    rng = np.random.default_rng(5)
    g = G.GateList(n)
    for _ in range(300):
        kind = int(rng.integers(0, 13))
        q = int(rng.integers(0, n))
        c = int(rng.integers(0, n))
        th = float(rng.random() * 6.3)
        if kind == 0: g.hadamard(q)
        elif kind == 1: g.pauli_x(q)
        elif kind == 2: g.pauli_y(q)
        elif kind == 3: g.pauli_z(q)
        elif kind == 4: g.phase(th, q)
        elif kind == 5: g.rotate_x(th, q)
        elif kind == 6: g.rotate_y(th, q)
        elif kind == 7: g.rotate_z(th, q)
        elif kind == 8: g.cnot(c, q)
        elif kind == 9: g.cphase(th, c, q)
        elif kind == 10: g.swap(q, c)
        elif kind == 11: g.mcphase(th, sorted(set(int(x) for x in rng.integers(0, n, 3))), q)
        else: g.gate(np.array([[math.cos(th), -1j * math.sin(th)], [0.3j, 0.7]]), q)
    ops = g.ops()
    psi = G.random_state(n, 6)
    for (k, L) in ((8, 4), (9, 0), (8, 2)):
        got, _ = run(n, k, L, ops, psi)
        assert np.max(np.abs(got - O.run_ops(psi, n, ops))) < 1e-11, (k, L)


def test_parameterised_qaoa():
    n = 12
    edges = G.random_regular_graph(n, 3, seed=1)
    ops = G.qaoa_maxcut(n, edges, 3).ops()
    params = np.random.default_rng(2).random(6) * math.pi
    got, _ = run(n, 9, 3, ops, O.basis(n), params)
    assert np.max(np.abs(got - O.run_ops(O.basis(n), n, ops, params))) < 1e-12


def test_pass_counts():
    from quantumsimevo_amd import capi
    lib = capi.load()
    p = C.c_int(0)

    def passes(n, prec, ops):
        capi.check(lib.qse_plan_passes(None, n, prec, ops.ctypes.data, len(ops), C.byref(p)))
        return p.value
    # QFT-30 complex128 (k=12, L=4): rows 0-11, 12-19, 20-27, rows 28-29 + the bit reversal
    assert passes(30, capi.C128, G.qft(30).ops()) == 4
    for rl in (0, 1):
        assert emu().emu_check_plan(30, 12, 4, G.qft(30).ops().ctypes.data, 480, None, rl) == 0

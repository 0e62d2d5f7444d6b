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


# planner variants: bit 0 relabelling SWAPs, bit 1 dependency-graph scheduling, bit 2
# lookahead growth of the qubit set, bit 3 out-of-place stores that make the next pass's
# tile contiguous; -1 is the library's choice among them
VARIANTS = (0, 1, 2, 3, 6, 7, 8, 10, 11, 14, 15, -1)


def run(n, k, L, ops, psi, params=None, relabel=None):
    """Emulated fused run; with relabel=None every planner variant must agree."""
    outs = []
    for rl in (VARIANTS if relabel is None else (relabel,)):
        s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        npass = C.c_int(0)
        pp = None if p is None else p.ctypes.data
        assert emu().emu_check_plan(n, k, L, ops.ctypes.data, len(ops), pp, rl) == 0
        rc = emu().emu_run_fused(n, k, L, ops.ctypes.data, len(ops), pp, s.ctypes.data, C.byref(npass), rl)
        assert rc == 0
        outs.append((s, npass.value))
    for o in outs[1:]:
        assert np.max(np.abs(outs[0][0] - o[0])) < 1e-12
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

    def info(n, prec, ops):
        v = [C.c_int(0) for _ in range(4)]
        capi.check(lib.qse_plan_info(None, n, prec, ops.ctypes.data, len(ops), *[C.byref(x) for x in v]))
        return tuple(x.value for x in v[:3])

    # QFT-30, tuned shape: complex128 12-qubit tiles with runs of 8 amplitudes:
    # rows 0-11, rows 12-20, rows 21-29 with the final bit reversal folded into
    # the last pass's store (relabelled SWAPs); complex64 13-qubit tiles, runs of 16
    qft = G.qft(30).ops()
    assert info(30, capi.C128, qft) == (3, 12, 3)
    assert info(30, capi.C64, qft) == (3, 13, 4)
    capi.check(lib.qse_plan_passes(None, 30, capi.C128, qft.ctypes.data, len(qft), C.byref(p)))
    assert p.value == 3
    # the 30-qubit random layered circuit (20 layers): dependency scheduling with
    # lookahead growth of each pass's qubit set (29 passes with earliest-gate growth)
    assert info(30, capi.C128, G.random_layered(30, 20, seed=7).ops()) == (14, 12, 3)
    for rl in (0, 1, 2, 3):
        assert emu().emu_check_plan(30, 12, 3, qft.ctypes.data, 480, None, rl) == 0


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_dependency_scheduling_random_mixed(seed):
    # out-of-order passes must respect every non-commuting pair: random gates of
    # every kind (controls, multi-controlled phases, SWAPs, non-unitary 2x2)
    n = 12
    ops = G.random_mixed(n, 400, seed).ops()
    psi = G.random_state(n, seed)
    got, _ = run(n, 9, 3, ops, psi)
    assert np.max(np.abs(got - O.run_ops(psi, n, ops))) < 1e-11


def test_dependency_scheduling_cuts_passes():
    # 20-qubit random layered circuit, 20 layers, 12-qubit tiles with 3 low qubits:
    # in gate order one layer needs about two passes; list scheduling covers
    # light cones of several layers per pass; growing a pass's qubit set by the
    # ready gate that unlocks the most gates per added qubit (flag 4) covers more
    lib = emu()
    lib.emu_plan_stats.argtypes = [C.c_int] * 3 + [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    ops = G.random_layered(20, 20, seed=7).ops()
    out = np.zeros((256, 6), np.int32)
    counts = {f: lib.emu_plan_stats(20, 12, 3, ops.ctypes.data, len(ops), None, out.ctypes.data, 256, f)
              for f in (0, 2, 6, -1)}
    assert counts[0] == 39 and counts[2] <= 14 and counts[6] <= 8, counts
    assert counts[-1] == min(counts[0], counts[2], counts[6]), counts


@pytest.mark.parametrize("n,k,L,seed", [(14, 9, 3, 1), (16, 10, 2, 2), (13, 8, 0, 3), (15, 11, 5, 4)])
def test_lookahead_scheduling_matches_oracle(n, k, L, seed):
    # the lookahead variant (flags 4 | 2, with and without relabelling) on
    # layered circuits and random mixed gates of every kind
    lib = emu()
    for ops in (G.random_layered(n, 12, seed=seed).ops(), G.random_mixed(n, 300, seed).ops()):
        psi = G.random_state(n, seed)
        want = O.run_ops(psi, n, ops)
        for flags in (6, 7):
            got, _ = run(n, k, L, ops, psi, relabel=flags)
            assert np.max(np.abs(got - want)) < 1e-11

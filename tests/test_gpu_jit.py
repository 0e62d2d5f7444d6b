"""Run-time specialised fused passes (hiprtc, qse_ctx_set_jit) on the GPU.

Bar: complex128 within 1e-12 max abs of the oracle and within 1e-13 of the
generic fused kernel (same products, possibly different FMA contraction);
complex64 within 2e-5 of the oracle."""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def run_mode(ctx, mode, n, ops, psi, params=None, prec=capi.C128):
    ctx.set_jit(mode)
    try:
        st = ctx.state(n, prec)
        st.upload(psi)
        st.run(ops, params)
        out = st.download()
        st.close()
    finally:
        ctx.set_jit(capi.JIT_AUTO)
    return out


def check(ctx, n, ops, params=None, prec=capi.C128, seed=11):
    psi = G.random_state(n, seed)
    want = O.run_ops(psi, n, ops, params)
    before = ctx.jit_stats()
    jit = run_mode(ctx, capi.JIT_ALWAYS, n, ops, psi, params, prec)
    after = ctx.jit_stats()
    assert after["launches"] > before["launches"], "compiled kernels were not launched"
    generic = run_mode(ctx, capi.JIT_OFF, n, ops, psi, params, prec)
    if prec == capi.C128:
        assert np.max(np.abs(jit - want)) < 1e-12
        assert np.max(np.abs(jit - generic)) < 1e-13
    else:
        assert np.max(np.abs(jit - want)) < 2e-5


def test_qft_iqft_default_shape(ctx):
    n = 14
    check(ctx, n, G.qft(n).ops())
    check(ctx, n, G.iqft(n).ops())


def test_random_layered_and_mixed(ctx):
    check(ctx, 16, G.random_layered(16, 8, seed=4).ops())
    check(ctx, 13, G.random_mixed(13, 300, 8).ops())


def test_small_tiles_with_external_slots(ctx):
    # 9-qubit tiles with 3 low qubits: terms on external qubits go through slots
    ctx.set_fusion(capi.C128, 9, 3)
    try:
        n = 14
        edges = G.random_regular_graph(n, 3, seed=3)
        ops = G.qaoa_maxcut(n, edges, 3).ops()
        params = np.random.default_rng(5).random(6) * math.pi
        check(ctx, n, ops, params)
        check(ctx, n, G.random_mixed(n, 200, 9).ops())
    finally:
        ctx.set_fusion(capi.C128, 0, 0)


def test_complex64(ctx):
    check(ctx, 16, G.random_layered(16, 6, seed=2).ops(), prec=capi.C64)
    check(ctx, 15, G.qft(15).ops(), prec=capi.C64)


def test_auto_compiles_on_second_run(ctx):
    n = 14
    ops = G.random_layered(n, 5, seed=123).ops()
    psi = G.random_state(n, 1)
    ctx.set_jit(capi.JIT_AUTO)
    st = ctx.state(n)
    st.upload(psi)
    s0 = ctx.jit_stats()
    st.run(ops)
    s1 = ctx.jit_stats()
    assert s1["compiled"] == s0["compiled"] and s1["launches"] == s0["launches"]
    st.run(ops)
    s2 = ctx.jit_stats()
    assert s2["compiled"] == s1["compiled"] + 1 and s2["launches"] > s1["launches"]
    st.run(ops)
    s3 = ctx.jit_stats()
    assert s3["compiled"] == s2["compiled"] and s3["launches"] > s2["launches"]
    want = psi
    for _ in range(3):
        want = O.run_ops(want, n, ops)
    assert np.max(np.abs(st.download() - want)) < 1e-12
    st.close()


def test_set_jit_rejects_bad_mode(ctx):
    with pytest.raises(capi.QSEError):
        ctx.set_jit(3)
    ctx.set_jit(capi.JIT_AUTO)


def test_same_program_both_precisions(ctx):
    # both precisions can plan a circuit identically; the compiled kernels of
    # one precision must never run on a state of the other
    n = 16
    ops = G.random_layered(n, 8, seed=31).ops()
    assert ctx.plan_info(n, capi.C128, ops)["tile_bits"] == ctx.plan_info(n, capi.C64, ops)["tile_bits"]
    want = O.run_ops(O.basis(n), n, ops)
    ctx.set_jit(capi.JIT_AUTO)
    for prec, tol in ((capi.C128, 1e-12), (capi.C64, 2e-5), (capi.C128, 1e-12)):
        for _ in range(2):
            st = ctx.state(n, prec)
            st.run(ops)
            assert np.max(np.abs(st.download() - want)) < tol, prec
            st.close()


def test_parameter_loop_uses_structure_keyed_kernels():
    """qse_run with other angles on every call (an optimiser loop): no program
    repeats, so the program-keyed kernels never compile; the structure seen
    twice is compiled on a worker thread and serves the later calls."""
    n = 13
    edges = G.random_regular_graph(n - 1, 3, seed=6)
    ops = G.qaoa_maxcut(n, edges, 2).ops()
    rng = np.random.default_rng(9)
    c = capi.Context(0)
    try:
        st = c.state(n, capi.C128)
        for it in range(8):
            x = rng.random(4) * math.pi
            st.set_basis(0)
            st.run(ops, x)
            want = O.run_ops(O.basis(n), n, ops, x)
            assert np.max(np.abs(st.download() - want)) < 1e-12, it
            if it == 1:
                c.jit_wait()
                before = c.jit_stats()["launches"]
        assert c.jit_stats()["launches"] > before
        st.close()
    finally:
        c.close()


@pytest.mark.parametrize("n, make", [
    (18, lambda n: G.qft(n).ops()),
    (18, lambda n: G.random_layered(n, 8, seed=2).ops()),
])
def test_autotune_measures_and_keeps_the_fastest_shape(n, make):
    """qse_autotune times the cost model's best shapes on the device and the
    fastest is used by later runs of the same gate list (also at other
    angles); results still match the oracle."""
    ops = make(n)
    c = capi.Context(0)
    try:
        st = c.state(n, capi.C128)
        st.set_random(3)
        r = c.autotune(st, ops, candidates=3)
        assert 1 <= len(r["tried"]) <= 3
        times = [t[2] for t in r["tried"]]
        assert r["ms"] == min(times) and (r["k"], r["L"]) == r["tried"][times.index(min(times))][:2]
        info = c.plan_info(n, capi.C128, ops)
        assert (info["tile_bits"], info["low_bits"]) == (r["k"], r["L"])
        # same structure, other angles: the measured shape is used too
        other = ops.copy()
        other["theta"] = other["theta"] * 0.5 + 0.1
        assert c.plan_info(n, capi.C128, other)["tile_bits"] == r["k"]
        st.set_random(3)
        psi0 = st.download()
        st.run(ops)
        assert np.max(np.abs(st.download() - O.run_ops(psi0, n, ops))) < 1e-12
        st.close()
    finally:
        c.close()


def check_vs_generic(ctx, n, ops, params=None, seed=5):
    psi = G.random_state(n, seed)
    jit = run_mode(ctx, capi.JIT_ALWAYS, n, ops, psi, params)
    generic = run_mode(ctx, capi.JIT_OFF, n, ops, psi, params)
    assert np.max(np.abs(jit - generic)) < 1e-13
    return jit, psi


def test_paired_tiles_from_13_tile_index_bits(ctx):
    """Complex128 passes with 13 or more tile-index bits take tiles in adjacent
    pairs (when the paired kernel fits the registers): QFT-25 (with its
    out-of-place bit reversal), a layered circuit, 11-qubit tiles with
    external-qubit slots, against the generic kernel; one short layered
    circuit against the oracle at 25 qubits."""
    check_vs_generic(ctx, 25, G.qft(25).ops())
    check_vs_generic(ctx, 25, G.random_layered(25, 6, seed=2).ops())
    ctx.set_fusion(capi.C128, 11, 3)
    try:
        edges = G.random_regular_graph(24, 3, seed=8)
        q = G.qaoa_maxcut(24, edges, 2).ops()
        check_vs_generic(ctx, 24, q, np.array([0.3, 1.1, 0.7, 0.4]))
    finally:
        ctx.set_fusion(capi.C128, 0, 0)
    ops = G.random_layered(25, 2, seed=6).ops()
    jit, psi = check_vs_generic(ctx, 25, ops)
    assert np.max(np.abs(jit - O.run_ops(psi, 25, ops))) < 1e-12


def test_paired_structure_keyed_kernels_in_the_adjoint():
    """The adjoint's internal runs at 26 qubits (structure-keyed kernels,
    paired) against the generic kernel: QAOA p = 1 gradients."""
    ctx = capi.Context(0)
    try:
        n = 26
        edges = G.random_regular_graph(n, 3, seed=12)
        ops = G.qaoa_maxcut(n, edges, 1).ops()
        masks, w = G.maxcut_observable(edges)
        params = np.array([[0.37, 0.81]])
        ctx.set_jit(capi.JIT_ALWAYS)
        E1, G1 = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
        ctx.set_jit(capi.JIT_OFF)
        E0, G0 = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
        assert abs(E1[0] - E0[0]) < 1e-11 and np.max(np.abs(G1 - G0)) < 1e-11
    finally:
        ctx.close()

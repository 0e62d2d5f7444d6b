"""FP64 renormalisation on the device (qse_state_renormalize) and its periodic
form in qse_run / qse_run_repeat / qse_run_sweeps (qse_ctx_set_renormalize).
Bars: the renormalised state equals psi / ||psi|| (numpy, FP64) to 1e-15
(complex128) or FP32 rounding (complex64), the returned norm^2 equals numpy's
sum to 1e-12 relative; a zero state stays zero; the period counts passes as the
plan reports them."""
import ctypes as C

import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def rand_state(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n))


@pytest.mark.parametrize("prec, tol", [(capi.C128, 1e-15), (capi.C64, 2e-7)])
def test_renormalize_matches_numpy(ctx, prec, tol):
    n = 14
    psi = rand_state(n, 1, scale=3.0)
    st = ctx.state(n, prec)
    st.upload(psi)
    stored = st.download()              # complex64 rounding of the input
    before = st.renormalize()
    got = st.download()
    st.close()
    want_n2 = float(np.sum(np.abs(stored) ** 2))
    assert abs(before - want_n2) <= 1e-12 * want_n2
    want = stored / np.sqrt(want_n2)
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want)) * 4
    assert abs(np.sum(np.abs(got) ** 2) - 1.0) < (1e-13 if prec == capi.C128 else 1e-6)


def test_stream_ordered_form_and_zero_state(ctx):
    n = 11
    psi = rand_state(n, 2, scale=0.01)
    st = ctx.state(n, capi.C128)
    st.upload(psi)
    assert ctx.lib.qse_state_renormalize(ctx.h, st.h, None) == 0   # no readback
    got = st.download()
    np.testing.assert_allclose(got, psi / np.linalg.norm(psi), rtol=0, atol=1e-15)
    st.upload(np.zeros(1 << n, dtype=np.complex128))
    assert st.renormalize() == 0.0
    assert not np.any(st.download())
    st.close()


def scaling_ops(n, factor, count):
    g = G.GateList(n)
    for i in range(count):
        g.gate(np.diag([factor, factor]).astype(complex), i % n)   # multiplies the whole state by factor
    return g.ops()


@pytest.mark.parametrize("n", [12, 8])   # fused passes, and the one-sweep-per-gate path below 10 qubits
def test_periodic_renormalisation_counts_passes(ctx, n):
    ops = scaling_ops(n, 1.1, 3)
    passes = ctx.plan_info(n, capi.C128, ops)["passes"] if n >= 10 else len(ops)
    psi = rand_state(n, 3)
    psi /= np.linalg.norm(psi)
    st = ctx.state(n, capi.C128)
    st.upload(psi)
    try:
        ctx.set_renormalize(2 * passes)
        st.run(ops)                          # passes counted: below the period
        assert abs(st.norm2() - 1.1 ** 6) < 1e-12
        st.run(ops)                          # reaches the period: renormalised
        assert abs(st.norm2() - 1.0) < 1e-13
        st.run_repeat(ops, 5)                # 5 runs counted over the call, renormalised at its end
        assert abs(st.norm2() - 1.0) < 1e-13
        ctx.set_renormalize(0)
        st.run(ops)
        assert abs(st.norm2() - 1.1 ** 6) < 1e-12
        np.testing.assert_allclose(st.download() / 1.1 ** 3, psi, rtol=0, atol=1e-13)
    finally:
        ctx.set_renormalize(0)
        st.close()
    with pytest.raises(capi.QSEError):
        ctx.set_renormalize(-1)


def test_complex64_drift_bounded(ctx):
    n = 20
    ops = G.random_layered(n, 20, seed=9).ops()
    drift = {}
    for every in (0, 1):
        st = ctx.state(n, capi.C64)
        st.set_random(4)
        ctx.set_renormalize(every)
        for _ in range(40):
            st.run(ops)
        ctx.set_renormalize(0)
        drift[every] = abs(st.norm2() - 1.0)
        st.close()
    print(f"complex64 |norm^2 - 1| after 40 runs: {drift[0]:.2e} without, {drift[1]:.2e} with renormalisation")
    assert drift[1] < 1e-6

"""The allocation-free expectation entry point (qse_expect_zstrings_noalloc)
against the allocating one (qse_expect_zstrings), through the C ABI: the same
kernels and summation order, so the values must be bit-identical. Two calls in
a row, the second on a smaller state with fewer terms (the shared scratch array
keeps the first call's contents), then the limit check."""
import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def _observable(n, nterms, seed):
    rng = np.random.default_rng(seed)
    masks = rng.integers(0, 1 << n, size=nterms, dtype=np.uint64)
    masks[0] = 0
    return masks, rng.standard_normal(nterms)


def test_noalloc_expectation_matches_allocating_entry_point():
    ctx = capi.Context(0)
    try:
        cases = [(capi.C128, 22, 64, 1), (capi.C128, 12, 5, 2), (capi.C64, 20, 37, 3), (capi.C64, 11, 1, 4)]
        for prec, n, nterms, seed in cases:
            print(f"noalloc case prec={prec} n={n} nterms={nterms}", flush=True)
            st = ctx.state(n, prec)
            st.run(G.random_layered(n, 4, seed=seed).ops())
            m, w = _observable(n, nterms, seed)
            want = st.expect_zstrings(m, w)
            got = st.expect_zstrings_noalloc(m, w)
            assert got == want, (prec, n, got, want)
            st.close()
        st = ctx.state(10, capi.C128)
        st.set_basis(3)
        m, w = _observable(10, 65, 9)
        with pytest.raises(capi.QSEError) as e:
            st.expect_zstrings_noalloc(m, w)
        assert e.value.status == 6  # QSE_ERR_LIMIT
        assert st.expect_zstrings_noalloc(m[:0], w[:0]) == 0.0
        assert st.expect_zstrings_noalloc([0b1000, 0b10], [2.0, 0.5]) == 1.5  # |3>: Z_3 = +1, Z_1 = -1
        st.close()
    finally:
        ctx.close()

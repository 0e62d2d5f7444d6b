"""The two pair-exchange kernels of qse_exchange_probe (LDS between register
layouts; DPP/ds_swizzle/ds_bpermute across wavefront lanes) give the state
the oracle gives for the same gates (RY(0.3 + 0.1 q) on qubits 0..5, in that
order, reps times), to 1e-12."""
import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as C


def ry_layers(n, reps):
    g = C.GateList(n)
    for _ in range(reps):
        for q in range(6):
            g.rotate_y(0.3 + 0.1 * q, q)
    return g.ops()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("n,reps", [(12, 1), (14, 3)])
def test_exchange_methods_match_oracle(ctx, method, n, reps):
    print(f"exchange method={method} n={n} reps={reps}", flush=True)
    psi0 = C.random_state(n, seed=5 + n)
    st = ctx.state(n, capi.C128)
    try:
        st.upload(psi0)
        ms = st.exchange_probe(method, reps)
        assert ms > 0
        psi = st.download()
    finally:
        st.close()
    ref = O.run_ops(psi0, n, ry_layers(n, reps))
    assert np.max(np.abs(psi - ref)) < 1e-12


@pytest.mark.gpu
def test_exchange_rejections(ctx):
    st = ctx.state(10, capi.C128)
    try:
        with pytest.raises(capi.QSEError):
            st.exchange_probe(0, 1)  # fewer than 12 qubits
    finally:
        st.close()
    st = ctx.state(12, capi.C64)
    try:
        with pytest.raises(capi.QSEError):
            st.exchange_probe(1, 1)  # complex64
    finally:
        st.close()

"""QFT at full size (33 qubits complex128, 34 qubits complex64: 128 GiB states)
checked against the FP64 DFT closed form evaluated on the GPU from the seed
(qse_qft_probe), which is itself pinned here to the CPU oracle at 12 qubits.

Bars: the probe matches the oracle's QFT to 1e-14 (12 qubits) and the fused GPU
run to 1e-12 (20 qubits); at full size, max |got - want| / max |want| over 16
probed amplitudes is below 1e-9 (complex128) and 1e-4 (complex64)."""
import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def probe_indices(n, count, seed):
    rng = np.random.default_rng(seed)
    ks = rng.integers(0, 1 << n, size=count - 2, dtype=np.uint64)
    return np.concatenate([np.array([0, (1 << n) - 1], dtype=np.uint64), ks])


def test_probe_pinned_to_oracle(ctx):
    n = 12
    st = ctx.state(n)
    st.set_random(5)
    psi = st.download()
    st.close()
    want = O.run_ops(psi, n, G.qft(n).ops())
    ks = probe_indices(n, 20, 1)
    got = ctx.qft_probe(n, 5, ks)
    assert np.max(np.abs(got - want[ks.astype(np.int64)])) < 1e-14
    # every output amplitude at a smaller size
    n = 8
    st = ctx.state(n)
    st.set_random(11)
    psi = st.download()
    st.close()
    want = O.run_ops(psi, n, G.qft(n).ops())
    got = ctx.qft_probe(n, 11, np.arange(1 << n, dtype=np.uint64))
    assert np.max(np.abs(got - want)) < 1e-14


def test_probe_vs_fused_run(ctx):
    n = 20
    st = ctx.state(n)
    st.set_random(9)
    st.run(G.qft(n).ops())
    out = st.download()
    st.close()
    ks = probe_indices(n, 16, 2)
    assert np.max(np.abs(ctx.qft_probe(n, 9, ks) - out[ks.astype(np.int64)])) < 1e-12


def full_size(n, prec, tol):
    bytes_per = 16 if prec == capi.C128 else 8
    if capi.device_info(0)["total_bytes"] < 2 * bytes_per * (1 << n) + (4 << 30):
        pytest.skip("device memory below two full-size state buffers")
    ctx = capi.Context(0)
    try:
        st = ctx.state(n, prec)
        st.set_random(2024)
        st.run(G.qft(n).ops())
        ks = probe_indices(n, 16, n)
        got = np.array([st.download_range(int(k), 1)[0] for k in ks])
        st.close()
        want = ctx.qft_probe(n, 2024, ks)
        rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
        assert rel < tol, rel
    finally:
        ctx.close()


def test_qft33_complex128_full_size():
    full_size(33, capi.C128, 1e-9)


def test_qft34_complex64_full_size():
    full_size(34, capi.C64, 1e-4)


def inverse_ops(ops):
    """The gate list of U^dagger: reversed order, angles negated, 2x2 matrices
    conjugate-transposed (H, X, Y, Z, CNOT and SWAP are their own inverses)."""
    inv = ops[::-1].copy()
    angle = np.isin(inv["kind"], [G.PHASE, G.RX, G.RY, G.RZ, G.CPHASE, G.MCPHASE])
    inv["theta"][angle] = -inv["theta"][angle]
    mat = inv["kind"] == G.MATRIX
    m = inv["m"][mat]
    inv["m"][mat] = np.stack([m[:, 0], -m[:, 1], m[:, 4], -m[:, 5], m[:, 2], -m[:, 3], m[:, 6], -m[:, 7]], axis=1)
    return inv


def test_random_layered_30q_round_trip():
    """The plain-method random circuit (one of H/RX/RY/RZ/phase per qubit per
    layer, CNOT brickwork) at 30 qubits complex128, 20 layers (29 fused
    passes): the FP64 norm stays 1 and U then U^dagger returns the seeded
    state at 4096 sampled amplitudes (size-independent properties; the
    oracle comparison of this circuit runs at 20 qubits)."""
    n = 30
    if capi.device_info(0)["total_bytes"] < 2 * 16 * (1 << n) + (4 << 30):
        pytest.skip("device memory below two 30-qubit state buffers")
    ctx = capi.Context(0)
    try:
        ops = G.random_layered(n, 20, seed=7).ops()
        st = ctx.state(n, capi.C128)
        st.set_random(31)
        rng = np.random.default_rng(5)
        ks = np.unique(rng.integers(0, 1 << n, size=4096))
        before = np.concatenate([st.download_range(int(k), 1) for k in ks[:64]] +
                                [st.download_range(0, 4032)])
        st.run(ops)
        assert abs(st.norm2() - 1.0) < 1e-12
        st.run(inverse_ops(ops))
        after = np.concatenate([st.download_range(int(k), 1) for k in ks[:64]] +
                               [st.download_range(0, 4032)])
        st.close()
        assert np.max(np.abs(after - before)) < 1e-12
    finally:
        ctx.close()

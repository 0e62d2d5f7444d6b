"""The benchmark configurations at their full sizes on the GPU, against the
oracle, through the C ABI:

  C2  512 Pauli-noise trajectories of a 20-qubit depth-30 random layered
      circuit, complex64 state, 10 000 shots each: frame masks bit-exact
      against the oracle's restatement of Noise.cpp:21-44, every sampled
      outcome bit-exact against the oracle's sequential-CDF sampler run on the
      downloaded complex64 amplitudes (so sampling exactness is separated from
      complex64 rounding), and the complex64 state's error against FP64
  C3  QAOA Max-Cut 24 qubits, p = 8, 64 parameter sets: adjoint energies and
      gradients against the oracle's adjoint sweep and parameter shift
      (tests/golden/fullsize_vectors.npz, 1e-10) and against central finite
      differences on the GPU for all 64 sets (h = 1e-5, 1e-6)
  C4  the 1024-circuit uneven ansatz batch (10..22 qubits) in one persistent
      launch: complex128 energies/gradients of 16 circuits (the deepest
      22-qubit ones among them) against the oracle (1e-11 / 1e-10), every
      gradient of the batch against the host-driven adjoint (1e-10), and the
      complex64 run's errors against FP64
  QFT-30 complex128 and complex64 on a seeded random 2^30-amplitude state
      against the FP64 DFT closed form, the round trip QFT then IQFT, and the
      complex64 state's amplitude error and fidelity against the complex128 run
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu

FIX = np.load(O.ROOT / "tests" / "golden" / "fullsize_vectors.npz", allow_pickle=False)


def report(**kw):
    print(" ".join(f"{k}={v:.3e}" if isinstance(v, float) else f"{k}={v}" for k, v in kw.items()), flush=True)


# ---------------------------------------------------------------- C2
def test_c2_noise_trajectories_full_size(ctx):
    from quantumsimevo_amd import _quantumsimevo as Q
    n, depth, ntraj, shots = 20, 30, 512, 10000
    ops = G.random_layered(n, depth, seed=2024).ops()
    seeds = np.arange(ntraj, dtype=np.uint32)
    x, z = Q.noise_frames(n, ops["kind"], ops["target"], seeds, 1e-3, 1.0)
    xo, zo = O.noise_trajectories(n, ops, 0, ntraj, rate=1e-3)
    assert np.array_equal(x, xo) and np.array_equal(z, zo)
    assert len(np.unique(x)) > 50  # p = 1e-3 over ~900 gates: most trajectories flip something

    want = O.run_ops(O.basis(n), n, ops)
    s128 = ctx.state(n, capi.C128)
    s128.run(ops)
    assert np.max(np.abs(s128.download() - want)) < 1e-12
    s128.close()
    st = ctx.state(n, capi.C64)
    st.run(ops)
    psi = st.download()  # the complex64 amplitudes, exactly, as FP64 pairs
    err = float(np.max(np.abs(psi - want)))
    fid = float(abs(np.vdot(want, psi)) ** 2 / np.vdot(psi, psi).real)
    report(c2_c64_max_amp_err=err, c2_c64_infidelity=1.0 - fid)
    assert err < 1e-5 and 1.0 - fid < 1e-9

    u = np.random.default_rng(11).random((ntraj, shots))
    got = st.sample_frames(x, u)
    st.close()
    for xm in np.unique(x):
        rows = np.nonzero(x == xm)[0]
        framed = O.pauli_frame(psi, n, int(xm), 0)
        want_s = O.sample(framed, n, u[rows].ravel())
        assert np.array_equal(got[rows].ravel(), want_s), int(xm)


# ---------------------------------------------------------------- C3
def c3_problem():
    n, p = 24, 8
    edges = G.random_regular_graph(n, 3, seed=2024)
    ops = G.qaoa_maxcut(n, edges, p).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(7).random((64, 2 * p)) * math.pi
    return n, edges, ops, masks, w, params


def test_c3_qaoa24_full_size(ctx):
    n, edges, ops, masks, w, params = c3_problem()
    assert np.array_equal(np.array(edges), FIX["c3_edges"]) and np.array_equal(params[:2], FIX["c3_params"])
    E, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    e_err = float(np.max(np.abs(E[:2] - FIX["c3_energy"])))
    g_err = float(np.max(np.abs(Gr[:2] - FIX["c3_grad"])))
    s_err = float(np.max(np.abs(Gr[0, FIX["c3_shift_params"]] - FIX["c3_shift_grad"])))
    report(c3_energy_err_vs_oracle=e_err, c3_grad_err_vs_oracle_adjoint=g_err, c3_grad_err_vs_oracle_shift=s_err)
    assert e_err < 1e-10 and g_err < 1e-10 and s_err < 1e-10
    # central finite differences on the GPU, every parameter of every set
    h = 1e-5
    fd_err = 0.0
    for j in range(params.shape[1]):
        Pp, Pm = params.copy(), params.copy()
        Pp[:, j] += h
        Pm[:, j] -= h
        fd = (ctx.batch_expect(n, capi.C128, ops, Pp, masks, w) -
              ctx.batch_expect(n, capi.C128, ops, Pm, masks, w)) / (2 * h)
        fd_err = max(fd_err, float(np.max(np.abs(Gr[:, j] - fd))))
    report(c3_grad_err_vs_fd=fd_err)
    assert fd_err < 1e-6
    # complex64 storage, FP64 accumulation: errors against the FP64 values
    E32, G32 = ctx.adjoint_expect(n, capi.C64, ops, params[:8], masks, w, ckpt_every=8)
    report(c3_c64_energy_err=float(np.max(np.abs(E32 - E[:8]))), c3_c64_grad_err=float(np.max(np.abs(G32 - Gr[:8]))))
    assert np.max(np.abs(E32 - E[:8])) < 1e-3 and np.max(np.abs(G32 - Gr[:8])) < 1e-2


# ---------------------------------------------------------------- C4
def c4_batch():
    circs = G.uneven_ansatz_batch(1024, seed=2024, n_range=(10, 22), depth_range=(4, 12))
    masks = [1 << q for q in range(22)]
    return circs, masks, [1.0] * 22


@pytest.fixture(scope="module")
def c4_runs(ctx):
    circs, masks, w = c4_batch()
    E, Gr = ctx.batch_run_small(capi.C128, circs, masks, w)
    E64, G64 = ctx.batch_run_small(capi.C64, circs, masks, w)
    return circs, E, Gr, E64, G64


def test_c4_batch_vs_oracle(c4_runs):
    circs, E, Gr, E64, G64 = c4_runs
    off = FIX["c4_goff"]
    e_err = g_err = e32 = g32 = 0.0
    for j, i in enumerate(FIX["c4_idx"]):
        n, ops, par = circs[int(i)]
        assert n == FIX["c4_n"][j] and len(ops) == FIX["c4_nops"][j] and abs(np.sum(par) - FIX["c4_param_sum"][j]) < 1e-9
        go = FIX["c4_grad"][off[j]:off[j + 1]]
        e_err = max(e_err, abs(E[i] - FIX["c4_energy"][j]))
        g_err = max(g_err, float(np.max(np.abs(Gr[i] - go))))
        e32 = max(e32, abs(E64[i] - FIX["c4_energy"][j]))
        g32 = max(g32, float(np.max(np.abs(G64[i] - go))))
    report(c4_energy_err_vs_oracle=e_err, c4_grad_err_vs_oracle=g_err, c4_c64_energy_err_vs_oracle=e32,
           c4_c64_grad_err_vs_oracle=g32)
    assert e_err < 1e-11 and g_err < 1e-10
    assert e32 < 1e-4 and g32 < 1e-3
    assert int(np.sum(FIX["c4_n"] == 22)) >= 6


def test_c4_all_gradients_vs_host_adjoint(ctx, c4_runs):
    """Every circuit of the batch against the host-driven adjoint
    (qse_adjoint_expect, one call per circuit, generic pass kernels), and the
    complex64 engine run against those FP64 values."""
    circs, E, Gr, E64, G64 = c4_runs
    ctx.set_jit(capi.JIT_OFF)  # 117 circuit shapes: no run-time compilation for one call each
    try:
        e_err = g_err = e32 = g32 = 0.0
        for i, (n, ops, par) in enumerate(circs):
            zm, zw = G.zsum_observable(n)
            Eh, Gh = ctx.adjoint_expect(n, capi.C128, ops, par, zm, zw)
            e_err = max(e_err, abs(E[i] - Eh[0]))
            g_err = max(g_err, float(np.max(np.abs(Gr[i] - Gh[0]))))
            e32 = max(e32, abs(E64[i] - Eh[0]))
            g32 = max(g32, float(np.max(np.abs(G64[i] - Gh[0]))))
    finally:
        ctx.set_jit(capi.JIT_AUTO)
    report(c4_energy_err_vs_host=e_err, c4_grad_err_vs_host=g_err, c4_c64_energy_err=e32, c4_c64_grad_err=g32)
    assert e_err < 1e-11 and g_err < 1e-10
    assert e32 < 1e-4 and g32 < 1e-3


def test_batch_deposit_levels_and_large_circuits(ctx):
    """21-22-qubit circuits and one 26-qubit circuit (more than 12 tile-index
    bits: three levels of the per-pass deposit tables) mixed with small ones,
    against the host-driven adjoint, and a 22-qubit energy against the oracle."""
    circs = G.uneven_ansatz_batch(3, seed=31, n_range=(21, 22), depth_range=(2, 3))
    circs += G.uneven_ansatz_batch(1, seed=32, n_range=(26, 26), depth_range=(2, 2))
    circs += G.uneven_ansatz_batch(8, seed=33, n_range=(10, 13), depth_range=(2, 4))
    masks = [1 << q for q in range(26)]
    w = [1.0] * 26
    for prec, te, tg in ((capi.C128, 1e-11, 1e-10), (capi.C64, 1e-4, 1e-3)):
        E, Gr = ctx.batch_run_small(prec, circs, masks, w)
        for i, (n, ops, par) in enumerate(circs):
            zm, zw = G.zsum_observable(n)
            Eh, Gh = ctx.adjoint_expect(n, capi.C128, ops, par, zm, zw)
            assert abs(E[i] - Eh[0]) < te, (prec, i, n)
            assert np.max(np.abs(Gr[i] - Gh[0])) < tg, (prec, i, n)
        if prec == capi.C128:
            j = next(i for i, c in enumerate(circs) if c[0] == 22)
            n, ops, par = circs[j]
            zm, zw = G.zsum_observable(n)
            psi = O.run_ops(O.basis(n), n, ops, par)
            assert abs(E[j] - O.expect_zstrings(psi, n, zm, zw)) < 1e-11


# ---------------------------------------------------------------- QFT-30
def _probe_ks(n, count, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([np.array([0, (1 << n) - 1], dtype=np.uint64),
                           rng.integers(0, 1 << n, size=count - 2, dtype=np.uint64)])


def _chunks(st, n, chunk=1 << 26):
    for first in range(0, 1 << n, chunk):
        yield st.download_range(first, chunk)


def test_qft30_full_size_both_precisions():
    """QFT-30 on the seeded random state (the benchmark's workload): 64 output
    amplitudes against the FP64 DFT closed form (complex128 1e-12 relative,
    complex64 1e-4), QFT^-1 returning the input (complex128, 1e-12 at 64
    sampled amplitudes), and the complex64 state's max amplitude error and
    fidelity against the complex128 run."""
    n = 30
    if capi.device_info(0)["total_bytes"] < (2 * 16 + 2 * 8) * (1 << n) + (8 << 30):
        pytest.skip("device memory below the four 30-qubit buffers")
    ctx = capi.Context(0)
    try:
        ops = G.qft(n).ops()
        ks = _probe_ks(n, 64, 30)
        want = ctx.qft_probe(n, 2024, ks)
        a = ctx.state(n, capi.C128)
        a.set_random(2024)
        rng = np.random.default_rng(6)
        idx = np.unique(rng.integers(0, 1 << n, size=64))
        before = np.array([a.download_range(int(k), 1)[0] for k in idx])
        a.run(ops)
        got = np.array([a.download_range(int(k), 1)[0] for k in ks])
        rel128 = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
        b = ctx.state(n, capi.C64)
        b.set_random(2024)
        b.run(ops)
        got64 = np.array([b.download_range(int(k), 1)[0] for k in ks])
        rel64 = float(np.max(np.abs(got64 - want)) / np.max(np.abs(want)))
        # complex64 against complex128, whole state in 2^26-amplitude chunks, FP64 sums
        ov, n64, err = 0j, 0.0, 0.0
        for x, y in zip(_chunks(a, n), _chunks(b, n)):
            ov += np.vdot(x, y)
            n64 += float(np.vdot(y, y).real)
            err = max(err, float(np.max(np.abs(x - y))))
        infid = 1.0 - abs(ov) ** 2 / n64
        report(qft30_c128_rel_err=rel128, qft30_c64_rel_err=rel64, qft30_c64_max_amp_err=err,
               qft30_c64_infidelity=infid)
        assert rel128 < 1e-12 and rel64 < 1e-4
        assert err < 1e-6 and infid < 1e-8
        b.close()
        a.run(G.iqft(n).ops())
        after = np.array([a.download_range(int(k), 1)[0] for k in idx])
        assert np.max(np.abs(after - before)) < 1e-12
        assert abs(a.norm2() - 1.0) < 1e-12
        a.close()
    finally:
        ctx.close()

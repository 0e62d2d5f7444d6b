"""GPU parity: the gfx950 kernels through the C ABI against the CPU oracle.

Bar: FP64 amplitudes within 1e-12 max-abs (complex128); permutation/sign gates
(X, Y, Z, CNOT, SWAP, Pauli frames) bit-exact; sample/measure outcome indices
bit-exact for the same uniforms on identical amplitudes; complex64 within 2e-5.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as C

pytestmark = pytest.mark.gpu

GOLD = np.load(O.ROOT / "tests" / "golden" / "ref_vectors.npz", allow_pickle=False)
EXACT = ("X", "Y", "Z", "CNOT", "SWAP")


def gold_ops(key):
    return GOLD[key].view(capi.OP_DTYPE)


def upload(ctx, psi, n, prec=capi.C128):
    st = ctx.state(n, prec)
    st.upload(psi)
    return st


@pytest.mark.parametrize("mode", ["single", "sweeps", "fused"])
def test_single_gates_vs_golden_and_oracle(ctx, mode):
    psi = GOLD["gate_input"]
    n = 6
    for name in GOLD["gate_names"]:
        name = str(name)
        ops = gold_ops(f"gate_{name}_ops")
        st = upload(ctx, psi, n)
        if mode == "single":
            o = ops[0]
            k, t, c = int(o["kind"]), int(o["target"]), int(o["control"])
            {capi.H: lambda: st.h_(t), capi.X: lambda: st.x(t), capi.Y: lambda: st.y(t),
             capi.Z: lambda: st.z(t), capi.PHASE: lambda: st.phase(t, o["theta"]),
             capi.RX: lambda: st.rx(t, o["theta"]), capi.RY: lambda: st.ry(t, o["theta"]),
             capi.RZ: lambda: st.rz(t, o["theta"]), capi.CNOT: lambda: st.cnot(c, t),
             capi.CPHASE: lambda: st.cphase(c, t, o["theta"]), capi.SWAP: lambda: st.swap(t, c),
             capi.MCPHASE: lambda: st.mcphase([q for q in range(64) if (int(o["controls"]) >> q) & 1], t,
                                               o["theta"]),
             capi.MATRIX: lambda: st.matrix(t, o["m"].view(np.complex128))}[k]()
        else:
            st.run(ops, fused=(mode == "fused"))
        got = st.download()
        want = O.run_ops(psi, n, ops)
        if name.rstrip("0123456789_").startswith(EXACT):
            assert np.array_equal(got, want), name
        assert np.max(np.abs(got - want)) < 1e-14, name
        assert np.max(np.abs(got - GOLD[f"gate_{name}_scalar"])) < 1e-13, name


@pytest.mark.parametrize("fused", [False, True])
def test_golden_circuits(ctx, fused):
    st = ctx.state(10)
    st.run(gold_ops("layered_ops"), fused=fused)
    assert np.max(np.abs(st.download() - GOLD["layered_out"])) < 1e-12
    st = upload(ctx, GOLD["qft_input"], 8)
    st.run(gold_ops("qft_ops"), fused=fused)
    assert np.max(np.abs(st.download() - GOLD["qft_out"])) < 1e-12


@pytest.mark.parametrize("n,layers,seed", [(12, 10, 1), (16, 12, 2), (20, 20, 3)])
def test_random_layered_circuit_c128(ctx, n, layers, seed):
    """Random layered circuit (20 qubits x 20 layers is the CPU-runnable config)."""
    ops = C.random_layered(n, layers, seed).ops()
    want = O.run_ops(O.basis(n), n, ops)
    for fused in (False, True):
        st = ctx.state(n)
        st.run(ops, fused=fused)
        got = st.download()
        assert np.max(np.abs(got - want)) < 1e-12, fused


def test_random_layered_sample_and_measure_c0(ctx):
    """Config C0: execute(), sample(100000), measure(); indices bit-exact vs the oracle
    on the same uniforms (the GPU samples the oracle's amplitudes, and its own)."""
    n = 20
    ops = C.random_layered(n, 20, 1).ops()
    want_state = O.run_ops(O.basis(n), n, ops)
    u = O.uniforms(1001, 100000)
    want = O.sample(want_state, n, u)
    st = upload(ctx, want_state, n)
    assert np.array_equal(st.sample(u), want)
    st2 = ctx.state(n)
    st2.run(ops)
    assert np.max(np.abs(st2.download() - want_state)) < 1e-12
    got2 = st2.sample(u)
    # On its own (1e-15-close) amplitudes: identical except for draws within that
    # band of a boundary; none occur for this seed.
    assert np.array_equal(got2, want)
    um = float(O.uniforms(1002, 1)[0])
    idx, collapsed = O.measure(want_state, n, um)
    assert st.measure(um) == idx
    assert np.array_equal(st.download(), collapsed)


@pytest.mark.parametrize("n", [1, 2, 5, 9, 14])
def test_sampling_exact_small_and_edge(ctx, n):
    psi = C.random_state(n, 40 + n)
    u = np.random.default_rng(n).random(20000)
    u[:3] = [0.0, np.nextafter(1.0, 0.0), 0.5]
    st = upload(ctx, psi, n)
    assert np.array_equal(st.sample(u), O.sample(psi, n, u))
    # states with zero-probability runs and exact ties
    psi2 = np.zeros(1 << n, dtype=np.complex128)
    psi2[:: max(1, (1 << n) // 4)] = 0.5
    psi2 /= np.sqrt(np.sum(np.abs(psi2) ** 2))
    st.upload(psi2)
    assert np.array_equal(st.sample(u), O.sample(psi2, n, u))


@pytest.mark.parametrize("n", [3, 9, 16])
def test_sampling_flagged_draws_resolve_exactly(ctx, n):
    """Targets placed on the sequential cumulative boundaries force the exact
    path (checkpoint chains of the sequential sum): states below one LDS chunk
    and below one checkpoint group of amplitudes, and a larger one."""
    psi = C.random_state(n, 77)
    p = psi.real * psi.real + psi.imag * psi.imag
    cum = np.cumsum(p)
    idx = np.random.default_rng(3).integers(0, (1 << n) - 1, 500) if n > 8 else np.arange((1 << n) - 1)
    u = cum[idx] / cum[-1]
    u = np.concatenate([u, np.nextafter(u, 0), np.nextafter(u, 2)])
    u = np.clip(u, 0, np.nextafter(1.0, 0))
    st = upload(ctx, psi, n)
    assert np.array_equal(st.sample(u), O.sample(psi, n, u))


def test_readout_reductions(ctx):
    n = 15
    psi = C.random_state(n, 5) * 0.9
    st = upload(ctx, psi, n)
    assert abs(st.norm2() - O.norm2(psi, n)) < 1e-13
    for q in (0, 7, 14):
        want = float(np.sum(np.abs(psi[[i for i in range(1 << n) if (i >> q) & 1]]) ** 2))
        assert abs(st.prob_one(q) - want) < 1e-13
    edges = C.random_regular_graph(14, 3, seed=2)
    m, w = C.maxcut_observable(edges)
    assert abs(st.expect_zstrings(m, w) - O.expect_zstrings(psi, n, m, w)) < 1e-11
    zm, zw = C.zsum_observable(n)
    assert abs(st.expect_zstrings(zm, zw) - O.expect_zstrings(psi, n, zm, zw)) < 1e-11


def test_collapse_qubit(ctx):
    n = 12
    psi = C.random_state(n, 6)
    for q, u, reset in ((3, 0.2, False), (11, 0.9, True), (0, 0.5, True)):
        o_want, want = O.collapse_qubit(psi, n, q, u, reset)
        st = upload(ctx, psi, n)
        assert st.collapse_qubit(q, u, reset) == o_want
        assert np.max(np.abs(st.download() - want)) < 1e-13


def test_pauli_frame_bit_exact(ctx):
    n = 13
    psi = C.random_state(n, 8)
    for x, z in ((0, 0b1010), (0b1001000000001, 0), (0b111, 0b1000000000110), (1 << 12, 1 << 12)):
        st = upload(ctx, psi, n)
        st.pauli_frame(x, z)
        assert np.array_equal(st.download(), O.pauli_frame(psi, n, x, z))


def test_noise_trajectory_sampling_bit_exact(ctx):
    """Config C2 semantics at test size: per-trajectory frames from the reference
    Noise draws; samples of each framed state bit-exact vs the oracle."""
    n = 12
    ops = C.random_layered(n, 30, 3).ops()
    ideal = O.run_ops(O.basis(n), n, ops)
    ntraj, shots = 16, 2000
    x, z = O.noise_trajectories(n, ops, 0, ntraj, rate=1e-2)
    assert np.any(x) and np.any(z)
    u = O.uniforms(5, ntraj * shots).reshape(ntraj, shots)
    st = upload(ctx, ideal, n)
    got = st.sample_frames(x, u)
    for t in range(ntraj):
        framed = O.pauli_frame(ideal, n, int(x[t]), int(z[t]))
        assert np.array_equal(got[t], O.sample(framed, n, u[t])), t


def test_complex64_mode(ctx):
    n = 16
    ops = C.random_layered(n, 8, 4).ops()
    want = O.run_ops(O.basis(n), n, ops)
    for fused in (False, True):
        st = ctx.state(n, capi.C64)
        st.run(ops, fused=fused)
        got = st.download()
        assert np.max(np.abs(got - want)) < 2e-5
        assert abs(st.norm2() - 1.0) < 1e-5


def test_qft_vs_oracle_and_closed_form(ctx):
    n = 14
    psi = C.random_state(n, 21)
    ops = C.qft(n).ops()
    want = O.run_ops(psi, n, ops)
    for fused in (False, True):
        st = upload(ctx, psi, n)
        st.run(ops, fused=fused)
        assert np.max(np.abs(st.download() - want)) < 1e-12


def test_edge_cases_and_errors(ctx):
    st = ctx.state(3)
    st.cnot(1, 1)  # reference loop is a no-op when control == target
    st.swap(2, 2)
    assert np.array_equal(st.download(), O.basis(3))
    st.run(C.GateList(3).ops())  # empty list
    with pytest.raises(capi.QSEError):
        st.h_(3)
    with pytest.raises(capi.QSEError):
        st.cnot(0, 7)
    bad = C.GateList(3).hadamard(0).ops()
    bad["kind"][0] = 13  # GateType::MEASURE has no dispatch in the reference
    with pytest.raises(capi.QSEError):
        st.run(bad)
    with pytest.raises(capi.QSEError):
        st.set_basis(8)
    st1 = ctx.state(1)
    st1.h_(0)
    assert np.max(np.abs(st1.download() - np.array([1, 1]) / math.sqrt(2))) < 1e-15

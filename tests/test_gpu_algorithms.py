"""The playground algorithms end to end on the HIP backend (QSEplayground/
Shor.cpp, Grover.cpp, QAOA.cpp): known answers, not just agreement.
  * Shor, N = 15: the 18-qubit order-finding circuit (6221 gates) through the
    Circuit API, loaded from a gate-list file; the phase register's
    distribution is exactly 1/r on the multiples of 2^8/r (1e-12), sampled
    phases give the order and the factors 3 x 5.
  * Grover, 14 qubits: success probability equals sin^2((2k+1) asin(2^-7))
    (1e-11) and a seeded measure() returns the target.
  * QAOA Max-Cut with its classical outer loop: gradient ascent on <C> with
    adjoint gradients raises the expected cut, and the best sampled partition of
    the optimised state is a maximum cut (brute force)."""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import algorithms as A
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Circuit():
    from quantumsimevo_amd import _quantumsimevo
    return _quantumsimevo.Circuit


def phase_marginal(state, L):
    p = np.abs(state) ** 2
    t = len(L.phase)
    k_of = np.array([A.phase_register_value(i, L) for i in range(1 << t)])
    marg = np.zeros(1 << t)
    np.add.at(marg, k_of[np.arange(len(p)) & ((1 << t) - 1)], p)
    return marg


@pytest.mark.parametrize("a, r", [(7, 4), (11, 2)])
def test_shor_15(Circuit, tmp_path, a, r):
    N = 15
    g, L = A.shor_order_finding(N, a)
    path = tmp_path / f"shor15_{a}.qsegates"
    g.save(path)
    qc = Circuit(L.qubits)
    qc.load_gates(str(path))
    qc.execute()
    psi = qc.state()
    marg = phase_marginal(psi, L)
    step = 256 // r
    want = np.zeros(256)
    want[::step] = 1.0 / r
    assert np.max(np.abs(marg - want)) < 1e-12
    if a == 7:
        ref = O.run_ops(O.basis(L.qubits), L.qubits, g.ops())
        assert np.max(np.abs(psi - ref)) < 1e-12
    qc.set_seed(1234)
    shots = qc.sample(2000)
    ks = {A.phase_register_value(int(s), L) for s in shots}
    assert ks == set(range(0, 256, step))
    found = set()
    for k in ks:
        rr = A.order_from_phase(k, 8, N, a)
        if rr:
            assert rr == r
            f = A.factors_from_order(N, a, rr)
            if f:
                found.add(f)
    assert found == {(3, 5)}


def test_grover_14(Circuit):
    n, target = 14, 0b10110011100101
    k = A.grover_iterations(n)
    g = A.grover(n, target, k)
    qc = Circuit(n)
    for o in g.ops():
        kind, t = int(o["kind"]), int(o["target"])
        if kind == capi.H: qc.hadamard(t)
        elif kind == capi.X: qc.pauli_x(t)
        elif kind == capi.Z: qc.pauli_z(t)
        else: qc.mcphase(float(o["theta"]), [q for q in range(n) if (int(o["controls"]) >> q) & 1], t)
    qc.execute()
    psi = qc.state()
    assert abs(abs(psi[target]) ** 2 - A.grover_success_probability(n, k)) < 1e-11
    qc.set_seed(5)
    assert qc.measure() == target


def test_qaoa_maxcut_outer_loop(ctx):
    n, p = 10, 2
    edges = G.random_regular_graph(n, 3, seed=12)
    x0 = np.array([0.1, 0.1, 0.4, 0.4])
    x, hist = A.optimize_maxcut(ctx, n, edges, p, x0, steps=40, lr=0.003)
    assert hist[-1] > hist[0] + 5.0  # 4.9 -> 11.2 of a maximum cut of 13 (CPU oracle, same steps)
    assert all(b >= a - 1e-9 for a, b in zip(hist, hist[1:]))  # small steps: <C> never drops
    best, _ = A.brute_force_max_cut(n, edges)
    assert hist[-1] < best + 1e-9
    st = ctx.state(n, capi.C128)
    st.run(G.qaoa_maxcut(n, edges, p).ops(), x)
    u = np.random.default_rng(3).random(4000)
    cut, mask = A.best_sampled_cut(st.sample(u), edges)
    assert cut == best and A.cut_size(mask, edges) == best


def test_tsp_xy_qaoa_4_cities(ctx):
    """TSP as the constraint-preserving QAOA of TVS.cpp on 4 cities (16
    qubits) on the GPU: state within 1e-12 of the oracle, 20 000 samples equal
    the oracle's on the same uniforms, and the best sampled feasible tour is
    the brute-force optimum (the known answer of TVS::runTVS)."""
    n = 4
    rng = np.random.default_rng(7)
    pts = rng.random((n, 2)) * 10
    dist = [[float(np.hypot(*(pts[a] - pts[b]))) for b in range(n)] for a in range(n)]
    gam, bet = A.tsp_schedule(4)
    ops = A.tsp_circuit(n, A.tsp_penalty(dist), 1.0, gam, bet, dist).ops()
    st = ctx.state(n * n)
    st.run(ops)
    want = O.run_ops(O.basis(n * n), n * n, ops)
    assert np.max(np.abs(st.download() - want)) < 1e-12
    u = O.uniforms(5, 20000)
    s = st.sample(u)
    assert np.array_equal(s, O.sample(want, n * n, u))
    order, length, count = A.best_sampled_tour(s, n, dist)
    opt, opt_len = A.tsp_brute_force(dist)
    assert abs(length - opt_len) < 1e-12 and count > 0
    st.close()

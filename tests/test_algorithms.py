"""Playground algorithms (QSEplayground/Grover.cpp, Shor.cpp, QAOA.cpp) as gate
lists, checked on the CPU oracle piece by piece: the Fourier transform and
constant adder, the doubly controlled modular adder and the controlled modular
multiplier exhaustively for N = 15, Grover's success probability against its
closed form, and the classical post-processing of Shor's algorithm."""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import algorithms as A
from quantumsimevo_amd import circuits as G


def basis_index(state):
    i = int(np.argmax(np.abs(state)))
    assert abs(abs(state[i]) - 1.0) < 1e-9, "not a basis state"
    return i


def test_fourier_transform_and_inverse():
    m = 5
    g = G.GateList(m)
    A.fourier(g, list(range(m)))
    for b in (0, 1, 7, 19, 31):
        out = O.run_ops(O.basis(m, b), m, g.ops())
        # product state: qubit j carries phase 2 pi b / 2^(j+1)
        want = np.ones(1, complex)
        for j in reversed(range(m)):
            want = np.kron(want, np.array([1, np.exp(2j * math.pi * b / 2 ** (j + 1))]) / math.sqrt(2))
        assert np.max(np.abs(out - want)) < 1e-12
    h = G.GateList(m)
    A.fourier(h, list(range(m)))
    A.fourier(h, list(range(m)), inverse=True)
    psi = G.random_state(m, 4)
    assert np.max(np.abs(O.run_ops(psi, m, h.ops()) - psi)) < 1e-12


@pytest.mark.parametrize("a", [0, 1, 5, 11, -3, 17])
def test_phase_adder(a):
    m = 4
    for b in range(16):
        g = G.GateList(m)
        A.fourier(g, list(range(m)))
        A.phase_add(g, list(range(m)), a)
        A.fourier(g, list(range(m)), inverse=True)
        assert basis_index(O.run_ops(O.basis(m, b), m, g.ops())) == (b + a) % 16


def test_modular_adder_exhaustive_n15():
    N, nb = 15, 5
    c1, c2, breg, anc = 0, 1, list(range(2, 7)), 7
    nq = 8
    for ctl in ((1, 1), (1, 0), (0, 1)):
        for a in range(N):
            g = G.GateList(nq)
            A.fourier(g, breg)
            A.mod_add(g, c1, c2, breg, anc, a, N)
            A.fourier(g, breg, inverse=True)
            ops = g.ops()
            for b in range(N):
                idx = ctl[0] | (ctl[1] << 1) | (b << 2)
                out = basis_index(O.run_ops(O.basis(nq, idx), nq, ops))
                want_b = (b + a) % N if ctl == (1, 1) else b
                assert out == (ctl[0] | (ctl[1] << 1) | (want_b << 2)), (ctl, a, b)


@pytest.mark.parametrize("a", [2, 7, 11, 13])
def test_controlled_modular_multiplier(a):
    N = 15
    c, xreg, breg, anc = 0, [1, 2, 3, 4], [5, 6, 7, 8, 9], 10
    nq = 11
    g = G.GateList(nq)
    A.controlled_modmul(g, c, xreg, breg, anc, a, N)
    ops = g.ops()
    for cv in (0, 1):
        for x in range(N):
            out = basis_index(O.run_ops(O.basis(nq, cv | (x << 1)), nq, ops))
            want = (a * x) % N if cv else x
            assert out == (cv | (want << 1)), (cv, x)


@pytest.mark.parametrize("n, target", [(3, 5), (6, 41), (8, 200)])
def test_grover_success_probability(n, target):
    k = A.grover_iterations(n)
    psi = O.run_ops(O.basis(n), n, A.grover(n, target, k).ops())
    p = abs(psi[target]) ** 2
    assert abs(p - A.grover_success_probability(n, k)) < 1e-12
    others = np.delete(np.abs(psi) ** 2, target)
    assert np.max(np.abs(others - others[0])) < 1e-12  # the rest stays uniform


def test_shor_classical_steps():
    L = A.ShorLayout(15)
    assert (L.n, L.qubits, len(L.phase)) == (4, 18, 8)
    # k / 2^8 = s / 4 for a = 7 (order 4)
    for k, s in ((64, 1), (192, 3)):
        assert A.order_from_phase(k, 8, 15, 7) == 4
    assert A.order_from_phase(128, 8, 15, 7) == 4  # 1/2 -> 2 fails, its multiple 4 works
    assert A.order_from_phase(0, 8, 15, 7) is None
    assert A.factors_from_order(15, 7, 4) == (3, 5)
    assert A.factors_from_order(21, 2, 6) == (3, 7)
    idx = 0
    for i, bit in enumerate([0, 0, 0, 0, 0, 0, 1, 0]):  # k = 64: bit 6 on phase qubit 8 - 1 - 6 = 1
        if bit:
            idx |= 1 << L.phase[len(L.phase) - 1 - i]
    assert A.phase_register_value(idx, L) == 64


def test_max_cut_helpers():
    edges = [(0, 1), (1, 2), (2, 3), (3, 0)]
    assert A.brute_force_max_cut(4, edges)[0] == 4
    assert A.best_sampled_cut([0, 0b0101, 0b0011], edges) == (4, 0b0101)


def _tsp_instance(n, seed):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2)) * 10
    return [[float(np.hypot(*(pts[a] - pts[b]))) for b in range(n)] for a in range(n)]


@pytest.mark.parametrize("n,rows,cols", [(3, True, True), (4, True, True), (4, False, True), (3, True, False)])
def test_tsp_ising_equals_qubo_up_to_a_constant(n, rows, cols):
    """The Ising coefficients (tsp_ising) against a brute-force evaluation of the
    QUBO (TVS::quboCost, TVS.cpp:232-262) over every assignment of the n^2
    variables: with z = 1 - 2x, QUBO(x) - (sum h z + sum J z z) is one constant.
    A penalty left out of the Ising form is subtracted from the QUBO."""
    dist = _tsp_instance(n, 3)
    pen, wgt = 17.5, 1.3
    h, pairs, coeffs = A.tsp_ising(n, pen, wgt, dist, rows=rows, cols=cols)
    q = n * n
    xs = np.arange(1 << q, dtype=np.int64)
    z = 1.0 - 2.0 * ((xs[:, None] >> np.arange(q)) & 1)
    ising = z @ h
    for (a, b), c in zip(pairs, coeffs):
        ising += c * z[:, a] * z[:, b]
    qubo = np.array([A.tsp_qubo_cost(int(x), n, dist, pen, wgt) for x in xs])
    bits = (xs[:, None] >> np.arange(q)) & 1
    if not rows:
        qubo -= pen * sum((1 - bits[:, [A.tsp_index(v, t, n) for t in range(n)]].sum(1)) ** 2 for v in range(n))
    if not cols:
        qubo -= pen * sum((1 - bits[:, [A.tsp_index(v, t, n) for v in range(n)]].sum(1)) ** 2 for t in range(n))
    d = qubo - ising
    assert np.max(np.abs(d - d[0])) < 1e-9


def test_tsp_classical_helpers():
    dist = _tsp_instance(5, 8)
    best, best_len = A.tsp_brute_force(dist)
    import itertools
    every = min(A.tour_length(dist, list(p)) for p in itertools.permutations(range(5)))
    assert abs(best_len - every) < 1e-12 and best[0] == 0
    n = 4
    ident = sum(1 << A.tsp_index(v, v, n) for v in range(n))
    assert A.tsp_decode(ident, n) == [0, 1, 2, 3]
    assert A.tsp_decode(ident | 1 << A.tsp_index(0, 1, n), n) is None  # two cities at step 1
    assert A.tsp_decode(ident & ~(1 << A.tsp_index(2, 2, n)), n) is None  # empty step
    # a feasible assignment's QUBO cost is its route length times the weight
    assert abs(A.tsp_qubo_cost(ident, n, dist, 99.0, 2.0) - 2.0 * A.tour_length(dist, [0, 1, 2, 3])) < 1e-12


def test_tsp_qaoa_on_the_oracle_finds_the_optimum():
    """4 cities (16 qubits), the reference's defaults (4 layers, gamma 0.8,
    beta 0.4, automatic penalty): the XY mixer keeps one excitation per city
    row (no amplitude elsewhere) and the best of 20 000 sampled feasible tours
    is the brute-force optimum."""
    n = 4
    dist = _tsp_instance(n, 7)
    gam, bet = A.tsp_schedule(4)
    g = A.tsp_circuit(n, A.tsp_penalty(dist), 1.0, gam, bet, dist)
    psi = O.run_ops(O.basis(n * n), n * n, g.ops())
    idx = np.arange(1 << (n * n))
    rows_ok = np.ones(len(idx), dtype=bool)
    for v in range(n):
        cnt = sum((idx >> A.tsp_index(v, t, n)) & 1 for t in range(n))
        rows_ok &= cnt == 1
    assert np.max(np.abs(psi[~rows_ok])) < 1e-12
    s = O.sample(psi, n * n, O.uniforms(5, 20000))
    order, length, count = A.best_sampled_tour(s, n, dist)
    assert abs(length - A.tsp_brute_force(dist)[1]) < 1e-12 and count > 0

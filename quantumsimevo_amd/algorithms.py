"""End-to-end algorithms of the reference's playground (QSEplayground/Grover.cpp,
Shor.cpp, QAOA.cpp) as gate lists for the HIP backend, with their classical
pre- and post-processing, plus the classical outer loop of a variational run.

The circuits are built from the mathematics:
  * Grover search: oracle = multi-controlled Z on the target pattern (X on its
    zero bits around an MCPHASE(pi)), diffusion = H X (multi-controlled Z) X H;
    success probability sin^2((2k+1) asin(2^(-n/2))) after k iterations.
  * Shor order finding: phase estimation of U_a |x> = |a x mod N> with a 2n-bit
    phase register; U_a from Fourier-basis constant adders (Draper), a doubly
    controlled modular adder (Beauregard) and a controlled modular multiplier,
    on 4n + 2 qubits (n = bits of N).
  * QAOA Max-Cut: sampling of the optimised state, best observed cut, brute
    force for comparison (QAOA.cpp:9-103); gradient ascent on <C> with adjoint
    gradients.

Qubit conventions: a register is a list of qubit indices, element 0 the least
significant bit of its integer value.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

from .circuits import GateList

PI = math.pi


# ---- Grover ----

def grover_iterations(n: int) -> int:
    """round(pi/4 * sqrt(2^n)), at least 1 (Grover.cpp:21-22)."""
    return max(1, int(round(PI / 4.0 * math.sqrt(2.0 ** n))))


def _mcz(g: GateList, qubits: list[int]):
    """Multi-controlled Z on all of `qubits` (phase -1 where every one is 1)."""
    if len(qubits) == 1:
        g.pauli_z(qubits[0])
    else:
        g.mcphase(PI, qubits[:-1], qubits[-1])


def grover(n: int, target: int, iterations: int | None = None) -> GateList:
    """Uniform superposition, then `iterations` rounds of oracle and diffusion."""
    if not 0 <= target < (1 << n):
        raise ValueError("target outside the search space")
    k = grover_iterations(n) if iterations is None else iterations
    g = GateList(n)
    qs = list(range(n))
    for q in qs:
        g.hadamard(q)
    zeros = [q for q in qs if not (target >> q) & 1]
    for _ in range(k):
        for q in zeros:
            g.pauli_x(q)
        _mcz(g, qs)
        for q in zeros:
            g.pauli_x(q)
        for q in qs:
            g.hadamard(q)
        for q in qs:
            g.pauli_x(q)
        _mcz(g, qs)
        for q in qs:
            g.pauli_x(q)
        for q in qs:
            g.hadamard(q)
    return g


def grover_success_probability(n: int, iterations: int) -> float:
    theta = math.asin(2.0 ** (-n / 2))
    return math.sin((2 * iterations + 1) * theta) ** 2


# ---- Fourier-basis arithmetic ----

def fourier(g: GateList, reg: list[int], inverse: bool = False):
    """|b> -> prod_j (|0> + exp(2 pi i b / 2^(j+1)) |1>)/sqrt(2) on reg[j]
    (a QFT without the final swaps); the inverse undoes it."""
    m = len(reg)
    order = range(m - 1, -1, -1)
    if not inverse:
        for j in order:
            g.hadamard(reg[j])
            for k in range(j - 1, -1, -1):
                g.cphase(PI / 2 ** (j - k), reg[k], reg[j])
    else:
        for j in reversed(list(order)):
            for k in range(0, j):
                g.cphase(-PI / 2 ** (j - k), reg[k], reg[j])
            g.hadamard(reg[j])


def _add_angles(a: int, m: int) -> list[float]:
    """Phase on reg[j] that adds a (mod 2^m) to a register in the Fourier basis."""
    a %= 1 << m
    return [2 * PI * a / 2 ** (j + 1) for j in range(m)]


def phase_add(g: GateList, reg: list[int], a: int, controls: tuple[int, ...] = ()):
    """b -> b + a (mod 2^len(reg)) in the Fourier basis, under 0, 1 or 2 controls."""
    for q, th in zip(reg, _add_angles(a, len(reg))):
        th = math.remainder(th, 2 * PI)
        if th == 0.0:
            continue
        if not controls:
            g.phase(th, q)
        elif len(controls) == 1:
            g.cphase(th, controls[0], q)
        else:
            g.mcphase(th, list(controls), q)


def toffoli(g: GateList, c1: int, c2: int, t: int):
    g.hadamard(t)
    g.mcphase(PI, [c1, c2], t)
    g.hadamard(t)


def controlled_swap(g: GateList, c: int, a: int, b: int):
    g.cnot(b, a)
    toffoli(g, c, a, b)
    g.cnot(b, a)


def mod_add(g: GateList, c1: int, c2: int, breg: list[int], anc: int, a: int, N: int):
    """Doubly controlled b -> (b + a) mod N with breg (n + 1 qubits, b < N) in the
    Fourier basis on entry and exit and the ancilla returned to |0>."""
    msb = breg[-1]
    phase_add(g, breg, a, (c1, c2))
    phase_add(g, breg, -N)
    fourier(g, breg, inverse=True)
    g.cnot(msb, anc)
    fourier(g, breg)
    phase_add(g, breg, N, (anc,))
    phase_add(g, breg, -a, (c1, c2))
    fourier(g, breg, inverse=True)
    g.pauli_x(msb)
    g.cnot(msb, anc)
    g.pauli_x(msb)
    fourier(g, breg)
    phase_add(g, breg, a, (c1, c2))


def controlled_mult_add(g: GateList, c: int, xreg: list[int], breg: list[int], anc: int, a: int, N: int):
    """|c>|x>|b> -> |c>|x>|(b + a x) mod N> when c = 1."""
    fourier(g, breg)
    f = a % N
    for xq in xreg:
        mod_add(g, c, xq, breg, anc, f, N)
        f = (2 * f) % N
    fourier(g, breg, inverse=True)


def controlled_modmul(g: GateList, c: int, xreg: list[int], breg: list[int], anc: int, a: int, N: int):
    """|c>|x>|0> -> |c>|a x mod N>|0> when c = 1 (gcd(a, N) = 1, x < N)."""
    n = len(xreg)
    controlled_mult_add(g, c, xreg, breg, anc, a, N)
    for i in range(n):
        controlled_swap(g, c, xreg[i], breg[i])
    inv = pow(a, -1, N)
    # b currently holds x; subtracting inv * (a x) mod N clears it
    fourier(g, breg)
    f = inv % N
    for xq in xreg:
        mod_add(g, c, xq, breg, anc, (N - f) % N, N)
        f = (2 * f) % N
    fourier(g, breg, inverse=True)


class ShorLayout:
    """Qubits of the order-finding circuit: 2n phase qubits, x (n), b (n + 1), ancilla."""

    def __init__(self, N: int):
        self.N = N
        self.n = max(1, math.ceil(math.log2(N)))
        n = self.n
        self.phase = list(range(0, 2 * n))
        self.x = list(range(2 * n, 3 * n))
        self.b = list(range(3 * n, 4 * n + 1))
        self.anc = 4 * n + 1
        self.qubits = 4 * n + 2


def shor_order_finding(N: int, a: int) -> tuple[GateList, ShorLayout]:
    """Phase estimation of x -> a x mod N from x = 1. After the circuit the
    phase register holds k with k / 2^(2n) close to s / r (r the order of a):
    bit i of k is on phase qubit 2n - 1 - i (phase_register_value)."""
    if math.gcd(a, N) != 1:
        raise ValueError("a must be coprime to N")
    L = ShorLayout(N)
    t = len(L.phase)
    g = GateList(L.qubits)
    g.pauli_x(L.x[0])
    for q in L.phase:
        g.hadamard(q)
    for j, c in enumerate(L.phase):
        controlled_modmul(g, c, L.x, L.b, L.anc, pow(a, 2 ** j, N), N)
    # phase qubit j now carries exp(2 pi i phi 2^j): the Fourier state of
    # k = 2^t phi on the reversed register, so invert it there
    fourier(g, [L.phase[t - 1 - i] for i in range(t)], inverse=True)
    return g, L


def phase_register_value(index: int, L: ShorLayout) -> int:
    """k from a measured basis index: bit i of k is on phase qubit t - 1 - i."""
    t = len(L.phase)
    return sum(((index >> L.phase[t - 1 - i]) & 1) << i for i in range(t))


def order_from_phase(k: int, t: int, N: int, a: int) -> int | None:
    """Smallest r <= N with a^r = 1 mod N among the continued-fraction
    convergents of k / 2^t and their small multiples, or None."""
    if k == 0:
        return None
    f = Fraction(k, 1 << t).limit_denominator(N)
    for m in range(1, N + 1):
        r = f.denominator * m
        if r > N:
            break
        if pow(a, r, N) == 1:
            return r
    return None


def factors_from_order(N: int, a: int, r: int) -> tuple[int, int] | None:
    if r % 2 or pow(a, r // 2, N) == N - 1:
        return None
    p = math.gcd(pow(a, r // 2, N) - 1, N)
    q = math.gcd(pow(a, r // 2, N) + 1, N)
    for d in (p, q):
        if 1 < d < N:
            return tuple(sorted((d, N // d)))
    return None


# ---- QAOA Max-Cut ----

def cut_size(mask: int, edges) -> int:
    return sum(((mask >> u) & 1) != ((mask >> v) & 1) for u, v in edges)


def brute_force_max_cut(n: int, edges) -> tuple[int, int]:
    """(best cut, one optimal mask) over all 2^(n-1) partitions (QAOA.cpp:9-48)."""
    best, arg = -1, 0
    for m in range(1 << (n - 1)):
        c = cut_size(m, edges)
        if c > best:
            best, arg = c, m
    return best, arg


def best_sampled_cut(samples, edges) -> tuple[int, int]:
    """(cut, mask) of the best distinct sampled partition (QAOA.cpp:65-90)."""
    best, arg = -1, 0
    for m in sorted(set(int(s) for s in samples)):
        c = cut_size(m, edges)
        if c > best:
            best, arg = c, m
    return best, arg


def optimize_maxcut(ctx, n: int, edges, p: int, params0, steps: int = 40, lr: float = 0.003, precision=None):
    """Gradient ascent on <C>(gamma, beta) with adjoint gradients: the classical
    outer loop of a variational run. Returns (params, history of <C>)."""
    from . import capi
    from .circuits import maxcut_observable, qaoa_maxcut
    prec = capi.C128 if precision is None else precision
    ops = qaoa_maxcut(n, edges, p).ops()
    masks, w = maxcut_observable(edges)
    x = np.array(params0, dtype=np.float64).reshape(1, -1)
    hist = []
    for _ in range(steps):
        E, G = ctx.adjoint_expect(n, prec, ops, x, masks, w)
        hist.append(float(E[0]))
        x = x + lr * G
    E, _ = ctx.adjoint_expect(n, prec, ops, x, masks, w)
    hist.append(float(E[0]))
    return x[0], hist


# ---- TSP as a constraint-preserving QAOA (QSEplayground/TVS.h:1-71, TVS.cpp:10-381) ----
#
# Binary variable x_{v,t} = 1: city v is visited at step t, on qubit v*n + t (n^2
# qubits). QUBO cost of a closed tour:
#   A sum_v (1 - sum_t x_{v,t})^2 + A sum_t (1 - sum_v x_{v,t})^2
#   + B sum_t sum_{u != v} w_uv x_{u,t} x_{v,t+1 mod n}.
# With x = (1 - Z)/2 a term c x_i becomes -c/2 Z_i (+ const) and q x_i x_j becomes
# q/4 (Z_i Z_j - Z_i - Z_j) (+ const). The circuit starts in the identity tour
# (city v at step v), applies p layers of the phase separator e^{-i gamma H_C}
# (route cost and the one-city-per-step penalty: the row constraint is kept by
# the mixer) and an XY ring mixer e^{-i beta (X_a X_b + Y_a Y_b)} on neighbouring
# steps of each city's row, which moves the row's single excitation and so never
# leaves the one-step-per-city subspace.

def tsp_index(city: int, step: int, n: int) -> int:
    """Qubit of x_{city,step} (TVS.cpp:9-12)."""
    return city * n + step


def tsp_ising(n: int, penalty: float, weight: float, dist, rows: bool = True, cols: bool = True):
    """Ising form h_i Z_i + J_ij Z_i Z_j (+ a constant) of the TSP QUBO above:
    returns (h, pairs, coeffs) with pairs (i < j) and J in matching order
    (TVS::buildIsing, TVS.cpp:30-101; rows / cols select the two penalties)."""
    q = n * n
    h = np.zeros(q)
    J: dict[tuple[int, int], float] = {}

    def linear(i, c):
        h[i] -= c / 2.0

    def quad(i, j, c):
        h[i] -= c / 4.0
        h[j] -= c / 4.0
        key = (min(i, j), max(i, j))
        J[key] = J.get(key, 0.0) + c / 4.0

    def one_hot(group):
        # A (1 - S)^2 = A (1 - sum x_i + 2 sum_{i<j} x_i x_j) for 0/1 variables
        for a in range(len(group)):
            linear(group[a], -penalty)
            for b in range(a + 1, len(group)):
                quad(group[a], group[b], 2.0 * penalty)

    if rows:
        for v in range(n):
            one_hot([tsp_index(v, t, n) for t in range(n)])
    if cols:
        for t in range(n):
            one_hot([tsp_index(v, t, n) for v in range(n)])
    for u in range(n):
        for v in range(n):
            if u != v:
                for t in range(n):
                    quad(tsp_index(u, t, n), tsp_index(v, (t + 1) % n, n), weight * dist[u][v])
    pairs = [k for k, c in J.items() if abs(c) >= 1e-15]
    return h, pairs, [J[k] for k in pairs]


def tsp_qubo_cost(bits: int, n: int, dist, penalty: float, weight: float) -> float:
    """QUBO cost of an assignment (bit v*n + t = x_{v,t}) (TVS.cpp:232-262)."""
    x = [(bits >> i) & 1 for i in range(n * n)]
    cost = 0.0
    for v in range(n):
        s = sum(x[tsp_index(v, t, n)] for t in range(n))
        cost += penalty * (1 - s) ** 2
    for t in range(n):
        s = sum(x[tsp_index(v, t, n)] for v in range(n))
        cost += penalty * (1 - s) ** 2
    for t in range(n):
        for u in range(n):
            for v in range(n):
                if u != v and x[tsp_index(u, t, n)] and x[tsp_index(v, (t + 1) % n, n)]:
                    cost += weight * dist[u][v]
    return cost


def tsp_circuit(n: int, penalty: float, weight: float, gammas, betas, dist) -> GateList:
    """The p-layer constraint-preserving QAOA of TVS::createCircuit
    (TVS.cpp:121-186): identity tour; per layer the ZZ phases (CNOT, RZ(2 gamma
    J), CNOT), the Z phases RZ(2 gamma h), then for every row v and step t the
    XY interaction on (v,t), (v,t+1 mod n) as Rxx(2 beta) (H-conjugated Rzz)
    followed by Ryy(2 beta) (RX(pi/2)-conjugated Rzz)."""
    h, pairs, coeffs = tsp_ising(n, penalty, weight, dist, rows=False, cols=True)
    g = GateList(n * n)
    for v in range(n):
        g.pauli_x(tsp_index(v, v, n))

    def rzz(a, b, theta):
        g.cnot(a, b).rotate_z(theta, b).cnot(a, b)

    for gamma, beta in zip(gammas, betas):
        for (a, b), c in zip(pairs, coeffs):
            rzz(a, b, 2.0 * gamma * c)
        for q in range(n * n):
            g.rotate_z(2.0 * gamma * h[q], q)
        for v in range(n):
            for t in range(n):
                a, b = tsp_index(v, t, n), tsp_index(v, (t + 1) % n, n)
                g.hadamard(a).hadamard(b)
                rzz(a, b, 2.0 * beta)
                g.hadamard(a).hadamard(b)
                g.rotate_x(PI / 2, a).rotate_x(PI / 2, b)
                rzz(a, b, 2.0 * beta)
                g.rotate_x(-PI / 2, a).rotate_x(-PI / 2, b)
    return g


def tsp_decode(sample: int, n: int) -> list[int] | None:
    """Visiting order [city at step 0, ...], or None when some step has not
    exactly one city or a city repeats (TVS::decodeSample, TVS.cpp:191-210)."""
    order, used = [], set()
    for t in range(n):
        cities = [v for v in range(n) if (sample >> tsp_index(v, t, n)) & 1]
        if len(cities) != 1 or cities[0] in used:
            return None
        used.add(cities[0])
        order.append(cities[0])
    return order


def tour_length(dist, order) -> float:
    """Closed tour length (TVS.cpp:267-274)."""
    return sum(dist[order[i]][order[(i + 1) % len(order)]] for i in range(len(order)))


def tsp_brute_force(dist) -> tuple[list[int], float]:
    """Shortest closed tour with city 0 first (TVS.cpp:276-294)."""
    import itertools
    n = len(dist)
    best, best_len = None, math.inf
    for perm in itertools.permutations(range(1, n)):
        order = [0, *perm]
        length = tour_length(dist, order)
        if length < best_len:
            best, best_len = order, length
    return best, best_len


def tsp_schedule(layers: int, gamma: float = 0.8, beta: float = 0.4):
    """Per-layer angles of TVS::runTVS (TVS.cpp:320-327): gamma ramps up,
    beta ramps down and keeps a quarter on the last layer."""
    gammas, betas = [], []
    for k in range(layers):
        frac = 1.0 if layers == 1 else (k + 1) / layers
        gammas.append(gamma * frac)
        betas.append(beta * (1.0 - frac) + beta * 0.25)
    return gammas, betas


def tsp_penalty(dist, weight: float = 1.0) -> float:
    """Automatic penalty A = B * max w * n + 1 (TVS.cpp:311-317)."""
    n = len(dist)
    return weight * max(max(r) for r in dist) * n + 1.0


def best_sampled_tour(samples, n: int, dist):
    """Shortest feasible tour among the samples and how often it was drawn
    (TVS.cpp:335-357): (order, length, count) or (None, inf, 0)."""
    vals, counts = np.unique(np.asarray(samples, dtype=np.uint64), return_counts=True)
    best, best_len, best_count = None, math.inf, 0
    for s, c in zip(vals, counts):
        order = tsp_decode(int(s), n)
        if order is not None:
            length = tour_length(dist, order)
            if length < best_len:
                best, best_len, best_count = order, length, int(c)
    return best, best_len, best_count

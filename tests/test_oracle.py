"""Pins the CPU oracle (oracle/qse_oracle.cpp) before the GPU tests compare against it.

Each oracle function is checked against something independent of its own formula:
dense Kronecker-product matrices (numpy), the reference's known-answer states
(QSETests/QSETests.cpp:76-170), golden vectors written by the compiled reference
(tests/golden/ref_vectors.npz), the DFT closed form of the QFT, brute-force
lower_bound sampling, the std::mt19937 known answer and the reference Noise
class's own flip patterns.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as C

GOLD = np.load(O.ROOT / "tests" / "golden" / "ref_vectors.npz", allow_pickle=False)
PI = math.pi
S2 = 1.0 / math.sqrt(2.0)


def kron_op(n, q, m):
    """Dense 2^n x 2^n matrix of 2x2 m on qubit q (qubit 0 = least significant)."""
    full = np.array([[1.0]], dtype=np.complex128)
    for b in range(n - 1, -1, -1):
        full = np.kron(full, m if b == q else np.eye(2))
    return full


def dense_diag(n, f):
    return np.diag([f(i) for i in range(1 << n)]).astype(np.complex128)


def dense_perm(n, f):
    N = 1 << n
    P = np.zeros((N, N), dtype=np.complex128)
    for i in range(N):
        P[f(i), i] = 1.0
    return P


def gate_matrix(n, kind, q, c=None, th=0.0, controls=()):
    if kind == "H":
        return kron_op(n, q, S2 * np.array([[1, 1], [1, -1]]))
    if kind == "X":
        return kron_op(n, q, np.array([[0, 1], [1, 0]]))
    if kind == "Y":
        return kron_op(n, q, np.array([[0, -1j], [1j, 0]]))
    if kind == "Z":
        return kron_op(n, q, np.diag([1, -1]))
    if kind == "P":
        return kron_op(n, q, np.diag([1, np.exp(1j * th)]))
    c2, s2 = math.cos(th / 2), math.sin(th / 2)
    if kind == "RX":
        return kron_op(n, q, np.array([[c2, -1j * s2], [-1j * s2, c2]]))
    if kind == "RY":
        return kron_op(n, q, np.array([[c2, -s2], [s2, c2]]))
    if kind == "RZ":
        return kron_op(n, q, np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)]))
    if kind == "CNOT":
        return dense_perm(n, lambda i: i ^ (1 << q) if (i >> c) & 1 else i)
    if kind == "CP":
        return dense_diag(n, lambda i: np.exp(1j * th) if ((i >> c) & 1 and (i >> q) & 1) else 1)
    if kind == "SWAP":
        def sw(i):
            a, b = (i >> q) & 1, (i >> c) & 1
            return (i & ~((1 << q) | (1 << c))) | (b << q) | (a << c)
        return dense_perm(n, sw)
    if kind == "MCP":
        mask = (1 << q) | sum(1 << x for x in controls)
        return dense_diag(n, lambda i: np.exp(1j * th) if (i & mask) == mask else 1)
    raise KeyError(kind)


CASES = [("H", 2, None, 0.0), ("X", 0, None, 0.0), ("Y", 4, None, 0.0), ("Z", 1, None, 0.0),
         ("P", 3, None, 0.9), ("RX", 0, None, 2.2), ("RY", 5, None, -1.3), ("RZ", 2, None, 0.6),
         ("CNOT", 1, 4, 0.0), ("CNOT", 4, 0, 0.0), ("CP", 0, 5, 1.1), ("SWAP", 1, 3, 0.0), ("SWAP", 5, 0, 0.0)]


@pytest.mark.parametrize("kind,q,c,th", CASES)
def test_gates_vs_dense_kron(kind, q, c, th):
    n = 6
    psi = C.random_state(n, 11)
    g = C.GateList(n)
    {"H": lambda: g.hadamard(q), "X": lambda: g.pauli_x(q), "Y": lambda: g.pauli_y(q), "Z": lambda: g.pauli_z(q),
     "P": lambda: g.phase(th, q), "RX": lambda: g.rotate_x(th, q), "RY": lambda: g.rotate_y(th, q),
     "RZ": lambda: g.rotate_z(th, q), "CNOT": lambda: g.cnot(c, q), "CP": lambda: g.cphase(th, c, q),
     "SWAP": lambda: g.swap(q, c)}[kind]()
    got = O.run_ops(psi, n, g.ops())
    want = gate_matrix(n, kind, q, c, th) @ psi
    assert np.max(np.abs(got - want)) < 1e-14


def test_mcphase_and_matrix_vs_dense():
    n = 6
    psi = C.random_state(n, 12)
    got = O.run_ops(psi, n, C.GateList(n).mcphase(0.7, [0, 3], 5).ops())
    want = gate_matrix(n, "MCP", 5, th=0.7, controls=(0, 3)) @ psi
    assert np.max(np.abs(got - want)) < 1e-14
    m = np.array([[0.6, 0.8j], [0.8j, 0.6]])
    got = O.run_ops(psi, n, C.GateList(n).gate(m, 2).ops())
    assert np.max(np.abs(got - kron_op(n, 2, m) @ psi)) < 1e-14


def test_qsetests_known_answers():
    """QSETests.cpp:76-118 (1 qubit) and :120-170 (SWAP cases), tolerance 1e-12."""
    def run1(f):
        g = C.GateList(1)
        f(g)
        return O.run_ops(O.basis(1), 1, g.ops())
    c, s = math.cos(PI / 4), math.sin(PI / 4)
    cases = [
        (lambda g: g.hadamard(0), [S2, S2]),
        (lambda g: g.pauli_x(0), [0, 1]),
        (lambda g: g.pauli_y(0), [0, 1j]),
        (lambda g: g.pauli_z(0), [1, 0]),
        (lambda g: g.phase(PI / 2, 0), [1, 0]),
        (lambda g: g.rotate_x(PI / 2, 0), [c, -1j * s]),
        (lambda g: g.rotate_y(PI / 2, 0), [c, s]),
        (lambda g: g.rotate_z(PI / 2, 0), [c - 1j * s, 0]),
    ]
    for f, want in cases:
        assert np.max(np.abs(run1(f) - np.array(want))) <= 1e-12
    swaps = [
        (lambda g: g.swap(0, 1), [1, 0, 0, 0]),
        (lambda g: (g.pauli_x(0), g.swap(0, 1)), [0, 0, 1, 0]),
        (lambda g: (g.pauli_x(1), g.swap(0, 1)), [0, 1, 0, 0]),
        (lambda g: (g.hadamard(0), g.swap(0, 1)), [S2, 0, S2, 0]),
        (lambda g: (g.pauli_x(0), g.swap(0, 1), g.swap(0, 1)), [0, 1, 0, 0]),
    ]
    for f, want in swaps:
        g = C.GateList(2)
        f(g)
        assert np.max(np.abs(O.run_ops(O.basis(2), 2, g.ops()) - np.array(want))) <= 1e-12


def bitrev(i, n):
    return int(format(i, f"0{n}b")[::-1], 2)


@pytest.mark.parametrize("n", [3, 6, 9])
def test_qft_closed_form(n):
    """QFT.cpp:9-22 applies H to qubit 0 first; with the final swaps the circuit is
    R F+ R with F+ the unitary DFT exp(+2 pi i x k / N) and R the n-bit reversal."""
    N = 1 << n
    psi = C.random_state(n, 5 + n)
    got = O.run_ops(psi, n, C.qft(n).ops())
    R = np.array([bitrev(i, n) for i in range(N)])
    x = psi[R]
    Fx = np.fft.ifft(x) * math.sqrt(N)  # (1/sqrt N) sum_x exp(+2 pi i x k/N) v_x
    want = Fx[R]
    assert np.max(np.abs(got - want)) < 1e-12
    back = O.run_ops(got, n, C.iqft(n).ops())
    assert np.max(np.abs(back - psi)) < 1e-12


def gold_ops(key):
    return GOLD[key].view(capi.OP_DTYPE)


def test_oracle_matches_reference_golden_gates():
    psi = GOLD["gate_input"]
    for name in GOLD["gate_names"]:
        ops = gold_ops(f"gate_{name}_ops")
        got = O.run_ops(psi, 6, ops)
        assert np.max(np.abs(got - GOLD[f"gate_{name}_scalar"])) < 1e-13, name
        if not name.startswith("Z0"):
            # AVX2 backend agrees except its known no-op PauliZ on qubit 0 (CircuitUnitaryOperationAVX.cpp:132-133)
            assert np.max(np.abs(got - GOLD[f"gate_{name}_avx"])) < 1e-13, name
        else:
            assert np.max(np.abs(GOLD[f"gate_{name}_avx"] - psi)) == 0.0


def test_oracle_matches_reference_golden_circuits():
    got = O.run_ops(O.basis(10), 10, gold_ops("layered_ops"))
    assert np.max(np.abs(got - GOLD["layered_out"])) < 1e-12
    got = O.run_ops(GOLD["qft_input"], 8, gold_ops("qft_ops"))
    assert np.max(np.abs(got - GOLD["qft_out"])) < 1e-12


@pytest.mark.skipif(not O.REF_PATH.exists(), reason="reference CPU backend (oracle/_ref) not built here")
def test_oracle_matches_live_reference():
    n = 11
    ops = C.random_layered(n, 12, seed=21).ops()
    psi = C.random_state(n, 3)
    got = O.run_ops(psi, n, ops)
    assert np.max(np.abs(got - O.ref_run_ops(psi, n, ops, backend=0))) < 1e-12
    ops = C.qft(n).ops()
    assert np.max(np.abs(O.run_ops(psi, n, ops) - O.ref_run_ops(psi, n, ops, backend=1))) < 1e-12


def probs_seq(psi):
    return psi.real * psi.real + psi.imag * psi.imag


def test_sample_is_lower_bound_of_sequential_cdf():
    n = 10
    psi = C.random_state(n, 8)
    u = np.random.default_rng(1).random(5000)
    u[:4] = [0.0, 1e-300, 0.5, np.nextafter(1.0, 0.0)]
    cum = np.cumsum(probs_seq(psi))  # numpy cumsum is the sequential order
    want = np.searchsorted(cum, u * cum[-1], side="left")
    assert np.array_equal(O.sample(psi, n, u), want.astype(np.uint64))


def test_sample_invariants():
    # basis state: every shot hits it; GHZ: only 0 and 2^n - 1
    n = 8
    u = np.random.default_rng(2).random(2000)
    assert np.all(O.sample(O.basis(n, 77), n, u) == 77)
    g = C.GateList(n).hadamard(0)
    for q in range(1, n):
        g.cnot(0, q)
    ghz = O.run_ops(O.basis(n), n, g.ops())
    s = O.sample(ghz, n, u)
    assert set(np.unique(s).tolist()) == {0, (1 << n) - 1}
    assert abs(np.mean(s == 0) - 0.5) < 0.05


def test_measure_rule_and_collapse():
    n = 7
    psi = C.random_state(n, 4)
    for u in (0.0, 0.3, 0.999):
        idx, st = O.measure(psi, n, u)
        cum = np.cumsum(probs_seq(psi))
        want = int(np.searchsorted(cum, u * cum[-1], side="left"))
        assert idx == want
        assert st[idx] == 1 and np.count_nonzero(st) == 1


def test_collapse_qubit_and_reset():
    n = 6
    psi = C.random_state(n, 9)
    p1 = sum(abs(psi[i]) ** 2 for i in range(1 << n) if (i >> 2) & 1)
    o, st = O.collapse_qubit(psi, n, 2, 0.0, reset=False)
    assert o == 1
    want = np.where([(i >> 2) & 1 for i in range(1 << n)], psi, 0) / math.sqrt(p1)
    assert np.max(np.abs(st - want)) < 1e-14
    o, st = O.collapse_qubit(psi, n, 2, 0.999999, reset=True)
    assert o == 0
    assert all(st[i] == 0 for i in range(1 << n) if (i >> 2) & 1)
    assert abs(O.norm2(st, n) - 1.0) < 1e-13


def test_expectation_brute_force():
    n = 7
    psi = C.random_state(n, 10)
    edges = [(0, 1), (1, 2), (2, 6), (3, 5)]
    masks, w = C.maxcut_observable(edges)
    got = O.expect_zstrings(psi, n, masks, w)
    p = np.abs(psi) ** 2
    cut = np.array([sum(((i >> a) & 1) != ((i >> b) & 1) for a, b in edges) for i in range(1 << n)])
    assert abs(got - float(np.sum(p * cut))) < 1e-13


def test_pauli_frame_closed_form():
    n = 6
    psi = C.random_state(n, 13)
    x, z = 0b101001, 0b011010
    got = O.pauli_frame(psi, n, x, z)
    want = np.array([(-1) ** bin(i & z).count("1") * psi[i ^ x] for i in range(1 << n)])
    assert np.max(np.abs(got - want)) == 0.0


def test_mt19937_known_answer():
    assert O.oracle().orc_mt19937_kat() == 4123659995


def test_noise_frames_match_reference_noise_class():
    ops = gold_ops("noise_ops")
    for row in GOLD["noise_frames"]:
        seed, mult = int(row["seed"]), float(row["mult"])
        x, z = O.noise_trajectories(12, ops, seed, 1, rate=-1.0, mult=mult)
        assert int(x[0]) == int(row["x"]) and int(z[0]) == int(row["z"]), (seed, mult)
        order = GOLD[f"noise_order_{seed}_{int(mult)}"].tolist()
        # the reference applies X_q then Z_q for q ascending (Noise.cpp:34-44)
        want = []
        for q in range(12):
            if (int(row["x"]) >> q) & 1:
                want.append(q + 1)
            if (int(row["z"]) >> q) & 1:
                want.append(-(q + 1))
        assert order == want
    assert any(int(r["x"]) or int(r["z"]) for r in GOLD["noise_frames"])


def test_product_noise_frames_match_reference_draws():
    """The Python module's trajectory frames (host Noise class) against the
    oracle's restatement of Noise.cpp (pinned to the reference Noise elsewhere)."""
    from quantumsimevo_amd import _quantumsimevo as Q
    n = 12
    ops = C.random_layered(n, 10, seed=3).ops()
    for rate, mult in ((-1.0, 1.0), (1e-2, 1.0), (-1.0, 25.0)):
        x, z = Q.noise_frames(n, ops["kind"], ops["target"], np.arange(40, dtype=np.uint32), rate, mult)
        xo, zo = O.noise_trajectories(n, ops, 0, 40, rate, mult)
        assert np.array_equal(x, xo) and np.array_equal(z, zo)


def adjoint_test_circuit(n, seed):
    """Every parameterised kind (RX, RY, RZ, PHASE, CPHASE, MCPHASE) between
    fixed H, X, Y, Z, CNOT, SWAP and a generic complex 2x2, some parameters
    shared by several gates with different scales."""
    rng = np.random.default_rng(seed)
    g = C.GateList(n)
    for q in range(n):
        g.hadamard(q)
    p = 0
    for layer in range(2):
        for q in range(n):
            k = (q + layer) % 6
            sc = float(rng.choice([1.0, -0.5, 2.0]))
            if k == 0: g.rotate_x(0.2, q, param=p, scale=sc)
            elif k == 1: g.rotate_y(-0.3, q, param=p, scale=sc)
            elif k == 2: g.rotate_z(0.1, q, param=p, scale=sc)
            elif k == 3: g.phase(0.4, q, param=p, scale=sc)
            elif k == 4: g.cphase(-0.2, q, (q + 2) % n, param=p, scale=sc)
            else: g._add(C.MCPHASE, q, theta=0.3, param=p, scale=sc, controls=(1 << ((q + 1) % n)) | (1 << ((q + 3) % n)))
            p += 1
        g.cnot(0, n - 1); g.rotate_z(0.0, n - 1, param=0, scale=1.5); g.cnot(0, n - 1)
        g.pauli_x(1); g.pauli_y(2); g.pauli_z(3); g.swap(0, n - 2)
        a, b = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        nrm = math.sqrt(abs(a) ** 2 + abs(b) ** 2)
        a, b = a / nrm, b / nrm
        g.gate(np.array([[a, -np.conj(b)], [b, np.conj(a)]]), layer + 1)
        for q in range(layer % 2, n - 1, 2):
            g.cnot(q, q + 1)
    return g.ops(), rng.random(p) * 2 * PI


def shift_gradient(n, ops, params, masks, w):
    """Parameter shift over every occurrence: d/dtheta_k = [E(+pi/2) - E(-pi/2)]/2
    (generators with eigenvalue gap 1), summed per parameter with the op's scale."""
    g = np.zeros(len(params))
    for k in np.nonzero(ops["param"] >= 0)[0]:
        for s in (+1.0, -1.0):
            sh = ops.copy()
            sh["theta"][k] += s * PI / 2
            psi = O.run_ops(O.basis(n), n, sh, params)
            g[int(ops[k]["param"])] += s * 0.5 * float(ops[k]["scale"]) * O.expect_zstrings(psi, n, masks, w)
    return g


def test_oracle_adjoint_vs_parameter_shift_and_fd():
    """The oracle's plain adjoint sweep against its own parameter shift (built on
    the pinned gate sweeps and expectation) and central finite differences."""
    n = 7
    ops, params = adjoint_test_circuit(n, 4)
    masks = np.array([1 << q for q in range(n)] + [0b1010, 0b1100001, 0], dtype=np.uint64)
    w = np.array([1.0] * n + [0.5, -0.75, 0.3])
    E, g = O.adjoint_expect(n, ops, params, masks, w)
    psi = O.run_ops(O.basis(n), n, ops, params)
    assert abs(E - O.expect_zstrings(psi, n, masks, w)) < 1e-13
    assert np.max(np.abs(g - shift_gradient(n, ops, params, masks, w))) < 1e-12
    h = 1e-5
    for j in range(len(params)):
        pp, pm = params.copy(), params.copy()
        pp[j] += h
        pm[j] -= h
        fd = (O.expect_zstrings(O.run_ops(O.basis(n), n, ops, pp), n, masks, w) -
              O.expect_zstrings(O.run_ops(O.basis(n), n, ops, pm), n, masks, w)) / (2 * h)
        assert abs(g[j] - fd) < 1e-8, j
    # a planted sign error in one generator term is caught by the shift gradient
    bad = ops.copy()
    ry = np.nonzero(bad["kind"] == C.RY)[0][0]
    bad["scale"][ry] = -bad["scale"][ry]
    _, gb = O.adjoint_expect(n, bad, params, masks, w)
    assert np.max(np.abs(gb - shift_gradient(n, ops, params, masks, w))) > 1e-3


def test_oracle_adjoint_qaoa_closed_form_p1():
    """QAOA p=1 Max-Cut on a triangle-free 3-regular graph has a closed form
    (Wang, Hadfield, Jiang, Rieffel, PRA 97, 022304 (2018), Thm. 1): with
    U_C = exp(-i g C), U_B = exp(-i b sum X), each edge gives
    <C_uv> = 1/2 + (1/2) sin(4b) sin(g) cos^2(g). Here the edge gate
    CNOT.RZ(4 gamma).CNOT = exp(-2i gamma Z_u Z_v) is U_C with g = -4 gamma and
    RX(2 beta) = exp(-i beta X), so <C> = m (1/2 - (1/2) sin(4 beta) sin(4 gamma)
    cos^2(4 gamma)); checked together with its derivatives."""
    n = 8
    edges = C.random_regular_graph(n, 3, seed=3)
    es = set(edges)
    tri = any((a, b) in es and (b, c) in es and (a, c) in es
              for a in range(n) for b in range(a + 1, n) for c in range(b + 1, n))
    assert not tri, "the seeded graph must be triangle-free for this closed form"
    ops = C.qaoa_maxcut(n, edges, 1).ops()
    masks, w = C.maxcut_observable(edges)

    def closed(gm, bt):
        return len(edges) * (0.5 - 0.5 * math.sin(4 * bt) * math.sin(4 * gm) * math.cos(4 * gm) ** 2)
    h = 1e-6
    for gam, bet in ((0.37, 0.91), (1.1, 0.2), (2.5, 2.9)):
        E, g = O.adjoint_expect(n, ops, np.array([gam, bet]), masks, w)
        assert abs(E - closed(gam, bet)) < 1e-12
        assert abs(g[0] - (closed(gam + h, bet) - closed(gam - h, bet)) / (2 * h)) < 1e-7
        assert abs(g[1] - (closed(gam, bet + h) - closed(gam, bet - h)) / (2 * h)) < 1e-7


def test_oracle_measured_replay():
    """Mid-circuit MEASURE / RESET_QUBIT replay (orc_run_measured) against closed
    forms: a Bell pair's two measurements agree, the outcome rule is u < P(1),
    reset leaves the qubit in |0> with the other qubits' amplitudes renormalised,
    and unitary stretches equal the plain replay."""
    n = 3
    g = C.GateList(n).hadamard(0).cnot(0, 1).measure(0, 0).measure(1, 1)
    for u0, want in ((0.2, 1), (0.7, 0), (0.4999, 1), (0.5001, 0)):  # P(q0 = 1) = 1/2 (to rounding)
        psi, res = O.run_measured(O.basis(n), n, g.ops(), [u0, 0.9], 2)
        assert res.tolist() == [want, want]
        assert abs(psi[want * 3] - 1) < 1e-15 and abs(np.sum(np.abs(psi) ** 2) - 1) < 1e-15
    # RY(t) on qubit 2, then reset: P(1) = sin^2(t/2); either branch ends in |0> on qubit 2,
    # and an H after it gives |+> there
    t = 1.0
    p1 = math.sin(t / 2) ** 2
    ops = C.GateList(n).pauli_x(0).rotate_y(t, 2).reset_qubit(2).hadamard(2).measure(2, 0).ops()
    for u in (p1 - 1e-9, p1 + 1e-9):
        psi, res = O.run_measured(O.basis(n), n, ops, [u, 0.25], 1)
        want = np.zeros(8, dtype=complex)
        want[0b001 | (4 if res[0] else 0)] = 1.0  # H then MEASURE(2) with u = 0.25 < 1/2 -> outcome 1
        assert res[0] == 1 and np.max(np.abs(psi - want)) < 1e-15
    # unitary stretches between measurements equal the plain replay
    lay = C.random_layered(6, 4, seed=3)
    psi, res = O.run_measured(O.basis(6), 6, lay.ops(), [], 0)
    assert np.max(np.abs(psi - O.run_ops(O.basis(6), 6, lay.ops()))) == 0.0

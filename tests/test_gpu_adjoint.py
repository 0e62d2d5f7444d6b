"""Adjoint gradients on the GPU (qse_adjoint_expect) against central finite
differences (h = 1e-5) and against parameter-shift gradients evaluated on the
CPU oracle. Bar: max |dg| <= 1e-6 against finite differences, <= 1e-10 against
the oracle's parameter shift in complex128."""
import math

import numpy as np
import pytest
import torch  # imported at collection time (the first import in a fresh environment can take minutes)

import oracle_lib as O
from quantumsimevo_amd import autograd as AG
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def oracle_energy(n, ops, params, masks, weights):
    psi = O.run_ops(O.basis(n), n, ops, params)
    return O.expect_zstrings(psi, n, masks, weights)


def oracle_shift(n, ops, params, masks, weights):
    g = np.zeros(len(params))
    for k in np.nonzero(ops["param"] >= 0)[0]:
        for s in (+1.0, -1.0):
            sh = ops.copy()
            sh["theta"][k] += s * math.pi / 2
            g[int(ops[k]["param"])] += s * 0.5 * float(ops[k]["scale"]) * oracle_energy(n, sh, params, masks, weights)
    return g


def mixed_circuit(n, seed):
    """Every parameterisable kind (RX, RY, RZ, PHASE, CPHASE) mixed with fixed
    H, CNOT, SWAP, MCPHASE, generic 2x2 and a CNOT-RZ-CNOT cost gate."""
    rng = np.random.default_rng(seed)
    g = G.GateList(n)
    for q in range(n):
        g.hadamard(q)
    p = 0
    for layer in range(3):
        for q in range(n):
            k = (q + layer) % 5
            if k == 0: g.rotate_x(0.1, q, param=p, scale=1.0)
            elif k == 1: g.rotate_y(-0.2, q, param=p, scale=2.0)
            elif k == 2: g.rotate_z(0.3, q, param=p, scale=1.0)
            elif k == 3: g.phase(0.0, q, param=p, scale=1.0)
            else: g.cphase(0.4, q, (q + 1) % n, param=p, scale=1.5)
            p += 1
        for q in range(0, n - 1, 2):
            g.cnot(q, q + 1)
        g.cnot(0, n - 1); g.rotate_z(0.0, n - 1, param=p, scale=3.0); g.cnot(0, n - 1); p += 1
        g.swap(1, n - 2)
        g.mcphase(0.7, [0, 2], 3)
        g.gate(np.array([[math.cos(0.3), -math.sin(0.3)], [math.sin(0.3), math.cos(0.3)]]), layer)
    return g.ops(), p


@pytest.fixture(scope="module")
def qaoa_small():
    n, p = 12, 2
    edges = G.random_regular_graph(n, 3, seed=4)
    ops = G.qaoa_maxcut(n, edges, p).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(1).random((3, 2 * p)) * math.pi
    return n, ops, masks, w, params


def test_qaoa_adjoint_vs_oracle_shift_and_fd(ctx, qaoa_small):
    n, ops, masks, w, params = qaoa_small
    E, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    for b in range(len(params)):
        assert abs(E[b] - oracle_energy(n, ops, params[b], masks, w)) < 1e-11
        gs = oracle_shift(n, ops, params[b], masks, w)
        assert np.max(np.abs(Gr[b] - gs)) < 1e-10
    # central finite differences on the GPU, h = 1e-5
    h = 1e-5
    for j in range(params.shape[1]):
        Pp, Pm = params.copy(), params.copy()
        Pp[:, j] += h
        Pm[:, j] -= h
        fd = (ctx.batch_expect(n, capi.C128, ops, Pp, masks, w) - ctx.batch_expect(n, capi.C128, ops, Pm, masks, w)) / (2 * h)
        assert np.max(np.abs(Gr[:, j] - fd)) < 1e-6


def test_mixed_kinds_vs_oracle_shift(ctx):
    n = 11
    ops, npar = mixed_circuit(n, 3)
    zm, zw = G.zsum_observable(n)
    zm = np.concatenate([zm, [0b101, 0b110000]]).astype(np.uint64)
    zw = np.concatenate([zw, [0.5, -0.25]])
    params = np.random.default_rng(2).random(npar) * 2 * math.pi
    E, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, zm, zw)
    assert abs(E[0] - oracle_energy(n, ops, params, zm, zw)) < 1e-11
    assert np.max(np.abs(Gr[0] - oracle_shift(n, ops, params, zm, zw))) < 1e-10
    # checkpointed reverse sweep gives the same gradient
    E2, G2 = ctx.adjoint_expect(n, capi.C128, ops, params, zm, zw, ckpt_every=2)
    assert abs(E2[0] - E[0]) < 1e-12 and np.max(np.abs(G2 - Gr)) < 1e-11


def test_complex64_gradients(ctx, qaoa_small):
    n, ops, masks, w, params = qaoa_small
    E, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    E32, G32 = ctx.adjoint_expect(n, capi.C64, ops, params, masks, w, ckpt_every=4)
    assert np.max(np.abs(E32 - E)) < 1e-4
    assert np.max(np.abs(G32 - Gr)) < 1e-3


def test_torch_autograd_function(ctx, qaoa_small):
    n, ops, masks, w, params = qaoa_small
    eng = AG.AdjointEngine(ctx, n, ops, masks, w)
    x = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    E = AG.expectation(x, eng)
    (E * torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)).sum().backward()
    _, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    want = Gr * np.array([1.0, -2.0, 0.5])[:, None]
    assert np.max(np.abs(x.grad.numpy() - want)) < 1e-12
    # parameter shift on the GPU agrees with the adjoint
    gs = eng.parameter_shift(params[:1])
    assert np.max(np.abs(gs - Gr[:1])) < 1e-10


def test_wide_zstring_reductions_qaoa16(ctx):
    """From 13 qubits the generator terms of a block are reduced 64 Z-strings
    per pass (Walsh-Hadamard reduction); checked against the oracle's
    parameter shift at 16 qubits."""
    n, p = 16, 2
    edges = G.random_regular_graph(n, 3, seed=9)
    ops = G.qaoa_maxcut(n, edges, p).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(5).random((2, 2 * p)) * math.pi
    E, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    for b in range(len(params)):
        assert abs(E[b] - oracle_energy(n, ops, params[b], masks, w)) < 1e-11
        assert np.max(np.abs(Gr[b] - oracle_shift(n, ops, params[b], masks, w))) < 1e-10


def test_more_than_64_terms_in_one_block(ctx):
    """A cost layer on the complete graph K_14 puts 91 Z_a Z_b generators in
    one block: two reduction passes (64 + 27 masks)."""
    n = 14
    edges = [(a, b) for a in range(n) for b in range(a + 1, n)]
    ops = G.qaoa_maxcut(n, edges, 1).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.array([0.37, 1.21])
    E, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    assert abs(E[0] - oracle_energy(n, ops, params, masks, w)) < 1e-11 * max(1.0, abs(E[0]))
    gs = oracle_shift(n, ops, params, masks, w)
    # gradients reach ~1e2 here: the bar is relative to their size
    assert np.max(np.abs(Gr[0] - gs)) < 1e-10 * max(1.0, float(np.max(np.abs(gs))))
    E32, G32 = ctx.adjoint_expect(n, capi.C64, ops, params, masks, w)
    assert abs(E32[0] - E[0]) < 1e-4 and np.max(np.abs(G32 - Gr)) < 1e-3


def test_structure_keyed_kernels():
    """The adjoint's internal runs through compiled kernels keyed by the plan
    structure (coefficients read from the program buffer): compiled on first
    use (JIT_ALWAYS) or on worker threads from the second sighting (JIT_AUTO);
    both agree with the generic kernel and with the oracle."""
    n, p = 14, 2
    edges = G.random_regular_graph(n, 3, seed=11)
    ops = G.qaoa_maxcut(n, edges, p).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(8).random((4, 2 * p)) * math.pi
    c = capi.Context(0)
    try:
        c.set_jit(capi.JIT_OFF)
        E0, G0 = c.adjoint_expect(n, capi.C128, ops, params, masks, w)
        assert c.jit_stats()["launches"] == 0
        c.set_jit(capi.JIT_ALWAYS)
        E1, G1 = c.adjoint_expect(n, capi.C128, ops, params, masks, w)
        st = c.jit_stats()
        assert st["compiled"] > 0 and st["launches"] > 0
        tol = 1e-10 * max(1.0, float(np.max(np.abs(G0))))
        assert np.max(np.abs(E1 - E0)) < 1e-11 and np.max(np.abs(G1 - G0)) < tol
        assert np.max(np.abs(G1[0] - oracle_shift(n, ops, params[0], masks, w))) < tol
        E32, G32 = c.adjoint_expect(n, capi.C64, ops, params, masks, w)
        assert np.max(np.abs(E32 - E0)) < 1e-4 and np.max(np.abs(G32 - G0)) < 1e-3
        Eb = c.batch_expect(n, capi.C128, ops, params, masks, w)
        assert np.max(np.abs(Eb - E0)) < 1e-11
    finally:
        c.close()
    c = capi.Context(0)
    try:
        c.set_jit(capi.JIT_AUTO)
        E2, G2 = c.adjoint_expect(n, capi.C128, ops, params, masks, w)  # starts the compilations
        c.jit_wait()
        launches = c.jit_stats()["launches"]
        E3, G3 = c.adjoint_expect(n, capi.C128, ops, params, masks, w)
        assert c.jit_stats()["launches"] > launches
        for E, Gr in ((E2, G2), (E3, G3)):
            assert np.max(np.abs(E - E0)) < 1e-11 and np.max(np.abs(Gr - G0)) < tol
    finally:
        c.close()


@pytest.mark.parametrize("n", [14, 16])
def test_pair_terms_vs_rotated_frames_and_oracle(ctx, n):
    """RX/RY generator terms summed over amplitude pairs (read-only passes over
    psi and lambda, terms on qubits below and above the 12-qubit tile) against
    the rotated-frame Z-strings and the oracle's parameter shift; complex64 and
    a checkpointed sweep in pair mode."""
    ops, npar = mixed_circuit(n, 7)
    zm, zw = G.zsum_observable(n)
    params = np.random.default_rng(n).random(npar) * 2 * math.pi
    res = {}
    try:
        for mode in (capi.ADJ_TERMS_ROTATE, capi.ADJ_TERMS_PAIRS):
            ctx.set_adjoint_terms(mode)
            res[mode] = ctx.adjoint_expect(n, capi.C128, ops, params, zm, zw)
        ctx.set_adjoint_terms(capi.ADJ_TERMS_PAIRS)
        Ec, Gc = ctx.adjoint_expect(n, capi.C128, ops, params, zm, zw, ckpt_every=2)
        E32, G32 = ctx.adjoint_expect(n, capi.C64, ops, params, zm, zw)
    finally:
        ctx.set_adjoint_terms(capi.ADJ_TERMS_AUTO)
    (Er, Gr), (Ep, Gp) = res[capi.ADJ_TERMS_ROTATE], res[capi.ADJ_TERMS_PAIRS]
    gs = oracle_shift(n, ops, params, zm, zw)
    assert abs(Ep[0] - oracle_energy(n, ops, params, zm, zw)) < 1e-11
    assert np.max(np.abs(Gp[0] - gs)) < 1e-10
    assert np.max(np.abs(Gp - Gr)) < 1e-11
    assert abs(Ec[0] - Ep[0]) < 1e-12 and np.max(np.abs(Gc - Gp)) < 1e-11
    assert np.max(np.abs(G32 - Gp)) < 1e-3


def test_pair_terms_two_high_tiles_qaoa22(ctx):
    """QAOA on 22 qubits: the mixer's X terms need the low tile and two tiles of
    high qubits (9 + 1). Pair mode against the rotated-frame schedule and central
    finite differences on the GPU; the automatic choice here is the joint
    schedule (run-time compiled), whose result it must reproduce (to rounding:
    the forward run is interpreted or specialised depending on whether its
    compilation on a worker thread has finished)."""
    n, p = 22, 1
    edges = G.random_regular_graph(n, 3, seed=3)
    ops = G.qaoa_maxcut(n, edges, p).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.array([[0.41, 1.07]])
    try:
        ctx.set_adjoint_terms(capi.ADJ_TERMS_ROTATE)
        Er, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
        ctx.set_adjoint_terms(capi.ADJ_TERMS_PAIRS)
        Ep, Gp = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
        ctx.set_adjoint_terms(capi.ADJ_TERMS_JOINT)
        Ej, Gj = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
        ctx.set_adjoint_terms(capi.ADJ_TERMS_AUTO)
        Ea, Ga = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    finally:
        ctx.set_adjoint_terms(capi.ADJ_TERMS_AUTO)
    assert np.max(np.abs(Gp - Gr)) < 1e-10 and abs(Ep[0] - Er[0]) < 1e-11
    assert np.max(np.abs(Gj - Gp)) < 1e-10 and abs(Ej[0] - Ep[0]) < 1e-11
    assert np.max(np.abs(Ga - Gj)) < 1e-12 and abs(Ea[0] - Ej[0]) < 1e-12
    h = 1e-5
    for j in range(2):
        Pp, Pm = params.copy(), params.copy()
        Pp[0, j] += h
        Pm[0, j] -= h
        fd = (ctx.batch_expect(n, capi.C128, ops, Pp, masks, w)[0] - ctx.batch_expect(n, capi.C128, ops, Pm, masks, w)[0]) / (2 * h)
        assert abs(Gp[0, j] - fd) < 1e-6


@pytest.mark.parametrize("n, complete", [(14, True), (22, False)])
def test_zstrings_carried_by_pair_passes(ctx, n, complete):
    """QAOA p = 2: the cost layer's Z_a Z_b terms share their evaluation point
    with the next mixer's X terms and ride along in its pair passes (12 per
    pass); on K_14 the 91 masks exceed what two pair passes take and the rest
    go through the Walsh-Hadamard reduction. Pair mode against the rotated
    frames, and K_14 against the oracle's parameter shift."""
    edges = ([(a, b) for a in range(n) for b in range(a + 1, n)] if complete
             else G.random_regular_graph(n, 3, seed=5))
    ops = G.qaoa_maxcut(n, edges, 2).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.array([[0.21, 0.93, 0.47, 1.31]])
    try:
        ctx.set_adjoint_terms(capi.ADJ_TERMS_ROTATE)
        Er, Gr = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
        ctx.set_adjoint_terms(capi.ADJ_TERMS_PAIRS)
        Ep, Gp = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    finally:
        ctx.set_adjoint_terms(capi.ADJ_TERMS_AUTO)
    scale = max(1.0, float(np.max(np.abs(Gr))))
    assert abs(Ep[0] - Er[0]) < 1e-11 * max(1.0, abs(Er[0]))
    assert np.max(np.abs(Gp - Gr)) < 1e-11 * scale
    if complete:
        gs = oracle_shift(n, ops, params[0], masks, w)
        assert np.max(np.abs(Gp[0] - gs)) < 1e-10 * scale

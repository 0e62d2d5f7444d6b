"""The adjoint's joint schedule (QSE_ADJ_TERMS_JOINT: the reverse sweep as
run-time compiled joint passes over psi and lambda, every generator term
summed in the tile at its point) against the oracle's parameter shift
(complex128, 1e-10), against the pair and rotated schedules of the same call
(1e-10), and in complex64 against complex128 (1e-4)."""
import math

import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G
from test_gpu_adjoint import mixed_circuit, oracle_energy, oracle_shift

pytestmark = pytest.mark.gpu


def run(ctx, mode, n, prec, ops, params, masks, w):
    ctx.set_adjoint_terms(mode)
    try:
        return ctx.adjoint_expect(n, prec, ops, params, masks, w)
    finally:
        ctx.set_adjoint_terms(capi.ADJ_TERMS_AUTO)


@pytest.mark.parametrize("n", [14, 16])
def test_joint_qaoa_vs_oracle_and_other_schedules(ctx, n):
    print(f"joint schedule QAOA n={n}", flush=True)
    edges = G.random_regular_graph(n, 3, seed=n)
    ops = G.qaoa_maxcut(n, edges, 2).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(n).random((3, 4)) * math.pi
    E, Gj = run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C128, ops, params, masks, w)
    for b in range(2):
        assert abs(E[b] - oracle_energy(n, ops, params[b], masks, w)) < 1e-11
        assert np.max(np.abs(Gj[b] - oracle_shift(n, ops, params[b], masks, w))) < 1e-10
    for mode in (capi.ADJ_TERMS_PAIRS, capi.ADJ_TERMS_ROTATE):
        _, Go = run(ctx, mode, n, capi.C128, ops, params, masks, w)
        assert np.max(np.abs(Gj - Go)) < 1e-10
    _, G64 = run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C64, ops, params, masks, w)
    assert np.max(np.abs(G64 - Gj)) < 1e-4


def test_joint_every_parameterised_kind(ctx):
    n = 14
    print(f"joint schedule mixed circuit n={n}", flush=True)
    ops, npar = mixed_circuit(n, 3)
    masks = [1 << q for q in range(n)] + [(1 << 1) | (1 << 12)]
    w = [1.0] * n + [-0.7]
    params = np.random.default_rng(4).random((2, npar)) * 2 * math.pi
    _, Gj = run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C128, ops, params, masks, w)
    assert np.max(np.abs(Gj[0] - oracle_shift(n, ops, params[0], masks, w))) < 1e-10
    _, Gp = run(ctx, capi.ADJ_TERMS_PAIRS, n, capi.C128, ops, params, masks, w)
    assert np.max(np.abs(Gj - Gp)) < 1e-10


def test_joint_qaoa22_vs_pairs(ctx):
    n = 22
    print(f"joint schedule QAOA n={n}", flush=True)
    edges = G.random_regular_graph(n, 3, seed=22)
    ops = G.qaoa_maxcut(n, edges, 3).ops()
    masks, w = G.maxcut_observable(edges)
    params = np.random.default_rng(22).random((2, 6)) * math.pi
    Ej, Gj = run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C128, ops, params, masks, w)
    Ep, Gp = run(ctx, capi.ADJ_TERMS_PAIRS, n, capi.C128, ops, params, masks, w)
    assert np.max(np.abs(Ej - Ep)) < 1e-10
    assert np.max(np.abs(Gj - Gp)) < 1e-10


def test_joint_needs_jit(ctx):
    n = 14
    edges = G.random_regular_graph(n, 3, seed=1)
    ops = G.qaoa_maxcut(n, edges, 1).ops()
    masks, w = G.maxcut_observable(edges)
    ctx.set_jit(capi.JIT_OFF)
    try:
        with pytest.raises(capi.QSEError):
            run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C128, ops, np.array([0.3, 0.9]), masks, w)
    finally:
        ctx.set_jit(capi.JIT_AUTO)


@pytest.mark.parametrize("n", [9, 14])
def test_joint_observable_groups(ctx, n):
    """The joint schedule's start, lambda = O psi and the energy in one pass
    (k_obs_energy: terms grouped by their three highest mask bits, expanded by
    an 8-point Walsh-Hadamard transform): a constant term (mask 0), terms on the
    top qubits alone, across groups and below them, against the oracle and the
    rotated-frame schedule; an empty observable gives zero energy and
    gradient."""
    print(f"joint schedule observable groups n={n}", flush=True)
    ops, npar = mixed_circuit(n, n + 3)
    top = n - 1
    masks = [0, 1 << top, (1 << top) | (1 << (top - 1)) | 1, 1 << (top - 2), (1 << 2) | (1 << (top - 1)), 1 << 3]
    w = [0.7, -1.2, 0.4, 0.9, -0.35, 0.5]
    params = np.random.default_rng(n + 1).random((2, npar)) * 2 * math.pi
    Ej, Gj = run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C128, ops, params, masks, w)
    for b in range(2):
        assert abs(Ej[b] - oracle_energy(n, ops, params[b], masks, w)) < 1e-11
    assert np.max(np.abs(Gj[0] - oracle_shift(n, ops, params[0], masks, w))) < 1e-10
    Er, Gr = run(ctx, capi.ADJ_TERMS_ROTATE, n, capi.C128, ops, params, masks, w)
    assert np.max(np.abs(Ej - Er)) < 1e-11 and np.max(np.abs(Gj - Gr)) < 1e-10
    E0, G0 = run(ctx, capi.ADJ_TERMS_JOINT, n, capi.C128, ops, params, [], [])
    assert np.all(E0 == 0.0) and np.all(G0 == 0.0)

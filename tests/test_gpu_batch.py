"""Batched small circuits in persistent launches (qse_batch_run_small) on the
GPU: energies and adjoint gradients of circuits of different sizes against the
oracle (energies 1e-12, gradients vs parameter shift 1e-10 in complex128;
1e-4 in complex64) and against the host-driven adjoint (qse_adjoint_expect)."""
import math

import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G
from test_batch_plan import many_probe_circuit, mixed_param_circuit, oracle_energy, oracle_shift

pytestmark = pytest.mark.gpu


def mixed_batch():
    circs = []
    for i, n in enumerate((10, 12, 11, 13)):
        ops, par = mixed_param_circuit(n, i)
        circs.append((n, ops, par))
    circs += G.uneven_ansatz_batch(4, seed=5, n_range=(10, 13), depth_range=(2, 4))
    masks = [1 << q for q in range(13)] + [0b101, 0b1100000000]
    weights = [1.0] * 13 + [0.5, -0.25]
    return circs, masks, weights


@pytest.mark.parametrize("prec,tol_e,tol_g", [(capi.C128, 1e-12, 1e-10), (capi.C64, 1e-4, 1e-4)])
def test_batch_vs_oracle(ctx, prec, tol_e, tol_g):
    circs, masks, weights = mixed_batch()
    E, Gr = ctx.batch_run_small(prec, circs, masks, weights)
    for (n, ops, par), e, g in zip(circs, E, Gr):
        assert abs(e - oracle_energy(n, ops, par, masks, weights)) < tol_e
        assert np.max(np.abs(g - oracle_shift(n, ops, par, masks, weights))) < tol_g
    E2 = ctx.batch_run_small(prec, circs, masks, weights, grads=False)
    assert np.max(np.abs(E2 - E)) < tol_e


def test_uneven_ansatz_batch_vs_host_adjoint(ctx):
    # 48 ansatz circuits of 10..18 qubits, depth 4..8, observable sum of Z_i over each circuit's qubits
    circs = G.uneven_ansatz_batch(48, seed=11, n_range=(10, 18), depth_range=(4, 8))
    masks = [1 << q for q in range(18)]
    weights = [1.0] * 18
    E, Gr = ctx.batch_run_small(capi.C128, circs, masks, weights)
    for i in (0, 7, 23, 47):
        n, ops, par = circs[i]
        zm, zw = G.zsum_observable(n)
        Eh, Gh = ctx.adjoint_expect(n, capi.C128, ops, par, zm, zw)
        assert abs(E[i] - Eh[0]) < 1e-11
        assert np.max(np.abs(Gr[i] - Gh[0])) < 1e-10
    E64, G64 = ctx.batch_run_small(capi.C64, circs, masks, weights)
    assert np.max(np.abs(E64 - E)) < 1e-4
    assert max(np.max(np.abs(a - b)) for a, b in zip(G64, Gr)) < 1e-3


def test_batch_rejects_bad_input(ctx):
    n = 10
    ops = G.ansatz(n, 1).ops()
    with pytest.raises(capi.QSEError):
        ctx.batch_run_small(capi.C128, [(7, G.ansatz(7, 1).ops(), np.zeros(14))], [1], [1.0])  # fewer than 8 qubits
    with pytest.raises(capi.QSEError):
        ctx.batch_run_small(capi.C128, [(n, ops, np.zeros(3))], [1], [1.0])  # parameter index out of range


def test_team_circuits_vs_host_adjoint(ctx):
    # circuits of 16 qubits and more run on teams of workgroups; mixed with small ones
    circs = G.uneven_ansatz_batch(6, seed=21, n_range=(18, 20), depth_range=(3, 5))
    circs += G.uneven_ansatz_batch(10, seed=22, n_range=(10, 14), depth_range=(2, 4))
    masks = [1 << q for q in range(20)]
    weights = [1.0] * 20
    E, Gr = ctx.batch_run_small(capi.C128, circs, masks, weights)
    for i in (0, 3, 5, 6, 15):
        n, ops, par = circs[i]
        zm, zw = G.zsum_observable(n)
        Eh, Gh = ctx.adjoint_expect(n, capi.C128, ops, par, zm, zw)
        assert abs(E[i] - Eh[0]) < 1e-11
        assert np.max(np.abs(Gr[i] - Gh[0])) < 1e-10
    # deterministic: the same batch twice gives identical values
    E2, G2 = ctx.batch_run_small(capi.C128, circs, masks, weights)
    assert np.array_equal(E, E2) and all(np.array_equal(a, b) for a, b in zip(Gr, G2))


@pytest.mark.parametrize("prec,tol_e,tol_g", [(capi.C128, 1e-11, 1e-10), (capi.C64, 1e-4, 1e-3)])
def test_autotuned_shape_keeps_results(ctx, prec, tol_e, tol_g):
    # the measured tile shape is one of the candidates and is kept by the context;
    # results under it agree with the default shape's (and so with the oracle's)
    circs = G.uneven_ansatz_batch(32, seed=5, n_range=(10, 16), depth_range=(3, 6))
    masks = [1 << q for q in range(16)]
    weights = [1.0] * 16
    E0, G0 = ctx.batch_run_small(prec, circs, masks, weights)
    tune = ctx.batch_autotune(prec, circs[:16], masks, weights)
    print(f"batch autotune prec={prec}: {tune}", flush=True)
    assert 10 <= tune["tile_qubits"] <= 12
    assert 2 <= tune["low_qubits"] <= min(5, tune["tile_qubits"] - 4)
    assert tune["team_workgroups"] in (4, 8, 16, 32)
    assert tune["ms"] > 0
    E1, G1 = ctx.batch_run_small(prec, circs, masks, weights)
    assert np.max(np.abs(E1 - E0)) < tol_e
    assert max(np.max(np.abs(a - b)) for a, b in zip(G1, G0)) < tol_g
    n, ops, par = circs[3]
    zm, zw = G.zsum_observable(n)
    assert abs(E1[3] - oracle_energy(n, ops, par, zm, zw)) < (1e-11 if prec == capi.C128 else 1e-4)


@pytest.mark.parametrize("prec,tol_g", [(capi.C128, 1e-10), (capi.C64, 1e-3)])
def test_reverse_sweep_in_probe_chunks(ctx, prec, tol_g):
    # 720 generator probes on one 12-qubit tile: the reverse sweep runs as several joint steps,
    # beside an ordinary circuit of the same batch
    n = 12
    ops, par = many_probe_circuit(n)
    other_ops, other_par = mixed_param_circuit(10, 7)
    masks = [1 << q for q in range(n)]
    weights = [1.0 + 0.1 * q for q in range(n)]
    E, Gr = ctx.batch_run_small(prec, [(n, ops, par), (10, other_ops, other_par)], masks, weights)
    assert np.max(np.abs(Gr[0] - oracle_shift(n, ops, par, masks, weights))) < tol_g
    assert np.max(np.abs(Gr[1] - oracle_shift(10, other_ops, other_par, masks, weights))) < tol_g
    assert abs(E[0] - oracle_energy(n, ops, par, masks, weights)) < (1e-12 if prec == capi.C128 else 1e-4)

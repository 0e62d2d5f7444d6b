"""CPU checks of the batched small-circuit engine's work lists
(quantumsimevo_amd/csrc/qse_batch_plan.cpp, shared schedule of
qse_adjoint_plan.cpp): the steps and programs it builds, executed in FP64 by the
host emulation of the kernel (tests/fused_emu.cpp: emu_batch), must reproduce
the oracle's energies (1e-12) and the oracle's parameter-shift gradients
(1e-10) for circuits of different sizes in one batch."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import circuits as G
from test_fused_plan import emu


def oracle_energy(n, ops, params, masks, weights):
    keep = [(int(m), w) for m, w in zip(masks, weights) if (int(m) >> n) == 0]
    psi = O.run_ops(O.basis(n), n, ops, params)
    if not keep:
        return 0.0
    return O.expect_zstrings(psi, n, np.array([k[0] for k in keep], dtype=np.uint64),
                             np.array([k[1] for k in keep]))


def oracle_shift(n, ops, params, masks, weights):
    g = np.zeros(len(params))
    for k in np.nonzero(ops["param"] >= 0)[0]:
        for s in (+1.0, -1.0):
            sh = ops.copy()
            sh["theta"][k] += s * math.pi / 2
            g[int(ops[k]["param"])] += s * 0.5 * float(ops[k]["scale"]) * oracle_energy(n, sh, params, masks, weights)
    return g


def run_emu(circuits, masks, weights, grads=True, kmax=12, L=3):
    nq, op_off, ops, par_off, params = G.pack_batch(circuits)
    lib = emu()
    lib.emu_batch.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    E = np.empty(len(circuits))
    Gr = np.empty(max(len(params), 1))
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    rc = lib.emu_batch(len(circuits), nq.ctypes.data, op_off.ctypes.data, ops.ctypes.data, par_off.ctypes.data,
                       params.ctypes.data if len(params) else None, m.ctypes.data, w.ctypes.data, len(m),
                       int(grads), kmax, L, E.ctypes.data, Gr.ctypes.data)
    assert rc == 0, rc
    return E, [Gr[par_off[i]:par_off[i + 1]] for i in range(len(circuits))]


def mixed_param_circuit(n, seed):
    """Every parameterised kind (RX, RY, RZ, PHASE, CPHASE, CNOT-RZ-CNOT) with
    fixed H, CNOT, SWAP, MCPHASE and a generic 2x2 between them."""
    rng = np.random.default_rng(seed)
    g = G.GateList(n)
    p = 0
    for q in range(n):
        g.hadamard(q)
    for layer in range(2):
        for q in range(n):
            k = (q + layer) % 5
            if k == 0: g.rotate_x(0.1, q, param=p, scale=1.0)
            elif k == 1: g.rotate_y(-0.2, q, param=p, scale=2.0)
            elif k == 2: g.rotate_z(0.3, q, param=p, scale=1.0)
            elif k == 3: g.phase(0.0, q, param=p, scale=1.0)
            else: g.cphase(0.4, q, (q + 1) % n, param=p, scale=1.5)
            p += 1
        for q in range(layer % 2, n - 1, 2):
            g.cnot(q, q + 1)
        g.cnot(0, n - 1); g.rotate_z(0.0, n - 1, param=p, scale=3.0); g.cnot(0, n - 1); p += 1
        g.swap(1, n - 2)
        g.mcphase(0.7, [0, 2], 3)
        g.gate(np.array([[math.cos(0.3), -math.sin(0.3)], [math.sin(0.3), math.cos(0.3)]]), layer)
    return g.ops(), rng.random(p) * 2 * math.pi


@pytest.mark.parametrize("kmax,L", [(12, 3), (9, 2)])
def test_uneven_batch_energies_and_gradients(kmax, L):
    circs = []
    for i, n in enumerate((10, 12, 11)):
        ops, par = mixed_param_circuit(n, i)
        circs.append((n, ops, par))
    circs += G.uneven_ansatz_batch(3, seed=5, n_range=(10, 12), depth_range=(2, 3))
    # Z on every qubit up to 12 plus two Z-strings: terms beyond a circuit's qubits are skipped
    masks = [1 << q for q in range(12)] + [0b101, 0b1100000000]
    weights = [1.0] * 12 + [0.5, -0.25]
    E, Gr = run_emu(circs, masks, weights, kmax=kmax, L=L)
    for (n, ops, par), e, g in zip(circs, E, Gr):
        assert abs(e - oracle_energy(n, ops, par, masks, weights)) < 1e-12
        assert np.max(np.abs(g - oracle_shift(n, ops, par, masks, weights))) < 1e-10


def test_energies_only_and_qaoa_cost_gate():
    n = 10
    edges = G.random_regular_graph(n, 3, seed=3)
    ops = G.qaoa_maxcut(n, edges, 2).ops()
    masks, w = G.maxcut_observable(edges)
    par = np.random.default_rng(1).random(4) * math.pi
    E = run_emu([(n, ops, par)], masks, w, grads=False)[0]
    assert abs(E[0] - oracle_energy(n, ops, par, masks, w)) < 1e-12
    E, Gr = run_emu([(n, ops, par), (n, ops, par[::-1].copy())], masks, w)
    assert np.max(np.abs(Gr[0] - oracle_shift(n, ops, par, masks, w))) < 1e-10
    assert np.max(np.abs(Gr[1] - oracle_shift(n, ops, par[::-1].copy(), masks, w))) < 1e-10


def many_probe_circuit(n=12, layers=30):
    """RZ and RY on every qubit, layer after layer, all on one tile: the reverse
    sweep holds 2 n layers probes, far more than one joint pass's 128
    accumulators, so it is planned in chunks."""
    g = G.GateList(n)
    p = 0
    for _ in range(layers):
        for q in range(n):
            g.rotate_z(0.0, q, param=p, scale=1.0)
            g.rotate_y(0.0, q, param=p + 1, scale=1.0)
            p += 2
    return g.ops(), np.random.default_rng(0).random(p)


def test_reverse_sweep_in_probe_chunks():
    n = 12
    ops, par = many_probe_circuit(n)
    masks = [1 << q for q in range(n)]
    weights = [1.0 + 0.1 * q for q in range(n)]
    E, Gr = run_emu([(n, ops, par)], masks, weights, kmax=12, L=3)
    assert abs(E[0] - oracle_energy(n, ops, par, masks, weights)) < 1e-12
    assert np.max(np.abs(Gr[0] - oracle_shift(n, ops, par, masks, weights))) < 1e-10
    lib = emu()
    nq, op_off, opsb, par_off, params = G.pack_batch([(n, ops, par)])
    out = np.zeros(5, dtype=np.int64)
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    lib.emu_batch_stats.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.c_int, C.c_int, C.c_void_p]
    assert lib.emu_batch_stats(1, nq.ctypes.data, op_off.ctypes.data, opsb.ctypes.data, par_off.ctypes.data,
                               params.ctypes.data, m.ctypes.data, len(m), 12, 3, out.ctypes.data) == 0
    # 720 probes over 128-accumulator passes: at least 6 joint steps besides the forward program
    assert out[3] >= 7

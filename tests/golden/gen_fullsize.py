"""Writes tests/golden/fullsize_vectors.npz: oracle values for the benchmark
configurations at their full sizes, where the oracle needs minutes per value
(too slow to run inside a GPU test).

Run in this container after `make -C oracle`:  python tests/golden/gen_fullsize.py
Everything is computed by the CPU oracle (oracle/liboracle.so, pinned in
tests/test_oracle.py); nothing here touches the reference tree.

  c3_*  QAOA Max-Cut, 24 qubits, seeded random 3-regular graph (36 edges), p = 8,
        16 parameters per set drawn as in bench.py (default_rng(7), [0, pi)):
        energy <C> and the full 16-component gradient of sets 0 and 1 from the
        oracle's adjoint sweep; for set 0 also the parameter-shift gradient of
        gamma_8 and beta_8 (every one of their 60 occurrences shifted by
        +-pi/2, each evaluation run from a checkpoint of the state before the
        occurrence)
  c4_*  16 circuits of the 1024-circuit uneven ansatz batch of bench.py
        (uneven_ansatz_batch(1024, seed=2024), 10..22 qubits, depth 4..12),
        among them the deepest 22-qubit ones: energy of sum_{i<n} Z_i and the
        adjoint gradient (concatenated, offsets in c4_goff)
"""
from __future__ import annotations

import math
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

from quantumsimevo_amd import circuits as C  # noqa: E402
import oracle_lib as O  # noqa: E402

C3_N, C3_P, C3_GRAPH_SEED, C3_PARAM_SEED = 24, 8, 2024, 7
C4_COUNT, C4_SEED = 1024, 2024


def c3_problem():
    edges = C.random_regular_graph(C3_N, 3, seed=C3_GRAPH_SEED)
    ops = C.qaoa_maxcut(C3_N, edges, C3_P).ops()
    masks, w = C.maxcut_observable(edges)
    params = np.random.default_rng(C3_PARAM_SEED).random((64, 2 * C3_P)) * np.pi
    return edges, ops, masks, w, params


def c4_batch():
    return C.uneven_ansatz_batch(C4_COUNT, seed=C4_SEED, n_range=(10, 22), depth_range=(4, 12))


def c4_pick(circs):
    """The 6 deepest 22-qubit circuits, the deepest 21-qubit one and 9 seeded
    others spread over the batch."""
    idx = sorted(range(len(circs)), key=lambda i: (-circs[i][0], -len(circs[i][1]), i))
    pick = [i for i in idx if circs[i][0] == 22][:6]
    pick += [i for i in idx if circs[i][0] == 21][:1]
    rest = [i for i in range(len(circs)) if i not in pick]
    pick += [int(i) for i in np.random.default_rng(99).choice(rest, size=9, replace=False)]
    return sorted(pick)


def job_c3_adjoint(b):
    _, ops, masks, w, params = c3_problem()
    t0 = time.time()
    e, g = O.adjoint_expect(C3_N, ops, params[b], masks, w)
    print(f"c3 adjoint set {b}: E = {e:.15g} ({time.time() - t0:.0f} s)", flush=True)
    return e, g


def job_c3_shift(pidx):
    """Parameter-shift derivative for parameter pidx of set 0 over its
    occurrences in the last layer (all of them for gamma_8 and beta_8)."""
    _, ops, masks, w, params = c3_problem()
    par = params[0]
    occ = [int(k) for k in np.nonzero(ops["param"] == pidx)[0]]
    t0 = time.time()
    start = occ[0]
    psi = O.run_ops(O.basis(C3_N), C3_N, ops[:start], par)  # checkpoint before the first occurrence
    pos = start
    d = 0.0
    for k in occ:
        psi = O.run_ops(psi, C3_N, ops[pos:k], par)  # advance the checkpoint to gate k
        pos = k
        for s in (+1.0, -1.0):
            sh = ops[k:].copy()
            sh["theta"][0] += s * math.pi / 2
            out = O.run_ops(psi, C3_N, sh, par)
            d += s * 0.5 * float(ops[k]["scale"]) * O.expect_zstrings(out, C3_N, masks, w)
    print(f"c3 shift param {pidx}: {d:.15g} over {len(occ)} occurrences ({time.time() - t0:.0f} s)", flush=True)
    return d


def job_c4(i):
    circs = c4_batch()
    n, ops, par = circs[i]
    zm, zw = C.zsum_observable(n)
    t0 = time.time()
    e, g = O.adjoint_expect(n, ops, par, zm, zw)
    print(f"c4 circuit {i} ({n}q, {len(ops)} gates): E = {e:.15g} ({time.time() - t0:.0f} s)", flush=True)
    return e, g


def main():
    edges, ops, _, _, params = c3_problem()
    occ_g8 = np.nonzero(ops["param"] == C3_P - 1)[0]
    occ_b8 = np.nonzero(ops["param"] == 2 * C3_P - 1)[0]
    assert len(occ_g8) == 36 and len(occ_b8) == 24
    circs = c4_batch()
    pick = c4_pick(circs)
    workers = int(os.environ.get("GEN_WORKERS", "6"))
    with ProcessPoolExecutor(max_workers=workers) as ex:
        f_adj = [ex.submit(job_c3_adjoint, b) for b in (0, 1)]
        f_shift = [ex.submit(job_c3_shift, p) for p in (C3_P - 1, 2 * C3_P - 1)]
        f_c4 = [ex.submit(job_c4, i) for i in pick]
        adj = [f.result() for f in f_adj]
        shift = [f.result() for f in f_shift]
        c4 = [f.result() for f in f_c4]
    goff = np.zeros(len(pick) + 1, dtype=np.int64)
    for j, (_, g) in enumerate(c4):
        goff[j + 1] = goff[j] + len(g)
    out = {
        "c3_edges": np.array(edges, dtype=np.int64),
        "c3_params": params[:2],
        "c3_energy": np.array([e for e, _ in adj]),
        "c3_grad": np.stack([g for _, g in adj]),
        "c3_shift_params": np.array([C3_P - 1, 2 * C3_P - 1], dtype=np.int64),
        "c3_shift_grad": np.array(shift),
        "c4_idx": np.array(pick, dtype=np.int64),
        "c4_n": np.array([circs[i][0] for i in pick], dtype=np.int64),
        "c4_nops": np.array([len(circs[i][1]) for i in pick], dtype=np.int64),
        "c4_param_sum": np.array([float(np.sum(circs[i][2])) for i in pick]),
        "c4_energy": np.array([e for e, _ in c4]),
        "c4_grad": np.concatenate([g for _, g in c4]),
        "c4_goff": goff,
    }
    np.savez_compressed(HERE / "fullsize_vectors.npz", **out)
    print("wrote", HERE / "fullsize_vectors.npz")


if __name__ == "__main__":
    main()

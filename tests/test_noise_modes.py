"""CPU checks of the interleaved Pauli-noise mode (no counterpart in the
reference): its insertions come from the same draw process as the
reference's deferred frame (Noise.cpp:21-44; noise_frames), so per trajectory
the parity of the inserted flips on each qubit is that frame, and each flip
sits right after the gate it was drawn for (X before Z, on the gate's target)."""
import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as C

Q = pytest.importorskip("quantumsimevo_amd._quantumsimevo")


def test_insertions_fold_to_the_deferred_frame():
    n = 12
    ops = C.random_layered(n, 10, seed=3).ops()
    seeds = np.arange(64, dtype=np.uint32)
    x, z = Q.noise_frames(n, ops["kind"], ops["target"], seeds, -1.0, 30.0)
    off, gate, code = Q.noise_insertions(n, ops["kind"], ops["target"], seeds, -1.0, 30.0)
    assert off[0] == 0 and off[-1] == len(gate) == len(code) and len(gate) > 0
    trajs = C.interleaved_trajectories(n, ops, seeds, -1.0, 30.0)
    for j in range(len(seeds)):
        g, c = gate[off[j]:off[j + 1]], code[off[j]:off[j + 1]]
        assert np.all(np.diff(g) > 0) and np.all((c >= 1) & (c <= 3))
        xm = zm = 0
        for gi, cc in zip(g, c):
            t = int(ops[int(gi)]["target"])
            xm ^= (1 << t) if cc & 1 else 0
            zm ^= (1 << t) if cc & 2 else 0
        assert xm == int(x[j]) and zm == int(z[j])
        # the trajectory's list: the circuit's gates in order, each drawn flip right after its gate
        t = trajs[j]
        assert len(t) == len(ops) + sum(bin(int(cc)).count("1") for cc in c)
        orig = t[np.isin(np.arange(len(t)), _inserted_positions(g, c), invert=True)]
        assert np.array_equal(orig, ops)


def test_insertions_fold_to_the_oracle_frames():
    # the oracle's restatement of Noise.cpp (pinned to the reference Noise class's own flip
    # patterns in test_oracle.py), trajectory j seeded with j
    import oracle_lib as O
    n = 10
    ops = C.random_layered(n, 8, seed=9).ops()
    x, z = O.noise_trajectories(n, ops, 0, 40, 0.02, 1.0)
    off, gate, code = Q.noise_insertions(n, ops["kind"], ops["target"], np.arange(40, dtype=np.uint32), 0.02, 1.0)
    for j in range(40):
        xm = zm = 0
        for gi, cc in zip(gate[off[j]:off[j + 1]], code[off[j]:off[j + 1]]):
            t = int(ops[int(gi)]["target"])
            xm ^= (1 << t) if cc & 1 else 0
            zm ^= (1 << t) if cc & 2 else 0
        assert xm == int(x[j]) and zm == int(z[j])


def _inserted_positions(g, c):
    pos, shift = [], 0
    for gi, cc in zip(g, c):
        k = bin(int(cc)).count("1")
        pos += [int(gi) + shift + 1 + i for i in range(k)]
        shift += k
    return np.array(pos, dtype=np.int64)


def test_every_draw_hits_at_rate_one():
    # rate 1: both flips on every eligible gate, X then Z on its target, right after it
    g = C.GateList(3)
    g.hadamard(0)
    g.cnot(0, 1)
    g.cphase(0.3, 1, 2)   # no draws for CPHASE (Noise.cpp:10-18)
    g.rotate_y(0.2, 2)
    ops = g.ops()
    t = C.interleaved_trajectories(3, ops, [5], 1.0, 1.0)[0]
    kinds = [int(k) for k in t["kind"]]
    targets = [int(q) for q in t["target"]]
    assert kinds == [capi.H, capi.X, capi.Z, capi.CNOT, capi.X, capi.Z, capi.CPHASE, capi.RY, capi.X, capi.Z]
    assert targets == [0, 0, 0, 1, 1, 1, 2, 2, 2, 2]
    x, z = Q.noise_frames(3, ops["kind"], ops["target"], np.array([5], dtype=np.uint32), 1.0, 1.0)
    assert int(x[0]) == 0b111 and int(z[0]) == 0b111


def test_no_draw_keeps_the_list():
    ops = C.random_layered(8, 3, seed=1).ops()
    t = C.interleaved_trajectories(8, ops, [0, 1, 2], 0.0, 1.0)
    assert all(np.array_equal(a, ops) for a in t)

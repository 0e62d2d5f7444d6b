"""Gate lists (qse_op arrays) and the synthetic workloads of the benchmark configs.

`GateList` mirrors the reference's Circuit builder methods (Circuit.cpp:9-62)
and produces the `qse_op` array the C ABI consumes. The circuit families are
the reference's own definitions:
  QFT / IQFT        QSEplayground/QFT.cpp:9-37
  QAOA Max-Cut      QSEplayground/QAOA.cpp:121-138 (H layer; per layer and edge
                    CNOT, RZ(4 gamma), CNOT; then RX(2 beta) on every qubit),
                    with float64 angles instead of the reference's float
and the seeded generators below define the synthetic inputs (random layered
circuits, random normalised states, random 3-regular graphs, uneven ansatz
batches). Generators use numpy's PCG64 (numpy.random.default_rng(seed)), so the
same seed yields the same gate list on every host.
"""
from __future__ import annotations

import math

import numpy as np

from .capi import (CNOT, CPHASE, H, IDENTITY, MATRIX, MCPHASE, MEASURE, OP_DTYPE, PHASE, RESET_QUBIT, RX, RY, RZ,
                   SWAP, X, Y, Z)

PI = 3.14159265358979323846  # QFT.cpp:7


class GateList:
    """Builder for a qse_op list with the reference Circuit's argument order."""

    def __init__(self, num_qubits: int):
        self.n = num_qubits
        self._ops: list[tuple] = []

    def _add(self, kind, target, control=-1, theta=0.0, param=-1, scale=0.0, controls=0, m=None):
        self._ops.append((kind, target, control, param, controls, theta, scale,
                          tuple(m) if m is not None else (0.0,) * 8))
        return self

    # single-qubit gates (Circuit.cpp:9-39)
    def hadamard(self, q): return self._add(H, q)
    def pauli_x(self, q): return self._add(X, q)
    def pauli_y(self, q): return self._add(Y, q)
    def pauli_z(self, q): return self._add(Z, q)
    def phase(self, theta, q, param=-1, scale=0.0): return self._add(PHASE, q, theta=theta, param=param, scale=scale)
    def rotate_x(self, theta, q, param=-1, scale=0.0): return self._add(RX, q, theta=theta, param=param, scale=scale)
    def rotate_y(self, theta, q, param=-1, scale=0.0): return self._add(RY, q, theta=theta, param=param, scale=scale)
    def rotate_z(self, theta, q, param=-1, scale=0.0): return self._add(RZ, q, theta=theta, param=param, scale=scale)

    # multi-qubit gates (Circuit.cpp:41-62)
    def cnot(self, control, target): return self._add(CNOT, target, control)

    def cphase(self, theta, control, target, param=-1, scale=0.0):
        return self._add(CPHASE, target, control, theta=theta, param=param, scale=scale)

    def mcphase(self, theta, controls, target):
        mask = 0
        for c in controls:
            mask |= 1 << int(c)
        return self._add(MCPHASE, target, -1, theta=theta, controls=mask)

    def swap(self, q0, q1): return self._add(SWAP, q0, q1)

    def gate(self, m, q):
        """Generic 2x2 (ICircuitUnitaryOperation::applyGate); m row-major complex."""
        mm = np.asarray(m, dtype=np.complex128).reshape(4)
        flat = []
        for z in mm:
            flat += [float(z.real), float(z.imag)]
        return self._add(MATRIX, q, m=flat)

    def identity(self, q=0): return self._add(IDENTITY, q)

    # mid-circuit measurement and reset (GateType::MEASURE / RESET_QUBIT, run by
    # State.run_measured; `slot` is the result slot, GateCommand::resultIdx)
    def measure(self, q, slot): return self._add(MEASURE, q, slot)
    def reset_qubit(self, q): return self._add(RESET_QUBIT, q)

    def extend(self, other: "GateList"):
        self._ops.extend(other._ops)
        return self

    def __len__(self):
        return len(self._ops)

    def ops(self) -> np.ndarray:
        a = np.zeros(len(self._ops), dtype=OP_DTYPE)
        for i, (k, t, c, p, cm, th, sc, m) in enumerate(self._ops):
            a[i] = (k, t, c, p, cm, th, sc, m)
        return a

    # gate-list files (text format of include/quantumsimevo/GateListFile.h)
    def to_text(self) -> str:
        return gates_to_text(self.n, self.ops())

    def save(self, path) -> None:
        save_gates(path, self.n, self.ops())

    @classmethod
    def from_ops(cls, n: int, ops: np.ndarray) -> "GateList":
        g = cls(n)
        for r in np.ascontiguousarray(ops, dtype=OP_DTYPE):
            g._ops.append((int(r["kind"]), int(r["target"]), int(r["control"]), int(r["param"]),
                           int(r["controls"]), float(r["theta"]), float(r["scale"]), tuple(float(x) for x in r["m"])))
        return g

    @classmethod
    def load(cls, path) -> "GateList":
        n, ops = load_gates(path)
        return cls.from_ops(n, ops)


# ---- gate-list files: one reader/writer, the C++ one (GateListFile.cpp) ----

def _native():
    from . import _quantumsimevo
    return _quantumsimevo


def gates_to_text(n: int, ops: np.ndarray) -> str:
    """Text of a gate list (qse-gates format, version 1)."""
    a = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    return _native().write_gate_list(int(n), a.tobytes())


def gates_from_text(text: str):
    """(n, ops) of a gate-list text; ValueError-like RuntimeError naming the line when malformed."""
    n, raw = _native().read_gate_list(text)
    return n, np.frombuffer(raw, dtype=OP_DTYPE).copy()


def save_gates(path, n: int, ops: np.ndarray) -> None:
    with open(path, "w", encoding="ascii", newline="\n") as f:
        f.write(gates_to_text(n, ops))


def load_gates(path):
    """(n, ops) from a gate-list file; usable directly as a batch entry (n, ops, params)."""
    with open(path, "r", encoding="ascii") as f:
        return gates_from_text(f.read())


# ---- circuit families ----

def qft(n: int, qubits=None) -> GateList:
    """QFT.cpp:9-22: H(q_i), CP(pi/2^(j-i), q_j -> q_i) for j > i, then swaps."""
    qs = list(range(n)) if qubits is None else list(qubits)
    g = GateList(n)
    m = len(qs)
    for i in range(m):
        g.hadamard(qs[i])
        for j in range(i + 1, m):
            g.cphase(PI / math.pow(2.0, j - i), qs[j], qs[i])
    for i in range(m // 2):
        g.swap(qs[i], qs[m - 1 - i])
    return g


def iqft(n: int, qubits=None) -> GateList:
    """QFT.cpp:24-37."""
    qs = list(range(n)) if qubits is None else list(qubits)
    g = GateList(n)
    m = len(qs)
    for i in range(m // 2):
        g.swap(qs[i], qs[m - 1 - i])
    for i in range(m - 1, -1, -1):
        for j in range(m - 1, i, -1):
            g.cphase(-PI / math.pow(2.0, j - i), qs[j], qs[i])
        g.hadamard(qs[i])
    return g


def random_layered(n: int, layers: int, seed: int) -> GateList:
    """Random layered circuit: per layer one of {H, RX, RY, RZ, phase} per qubit
    (uniform choice, angle uniform in [0, 2pi), always drawn), then a CNOT
    brickwork on (q, q+1) starting at q = 0 on even layers and q = 1 on odd ones."""
    rng = np.random.default_rng(seed)
    g = GateList(n)
    for layer in range(layers):
        for q in range(n):
            kind = int(rng.integers(0, 5))
            theta = float(rng.random() * 2.0 * math.pi)
            if kind == 0:
                g.hadamard(q)
            elif kind == 1:
                g.rotate_x(theta, q)
            elif kind == 2:
                g.rotate_y(theta, q)
            elif kind == 3:
                g.rotate_z(theta, q)
            else:
                g.phase(theta, q)
        for q in range(layer % 2, n - 1, 2):
            g.cnot(q, q + 1)
    return g


def random_mixed(n: int, count: int, seed: int) -> GateList:
    """`count` gates drawn uniformly from every kind the gate list supports (H,
    X, Y, Z, phase, RX, RY, RZ, CNOT, CPhase, SWAP, MCPhase with up to three
    controls, and the generic 2x2 [[cos t, -i sin t], [0.3i, 0.7]], which is not
    unitary) on uniform random qubits; control == target draws are kept (they
    exercise the degenerate cases the reference allows)."""
    rng = np.random.default_rng(seed)
    g = GateList(n)
    for _ in range(count):
        kind = int(rng.integers(0, 13))
        q = int(rng.integers(0, n))
        c = int(rng.integers(0, n))
        th = float(rng.random() * 6.3)
        if kind == 0: g.hadamard(q)
        elif kind == 1: g.pauli_x(q)
        elif kind == 2: g.pauli_y(q)
        elif kind == 3: g.pauli_z(q)
        elif kind == 4: g.phase(th, q)
        elif kind == 5: g.rotate_x(th, q)
        elif kind == 6: g.rotate_y(th, q)
        elif kind == 7: g.rotate_z(th, q)
        elif kind == 8: g.cnot(c, q)
        elif kind == 9: g.cphase(th, c, q)
        elif kind == 10: g.swap(q, c)
        elif kind == 11: g.mcphase(th, sorted(set(int(x) for x in rng.integers(0, n, 3))), q)
        else: g.gate(np.array([[math.cos(th), -1j * math.sin(th)], [0.3j, 0.7]]), q)
    return g


def random_state(n: int, seed: int) -> np.ndarray:
    """Complex Gaussian amplitudes, normalised in FP64."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    a /= np.sqrt(np.sum(np.abs(a) ** 2))
    return a


def random_regular_graph(n: int, degree: int, seed: int) -> list[tuple[int, int]]:
    """Random d-regular simple graph by the configuration model with rejection."""
    if (n * degree) % 2 or degree >= n:
        raise ValueError(f"no simple {degree}-regular graph on {n} vertices")
    rng = np.random.default_rng(seed)
    while True:
        stubs = np.repeat(np.arange(n), degree)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        edges = set()
        ok = True
        for a, b in pairs:
            a, b = int(a), int(b)
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            return sorted(edges)


def qaoa_maxcut(n: int, edges, p: int) -> GateList:
    """QAOA.cpp:121-138 with parameters: gamma_l = params[l], beta_l = params[p + l].
    RZ angle = 4 * gamma (scale 4), RX angle = 2 * beta (scale 2)."""
    g = GateList(n)
    for q in range(n):
        g.hadamard(q)
    for layer in range(p):
        for (a, b) in edges:
            g.cnot(a, b)
            g.rotate_z(0.0, b, param=layer, scale=4.0)
            g.cnot(a, b)
        for q in range(n):
            g.rotate_x(0.0, q, param=p + layer, scale=2.0)
    return g


def maxcut_observable(edges):
    """<C> = sum_edges (1 - Z_a Z_b)/2 as Z-strings: masks and weights
    (constant term as mask 0)."""
    masks = [0] + [(1 << a) | (1 << b) for (a, b) in edges]
    weights = [0.5 * len(edges)] + [-0.5] * len(edges)
    return np.array(masks, dtype=np.uint64), np.array(weights, dtype=np.float64)


def zsum_observable(n: int):
    """sum_i Z_i."""
    return np.array([1 << i for i in range(n)], dtype=np.uint64), np.ones(n)


def ansatz(n: int, depth: int) -> GateList:
    """Layers of RY(theta) RZ(phi) per qubit then a linear CNOT ladder; the
    parameters are numbered in gate order (2 per qubit per layer)."""
    g = GateList(n)
    k = 0
    for _ in range(depth):
        for q in range(n):
            g.rotate_y(0.0, q, param=k, scale=1.0)
            g.rotate_z(0.0, q, param=k + 1, scale=1.0)
            k += 2
        for q in range(n - 1):
            g.cnot(q, q + 1)
    return g


def pack_batch(circuits):
    """Flatten [(n, ops, params), ...] into the arrays of qse_batch_run_small:
    nq (int32), op_off (int64, ncirc+1), ops, par_off (int64, ncirc+1), params."""
    nq = np.array([c[0] for c in circuits], dtype=np.int32)
    op_off = np.zeros(len(circuits) + 1, dtype=np.int64)
    par_off = np.zeros(len(circuits) + 1, dtype=np.int64)
    for i, (_, ops, par) in enumerate(circuits):
        op_off[i + 1] = op_off[i] + len(ops)
        par_off[i + 1] = par_off[i] + len(par)
    ops = np.concatenate([np.asarray(c[1]) for c in circuits]) if circuits else np.zeros(0)
    params = np.concatenate([np.asarray(c[2], dtype=np.float64) for c in circuits]) if circuits else np.zeros(0)
    return nq, op_off, ops, par_off, params


def uneven_ansatz_batch(count: int, seed: int, n_range=(10, 22), depth_range=(4, 12)):
    """`count` ansatz circuits (see ansatz()): qubit counts uniform in n_range,
    depths uniform in depth_range, angles uniform in [0, 2pi); returns
    [(n, ops, params), ...] for pack_batch."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        ops = ansatz(n, depth).ops()
        out.append((n, ops, rng.random(2 * n * depth) * 2.0 * math.pi))
    return out


def interleaved_trajectories(n: int, ops: np.ndarray, seeds, rate: float = -1.0, multiplier: float = 1.0):
    """Pauli-noise trajectories in the interleaved mode: for each seed, the
    gate list with each drawn flip inserted right after the gate it was drawn
    for (X before Z, on that gate's target). The draws are the Noise class's
    (two per eligible gate, the reference's order, Noise.cpp:21-32), the ones
    noise_frames folds into one frame after the last gate in the reference's
    deferred mode. Returns [ops_j, ...] (an unchanged op array when nothing
    was drawn)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    off, gate, code = _native().noise_insertions(n, ops["kind"], ops["target"],
                                                 np.ascontiguousarray(seeds, dtype=np.uint32), rate, multiplier)
    out = []
    for j in range(len(off) - 1):
        a, b = int(off[j]), int(off[j + 1])
        if a == b:
            out.append(ops)
            continue
        pieces, prev = [], 0
        for g, c in zip(gate[a:b], code[a:b]):
            g = int(g)
            pieces.append(ops[prev:g + 1])
            extra = np.zeros(2, dtype=OP_DTYPE)
            k = 0
            for bit, kind in ((1, X), (2, Z)):
                if int(c) & bit:
                    extra[k]["kind"] = kind
                    extra[k]["target"] = ops[g]["target"]
                    extra[k]["control"] = -1
                    extra[k]["param"] = -1
                    k += 1
            pieces.append(extra[:k])
            prev = g + 1
        pieces.append(ops[prev:])
        out.append(np.concatenate(pieces))
    return out


def amp_updates(ops: np.ndarray, n: int) -> float:
    """Algorithmic amplitude updates sum_g T(g): N for H, X, Y, RX, RY, RZ,
    generic 2x2; N/2 for Z, phase, CNOT, SWAP; N/4 for CPhase; N/2^m for
    MCPhase with m distinct pinned qubits (0 for identity, self-swap, c == t CNOT)."""
    N = float(1 << n)
    tot = 0.0
    for o in ops:
        k = int(o["kind"])
        if k in (H, X, Y, RX, RY, RZ, MATRIX):
            tot += N
        elif k in (Z, PHASE):
            tot += N / 2
        elif k == CNOT:
            tot += 0.0 if o["control"] == o["target"] else N / 2
        elif k == SWAP:
            tot += 0.0 if o["control"] == o["target"] else N / 2
        elif k == CPHASE:
            m = 1 if o["control"] == o["target"] else 2
            tot += N / (1 << m)
        elif k == MCPHASE:
            mask = int(o["controls"]) | (1 << int(o["target"]))
            tot += N / (1 << bin(mask).count("1"))
    return tot

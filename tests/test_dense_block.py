"""qse_compose_dense (host): the 2^m x 2^m unitary of a gate list on qubits
0..m-1, checked without a GPU against the oracle's state-vector run of the same
gates (every column, and a random state), against the reference's compiled CPU
backend when it is built here, and against closed forms: the QFT of QFT.cpp is
the bit-reversed unitary DFT, a single Hadamard on qubit q is I (x) H (x) I, and
every composed matrix is unitary."""
import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G


def bitrev(v, n):
    return int(format(v, f"0{n}b")[::-1], 2)


@pytest.mark.parametrize("m", [1, 2, 4, 5, 6])
def test_columns_equal_oracle_runs(m):
    ops = G.random_layered(m, 5, seed=11 + m).ops()
    u = capi.compose_dense(m, ops)
    for c in range(1 << m):
        np.testing.assert_allclose(u[:, c], O.run_ops(O.basis(m, c), m, ops), rtol=0, atol=1e-13)
    np.testing.assert_allclose(u.conj().T @ u, np.eye(1 << m), rtol=0, atol=1e-12)


def test_random_state_and_reference():
    m = 6
    rng = np.random.default_rng(5)
    ops = G.random_layered(m, 8, seed=3).ops()
    psi = rng.normal(size=1 << m) + 1j * rng.normal(size=1 << m)
    psi /= np.linalg.norm(psi)
    u = capi.compose_dense(m, ops)
    np.testing.assert_allclose(u @ psi, O.run_ops(psi, m, ops), rtol=0, atol=1e-13)
    if O.ref() is None:
        pytest.skip("reference CPU backend not built here")
    np.testing.assert_allclose(u @ psi, O.ref_run_ops(psi, m, ops), rtol=0, atol=1e-13)


def test_every_gate_kind_and_parameters():
    m = 4
    g = G.GateList(m)
    g.hadamard(0); g.pauli_x(1); g.pauli_y(2); g.pauli_z(3); g.phase(0.3, 1)
    g.rotate_x(0.7, 2, param=0, scale=2.0); g.rotate_y(-1.1, 3); g.rotate_z(2.2, 0); g.cnot(0, 3)
    g.cphase(0.9, 1, 2); g.mcphase(1.3, [0, 1, 2], 3); g.swap(0, 2)
    g.gate(np.array([[0.6, 0.8j], [0.8j, 0.6]]), 1)
    ops = g.ops()   # RX angle 0.7 + 2 * params[0]
    params = np.array([0.25])
    u = capi.compose_dense(m, ops, params)
    for c in range(1 << m):
        np.testing.assert_allclose(u[:, c], O.run_ops(O.basis(m, c), m, ops, params), rtol=0, atol=1e-13)


@pytest.mark.parametrize("m", [3, 5])
def test_qft_is_bit_reversed_dft(m):
    N = 1 << m
    u = capi.compose_dense(m, G.qft(m).ops())
    k = np.array([bitrev(i, m) for i in range(N)])
    want = np.exp(2j * np.pi * np.outer(k, k) / N) / np.sqrt(N)
    np.testing.assert_allclose(u, want, rtol=0, atol=1e-13)


def test_single_gate_kron():
    m, q = 3, 1
    g = G.GateList(m)
    g.hadamard(q)
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    want = np.kron(np.kron(np.eye(2), h), np.eye(2))   # qubit 2 (x) qubit 1 (x) qubit 0
    np.testing.assert_allclose(capi.compose_dense(m, g.ops()), want, rtol=0, atol=1e-15)


def test_rejects_qubits_outside_block():
    g = G.GateList(6)
    g.hadamard(5)
    with pytest.raises(capi.QSEError):
        capi.compose_dense(4, g.ops())

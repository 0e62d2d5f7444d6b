"""Gate-list files (include/quantumsimevo/GateListFile.h): the text format is
read and written by one C++ implementation (GateListFile.cpp), reached here
through the Python module. Round trips keep every field bit for bit, the
committed example pins the format, and malformed input is rejected with the
line number."""
import math

import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

ROOT = __import__("pathlib").Path(__file__).resolve().parent


def example_gates() -> G.GateList:
    g = G.GateList(5)
    g.hadamard(0).pauli_x(1).pauli_y(2).pauli_z(3).identity(4)
    g.phase(0.5, 0).rotate_x(-1.25, 1).rotate_y(3.0000000000000004, 2)
    g.rotate_z(0.125, 3, param=2, scale=-2.0)
    g.cnot(0, 4)
    g.cphase(0.78539816339744828, 1, 3, param=0, scale=1.5)
    g.mcphase(3.1415926535897931, [0, 2, 4], 1)
    g.swap(4, 2)
    g.gate(np.array([[0.6, 0.8j], [0.8j, 0.6]]), 3)
    return g


def test_example_file_reads_to_the_expected_records():
    n, ops = G.load_gates(ROOT / "golden" / "gatelist_example.qsegates")
    want = example_gates().ops()
    assert n == 5 and len(ops) == len(want)
    assert ops.tobytes() == want.tobytes()


def test_canonical_text_is_stable():
    text = example_gates().to_text()
    assert text.splitlines()[:3] == ["qse-gates 1", "qubits 5", "H 0"]
    assert "RZ 3 0.125 param 2 scale -2" in text
    assert "CPHASE 1 3 0.78539816339744828 param 0 scale 1.5" in text
    assert "MCPHASE 0,2,4 1 3.1415926535897931" in text
    assert "MATRIX 3 0.59999999999999998 0 0 0.80000000000000004 0 0.80000000000000004 0.59999999999999998 0" in text
    mr = G.GateList(4).measure(2, 7).reset_qubit(1).to_text()
    assert "MEASURE 2 7\nRESET 1\n" in mr
    n, ops = G.gates_from_text(text)
    assert G.gates_to_text(n, ops) == text


@pytest.mark.parametrize("make", [
    lambda: (12, G.random_mixed(12, 400, 7).ops()),
    lambda: (16, G.qaoa_maxcut(16, G.random_regular_graph(16, 3, seed=3), 4).ops()),
    lambda: (10, G.ansatz(10, 6).ops()),
    lambda: (9, G.qft(9).ops()),
    lambda: (6, G.GateList(6).hadamard(0).measure(0, 0).reset_qubit(3).cnot(0, 5).measure(5, 1).ops()),
])
def test_round_trip_bit_exact(make, tmp_path):
    n, ops = make()
    path = tmp_path / "c.qsegates"
    G.save_gates(path, n, ops)
    n2, ops2 = G.load_gates(path)
    assert n2 == n and ops2.tobytes() == ops.tobytes()
    if np.any(ops2["kind"] >= capi.MEASURE) and np.all(ops2["kind"] != capi.MATRIX):
        return
    # the loaded list is a batch entry as it stands
    entry = (n2, ops2, np.zeros(max(int(ops2["param"].max()) + 1, 0)))
    nq, op_off, packed, par_off, params = G.pack_batch([entry, entry])
    assert list(nq) == [n, n] and len(packed) == 2 * len(ops)


@pytest.mark.parametrize("text, needle", [
    ("", "empty input"),
    ("qse-gates 2\nqubits 3\n", "unsupported format version"),
    ("qubits 3\nH 0\n", "expected 'qse-gates 1'"),
    ("qse-gates 1\nH 0\n", "expected 'qubits <n>'"),
    ("qse-gates 1\nqubits 3\nFOO 1\n", "line 3: unknown gate 'FOO'"),
    ("qse-gates 1\nqubits 3\nH 3\n", "line 3: qubit 3 outside [0, 3)"),
    ("qse-gates 1\nqubits 3\nRX 1\n", "RX takes 2 fields"),
    ("qse-gates 1\nqubits 3\nCNOT 0 1 param 0 scale 1\n", "CNOT cannot carry a parameter"),
    ("qse-gates 1\nqubits 3\nRZ 0 zero\n", "not a number: 'zero'"),
    ("qse-gates 1\nqubits 3\nMCPHASE 0,0 1 1.0\n", "repeated control qubit 0"),
    ("qse-gates 1\nqubits 3\nH -1\n", "qubit -1 outside"),
    ("qse-gates 1\nqubits 3\nMEASURE 1\n", "MEASURE takes 2 fields"),
    ("qse-gates 1\nqubits 3\nMEASURE 1 -2\n", "result slot outside"),
    ("qse-gates 1\nqubits 3\nRESET 1 0\n", "RESET takes 1 fields"),
])
def test_malformed_input_names_the_problem(text, needle):
    with pytest.raises(RuntimeError, match=__import__("re").escape(needle)):
        G.gates_from_text(text)


def test_writer_rejects_what_the_format_cannot_hold():
    bad = np.zeros(1, dtype=capi.OP_DTYPE)
    bad[0] = (capi.CNOT, 1, 0, 0, 0, 0.0, 1.0, (0.0,) * 8)  # parameter on a CNOT
    with pytest.raises(RuntimeError, match="cannot carry a parameter"):
        G.gates_to_text(3, bad)
    angles = G.GateList(3).rotate_x(math.pi / 3, 0).ops()
    assert G.gates_from_text(G.gates_to_text(3, angles))[1]["theta"][0] == math.pi / 3

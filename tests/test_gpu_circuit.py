"""The reference's Python API (QSEPythonModule/QuantumSimEvoPy.cpp) on the HIP backend:
known answers of QSETests.cpp, the flush/markHostDirty contract, noise draws and
seeded device readout, each against the CPU oracle."""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Circuit():
    from quantumsimevo_amd import _quantumsimevo
    return _quantumsimevo.Circuit


def test_qsetests_known_answers(Circuit):
    """QSETests.cpp:76-170 through the Circuit API (1e-12)."""
    s2 = 1 / math.sqrt(2)
    c, s = math.cos(math.pi / 4), math.sin(math.pi / 4)
    cases = [
        (lambda q: q.hadamard(0), [s2, s2]),
        (lambda q: q.pauli_x(0), [0, 1]),
        (lambda q: q.pauli_y(0), [0, 1j]),
        (lambda q: q.pauli_z(0), [1, 0]),
        (lambda q: q.phase(math.pi / 2, 0), [1, 0]),
        (lambda q: q.rotate_x(math.pi / 2, 0), [c, -1j * s]),
        (lambda q: q.rotate_y(math.pi / 2, 0), [c, s]),
        (lambda q: q.rotate_z(math.pi / 2, 0), [c - 1j * s, 0]),
    ]
    qc = Circuit(1)
    for f, want in cases:
        qc.reset()
        f(qc)
        qc.execute()
        assert np.max(np.abs(qc.state() - np.array(want))) <= 1e-12
    for f, want in [
        (lambda q: q.swap(0, 1), [1, 0, 0, 0]),
        (lambda q: (q.pauli_x(0), q.swap(0, 1)), [0, 0, 1, 0]),
        (lambda q: (q.pauli_x(1), q.swap(0, 1)), [0, 1, 0, 0]),
        (lambda q: (q.hadamard(0), q.swap(0, 1)), [s2, 0, s2, 0]),
        (lambda q: (q.pauli_x(0), q.swap(0, 1), q.swap(0, 1)), [0, 1, 0, 0]),
    ]:
        q2 = Circuit(2)
        f(q2)
        q2.execute(True)  # print_steps: per-gate timed path
        assert np.max(np.abs(q2.state() - np.array(want))) <= 1e-12


def build(qc, gl):
    for o in gl.ops():
        k, t, c, th = int(o["kind"]), int(o["target"]), int(o["control"]), float(o["theta"])
        {capi.H: lambda: qc.hadamard(t), capi.X: lambda: qc.pauli_x(t), capi.Y: lambda: qc.pauli_y(t),
         capi.Z: lambda: qc.pauli_z(t), capi.PHASE: lambda: qc.phase(th, t), capi.RX: lambda: qc.rotate_x(th, t),
         capi.RY: lambda: qc.rotate_y(th, t), capi.RZ: lambda: qc.rotate_z(th, t), capi.CNOT: lambda: qc.cnot(c, t),
         capi.CPHASE: lambda: qc.cphase(th, c, t), capi.SWAP: lambda: qc.swap(t, c),
         capi.MCPHASE: lambda: qc.mcphase(th, [q for q in range(64) if (int(o["controls"]) >> q) & 1], t)}[k]()


def test_circuit_execute_replays_from_current_state(Circuit):
    n = 14
    gl = G.random_layered(n, 8, 5)
    qc = Circuit(n)
    build(qc, gl)
    qc.execute()
    once = O.run_ops(O.basis(n), n, gl.ops())
    assert np.max(np.abs(qc.state() - once)) < 1e-12
    qc.execute()  # Circuit.cpp:67: replays the list from the current state
    twice = O.run_ops(once, n, gl.ops())
    assert np.max(np.abs(qc.state() - twice)) < 1e-12
    qc.reset()
    assert np.array_equal(qc.state(), O.basis(n))


def test_seeded_sample_and_measure_match_oracle(Circuit):
    n = 16
    gl = G.random_layered(n, 10, 9)
    qc = Circuit(n)
    build(qc, gl)
    qc.execute()
    want = O.run_ops(O.basis(n), n, gl.ops())
    qc.set_seed(77)
    got = np.array(qc.sample(5000), dtype=np.uint64)
    u = O.uniforms(77, 5001)
    assert np.array_equal(got, O.sample(want, n, u[:5000]))
    idx = qc.measure()
    o_idx, collapsed = O.measure(want, n, u[5000])
    assert idx == o_idx
    assert np.array_equal(qc.state(), collapsed)


def test_noise_frame_matches_reference_draws(Circuit):
    """Noise.cpp semantics: seed 42, draws before each gate, X then Z at the end."""
    n = 12
    gl = G.random_layered(n, 30, 3)
    qc = Circuit(n, True)
    build(qc, gl)
    qc.execute()
    x, z = O.noise_trajectories(n, gl.ops(), 42, 1)
    stats = qc.noise_stats()
    assert stats["x_mask"] == int(x[0]) and stats["z_mask"] == int(z[0])
    want = O.pauli_frame(O.run_ops(O.basis(n), n, gl.ops()), n, int(x[0]), int(z[0]))
    assert np.max(np.abs(qc.state() - want)) < 1e-12


def test_flush_contract_and_printing(Circuit, capfd):
    qc = Circuit(3)
    qc.hadamard(0)
    qc.cnot(0, 2)
    qc.execute()
    qc.print_state()  # flushes, prints |000> and |101>
    out = capfd.readouterr().out
    assert "|000>" in out and "|101>" in out


def test_state_io_and_expectation(Circuit):
    n = 12
    psi = G.random_state(n, 2)
    qc = Circuit(n)
    qc.set_state(psi)
    assert np.array_equal(qc.state(), psi)
    m, w = G.zsum_observable(n)
    assert abs(qc.expectation_z(list(map(int, m)), list(w)) - O.expect_zstrings(psi, n, m, w)) < 1e-12


def test_generic_gate_and_errors(Circuit):
    qc = Circuit(3)
    qc.gate(np.array([[0, 1], [1, 0]], dtype=complex), 1)
    qc.execute()
    assert np.array_equal(qc.state(), O.basis(3, 2))
    with pytest.raises(IndexError):
        qc.hadamard(5)


def test_complex64_circuit(Circuit):
    n = 15
    gl = G.random_layered(n, 6, 1)
    qc = Circuit(n, False, "c64")
    build(qc, gl)
    qc.execute()
    assert np.max(np.abs(qc.state() - O.run_ops(O.basis(n), n, gl.ops()))) < 2e-5


def test_gate_list_files_through_the_circuit(Circuit, tmp_path):
    """Circuit.save_gates writes the same text as the Python writer for the
    same gates, and a Circuit built from the file (load_gates) reaches the
    same state, checked against the oracle (1e-12)."""
    n = 12
    gl = G.random_mixed(n, 300, 11)
    ops = gl.ops()
    ops = ops[ops["kind"] != capi.IDENTITY]  # the Circuit API has no identity gate
    keep = G.GateList.from_ops(n, ops)
    qc = Circuit(n)
    for o in ops:
        if int(o["kind"]) == capi.MATRIX:
            m = np.array(o["m"]).reshape(2, 2, 2)
            qc.gate(m[..., 0] + 1j * m[..., 1], int(o["target"]))
        else:
            build(qc, G.GateList.from_ops(n, o[None]))
    path = tmp_path / "mixed.qsegates"
    qc.save_gates(str(path))
    assert path.read_text() == keep.to_text()
    assert np.frombuffer(qc.gate_list(), dtype=capi.OP_DTYPE).tobytes() == ops.tobytes()
    qc.execute()
    qd = Circuit(n)
    qd.load_gates(str(path))
    qd.execute()
    want = O.run_ops(O.basis(n), n, ops)
    assert np.array_equal(qd.state(), qc.state())
    assert np.max(np.abs(qd.state() - want)) < 1e-12
    # the committed example file, through the Circuit (parameter bindings refused there)
    with pytest.raises(RuntimeError, match="parameter bindings"):
        Circuit(5).load_gates(str(Path(__file__).resolve().parent / "golden" / "gatelist_example.qsegates"))

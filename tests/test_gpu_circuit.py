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


def measured_list(n, seed, layers=6):
    """A random layered circuit with MEASURE and RESET_QUBIT ops inserted after
    random gates (result slots in order); returns (ops, number of measurement
    ops, number of result slots)."""
    rng = np.random.default_rng(seed)
    base = G.random_layered(n, layers, seed).ops()
    g = G.GateList.from_ops(n, base)
    rows = list(g._ops)
    out, nmeas, nslot = [], 0, 0
    for r in rows:
        out.append(r)
        x = rng.random()
        if x < 0.08:
            out.append((capi.MEASURE, int(rng.integers(n)), nslot, -1, 0, 0.0, 0.0, (0.0,) * 8))
            nslot += 1
            nmeas += 1
        elif x < 0.12:
            out.append((capi.RESET_QUBIT, int(rng.integers(n)), -1, -1, 0, 0.0, 0.0, (0.0,) * 8))
            nmeas += 1
    g._ops = out
    return g.ops(), nmeas, nslot


@pytest.mark.parametrize("n,prec", [(4, capi.C128), (12, capi.C128), (16, capi.C128), (14, capi.C64)])
def test_run_measured_vs_oracle_replay(ctx, n, prec):
    """GateType::MEASURE / RESET_QUBIT mid-circuit (qse_run_measured) against the
    oracle's replay with the same uniforms: outcomes equal, final state 1e-12
    (complex64: 1e-5; uniforms within 1e-6 of P(1) are avoided there)."""
    ops, nmeas, nslot = measured_list(n, 40 + n)
    assert nmeas >= 2
    u = np.random.default_rng(n).random(nmeas)
    want, res_o = O.run_measured(O.basis(n), n, ops, u, nslot)
    st = ctx.state(n, prec)
    res = st.run_measured(ops, u, nslot)
    assert np.array_equal(res, res_o)
    assert np.max(np.abs(st.download() - want)) < (1e-12 if prec == capi.C128 else 1e-5)
    st.close()


def test_run_measured_rejects_bad_lists(ctx):
    st = ctx.state(5)
    ops = G.GateList(5).hadamard(0).measure(0, 0).reset_qubit(1).ops()
    with pytest.raises(capi.QSEError):
        st.run_measured(ops, [0.5], 1)  # two measurement ops, one uniform
    with pytest.raises(capi.QSEError):
        st.run_measured(ops, [0.5, 0.5], 0)  # result slot 0 outside [0, 0)
    with pytest.raises(capi.QSEError):
        st.run(ops)  # the unitary entry point rejects measurement ops
    assert np.array_equal(st.download(), O.basis(5))  # nothing ran
    st.close()


def test_circuit_measure_and_reset_commands(Circuit):
    """Circuit.queue_measure / queue_reset: executed in command order with the
    circuit's seeded readout uniforms, outcomes in their result slots; equal to
    the oracle replay of the circuit's own gate list with those uniforms, and
    the gate list round-trips through a file."""
    n = 11
    qc = Circuit(n)
    qc.set_seed(17)
    ops, nmeas, nslot = measured_list(n, 5)
    slots = []
    for o in ops:
        k, t = int(o["kind"]), int(o["target"])
        if k == capi.MEASURE:
            slots.append(qc.queue_measure(t))
        elif k == capi.RESET_QUBIT:
            qc.queue_reset(t)
        else:
            build(qc, G.GateList.from_ops(n, o[None]))
    assert slots == list(range(nslot))
    qc.execute()
    u = O.uniforms(17, nmeas)
    want, res_o = O.run_measured(O.basis(n), n, ops, u, nslot)
    assert list(qc.measurement_results) == res_o.tolist()
    assert np.max(np.abs(qc.state() - want)) < 1e-12
    assert np.frombuffer(qc.gate_list(), dtype=capi.OP_DTYPE).tobytes() == ops.tobytes()


def test_device_mirror_contract():
    """The device-mirror contract of the backend interface
    (CircuitUnitaryOperationGPU.cu:158-192, ICircuitUnitaryOperation.h:12-17),
    driven through StateVector and the HIP backend directly:
      - the first gate uploads the host state; flush() brings the device state back;
      - flush() is a no-op while the host is marked dirty (GPU.cu:183): host
        writes after mark_host_dirty() survive it, and the next gate re-uploads them;
      - gates queued before mark_host_dirty() act on content that the re-upload
        replaces, so they are lost (as on the reference's device);
      - binding another StateVector flushes the previous one first (GPU.cu:163)."""
    from quantumsimevo_amd import _quantumsimevo as Q
    n = 3
    b = Q.HIPBackend(0, "c128")
    a = Q.StateVector(n)
    b.apply_pauli_x(a, 0)
    assert np.array_equal(a.host_data(), O.basis(n, 0))  # no flush yet: host still |000>
    b.flush()
    assert np.array_equal(a.host_data(), O.basis(n, 1))
    # host write + dirty mark: flush keeps the host copy
    a.set_host_data(O.basis(n, 2))
    b.mark_host_dirty()
    b.flush()
    assert np.array_equal(a.host_data(), O.basis(n, 2))
    b.apply_pauli_x(a, 2)  # re-upload |010>, then X(2)
    b.flush()
    assert np.array_equal(a.host_data(), O.basis(n, 6))
    # a gate queued before the dirty mark is lost with the device content it acted on
    b.apply_pauli_x(a, 0)
    b.mark_host_dirty()
    b.apply_pauli_x(a, 1)  # re-upload |110> (host), X(1) -> |100>
    b.flush()
    assert np.array_equal(a.host_data(), O.basis(n, 4))
    # rebinding flushes the previous state
    b.apply_hadamard(a, 0)
    c = Q.StateVector(n)
    b.apply_pauli_x(c, 1)  # binds c: a is flushed first
    want = O.run_ops(O.basis(n, 4), n, G.GateList(n).hadamard(0).ops())
    assert np.max(np.abs(a.host_data() - want)) < 1e-15
    b.flush()
    assert np.array_equal(c.host_data(), O.basis(n, 2))


def test_get_state_vector_does_not_flush(Circuit):
    """Circuit.execute() and getStateVector() leave the host copy as it was
    (Circuit.cpp:64-120, Circuit.h:53); print_state() flushes (Circuit.cpp:135-138)."""
    n = 4
    qc = Circuit(n)
    qc.pauli_x(2)
    qc.execute()
    assert np.array_equal(qc.host_state(), O.basis(n, 0))  # stale until a flush
    assert np.array_equal(qc.state(), O.basis(n, 4))       # device copy, read without flushing
    assert np.array_equal(qc.host_state(), O.basis(n, 0))
    qc.print_state()
    assert np.array_equal(qc.host_state(), O.basis(n, 4))


def test_hardware_check_report(capfd):
    """HardwareCheck (HardwareCheck.cpp:35-98) on the HIP device: the largest
    state follows the reference's maxQubits rule floor(log2(0.8 * bytes / 16))
    (HardwareCheck.cpp:56-60), and 8 bytes per amplitude for complex64."""
    from quantumsimevo_amd import _quantumsimevo as Q
    r = Q.device_report(0)
    assert r["available"] and r["name"] and r["compute_units"] > 0
    info = capi.device_info(0)
    assert r["total_bytes"] == info["total_bytes"] and r["compute_units"] == info["compute_units"]
    assert 0 < r["free_bytes"] <= r["total_bytes"]
    assert r["max_qubits_c128"] == math.floor(math.log2(0.8 * r["total_bytes"] / 16.0))
    assert r["max_qubits_c64"] == r["max_qubits_c128"] + 1
    # a state of max_qubits_c128 qubits fits the device at 16 bytes per amplitude (MI355X: 33)
    assert 16 * 2 ** r["max_qubits_c128"] <= r["total_bytes"]
    Q.print_report(0)
    out = capfd.readouterr().out
    assert f"Max qubits: {r['max_qubits_c128']} (complex128), {r['max_qubits_c64']} (complex64)" in out
    assert r["name"] in out


def test_reference_import_name():
    """`import QuantumSimEvo` (the reference's module name,
    QuantumSimEvoPy.cpp:8) gives the Circuit API on the HIP backend."""
    import QuantumSimEvo
    c = QuantumSimEvo.Circuit(20)
    c.hadamard(0)
    for q in range(19):
        c.cnot(q, q + 1)
    c.execute()
    assert set(int(x) for x in c.sample(500)) == {0, (1 << 20) - 1}

"""The drop-in claim on the GPU: the module built from the reference's own
Circuit.cpp, Noise.cpp and QuantumSimEvoPy.cpp (oracle/Makefile `dropin`) runs
on the HIP backend through the unchanged interface. Its results are checked
against the oracle: a 4-qubit circuit of every gate kind as the reference's
printState prints it (6 significant digits, so 1e-5), GHZ-20 sampling and
measurement, and GHZ-20 under the reference's own Pauli noise (seed 42,
its rates), whose outcomes must be the GHZ pair moved by the X frame that the
oracle's restatement of the same draws predicts."""
import json
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu

DROPIN = O.ROOT / "oracle" / "_ref" / "dropin"


@pytest.mark.skipif(not any(DROPIN.glob("QuantumSimEvo*.so")),
                    reason="drop-in module not built (needs the reference tree at build time)")
def test_reference_circuit_code_on_the_hip_backend():
    r = subprocess.run([sys.executable, str(O.ROOT / "tests" / "dropin_driver.py"), str(DROPIN), "run"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    # 1. every gate kind, 4 qubits
    g = G.GateList(4)
    g.hadamard(0).rotate_x(0.3, 1).rotate_y(-1.1, 2).rotate_z(0.7, 3).cnot(0, 1)
    g.phase(0.4, 2).pauli_y(3).cphase(0.9, 1, 2).mcphase(1.3, [0, 1], 3).swap(0, 3)
    g.pauli_x(1).pauli_z(2).hadamard(3)
    want = O.run_ops(O.basis(4), 4, g.ops())
    got = np.zeros(16, dtype=complex)
    for k, (re_, im_) in d["state4"].items():
        got[int(k)] = complex(re_, im_)
    assert np.max(np.abs(got - want)) < 1e-5
    # 2. GHZ-20
    top = (1 << 20) - 1
    assert d["ghz_samples"] == [0, top] and d["ghz_measure"] in (0, top)
    # 3. GHZ-20 with the reference's noise: the X frame moves both outcomes
    z = G.GateList(20).hadamard(0)
    for q in range(19):
        z.cnot(q, q + 1)
    for _ in range(250):
        for q in (2, 9, 15):
            z.cnot(0, q).cnot(0, q)
    x, _ = O.noise_trajectories(20, z.ops(), 42, 1)
    frame = int(x[0])
    assert frame == (1 << 2) | (1 << 9) | (1 << 15)  # the draws of seed 42 flip these three
    assert d["noisy_samples"] == sorted({frame, frame ^ top})

"""Runs circuits through the module built from the reference's own Circuit.cpp,
Noise.cpp and QuantumSimEvoPy.cpp against this repository's boundary headers and
HIP backend (oracle/Makefile `dropin`), in a process of its own so its module
name (QuantumSimEvo) does not meet the package's. Prints one JSON object.

  python tests/dropin_driver.py <module dir> describe
  python tests/dropin_driver.py <module dir> run
"""
import json
import os
import sys
import tempfile

sys.path.insert(0, sys.argv[1])
import QuantumSimEvo as Q  # noqa: E402


def parse_state(text):
    """printState lines '|b..b> : (re,im)' (StateVector.cpp:250-265) -> {index: complex}."""
    amps = {}
    for line in text.splitlines():
        if line.startswith("|") and " : (" in line:
            ket, val = line.split(" : ")
            re_, im_ = val.strip()[1:-1].split(",")
            amps[int(ket[1:-1], 2)] = complex(float(re_), float(im_))
    return amps


def main():
    out = {"module": Q.__file__, "methods": sorted(m for m in dir(Q.Circuit) if not m.startswith("_"))}
    if sys.argv[2] == "describe":
        try:
            Q.Circuit(3)
            out["created"] = True
        except Exception as e:  # no device: the factory refuses
            out["created"] = False
            out["error"] = f"{type(e).__name__}: {e}"
        print(json.dumps(out))
        return
    # 1. a 4-qubit circuit of every gate kind, printed by the reference's printState
    c = Q.Circuit(4)
    c.hadamard(0); c.rotate_x(0.3, 1); c.rotate_y(-1.1, 2); c.rotate_z(0.7, 3); c.cnot(0, 1)
    c.phase(0.4, 2); c.pauli_y(3); c.cphase(0.9, 1, 2); c.mcphase(1.3, [0, 1], 3); c.swap(0, 3)
    c.pauli_x(1); c.pauli_z(2); c.hadamard(3)
    c.execute()
    # printState writes through std::cout: capture file descriptor 1
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile(mode="w+") as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            c.print_state()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        text = tmp.read()
    out["state4"] = {str(k): [v.real, v.imag] for k, v in parse_state(text).items()}
    # 2. GHZ on 20 qubits: samples and one measurement
    g = Q.Circuit(20)
    g.hadamard(0)
    for q in range(19):
        g.cnot(q, q + 1)
    g.execute()
    out["ghz_samples"] = sorted(set(int(x) for x in g.sample(2000)))
    out["ghz_measure"] = int(g.measure())
    # 3. the same GHZ with the reference's Pauli noise (seed 42, its rates) and 250
    # rounds of CNOT pairs (identities) onto qubits 2, 9, 15 to draw flips: the
    # frame is applied after the last gate
    z = Q.Circuit(20, True)
    z.hadamard(0)
    for q in range(19):
        z.cnot(q, q + 1)
    for _ in range(250):
        for q in (2, 9, 15):
            z.cnot(0, q)
            z.cnot(0, q)
    z.execute()
    out["noisy_samples"] = sorted(set(int(x) for x in z.sample(2000)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()

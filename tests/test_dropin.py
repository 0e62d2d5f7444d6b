"""The drop-in claim on the CPU: the reference's own Circuit.cpp, Noise.cpp and
pybind11 module (QSEPythonModule/QuantumSimEvoPy.cpp) compile against
include/quantumsimevo and link with the HIP backend (oracle/Makefile `dropin`,
built by build() where the reference tree is present). The module imports,
binds the reference's Python API, and its factory refuses to run without a HIP
device (there is no CPU fallback)."""
import json
import subprocess
import sys

import pytest

import oracle_lib as O
from quantumsimevo_amd import capi

DROPIN = O.ROOT / "oracle" / "_ref" / "dropin"
# QuantumSimEvoPy.cpp:11-40
REFERENCE_METHODS = sorted(["hadamard", "pauli_x", "pauli_y", "pauli_z", "phase", "rotate_x", "rotate_y", "rotate_z",
                            "cnot", "cphase", "mcphase", "swap", "execute", "measure", "sample", "reset",
                            "print_state"])


def built():
    return any(DROPIN.glob("QuantumSimEvo*.so"))


@pytest.mark.skipif(not built(), reason="drop-in module not built (needs the reference tree at build time)")
def test_reference_module_builds_against_the_boundary_headers():
    r = subprocess.run([sys.executable, str(O.ROOT / "tests" / "dropin_driver.py"), str(DROPIN), "describe"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["methods"] == REFERENCE_METHODS
    if capi.device_count() == 0:
        assert not d["created"] and "no HIP device" in d["error"]


def test_symlinked_headers_are_the_boundary_ones():
    """Where the drop-in tree exists, the headers this repository replaces
    resolve into include/quantumsimevo and only Circuit.h and the CPU-backend
    declarations into the reference."""
    if not (DROPIN / "QuantumSimEvo").is_dir():
        pytest.skip("drop-in tree not built here")
    for h in ("StateVector.h", "Structurs.h", "ICircuitUnitaryOperation.h", "Noise.h",
              "CircuitUnitaryOperationFactory.h"):
        assert (DROPIN / "QuantumSimEvo" / h).resolve() == (O.ROOT / "include" / "quantumsimevo" / h).resolve()

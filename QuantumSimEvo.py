"""The reference's Python module name (PYBIND11_MODULE(QuantumSimEvo, m),
QSEPythonModule/QuantumSimEvoPy.cpp:8) for the MI355X build: `import
QuantumSimEvo; QuantumSimEvo.Circuit(20)` gives the same Circuit API, running
on the HIP backend. Everything is the compiled module
quantumsimevo_amd._quantumsimevo."""
from quantumsimevo_amd._quantumsimevo import *  # noqa: F401,F403
from quantumsimevo_amd._quantumsimevo import Circuit, HIPBackend, StateVector  # noqa: F401

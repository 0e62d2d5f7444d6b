"""quantumsimevo_amd — MI355X (gfx950) state-vector backend for the QuantumSimEvo simulator.

Layers:
  libqse_hip.so            HIP kernels + the C ABI (include/qse_hip.h); Python view in `capi`
  _quantumsimevo (pybind)  the C++ Circuit / StateVector / Noise / CircuitUnitaryOperationHIP
                           classes with the reference module's snake_case API
  circuits                 gate-list builders (QFT, QAOA, random layered circuits)
  autograd                 adjoint-method gradients behind torch.autograd.Function
"""
from . import capi, circuits  # noqa: F401

__all__ = ["capi", "circuits", "Circuit"]


def __getattr__(name):
    if name == "Circuit":
        from ._quantumsimevo import Circuit  # the pybind11 module (C++ host classes)

        return Circuit
    raise AttributeError(name)

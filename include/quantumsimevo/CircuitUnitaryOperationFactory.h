// CircuitUnitaryOperationFactory.h — backend selection (reference:
// QuantumSimEvo/CircuitUnitaryOperationFactory.h:4-10, .cpp:9-109).
// create() probes the HIP devices and binds the one with the most free memory
// that holds the state (2^n amplitudes + 64 MiB, as the reference's CUDA
// probe). There is no CPU backend: without a fitting device create() throws.
#pragma once
#include <memory>

#include "../qse_hip.h"
#include "ICircuitUnitaryOperation.h"

class CircuitUnitaryOperationFactory {
public:
    static std::unique_ptr<ICircuitUnitaryOperation> create(size_t numQubits, int precision = QSE_C128);
    // Device index create() would choose, or -1.
    static int selectDevice(size_t numQubits, int precision = QSE_C128);
};

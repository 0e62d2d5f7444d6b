// CircuitUnitaryOperationFactory.cpp — HIP device probe replacing the CUDA
// probe of the reference (CircuitUnitaryOperationFactory.cpp:17-93): the
// device with the most free memory that holds state + 64 MiB is bound.
#include "quantumsimevo/CircuitUnitaryOperationFactory.h"

#include <iostream>
#include <stdexcept>
#include <string>

#include "quantumsimevo/CircuitUnitaryOperationHIP.h"

int CircuitUnitaryOperationFactory::selectDevice(size_t numQubits, int precision) {
    int count = 0;
    if (qse_device_count(&count) != QSE_OK || count == 0) return -1;
    const size_t amp = (precision == QSE_C64) ? 8 : 16;
    const size_t required = (amp << numQubits) + (size_t(64) << 20);
    int best = -1;
    size_t bestFree = 0;
    for (int d = 0; d < count; ++d) {
        size_t f = 0, t = 0;
        if (qse_device_mem_info(d, &f, &t) != QSE_OK) continue;
        if (f >= required && f > bestFree) {
            best = d;
            bestFree = f;
        }
    }
    return best;
}

std::unique_ptr<ICircuitUnitaryOperation> CircuitUnitaryOperationFactory::create(size_t numQubits, int precision) {
    int count = 0;
    qse_device_count(&count);
    if (count == 0) throw std::runtime_error("CircuitUnitaryOperationFactory: no HIP device visible");
    const int dev = selectDevice(numQubits, precision);
    if (dev < 0)
        throw std::runtime_error("CircuitUnitaryOperationFactory: no HIP device has room for a " +
                                 std::to_string(numQubits) + "-qubit state");
    return std::make_unique<CircuitUnitaryOperationHIP>(dev, precision);
}

// StateVector.h — the simulator's state storage (API of the reference's
// QuantumSimEvo/StateVector.h:17-47).
//
// Differences, all deliberate for states of up to 34 qubits on one MI355X:
//  - the host copy (2^n complex<double>) is allocated lazily, on the first
//    data()/getRawPtr() call: a 33-qubit state never needs its 128 GiB host
//    mirror unless the caller reads it. Until then the logical content is the
//    basis state |basisIndex()> (|0...0> after construction or reset());
//  - measure/sample/measureQubit/resetQubit/applyUnitaryOperation run on the
//    device through the backend currently mirroring this state (the reference
//    runs them on the host copy, StateVector.cpp:77-248); with no device
//    backend bound they throw std::runtime_error.
#ifndef QSE_STATE_VECTOR_H
#define QSE_STATE_VECTOR_H
#include <complex>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "Structurs.h"

class CircuitUnitaryOperationHIP;

class StateVector {
public:
    explicit StateVector(size_t num);

    // Host buffer (allocated and filled with the basis state on first use).
    Complex* getRawPtr();
    std::vector<Complex>& data();

    size_t size() const { return stateSize; }
    size_t qubits() const { return numQubits; }
    bool hostAllocated() const { return materialized; }
    // Logical content while the host buffer is not allocated.
    uint64_t basisIndex() const { return basis; }

    void reset();
    size_t measure();
    int measureQubit(size_t qubit);
    void resetQubit(size_t qubit);
    std::vector<size_t> sample(int numShots);
    void applyUnitaryOperation(const Gate2x2& gate, size_t targetQubit);
    void printState();
    std::string stateAsString(size_t index);

    // Backend mirroring this state on a device (set by the backend on upload).
    void attach(CircuitUnitaryOperationHIP* b) { device = b; }
    CircuitUnitaryOperationHIP* attached() const { return device; }

private:
    void materialize();
    size_t numQubits, stateSize;
    std::vector<Complex> stateVector;
    bool materialized = false;
    uint64_t basis = 0;
    CircuitUnitaryOperationHIP* device = nullptr;
};

#endif

// StateVector.cpp — state storage with a lazily allocated host copy; readout
// runs on the device backend the state is attached to (see StateVector.h).
#include "quantumsimevo/StateVector.h"

#include <bitset>
#include <iostream>
#include <random>
#include <stdexcept>

#include "quantumsimevo/CircuitUnitaryOperationHIP.h"

StateVector::StateVector(size_t num) : numQubits(num) {
    if (num < 1 || num > 40) throw std::invalid_argument("StateVector: qubit count must be in [1, 40]");
    stateSize = 1ULL << numQubits;
}

void StateVector::materialize() {
    if (materialized) return;
    try {
        stateVector.assign(stateSize, Complex(0.0, 0.0));
    } catch (const std::bad_alloc&) {
        throw std::runtime_error("StateVector: not enough host memory for the " + std::to_string(numQubits) +
                                 "-qubit host copy");
    }
    stateVector[basis] = 1.0;
    materialized = true;
}

Complex* StateVector::getRawPtr() {
    materialize();
    return stateVector.data();
}

std::vector<Complex>& StateVector::data() {
    materialize();
    return stateVector;
}

// StateVector.cpp:34-46: |0...0>. Without a host copy only the basis index changes.
void StateVector::reset() {
    basis = 0;
    if (materialized) {
        std::fill(stateVector.begin(), stateVector.end(), Complex(0.0, 0.0));
        stateVector[0] = 1.0;
    }
}

static CircuitUnitaryOperationHIP& need_device(CircuitUnitaryOperationHIP* d) {
    if (!d) throw std::runtime_error("StateVector readout needs a HIP backend mirroring this state");
    return *d;
}

size_t StateVector::measure() {
    std::random_device rd;
    std::mt19937 gen(rd());
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    return need_device(device).measure(*this, dist(gen));
}

std::vector<size_t> StateVector::sample(int numShots) {
    std::random_device rd;
    std::mt19937 gen(rd());
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    std::vector<double> u((size_t)std::max(numShots, 0));
    for (auto& x : u) x = dist(gen);
    return need_device(device).sample(*this, u);
}

int StateVector::measureQubit(size_t qubit) {
    std::random_device rd;
    std::mt19937 gen(rd());
    return need_device(device).collapseQubit(*this, qubit, std::uniform_real_distribution<double>(0.0, 1.0)(gen),
                                             false);
}

void StateVector::resetQubit(size_t qubit) {
    std::random_device rd;
    std::mt19937 gen(rd());
    need_device(device).collapseQubit(*this, qubit, std::uniform_real_distribution<double>(0.0, 1.0)(gen), true);
}

void StateVector::applyUnitaryOperation(const Gate2x2& gate, size_t targetQubit) {
    need_device(device).applyGate(*this, gate, targetQubit);
}

// StateVector.cpp:250-260: amplitudes above 1e-10 of the first 32 indices
// (all of them up to 10 qubits).
void StateVector::printState() {
    const size_t limit = (numQubits > 10) ? 32 : stateSize;
    materialize();
    for (size_t i = 0; i < limit; ++i)
        if (std::abs(stateVector[i]) > 1e-10) std::cout << stateAsString(i) << " : " << stateVector[i] << std::endl;
    if (numQubits > 10) std::cout << "... (truncated for large N)\n";
}

// StateVector.cpp:262-265: |b_{n-1} ... b_0>, most significant qubit first.
std::string StateVector::stateAsString(size_t index) {
    const std::string s = std::bitset<64>(index).to_string();
    return "|" + s.substr(64 - numQubits) + ">";
}

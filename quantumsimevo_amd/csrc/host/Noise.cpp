// Noise.cpp — Pauli noise with the reference's draw order (Noise.cpp:3-44).
#include "quantumsimevo/Noise.h"

#include <algorithm>

// Noise.cpp:3-19: rates per gate type; CPHASE, MCPHASE and SWAP draw nothing.
Noise::Noise(size_t N, unsigned int seed, double errorRateMultiplier)
    : bitFlipCounts(N, 0), phaseFlipCounts(N, 0), errorMultiplier(errorRateMultiplier), gen(seed), dist(0.0, 1.0) {
    errorRates[GateType::X] = 0.001;
    errorRates[GateType::Y] = 0.001;
    errorRates[GateType::Z] = 0.001;
    errorRates[GateType::H] = 0.002;
    errorRates[GateType::CNOT] = 0.005;
    errorRates[GateType::PHASE] = 0.001;
    errorRates[GateType::RX] = 0.001;
    errorRates[GateType::RY] = 0.001;
    errorRates[GateType::RZ] = 0.001;
}

// Noise.cpp:21-32: two draws per eligible gate, always both.
void Noise::accumulateError(GateType type, size_t target) { drawError(type, target); }

int Noise::drawError(GateType type, size_t target) {
    auto it = errorRates.find(type);
    if (it == errorRates.end()) return 0;
    const double p = it->second * errorMultiplier;
    int drawn = 0;
    if (dist(gen) < p) {
        bitFlipCounts[target]++;
        drawn |= 1;
    }
    if (dist(gen) < p) {
        phaseFlipCounts[target]++;
        drawn |= 2;
    }
    return drawn;
}

// Noise.cpp:34-44: X_q then Z_q for every qubit with an odd count, ascending q.
void Noise::applyPauliNoise(StateVector& sv, ICircuitUnitaryOperation& op) {
    for (size_t q = 0; q < bitFlipCounts.size(); ++q) {
        if (bitFlipCounts[q] % 2 != 0) op.applyPauliX(sv, q);
        if (phaseFlipCounts[q] % 2 != 0) op.applyPauliZ(sv, q);
    }
}

void Noise::resetStats() {
    std::fill(bitFlipCounts.begin(), bitFlipCounts.end(), 0);
    std::fill(phaseFlipCounts.begin(), phaseFlipCounts.end(), 0);
}

size_t Noise::getTotalBitFlips() const {
    size_t t = 0;
    for (int c : bitFlipCounts) t += c;
    return t;
}

size_t Noise::getTotalPhaseFlips() const {
    size_t t = 0;
    for (int c : phaseFlipCounts) t += c;
    return t;
}

uint64_t Noise::bitFlipMask() const {
    uint64_t m = 0;
    for (size_t q = 0; q < bitFlipCounts.size(); ++q)
        if (bitFlipCounts[q] % 2) m |= 1ULL << q;
    return m;
}

uint64_t Noise::phaseFlipMask() const {
    uint64_t m = 0;
    for (size_t q = 0; q < phaseFlipCounts.size(); ++q)
        if (phaseFlipCounts[q] % 2) m |= 1ULL << q;
    return m;
}

void Noise::setRate(GateType type, double rate) {
    if (rate < 0) errorRates.erase(type);
    else errorRates[type] = rate;
}

// Noise.h — Pauli bit-/phase-flip noise (reference QuantumSimEvo/Noise.h:13-34,
// Noise.cpp:3-64): per-gate-type error rates, two std::mt19937 uniform draws per
// eligible gate (bit flip, then phase flip, on the gate's target), and at the
// end of an execution X_q then Z_q on every qubit with an odd flip count,
// applied through the backend.
#ifndef QSE_NOISE_H
#define QSE_NOISE_H
#include <cstdint>
#include <map>
#include <random>
#include <vector>

#include "ICircuitUnitaryOperation.h"
#include "Structurs.h"

// Where the drawn flips act. Deferred: the reference's model (Circuit.cpp:64-120,
// Noise.cpp:34-44): the flips are counted during the run and their parities
// applied as one Pauli frame after the last gate. Interleaved (no counterpart in
// the reference): each drawn flip acts right after the gate it was drawn for,
// X before Z. Both modes make the same draws in the same order.
enum class NoiseMode { Deferred, Interleaved };

class Noise {
public:
    explicit Noise(size_t N, unsigned int seed = 42, double errorRateMultiplier = 1);

    void applyPauliNoise(StateVector& sv, ICircuitUnitaryOperation& op);
    void accumulateError(GateType type, size_t target);
    void resetStats();
    size_t getTotalBitFlips() const;
    size_t getTotalPhaseFlips() const;

    // ---- extensions ----
    // The frame applyPauliNoise would apply: qubits with odd bit/phase flip counts.
    uint64_t bitFlipMask() const;
    uint64_t phaseFlipMask() const;
    void setRate(GateType type, double rate);  // rate < 0 removes the gate type
    // accumulateError's two draws, also returned: bit 0 a bit flip, bit 1 a
    // phase flip on `target` (the interleaved mode applies them at once).
    int drawError(GateType type, size_t target);

private:
    std::vector<int> bitFlipCounts;
    std::vector<int> phaseFlipCounts;
    std::map<GateType, double> errorRates;
    double errorMultiplier;
    std::mt19937 gen;
    std::uniform_real_distribution<double> dist;
};

#endif

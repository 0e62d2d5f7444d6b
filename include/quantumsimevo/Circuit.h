// Circuit.h — the command-list circuit API of the reference
// (QuantumSimEvo/Circuit.h:19-54, Circuit.cpp:5-143), on the HIP backend.
//
// Gates are queued as GateCommands; execute() replays them through the backend
// (fused on the device), drawing noise before each gate when Pauli noise is on.
// measure()/sample() run on the device with canonical uniforms drawn from the
// circuit's std::mt19937 (seeded from std::random_device like the reference's
// readout, or fixed with setSeed()). measure() collapses the device state.
#ifndef QSE_CIRCUIT_H
#define QSE_CIRCUIT_H
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "ICircuitUnitaryOperation.h"
#include "Noise.h"
#include "StateVector.h"
#include "Structurs.h"
#include "../qse_hip.h"

class CircuitUnitaryOperationHIP;

class Circuit {
public:
    Circuit(size_t N, bool simulatePauliNoise = false, int precision = 0);

    void hadamard(size_t q);
    void pauliX(size_t q);
    void pauliZ(size_t q);
    void pauliY(size_t q);
    void phase(double theta, size_t q);
    void rotateX(double theta, size_t q);
    void rotateY(double theta, size_t q);
    void rotateZ(double theta, size_t q);
    void cnot(size_t control, size_t target);
    void cphase(double theta, size_t control, size_t target);
    void mcphase(double theta, const std::vector<size_t>& controls, size_t target);
    void swap(size_t q0, size_t q1);

    void execute(bool print_steps = false);
    size_t measure();
    std::vector<size_t> sample(int numShots);
    void reset();

    void printState();
    StateVector& getStateVector() { return stateVector; }

    // ---- extensions ----
    void gate(const Gate2x2& g, size_t q);                 // generic 2x2 (applyGate)
    void setSeed(uint64_t seed);                           // readout RNG
    std::vector<Complex> state();                          // flushed copy of the state
    void setState(const std::vector<Complex>& amps);       // replace the state
    double expectationZ(const std::vector<uint64_t>& masks, const std::vector<double>& weights);
    int measureQubit(size_t q);                            // mid-circuit collapse, now
    void resetQubit(size_t q);
    // Mid-circuit measurement and reset as commands (GateType::MEASURE /
    // RESET_QUBIT, which the reference declares but never executes): run by
    // execute() in command order, each with the next uniform of the readout
    // generator. queueMeasure returns the command's result slot; the outcome of
    // the last execute() is measurementResults()[slot] (-1 before).
    size_t queueMeasure(size_t q);
    void queueReset(size_t q);
    const std::vector<int>& measurementResults() const { return results; }
    const std::vector<GateCommand>& commands() const { return commandList; }
    // Gate-list files (GateListFile.h): the queued gates as C-ABI records, and
    // saving them / appending a file's gates (same qubit count; parameter
    // bindings are for the batched and adjoint APIs and are rejected here).
    std::vector<qse_op> gateList() const;
    void saveGates(const std::string& path) const;
    void loadGates(const std::string& path);
    size_t numQubits() const { return stateVector.qubits(); }
    Noise& noise() { return pauliNoise; }
    // Deferred (default, the reference's frame after the last gate) or
    // Interleaved (each drawn flip right after its gate); see NoiseMode.
    void setNoiseMode(NoiseMode mode) { noiseMode = mode; }
    NoiseMode getNoiseMode() const { return noiseMode; }
    CircuitUnitaryOperationHIP& backend();

private:
    StateVector stateVector;
    std::vector<GateCommand> commandList;
    std::unique_ptr<ICircuitUnitaryOperation> circuitOp;
    bool simulateNoise;
    Noise pauliNoise;
    NoiseMode noiseMode = NoiseMode::Deferred;
    std::mt19937 readoutGen;
    std::vector<Gate2x2> matrices;  // generic 2x2 gates queued with gate()
    std::vector<int> results;       // outcomes of MEASURE commands by result slot
};

#endif

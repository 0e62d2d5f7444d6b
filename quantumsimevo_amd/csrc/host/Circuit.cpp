// Circuit.cpp — command list and replay (reference Circuit.cpp:5-143) on the
// HIP backend.
#include "quantumsimevo/Circuit.h"

#include <iostream>
#include <stdexcept>

#include "quantumsimevo/CircuitUnitaryOperationFactory.h"
#include "quantumsimevo/CircuitUnitaryOperationHIP.h"
#include "quantumsimevo/GateListFile.h"

Circuit::Circuit(size_t N, bool simulatePauliNoise, int precision)
    : stateVector(N), simulateNoise(simulatePauliNoise), pauliNoise(N, 42, 1) {
    circuitOp = CircuitUnitaryOperationFactory::create(N, precision);
    std::random_device rd;
    readoutGen.seed(rd());
}

static void check_q(const StateVector& sv, size_t q) {
    if (q >= sv.qubits()) throw std::out_of_range("qubit index " + std::to_string(q) + " out of range");
}

// Circuit.cpp:9-62: one GateCommand per call (argument order as the reference).
void Circuit::hadamard(size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::H, q, 0, 0.0, "Hadamard on qubit " + std::to_string(q)});
}
void Circuit::pauliX(size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::X, q, 0, 0.0, "PauliX on qubit " + std::to_string(q)});
}
void Circuit::pauliZ(size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::Z, q, 0, 0.0, "PauliZ on qubit " + std::to_string(q)});
}
void Circuit::pauliY(size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::Y, q, 0, 0.0, "PauliY on qubit " + std::to_string(q)});
}
void Circuit::phase(double theta, size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::PHASE, q, 0, theta, "Phase shift on qubit " + std::to_string(q)});
}
void Circuit::rotateX(double theta, size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::RX, q, 0, theta, "RX on qubit " + std::to_string(q)});
}
void Circuit::rotateY(double theta, size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::RY, q, 0, theta, "RY on qubit " + std::to_string(q)});
}
void Circuit::rotateZ(double theta, size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::RZ, q, 0, theta, "RZ on qubit " + std::to_string(q)});
}
void Circuit::cnot(size_t control, size_t target) {
    check_q(stateVector, control);
    check_q(stateVector, target);
    commandList.push_back({GateType::CNOT, target, control, 0.0,
                           "CNOT (control: " + std::to_string(control) + ", target: " + std::to_string(target) + ")"});
}
void Circuit::cphase(double theta, size_t control, size_t target) {
    check_q(stateVector, control);
    check_q(stateVector, target);
    commandList.push_back({GateType::CPHASE, target, control, theta,
                           "CPhase (theta: " + std::to_string(theta) + ", control: " + std::to_string(control) +
                               ", target: " + std::to_string(target) + ")"});
}
void Circuit::mcphase(double theta, const std::vector<size_t>& controls, size_t target) {
    check_q(stateVector, target);
    std::string desc = "MCPhase (theta: " + std::to_string(theta) + ", target: " + std::to_string(target) + ", controls:";
    for (size_t c : controls) {
        check_q(stateVector, c);
        desc += " " + std::to_string(c);
    }
    desc += ")";
    GateCommand cmd{GateType::MCPHASE, target, 0, theta, desc, 0, controls};
    commandList.push_back(cmd);
}
void Circuit::swap(size_t q0, size_t q1) {
    check_q(stateVector, q0);
    check_q(stateVector, q1);
    commandList.push_back({GateType::SWAP, q0, q1, 0.0,
                           "SWAP (q0: " + std::to_string(q0) + ", q1: " + std::to_string(q1) + ")"});
}

void Circuit::gate(const Gate2x2& g, size_t q) {
    check_q(stateVector, q);
    GateCommand cmd{GateType::MATRIX, q, 0, 0.0, "Gate2x2 on qubit " + std::to_string(q)};
    cmd.resultIdx = matrices.size();  // the matrix rides in a side table
    matrices.push_back(g);
    commandList.push_back(cmd);
}

size_t Circuit::queueMeasure(size_t q) {
    check_q(stateVector, q);
    GateCommand cmd{GateType::MEASURE, q, 0, 0.0, "Measure qubit " + std::to_string(q)};
    cmd.resultIdx = results.size();
    results.push_back(-1);
    commandList.push_back(cmd);
    return cmd.resultIdx;
}

void Circuit::queueReset(size_t q) {
    check_q(stateVector, q);
    commandList.push_back({GateType::RESET_QUBIT, q, 0, 0.0, "Reset qubit " + std::to_string(q)});
}

// Circuit.cpp:64-120: replay from the current state, noise draws before each gate,
// noise frame after the last one (Deferred); Interleaved applies each gate's drawn
// flips right after the gate instead.
void Circuit::execute(bool print_steps) {
    if (print_steps) std::cout << "Initial state: |0...0>\n";
    circuitOp->setPrintSteps(print_steps);
    const bool interleaved = simulateNoise && noiseMode == NoiseMode::Interleaved;
    for (const auto& cmd : commandList) {
        if (print_steps) std::cout << "Executing: " << cmd.description << std::endl;
        const int drawn = simulateNoise ? pauliNoise.drawError(cmd.type, cmd.target) : 0;
        switch (cmd.type) {
            case GateType::H: circuitOp->applyHadamard(stateVector, cmd.target); break;
            case GateType::X: circuitOp->applyPauliX(stateVector, cmd.target); break;
            case GateType::RX: circuitOp->applyRotateX(stateVector, cmd.target, cmd.theta); break;
            case GateType::RY: circuitOp->applyRotateY(stateVector, cmd.target, cmd.theta); break;
            case GateType::RZ: circuitOp->applyRotateZ(stateVector, cmd.target, cmd.theta); break;
            case GateType::CNOT: circuitOp->applyCNOT(stateVector, cmd.control, cmd.target); break;
            case GateType::CPHASE: circuitOp->applyCPhase(stateVector, cmd.control, cmd.target, cmd.theta); break;
            case GateType::MCPHASE:
                circuitOp->applyMCPhase(stateVector, cmd.controls_list, cmd.target, cmd.theta);
                break;
            case GateType::SWAP: circuitOp->applySwap(stateVector, cmd.target, cmd.control); break;
            case GateType::Y: circuitOp->applyPauliY(stateVector, cmd.target); break;
            case GateType::PHASE: circuitOp->applyPhase(stateVector, cmd.target, cmd.theta); break;
            case GateType::Z: circuitOp->applyPauliZ(stateVector, cmd.target); break;
            case GateType::IDENTITY: break;
            case GateType::MATRIX: circuitOp->applyGate(stateVector, matrices[cmd.resultIdx], cmd.target); break;
            case GateType::MEASURE:
            case GateType::RESET_QUBIT: {
                // StateVector::measureQubit / resetQubit (StateVector.cpp:158-221) on the device, with the
                // next uniform of the circuit's readout generator
                std::uniform_real_distribution<double> dist(0.0, 1.0);
                const int o = backend().collapseQubit(stateVector, cmd.target, dist(readoutGen),
                                                      cmd.type == GateType::RESET_QUBIT);
                if (cmd.type == GateType::MEASURE) results[cmd.resultIdx] = o;
                break;
            }
            default: std::cerr << "Unsupported gate type: " << static_cast<int>(cmd.type) << std::endl;
        }
        if (interleaved) {
            if (drawn & 1) circuitOp->applyPauliX(stateVector, cmd.target);
            if (drawn & 2) circuitOp->applyPauliZ(stateVector, cmd.target);
        }
    }
    if (simulateNoise && !interleaved) pauliNoise.applyPauliNoise(stateVector, *circuitOp);
    if (print_steps)
        std::cout << "Total of occured errors: \n BitFlips:" << pauliNoise.getTotalBitFlips()
                  << " times\n PhaseFlips:" << pauliNoise.getTotalPhaseFlips() << " times.\n";
}

CircuitUnitaryOperationHIP& Circuit::backend() { return static_cast<CircuitUnitaryOperationHIP&>(*circuitOp); }

// Circuit.cpp:122-125 with the collapse on the device: the device copy is the
// one later gates act on, so it is the one that collapses.
size_t Circuit::measure() {
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    return backend().measure(stateVector, dist(readoutGen));
}

// Circuit.cpp:140-143, sampled on the device.
std::vector<size_t> Circuit::sample(int numShots) {
    if (numShots < 0) throw std::invalid_argument("sample: negative shot count");
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    std::vector<double> u((size_t)numShots);
    for (auto& x : u) x = dist(readoutGen);
    return backend().sample(stateVector, u);
}

// Circuit.cpp:127-133.
void Circuit::reset() {
    stateVector.reset();
    circuitOp->markHostDirty();
    if (simulateNoise) pauliNoise.resetStats();
    commandList.clear();
    matrices.clear();
    results.clear();
}

void Circuit::printState() {
    circuitOp->flush();
    stateVector.printState();
}

void Circuit::setSeed(uint64_t seed) { readoutGen.seed((std::mt19937::result_type)seed); }

std::vector<Complex> Circuit::state() { return backend().readState(stateVector); }

void Circuit::setState(const std::vector<Complex>& amps) { backend().setState(stateVector, amps); }

double Circuit::expectationZ(const std::vector<uint64_t>& masks, const std::vector<double>& weights) {
    return backend().expectZStrings(stateVector, masks, weights);
}

int Circuit::measureQubit(size_t q) {
    check_q(stateVector, q);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    return backend().collapseQubit(stateVector, q, dist(readoutGen), false);
}

void Circuit::resetQubit(size_t q) {
    check_q(stateVector, q);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    backend().collapseQubit(stateVector, q, dist(readoutGen), true);
}

std::vector<qse_op> Circuit::gateList() const {
    std::vector<qse_op> out;
    out.reserve(commandList.size());
    for (const GateCommand& c : commandList) {
        qse_op g{};
        g.target = (int32_t)c.target;
        g.control = -1;
        g.param = -1;
        g.theta = c.theta;
        switch (c.type) {
            case GateType::IDENTITY:
                g.kind = QSE_GATE_IDENTITY;
                g.theta = 0.0;
                break;
            case GateType::MATRIX: {
                g.kind = QSE_GATE_MATRIX;
                const Gate2x2& m = matrices[c.resultIdx];
                for (int r = 0; r < 2; ++r)
                    for (int k = 0; k < 2; ++k) {
                        g.m[4 * r + 2 * k] = m.data[r][k].real();
                        g.m[4 * r + 2 * k + 1] = m.data[r][k].imag();
                    }
                g.theta = 0.0;
                break;
            }
            case GateType::MEASURE:
                g.kind = QSE_GATE_MEASURE;
                g.control = (int32_t)c.resultIdx;
                g.theta = 0.0;
                break;
            case GateType::RESET_QUBIT:
                g.kind = QSE_GATE_RESET_QUBIT;
                g.theta = 0.0;
                break;
            case GateType::X: g.kind = QSE_GATE_X; break;
            case GateType::Y: g.kind = QSE_GATE_Y; break;
            case GateType::Z: g.kind = QSE_GATE_Z; break;
            case GateType::H: g.kind = QSE_GATE_H; break;
            case GateType::PHASE: g.kind = QSE_GATE_PHASE; break;
            case GateType::RX: g.kind = QSE_GATE_RX; break;
            case GateType::RY: g.kind = QSE_GATE_RY; break;
            case GateType::RZ: g.kind = QSE_GATE_RZ; break;
            case GateType::CNOT: g.kind = QSE_GATE_CNOT; g.control = (int32_t)c.control; break;
            case GateType::CPHASE: g.kind = QSE_GATE_CPHASE; g.control = (int32_t)c.control; break;
            case GateType::MCPHASE:
                g.kind = QSE_GATE_MCPHASE;
                for (size_t q : c.controls_list) g.controls |= 1ULL << q;
                break;
            case GateType::SWAP: g.kind = QSE_GATE_SWAP; g.control = (int32_t)c.control; break;
            default: throw std::runtime_error("gate list: " + c.description + " has no gate-list form");
        }
        if (g.kind == QSE_GATE_X || g.kind == QSE_GATE_Y || g.kind == QSE_GATE_Z || g.kind == QSE_GATE_H ||
            g.kind == QSE_GATE_CNOT || g.kind == QSE_GATE_SWAP)
            g.theta = 0.0;
        out.push_back(g);
    }
    return out;
}

void Circuit::saveGates(const std::string& path) const {
    const std::vector<qse_op> ops = gateList();
    qse_gates::save(path, (int)numQubits(), ops.data(), ops.size());
}

void Circuit::loadGates(const std::string& path) {
    int n = 0;
    std::vector<qse_op> ops;
    qse_gates::load(path, &n, ops);
    if ((size_t)n != numQubits())
        throw std::runtime_error("gate list " + path + " is for " + std::to_string(n) + " qubits, the circuit has " +
                                 std::to_string(numQubits()));
    const size_t slot0 = results.size();  // the file's result slots follow the circuit's own
    for (const qse_op& g : ops) {
        if (g.param >= 0)
            throw std::runtime_error("gate list " + path + ": parameter bindings need the batched or adjoint API");
        switch (g.kind) {
            case QSE_GATE_IDENTITY: break;
            case QSE_GATE_X: pauliX(g.target); break;
            case QSE_GATE_Y: pauliY(g.target); break;
            case QSE_GATE_Z: pauliZ(g.target); break;
            case QSE_GATE_H: hadamard(g.target); break;
            case QSE_GATE_PHASE: phase(g.theta, g.target); break;
            case QSE_GATE_RX: rotateX(g.theta, g.target); break;
            case QSE_GATE_RY: rotateY(g.theta, g.target); break;
            case QSE_GATE_RZ: rotateZ(g.theta, g.target); break;
            case QSE_GATE_CNOT: cnot(g.control, g.target); break;
            case QSE_GATE_CPHASE: cphase(g.theta, g.control, g.target); break;
            case QSE_GATE_MCPHASE: {
                std::vector<size_t> cs;
                for (int q = 0; q < 64; ++q)
                    if ((g.controls >> q) & 1) cs.push_back((size_t)q);
                mcphase(g.theta, cs, g.target);
                break;
            }
            case QSE_GATE_SWAP: swap(g.target, g.control); break;
            case QSE_GATE_MATRIX:
                gate(Gate2x2(Complex(g.m[0], g.m[1]), Complex(g.m[2], g.m[3]), Complex(g.m[4], g.m[5]),
                             Complex(g.m[6], g.m[7])),
                     g.target);
                break;
            case QSE_GATE_MEASURE: {
                check_q(stateVector, g.target);
                GateCommand cmd{GateType::MEASURE, (size_t)g.target, 0, 0.0,
                                "Measure qubit " + std::to_string(g.target)};
                cmd.resultIdx = slot0 + (size_t)g.control;
                if (results.size() <= cmd.resultIdx) results.resize(cmd.resultIdx + 1, -1);
                commandList.push_back(cmd);
                break;
            }
            case QSE_GATE_RESET_QUBIT: queueReset(g.target); break;
            default: throw std::runtime_error("gate list: unsupported gate kind");
        }
    }
}

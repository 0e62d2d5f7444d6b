/*
 * ref_driver.cpp — ctypes driver for the reference simulator's own CPU code.
 *
 * TEST INFRASTRUCTURE ONLY (the reference oracle and the CPU timing baseline). Built by
 * oracle/Makefile into oracle/_ref/libqse_ref.so, together with the reference's
 * CPU sources compiled where they lie under the reference tree:
 *   QuantumSimEvo/CircuitUnitaryOperation.cpp     (scalar OpenMP backend)
 *   QuantumSimEvo/CircuitUnitaryOperationAVX.cpp  (AVX2 + OpenMP backend)
 *   QuantumSimEvo/StateVector.cpp, Noise.cpp, QubitUnitaryOperation.cpp
 * Circuit.cpp is not linked: its CircuitUnitaryOperationFactory::create needs the
 * CUDA runtime. This driver replays a gate list the way Circuit::execute does
 * (Circuit.cpp:73-112), calling the backends directly.
 *
 * Nothing here is part of the product path.
 */
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "CircuitUnitaryOperation.h"
#include "CircuitUnitaryOperationAVX.h"
#include "Noise.h"
#include "StateVector.h"

#include "../include/qse_hip.h"

namespace {

double op_angle(const qse_op& op, const double* params) {
    if (op.param >= 0 && params) return op.theta + op.scale * params[op.param];
    return op.theta;
}

std::unique_ptr<ICircuitUnitaryOperation> make_backend(int backend) {
    if (backend == 1) return std::make_unique<CircuitUnitaryOperationAVX>();
    return std::make_unique<CircuitUnitaryOperation>();
}

int replay(ICircuitUnitaryOperation& b, StateVector& sv, const qse_op* ops, int nops, const double* params) {
    for (int k = 0; k < nops; ++k) {
        const qse_op& op = ops[k];
        const double th = op_angle(op, params);
        switch (op.kind) {
            case QSE_GATE_IDENTITY: break;
            case QSE_GATE_H: b.applyHadamard(sv, op.target); break;
            case QSE_GATE_X: b.applyPauliX(sv, op.target); break;
            case QSE_GATE_Y: b.applyPauliY(sv, op.target); break;
            case QSE_GATE_Z: b.applyPauliZ(sv, op.target); break;
            case QSE_GATE_PHASE: b.applyPhase(sv, op.target, th); break;
            case QSE_GATE_RX: b.applyRotateX(sv, op.target, th); break;
            case QSE_GATE_RY: b.applyRotateY(sv, op.target, th); break;
            case QSE_GATE_RZ: b.applyRotateZ(sv, op.target, th); break;
            case QSE_GATE_CNOT: b.applyCNOT(sv, op.control, op.target); break;
            case QSE_GATE_CPHASE: b.applyCPhase(sv, op.control, op.target, th); break;
            case QSE_GATE_MCPHASE: {
                std::vector<size_t> c;
                for (int q = 0; q < 64; ++q)
                    if ((op.controls >> q) & 1) c.push_back((size_t)q);
                b.applyMCPhase(sv, c, op.target, th);
                break;
            }
            case QSE_GATE_SWAP: b.applySwap(sv, op.target, op.control); break;
            case QSE_GATE_MATRIX: {
                Gate2x2 g(Complex(op.m[0], op.m[1]), Complex(op.m[2], op.m[3]), Complex(op.m[4], op.m[5]),
                          Complex(op.m[6], op.m[7]));
                b.applyGate(sv, g, op.target);
                break;
            }
            default: return -1;
        }
    }
    return 0;
}

/* Records the X/Z calls Noise::applyPauliNoise makes through the backend
 * interface, so the reference's own flip pattern can be read out. */
class RecordingOp : public ICircuitUnitaryOperation {
public:
    uint64_t x = 0, z = 0;
    std::vector<int> order;  // +q+1 for X_q, -(q+1) for Z_q, in call order
    void applyGate(StateVector&, const Gate2x2&, size_t) override {}
    void applyCNOT(StateVector&, size_t, size_t) override {}
    void applyPauliX(StateVector&, size_t q) override { x ^= 1ULL << q; order.push_back((int)q + 1); }
    void applyPauliY(StateVector&, size_t) override {}
    void applyPauliZ(StateVector&, size_t q) override { z ^= 1ULL << q; order.push_back(-(int)q - 1); }
    void applyPhase(StateVector&, size_t, double) override {}
    void applyRotateX(StateVector&, size_t, double) override {}
    void applyRotateY(StateVector&, size_t, double) override {}
    void applyRotateZ(StateVector&, size_t, double) override {}
    void applyHadamard(StateVector&, size_t) override {}
    void applyCPhase(StateVector&, size_t, size_t, double) override {}
    void applyMCPhase(StateVector&, const std::vector<size_t>&, size_t, double) override {}
    void applySwap(StateVector&, size_t, size_t) override {}
};

}  // namespace

extern "C" {

/* Replay ops on psi (FP64 (re,im) pairs, 2^n amplitudes) with the reference's
 * scalar (backend 0) or AVX2 (backend 1) class. Returns 0, or -1 on an unknown kind. */
int ref_run_ops(int backend, double* psi, int n, const qse_op* ops, int nops, const double* params) {
    StateVector sv((size_t)n);
    std::memcpy(sv.getRawPtr(), psi, sizeof(double) * 2 * sv.size());
    auto b = make_backend(backend);
    const int rc = replay(*b, sv, ops, nops, params);
    std::memcpy(psi, sv.getRawPtr(), sizeof(double) * 2 * sv.size());
    return rc;
}

/* Time the replay of ops on a 2^n state initialised from psi_init (or |0> when
 * NULL) with `backend`; returns wall seconds of the replay alone (steady_clock,
 * the timing StateVector.cpp and CircuitUnitaryOperation.cpp use). */
double ref_time_ops(int backend, int n, const double* psi_init, const qse_op* ops, int nops, const double* params) {
    StateVector sv((size_t)n);
    if (psi_init) std::memcpy(sv.getRawPtr(), psi_init, sizeof(double) * 2 * sv.size());
    auto b = make_backend(backend);
    const auto t0 = std::chrono::steady_clock::now();
    replay(*b, sv, ops, nops, params);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

/* Timing protocol of the CPU baseline: one StateVector (allocated and zero-filled
 * once by its constructor, StateVector.cpp:14-32), `warmups` untimed replays of
 * ops[0..warm_ops) (page faults, thread pools), then `reps` timed replays of the
 * whole list, each time written to times[]. The state keeps evolving between
 * replays; the work per replay is the same. */
int ref_time_ops_repeat(int backend, int n, const qse_op* ops, int nops, int warm_ops, int warmups, int reps,
                        double* times) {
    StateVector sv((size_t)n);
    auto b = make_backend(backend);
    for (int w = 0; w < warmups; ++w) replay(*b, sv, ops, std::min(warm_ops, nops), nullptr);
    for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        replay(*b, sv, ops, nops, nullptr);
        const auto t1 = std::chrono::steady_clock::now();
        times[r] = std::chrono::duration<double>(t1 - t0).count();
    }
    return 0;
}

/* Drive the reference Noise class over a gate list exactly as Circuit::execute
 * does (accumulateError(cmd.type, cmd.target) before each gate, then
 * applyPauliNoise) and return the frame it applies: xmask/zmask of the qubits
 * that receive X/Z, the call order (2n ints max, +q+1 = X_q, -(q+1) = Z_q), and
 * the total bit/phase flip counts. */
int ref_noise_frame(int n, const qse_op* ops, int nops, uint32_t seed, double mult, uint64_t* xmask,
                    uint64_t* zmask, int* order, int* norder, int64_t* totals) {
    Noise noise((size_t)n, seed, mult);
    for (int k = 0; k < nops; ++k) noise.accumulateError(static_cast<GateType>(ops[k].kind), (size_t)ops[k].target);
    StateVector sv(1);
    RecordingOp rec;
    noise.applyPauliNoise(sv, rec);
    *xmask = rec.x;
    *zmask = rec.z;
    *norder = (int)rec.order.size();
    for (size_t i = 0; i < rec.order.size(); ++i) order[i] = rec.order[i];
    totals[0] = (int64_t)noise.getTotalBitFlips();
    totals[1] = (int64_t)noise.getTotalPhaseFlips();
    return 0;
}

}  // extern "C"

// CircuitUnitaryOperationHIP.cpp — see the header for the device-mirror
// contract. Every apply* becomes one qse_op (the C form of GateCommand) on the
// backend's queue; the queue runs through qse_run (fused passes).
#include "quantumsimevo/CircuitUnitaryOperationHIP.h"

#include <chrono>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>

namespace {

constexpr size_t QUEUE_LIMIT = 1 << 16;

qse_op make_op(int kind, size_t target, long control = -1, double theta = 0.0) {
    qse_op o{};
    o.kind = kind;
    o.target = (int32_t)target;
    o.control = (int32_t)control;
    o.param = -1;
    o.theta = theta;
    return o;
}

const char* kind_name(int k) {
    switch (k) {
        case QSE_GATE_X: return "PauliX";
        case QSE_GATE_Y: return "PauliY";
        case QSE_GATE_Z: return "PauliZ";
        case QSE_GATE_H: return "Hadamard";
        case QSE_GATE_PHASE: return "Phase";
        case QSE_GATE_RX: return "RotateX";
        case QSE_GATE_RY: return "RotateY";
        case QSE_GATE_RZ: return "RotateZ";
        case QSE_GATE_CNOT: return "CNOT";
        case QSE_GATE_CPHASE: return "CPhase";
        case QSE_GATE_MCPHASE: return "MCPhase";
        case QSE_GATE_SWAP: return "Swap";
        case QSE_GATE_MATRIX: return "Gate";
        default: return "Gate";
    }
}

}  // namespace

CircuitUnitaryOperationHIP::CircuitUnitaryOperationHIP(int device, int precision) : prec(precision) {
    check(qse_ctx_create(device, &ctx), "qse_ctx_create");
}

CircuitUnitaryOperationHIP::~CircuitUnitaryOperationHIP() {
    if (bound) bound->attach(nullptr);
    if (d_state) qse_state_destroy(ctx, d_state);
    if (ctx) qse_ctx_destroy(ctx);
}

void CircuitUnitaryOperationHIP::check(int rc, const char* what) const {
    if (rc != QSE_OK) throw std::runtime_error(std::string(what) + ": " + qse_status_string(rc) + ": " + qse_last_error());
}

// CircuitUnitaryOperationGPU.cu:158-180 semantics.
void CircuitUnitaryOperationHIP::ensureOnDevice(StateVector& sv) {
    const bool same = (bound == &sv) && d_state;
    if (same && !hostDirty) return;
    if (bound && bound != &sv && d_state) {
        flush();
        bound->attach(nullptr);
    }
    queue.clear();
    int n = 0;
    if (d_state) qse_state_info(d_state, &n, nullptr, nullptr, nullptr);
    if (d_state && (size_t)n != sv.qubits()) {
        qse_state_destroy(ctx, d_state);
        d_state = nullptr;
    }
    if (!d_state) check(qse_state_create(ctx, (int)sv.qubits(), prec, &d_state), "qse_state_create");
    if (sv.hostAllocated())
        check(qse_state_upload(ctx, d_state, reinterpret_cast<const double*>(sv.getRawPtr())), "qse_state_upload");
    else
        check(qse_state_set_basis(ctx, d_state, sv.basisIndex()), "qse_state_set_basis");
    bound = &sv;
    hostDirty = false;
    sv.attach(this);
}

void CircuitUnitaryOperationHIP::execute() {
    if (!d_state || queue.empty()) return;
    const int rc = fuse ? qse_run(ctx, d_state, queue.data(), (int)queue.size(), nullptr)
                        : qse_run_sweeps(ctx, d_state, queue.data(), (int)queue.size(), nullptr);
    queue.clear();
    check(rc, fuse ? "qse_run" : "qse_run_sweeps");
}

void CircuitUnitaryOperationHIP::enqueue(StateVector& sv, const qse_op& op, const char* label) {
    ensureOnDevice(sv);
    if (print_steps) {
        // Per-gate timing (the reference's print_steps mode,
        // CircuitUnitaryOperationGPU.cu:194-207): run this gate alone.
        execute();
        check(qse_ctx_sync(ctx), "qse_ctx_sync");
        const auto t0 = std::chrono::steady_clock::now();
        check(qse_run_sweeps(ctx, d_state, &op, 1, nullptr), label);
        check(qse_ctx_sync(ctx), "qse_ctx_sync");
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::cout << label << " applied in: " << ms << " ms (GPU)\n";
        return;
    }
    queue.push_back(op);
    if (queue.size() >= QUEUE_LIMIT) execute();
}

void CircuitUnitaryOperationHIP::applyGate(StateVector& sv, const Gate2x2& g, size_t q) {
    qse_op o = make_op(QSE_GATE_MATRIX, q);
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            o.m[2 * (2 * r + c)] = g.data[r][c].real();
            o.m[2 * (2 * r + c) + 1] = g.data[r][c].imag();
        }
    enqueue(sv, o, kind_name(QSE_GATE_MATRIX));
}
void CircuitUnitaryOperationHIP::applyCNOT(StateVector& sv, size_t c, size_t t) {
    enqueue(sv, make_op(QSE_GATE_CNOT, t, (long)c), "CNOT");
}
void CircuitUnitaryOperationHIP::applyPauliX(StateVector& sv, size_t q) { enqueue(sv, make_op(QSE_GATE_X, q), "PauliX"); }
void CircuitUnitaryOperationHIP::applyPauliY(StateVector& sv, size_t q) { enqueue(sv, make_op(QSE_GATE_Y, q), "PauliY"); }
void CircuitUnitaryOperationHIP::applyPauliZ(StateVector& sv, size_t q) { enqueue(sv, make_op(QSE_GATE_Z, q), "PauliZ"); }
void CircuitUnitaryOperationHIP::applyPhase(StateVector& sv, size_t q, double th) {
    enqueue(sv, make_op(QSE_GATE_PHASE, q, -1, th), "Phase");
}
void CircuitUnitaryOperationHIP::applyRotateX(StateVector& sv, size_t q, double th) {
    enqueue(sv, make_op(QSE_GATE_RX, q, -1, th), "RotateX");
}
void CircuitUnitaryOperationHIP::applyRotateY(StateVector& sv, size_t q, double th) {
    enqueue(sv, make_op(QSE_GATE_RY, q, -1, th), "RotateY");
}
void CircuitUnitaryOperationHIP::applyRotateZ(StateVector& sv, size_t q, double th) {
    enqueue(sv, make_op(QSE_GATE_RZ, q, -1, th), "RotateZ");
}
void CircuitUnitaryOperationHIP::applyHadamard(StateVector& sv, size_t q) {
    enqueue(sv, make_op(QSE_GATE_H, q), "Hadamard");
}
void CircuitUnitaryOperationHIP::applyCPhase(StateVector& sv, size_t c, size_t t, double th) {
    enqueue(sv, make_op(QSE_GATE_CPHASE, t, (long)c, th), "CPhase");
}
void CircuitUnitaryOperationHIP::applyMCPhase(StateVector& sv, const std::vector<size_t>& controls, size_t t,
                                              double th) {
    qse_op o = make_op(QSE_GATE_MCPHASE, t, -1, th);
    for (size_t c : controls) {
        if (c >= 64) throw std::invalid_argument("applyMCPhase: control qubit out of range");
        o.controls |= 1ULL << c;
    }
    enqueue(sv, o, "MCPhase");
}
void CircuitUnitaryOperationHIP::applySwap(StateVector& sv, size_t q0, size_t q1) {
    enqueue(sv, make_op(QSE_GATE_SWAP, q0, (long)q1), "Swap");
}

// CircuitUnitaryOperationGPU.cu:182-188: no-op while the host is dirty.
void CircuitUnitaryOperationHIP::flush() {
    if (!bound || !d_state || hostDirty) return;
    execute();
    check(qse_state_download(ctx, d_state, reinterpret_cast<double*>(bound->getRawPtr())), "qse_state_download");
}

// CircuitUnitaryOperationGPU.cu:190-192. Queued gates act on device content the
// next gate replaces by a re-upload, so they are dropped.
void CircuitUnitaryOperationHIP::markHostDirty() {
    hostDirty = true;
    queue.clear();
}

void CircuitUnitaryOperationHIP::run(StateVector& sv, const std::vector<qse_op>& ops, const std::vector<double>& params) {
    ensureOnDevice(sv);
    execute();
    check(qse_run(ctx, d_state, ops.data(), (int)ops.size(), params.empty() ? nullptr : params.data()), "qse_run");
}

std::vector<size_t> CircuitUnitaryOperationHIP::sample(StateVector& sv, const std::vector<double>& u) {
    ensureOnDevice(sv);
    execute();
    std::vector<uint64_t> out(u.size());
    if (!u.empty()) check(qse_sample(ctx, d_state, u.data(), (int64_t)u.size(), out.data()), "qse_sample");
    return std::vector<size_t>(out.begin(), out.end());
}

size_t CircuitUnitaryOperationHIP::measure(StateVector& sv, double u) {
    ensureOnDevice(sv);
    execute();
    uint64_t idx = 0;
    check(qse_measure(ctx, d_state, u, &idx), "qse_measure");
    return (size_t)idx;
}

int CircuitUnitaryOperationHIP::collapseQubit(StateVector& sv, size_t q, double u, bool reset) {
    ensureOnDevice(sv);
    execute();
    int o = 0;
    check(qse_collapse_qubit(ctx, d_state, (int)q, u, reset ? 1 : 0, &o), "qse_collapse_qubit");
    return o;
}

double CircuitUnitaryOperationHIP::expectZStrings(StateVector& sv, const std::vector<uint64_t>& masks,
                                                  const std::vector<double>& weights) {
    if (masks.size() != weights.size()) throw std::invalid_argument("expectZStrings: masks/weights size mismatch");
    ensureOnDevice(sv);
    execute();
    double v = 0.0;
    check(qse_expect_zstrings(ctx, d_state, masks.data(), weights.data(), (int)masks.size(), &v),
          "qse_expect_zstrings");
    return v;
}

double CircuitUnitaryOperationHIP::norm2(StateVector& sv) {
    ensureOnDevice(sv);
    execute();
    double v = 0.0;
    check(qse_norm2(ctx, d_state, &v), "qse_norm2");
    return v;
}

double CircuitUnitaryOperationHIP::renormalize(StateVector& sv) {
    ensureOnDevice(sv);
    execute();
    double v = 0.0;
    check(qse_state_renormalize(ctx, d_state, &v), "qse_state_renormalize");
    return v;
}

void CircuitUnitaryOperationHIP::setRenormalizePeriod(int passes) {
    check(qse_ctx_set_renormalize(ctx, passes), "qse_ctx_set_renormalize");
}

void CircuitUnitaryOperationHIP::setState(StateVector& sv, const std::vector<Complex>& amps) {
    if (amps.size() != sv.size()) throw std::invalid_argument("setState: size mismatch");
    ensureOnDevice(sv);
    queue.clear();
    check(qse_state_upload(ctx, d_state, reinterpret_cast<const double*>(amps.data())), "qse_state_upload");
}

std::vector<Complex> CircuitUnitaryOperationHIP::readState(StateVector& sv) {
    ensureOnDevice(sv);
    execute();
    std::vector<Complex> out(sv.size());
    check(qse_state_download(ctx, d_state, reinterpret_cast<double*>(out.data())), "qse_state_download");
    return out;
}

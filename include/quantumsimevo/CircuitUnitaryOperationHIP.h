// CircuitUnitaryOperationHIP.h — the MI355X backend behind the unchanged
// ICircuitUnitaryOperation interface; replaces the reference's
// CircuitUnitaryOperationGPU (QuantumSimEvo/CircuitUnitaryOperationGPU.h:10-48).
//
// Device mirror contract (CircuitUnitaryOperationGPU.cu:158-192): the first gate
// on a StateVector uploads it (or, when its host copy was never allocated,
// writes its basis state on the device); gates then run on the device only;
// flush() brings the device state back into the bound StateVector (a no-op
// while the host is marked dirty); markHostDirty() makes the next gate
// re-upload. Binding a different StateVector flushes the previous one.
//
// Gates are queued and executed in fused LDS-tiled passes (libqse_hip's
// qse_run) when the state is observed: flush(), device readout, binding
// change, or when the queue reaches its limit. With print_steps on (the
// reference's per-gate timing mode) every gate runs and is timed on its own.
// Failures of the HIP layer raise std::runtime_error (the reference aborts,
// CircuitUnitaryOperationGPU.cu:13-21).
#ifndef QSE_CIRCUIT_UNITARY_OPERATION_HIP_H
#define QSE_CIRCUIT_UNITARY_OPERATION_HIP_H
#include <cstdint>
#include <vector>

#include "../qse_hip.h"
#include "ICircuitUnitaryOperation.h"

class CircuitUnitaryOperationHIP : public ICircuitUnitaryOperation {
public:
    explicit CircuitUnitaryOperationHIP(int device = 0, int precision = QSE_C128);
    ~CircuitUnitaryOperationHIP() override;
    CircuitUnitaryOperationHIP(const CircuitUnitaryOperationHIP&) = delete;
    CircuitUnitaryOperationHIP& operator=(const CircuitUnitaryOperationHIP&) = delete;

    void applyGate(StateVector& sv, const Gate2x2& gate, size_t q) override;
    void applyCNOT(StateVector& sv, size_t control, size_t target) override;
    void applyPauliX(StateVector& sv, size_t q) override;
    void applyPauliY(StateVector& sv, size_t q) override;
    void applyPauliZ(StateVector& sv, size_t q) override;
    void applyPhase(StateVector& sv, size_t q, double theta) override;
    void applyRotateX(StateVector& sv, size_t q, double theta) override;
    void applyRotateY(StateVector& sv, size_t q, double theta) override;
    void applyRotateZ(StateVector& sv, size_t q, double theta) override;
    void applyHadamard(StateVector& sv, size_t targetQubit) override;
    void applyCPhase(StateVector& sv, size_t control, size_t target, double theta) override;
    void applyMCPhase(StateVector& sv, const std::vector<size_t>& controls, size_t target, double theta) override;
    void applySwap(StateVector& sv, size_t q0, size_t q1) override;

    void flush() override;
    void markHostDirty() override;

    // ---- device-side extensions (no counterpart in the reference) ----
    // Run the queued gates now.
    void execute();
    // Gates on a bound state, in one fused call (params for param >= 0 ops).
    void run(StateVector& sv, const std::vector<qse_op>& ops, const std::vector<double>& params = {});
    // Readout on the device with canonical uniforms (StateVector.cpp:77-156 rules).
    std::vector<size_t> sample(StateVector& sv, const std::vector<double>& uniforms);
    size_t measure(StateVector& sv, double u);
    int collapseQubit(StateVector& sv, size_t q, double u, bool reset);
    double expectZStrings(StateVector& sv, const std::vector<uint64_t>& masks, const std::vector<double>& weights);
    double norm2(StateVector& sv);
    // psi / ||psi|| in FP64 on the device; returns the norm^2 before.
    double renormalize(StateVector& sv);
    // Renormalise once gate execution has made `passes` passes over the state
    // since the last time (0: never; qse_ctx_set_renormalize).
    void setRenormalizePeriod(int passes);
    // Replace the state's content (host layout: complex<double>).
    void setState(StateVector& sv, const std::vector<Complex>& amps);
    // Copy of the device state without touching the StateVector's host buffer.
    std::vector<Complex> readState(StateVector& sv);
    // Fused passes (default) or one sweep per gate.
    void setFusion(bool fused) { fuse = fused; }
    int precision() const { return prec; }
    qse_ctx* context() const { return ctx; }
    qse_state* deviceState() const { return d_state; }

private:
    void ensureOnDevice(StateVector& sv);
    void enqueue(StateVector& sv, const qse_op& op, const char* label);
    void check(int rc, const char* what) const;

    qse_ctx* ctx = nullptr;
    qse_state* d_state = nullptr;
    StateVector* bound = nullptr;
    bool hostDirty = false;
    int prec = QSE_C128;
    bool fuse = true;
    std::vector<qse_op> queue;
};

#endif

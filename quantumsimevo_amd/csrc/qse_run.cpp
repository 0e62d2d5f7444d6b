// qse_run.cpp — gate-list execution entry points: resolves each qse_op into
// the single-gate sweeps (the replay loop of Circuit::execute,
// Circuit.cpp:73-112), or hands the list to the fusion planner.
#include <algorithm>
#include <cmath>
#include <string>

#include "qse_internal.hpp"

namespace qse {

int run_fused(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int* passes_out,
              bool dry_run, bool allow_jit);

double op_angle(const qse_op& op, const double* params);
int validate_ops(int n, const qse_op* ops, int nops, const double* params);

// One op through the single-gate sweeps.
int apply_one(qse_ctx* ctx, qse_state* st, const qse_op& o, const double* params) {
    const double th = op_angle(o, params);
    switch (o.kind) {
        case QSE_GATE_IDENTITY: return QSE_OK;
        case QSE_GATE_H: return qse_apply_hadamard(ctx, st, o.target);
        case QSE_GATE_X: return qse_apply_x(ctx, st, o.target);
        case QSE_GATE_Y: return qse_apply_y(ctx, st, o.target);
        case QSE_GATE_Z: return qse_apply_z(ctx, st, o.target);
        case QSE_GATE_PHASE: return qse_apply_phase(ctx, st, o.target, th);
        case QSE_GATE_RX: return qse_apply_rx(ctx, st, o.target, th);
        case QSE_GATE_RY: return qse_apply_ry(ctx, st, o.target, th);
        case QSE_GATE_RZ: return qse_apply_rz(ctx, st, o.target, th);
        case QSE_GATE_CNOT: return qse_apply_cnot(ctx, st, o.control, o.target);
        case QSE_GATE_CPHASE: return qse_apply_cphase(ctx, st, o.control, o.target, th);
        case QSE_GATE_MCPHASE: {
            int c[64], nc = 0;
            for (int q = 0; q < 64; ++q)
                if ((o.controls >> q) & 1) c[nc++] = q;
            return qse_apply_mcphase(ctx, st, c, nc, o.target, th);
        }
        case QSE_GATE_SWAP: return qse_apply_swap(ctx, st, o.target, o.control);
        case QSE_GATE_MATRIX: return qse_apply_matrix(ctx, st, o.target, o.m);
        default: return set_error(QSE_ERR_UNSUPPORTED, "unsupported gate kind");
    }
}

int run_sweeps(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params) {
    for (int k = 0; k < nops; ++k) {
        int rc = apply_one(ctx, st, ops[k], params);
        if (rc) return rc;
    }
    return QSE_OK;
}

// Highest qubit an op touches.
static int op_top_qubit(const qse_op& o) {
    int top = o.target;
    if ((o.kind == QSE_GATE_CNOT || o.kind == QSE_GATE_CPHASE || o.kind == QSE_GATE_SWAP) && o.control > top)
        top = o.control;
    if (o.kind == QSE_GATE_MCPHASE && o.controls) top = std::max(top, 63 - __builtin_clzll(o.controls));
    return top;
}

// Stretches of the list that go to the FP64 matrix cores: maximal runs of
// consecutive ops confined to qubits 0..m-1 (m <= 6, at least 5) with at least
// dense_per_qubit * m ops, on complex128 states of m + 4 qubits or more. Such a
// run becomes one pass with its composed 2^m x 2^m unitary (qse_apply_dense),
// which takes the same time whatever the depth (30 qubits: 6.9 ms for m = 5,
// 9.1 ms for m = 6), while a fused pass slows with the gates it carries (64
// layers on 4-6 qubits: 18-33 ms; 16 layers on 5-6 qubits, about 22 m gates:
// 9.2-10.3 ms). A run on qubits 0..3 is composed as a 5-qubit unitary: the
// m = 5 kernel is faster than the m = 4 one (6.9 against 7.3 ms), and 16
// layers on 4 qubits (88 ops) run faster fused (6.65 ms). Returns [begin, end,
// m] triples in list order.
static std::vector<int> dense_runs(const qse_ctx* ctx, const qse_state* st, const qse_op* ops, int nops) {
    std::vector<int> out;
    const int per = ctx->dense_per_qubit;
    if (per <= 0 || st->prec != QSE_C128 || st->n < 8) return out;
    int i = 0;
    while (i < nops) {
        int j = i, top = -1;
        while (j < nops && op_top_qubit(ops[j]) < 6) top = std::max(top, op_top_qubit(ops[j++]));
        const int m = std::max(5, top + 1);
        if (j - i >= per * m && st->n >= m + 4) {
            out.push_back(i);
            out.push_back(j);
            out.push_back(m);
        }
        i = std::max(j, i + 1);
    }
    return out;
}

}  // namespace qse

using namespace qse;

extern "C" {

int qse_run(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params) {
    if (!ctx || !st || !st->d) return set_error(QSE_ERR_INVALID_ARG, "null context or state");
    if (nops < 0 || (nops > 0 && !ops)) return set_error(QSE_ERR_INVALID_ARG, "bad op list");
    int rc = validate_ops(st->n, ops, nops, params);
    if (rc) return rc;
    if (nops == 0) return QSE_OK;
    int passes = 0;
    const std::vector<int> dense = dense_runs(ctx, st, ops, nops);
    if (dense.empty()) {
        rc = run_fused(ctx, st, ops, nops, params, &passes, false, true);
        if (rc) return rc;
        return renorm_tick(ctx, st, passes);
    }
    // fused stretches around the dense blocks
    int at = 0;
    std::vector<double> u;
    for (size_t r = 0; r <= dense.size(); r += 3) {
        const int b = r < dense.size() ? dense[r] : nops;
        if (b > at) {
            int p = 0;
            rc = run_fused(ctx, st, ops + at, b - at, params, &p, false, true);
            if (rc) return rc;
            passes += p;
        }
        if (r == dense.size()) break;
        const int e = dense[r + 1], m = dense[r + 2];
        u.assign((size_t)2 << (2 * m), 0.0);
        rc = qse_compose_dense(m, ops + b, e - b, params, u.data());
        if (rc) return rc;
        rc = qse_apply_dense(ctx, st, m, u.data());
        if (rc) return rc;
        ++passes;
        at = e;
    }
    return renorm_tick(ctx, st, passes);
}

int qse_ctx_set_dense_blocks(qse_ctx* ctx, int ops_per_qubit) {
    if (!ctx || ops_per_qubit < 0) return set_error(QSE_ERR_INVALID_ARG, "bad dense-block threshold");
    ctx->dense_per_qubit = ops_per_qubit;
    return QSE_OK;
}

int qse_run_repeat(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int reps) {
    if (!ctx || !st || !st->d) return set_error(QSE_ERR_INVALID_ARG, "null context or state");
    if (nops < 0 || (nops > 0 && !ops) || reps < 0) return set_error(QSE_ERR_INVALID_ARG, "bad op list or count");
    int rc = validate_ops(st->n, ops, nops, params);
    if (rc) return rc;
    if (nops == 0 || reps == 0) return QSE_OK;
    // Eager runs first: they cache the plan, place the program on the device,
    // size the scratch buffers and (under QSE_JIT_AUTO) compile the passes, so
    // the captured run makes no allocation, upload, compile or host wait.
    int done = 0, passes = 0;
    for (; done < std::min(reps, 2); ++done) {
        rc = run_fused(ctx, st, ops, nops, params, &passes, false, true);
        if (rc) return rc;
    }
    // periodic renormalisation is counted over the whole call and applied at
    // its end, so the captured runs hold no renormalisation
    if (done == reps) return renorm_tick(ctx, st, (int64_t)passes * reps);
    // Capture one run, or two when a run exchanges the state's buffers (an
    // out-of-place last pass), so that the replay leaves the pointers as the
    // capture found them.
    const bool timing = ctx->timing;
    ctx->timing = false;
    void* const d0 = st->d;
    hipGraph_t graph = nullptr;
    QSE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int per = 1;
    rc = run_fused(ctx, st, ops, nops, params, nullptr, false, true);
    if (rc == QSE_OK && st->d != d0) {
        per = 2;
        rc = run_fused(ctx, st, ops, nops, params, nullptr, false, true);
    }
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    ctx->timing = timing;
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return hip_error(ec, "hipStreamEndCapture");
    if (st->d != d0) {
        hipGraphDestroy(graph);
        return set_error(QSE_ERR_UNSUPPORTED, "run_repeat: the captured runs did not restore the state buffers");
    }
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
    for (; done + per <= reps && e == hipSuccess; done += per) e = hipGraphLaunch(exec, ctx->stream);
    hipGraphExecDestroy(exec);
    if (e != hipSuccess) return hip_error(e, "hipGraphLaunch");
    for (; done < reps; ++done) {
        rc = run_fused(ctx, st, ops, nops, params, nullptr, false, true);
        if (rc) return rc;
    }
    return renorm_tick(ctx, st, (int64_t)passes * reps);
}

int qse_run_measured(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, const double* u,
                     int nu, int32_t* results, int nresults) {
    if (!ctx || !st || !st->d) return set_error(QSE_ERR_INVALID_ARG, "null context or state");
    if (nops < 0 || (nops > 0 && !ops) || nu < 0 || (nu > 0 && !u) || nresults < 0 || (nresults > 0 && !results))
        return set_error(QSE_ERR_INVALID_ARG, "bad op list, uniforms or result slots");
    // validate everything before touching the state: unitary stretches as
    // qse_run, measurement ops by qubit, slot and uniform count
    int count = 0, begin = 0;
    for (int k = 0; k <= nops; ++k) {
        const bool meas = k < nops && (ops[k].kind == QSE_GATE_MEASURE || ops[k].kind == QSE_GATE_RESET_QUBIT);
        if (k < nops && !meas) continue;
        int rc = validate_ops(st->n, ops + begin, k - begin, params);
        if (rc) return rc;
        if (meas) {
            const qse_op& o = ops[k];
            if (o.target < 0 || o.target >= st->n)
                return set_error(QSE_ERR_INVALID_ARG, "op " + std::to_string(k) + ": bad target");
            if (o.kind == QSE_GATE_MEASURE && (o.control < 0 || o.control >= nresults))
                return set_error(QSE_ERR_INVALID_ARG, "op " + std::to_string(k) + ": result slot " +
                                                          std::to_string(o.control) + " outside [0, nresults)");
            ++count;
        }
        begin = k + 1;
    }
    if (count != nu)
        return set_error(QSE_ERR_INVALID_ARG, std::to_string(count) + " measurements but " + std::to_string(nu) +
                                                  " uniforms");
    begin = 0;
    int next = 0;
    for (int k = 0; k <= nops; ++k) {
        const bool meas = k < nops && (ops[k].kind == QSE_GATE_MEASURE || ops[k].kind == QSE_GATE_RESET_QUBIT);
        if (k < nops && !meas) continue;
        if (k > begin) {
            int rc = qse_run(ctx, st, ops + begin, k - begin, params);
            if (rc) return rc;
        }
        if (meas) {
            const qse_op& o = ops[k];
            int outcome = 0;
            int rc = qse_collapse_qubit(ctx, st, o.target, u[next++], o.kind == QSE_GATE_RESET_QUBIT, &outcome);
            if (rc) return rc;
            if (o.kind == QSE_GATE_MEASURE) results[o.control] = outcome;
        }
        begin = k + 1;
    }
    return QSE_OK;
}

int qse_run_sweeps(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params) {
    if (!ctx || !st || !st->d) return set_error(QSE_ERR_INVALID_ARG, "null context or state");
    if (nops < 0 || (nops > 0 && !ops)) return set_error(QSE_ERR_INVALID_ARG, "bad op list");
    int rc = validate_ops(st->n, ops, nops, params);
    if (rc) return rc;
    rc = run_sweeps(ctx, st, ops, nops, params);
    if (rc) return rc;
    return renorm_tick(ctx, st, nops);
}

}  // extern "C"

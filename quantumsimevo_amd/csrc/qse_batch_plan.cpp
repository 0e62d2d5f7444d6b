// qse_batch_plan.cpp — host builder of the batched small-circuit engine's
// work (qse_batch.hpp): per circuit, the fused program of its gate list, its
// observable and energy, and the adjoint reverse sweep as joint (psi, lambda)
// passes whose probes evaluate each parameter's generator term in the tile
// (build_reverse_plan, qse_plan.cpp); every program's passes are appended to
// one shared program with rebased indices.
#include <algorithm>
#include <numeric>
#include <string>
#include <thread>

#include "qse_batch.hpp"

namespace qse {

int set_error(int status, const std::string& msg);
int validate_ops(int n, const qse_op* ops, int nops, const double* params);
int build_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P, int flags);

namespace {

// Append P to G with every index rebased; returns P's pass range in G.
void append_plan(HostPlan& G, const HostPlan& P, int32_t* pass_begin, int32_t* pass_end) {
    const uint32_t st0 = (uint32_t)G.stages.size(), op0 = (uint32_t)G.ops.size(), te0 = (uint32_t)G.terms.size();
    const uint32_t sl0 = (uint32_t)G.slots.size(), et0 = (uint32_t)G.eterms.size(), al0 = (uint32_t)G.alphas.size();
    *pass_begin = (int32_t)G.passes.size();
    for (FPass fp : P.passes) {
        fp.stage_begin += st0;
        fp.slot_begin += sl0;
        G.passes.push_back(fp);
    }
    *pass_end = (int32_t)G.passes.size();
    for (FStage s : P.stages) {
        s.op_begin += op0;
        s.op_end += op0;
        G.stages.push_back(s);
    }
    for (FOp o : P.ops) {
        if (o.type == FOP_MONO) {
            o.term_begin += te0;
            o.term_end += te0;
            if (o.alpha_begin != FUSE_NO_ALPHA) o.alpha_begin += al0;
        }
        G.ops.push_back(o);
    }
    G.terms.insert(G.terms.end(), P.terms.begin(), P.terms.end());  // FTerm::slot is relative to its pass
    for (FSlot sl : P.slots) {
        sl.begin += et0;
        sl.end += et0;
        G.slots.push_back(sl);
    }
    G.eterms.insert(G.eterms.end(), P.eterms.begin(), P.eterms.end());
    G.alphas.insert(G.alphas.end(), P.alphas.begin(), P.alphas.end());
}

}  // namespace

namespace {

// One circuit's work before it is merged into the batch: its own program, its
// steps (pass ranges and value slots relative to itself) and gradient terms.
struct CircWork {
    BCirc circ{};
    HostPlan prog;
    std::vector<BStep> steps;
    std::vector<BatchGradRef> grads;  // slot relative to the circuit
    int nvals = 0;
    int max_probes = 0;
    int energy_slot = 0;
    double cost = 0.0, cost_rev = 0.0;  // estimated bytes per amplitude byte: all steps, joint steps
    int rc = QSE_OK;
};

int plan_circuit(int prec, int c, int n, const qse_op* cops, int nops, const double* cpar, int npar,
                 const uint64_t* obs_masks, int nterms, bool with_grads, int kmax, int default_L, CircWork& W) {
    for (int i = 0; i < nops; ++i)
        if (cops[i].param >= npar) return set_error(QSE_ERR_INVALID_ARG, "batch: op parameter out of range");
    int rc = validate_ops(n, cops, nops, cpar);
    if (rc) return rc;
    const int k = std::min(kmax, n);
    const int L = std::max(0, std::min(default_L, k - FUSE_REG));
    W.circ.n = n;
    W.circ.k = k;
    for (int t = 0; t < nterms; ++t)
        if ((obs_masks[t] >> n) == 0) W.circ.obs_keep |= 1ULL << t;
    BStep s{};
    s.kind = BS_BASIS;
    W.steps.push_back(s);
    int npasses = 0, njoint = 0;
    auto add_prog = [&](const std::vector<qse_op>& list, const double* pr, int target) -> int {
        if (list.empty()) return QSE_OK;
        // dependency scheduling, in place (the engine has no second buffer per state)
        HostPlan P;
        const int r = build_plan(n, k, L, list.data(), (int)list.size(), pr, P, PLAN_DAG);
        if (r) return r;
        BStep ps{};
        ps.kind = BS_PROG;
        ps.target = target;
        append_plan(W.prog, P, &ps.pass_begin, &ps.pass_end);
        npasses += (ps.pass_end - ps.pass_begin) * ((target == 3) ? 2 : 1);
        W.steps.push_back(ps);
        return QSE_OK;
    };
    rc = add_prog(std::vector<qse_op>(cops, cops + nops), cpar, 1);
    if (rc) return rc;
    s = BStep{};
    s.kind = BS_ENERGY;
    s.out = W.nvals++;
    W.energy_slot = s.out;
    W.steps.push_back(s);
    double fwd_units = 2.0 * npasses + 2.0;  // passes read and write; basis write, energy read
    if (with_grads && npar > 0) {
        s = BStep{};
        s.kind = BS_OBS;
        W.steps.push_back(s);
        fwd_units += 2.0;  // psi read, lambda written
        // reverse sweep: joint passes over (psi, lambda) with one probe per generator term
        const int kj = std::min(k, batch_joint_kmax(prec));
        const int Lj = std::max(0, std::min(default_L, kj - FUSE_REG));
        std::vector<HostPlan> chunks;
        std::vector<ProbeRef> refs;
        rc = build_reverse_plan(n, kj, Lj, cops, nops, cpar, chunks, refs);
        if (rc) return rc;
        for (const HostPlan& C : chunks) {
            BStep js{};
            js.kind = BS_JOINT;
            append_plan(W.prog, C, &js.pass_begin, &js.pass_end);
            js.nprobes = (int32_t)C.probe_ids.size();
            js.out = W.nvals;
            W.nvals += js.nprobes;
            for (int v = 0; v < js.nprobes; ++v) {
                const ProbeRef& r = refs[C.probe_ids[v]];
                W.grads.push_back({js.out + v, c, r.param, r.coeff});
            }
            njoint += js.pass_end - js.pass_begin;
            W.max_probes = std::max(W.max_probes, js.nprobes);
            W.steps.push_back(js);
        }
    }
    // estimated HBM bytes per amplitude byte; a joint pass reads and writes psi and lambda
    W.cost_rev = (double)(1ULL << n) * 4.0 * njoint;
    W.cost = (double)(1ULL << n) * fwd_units + W.cost_rev;
    return QSE_OK;
}

}  // namespace

int build_batch_plan(int prec, int ncirc, const int* nq, const int64_t* op_off, const qse_op* ops,
                     const int64_t* par_off, const double* params, const uint64_t* obs_masks, int nterms,
                     bool with_grads, int kmax, int default_L, BatchPlan& B) {
    B = BatchPlan();
    if (nterms > BATCH_MAX_TERMS) return set_error(QSE_ERR_LIMIT, "batch: at most 64 observable terms");
    for (int c = 0; c < ncirc; ++c) {
        if (nq[c] < FUSE_MIN_K || nq[c] > 30) return set_error(QSE_ERR_INVALID_ARG, "batch: circuits need 8..30 qubits");
        if (op_off[c + 1] < op_off[c]) return set_error(QSE_ERR_INVALID_ARG, "batch: op offsets must not decrease");
        if (par_off && par_off[c + 1] < par_off[c])
            return set_error(QSE_ERR_INVALID_ARG, "batch: parameter offsets must not decrease");
    }
    // circuits are planned independently on host threads, then merged in order
    std::vector<CircWork> work(ncirc);
    auto plan_into = [&](int c, CircWork& W) {
        const int64_t p0 = par_off ? par_off[c] : 0, p1 = par_off ? par_off[c + 1] : 0;
        const int npar = (int)(p1 - p0);
        return plan_circuit(prec, c, nq[c], ops + op_off[c], (int)(op_off[c + 1] - op_off[c]),
                            (params && npar > 0) ? params + p0 : nullptr, npar, obs_masks, nterms, with_grads, kmax,
                            default_L, W);
    };
    auto plan_one = [&](int c) { work[c].rc = plan_into(c, work[c]); };
    const int nthreads = std::max(1, std::min<int>(16, std::min<int>((int)std::thread::hardware_concurrency(),
                                                                      (ncirc + 7) / 8)));
    if (nthreads > 1) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t)
            pool.emplace_back([&, t] {
                for (int c = t; c < ncirc; c += nthreads) plan_one(c);
            });
        for (auto& th : pool) th.join();
    } else {
        for (int c = 0; c < ncirc; ++c) plan_one(c);
    }
    for (int c = 0; c < ncirc; ++c)
        if (work[c].rc) {
            CircWork again;  // re-plan on this thread: error messages are per thread
            return plan_into(c, again);
        }
    for (int c = 0; c < ncirc; ++c) {
        CircWork& W = work[c];
        BCirc bc = W.circ;
        bc.psi_off = B.arena_amps;
        B.arena_amps += 1ULL << bc.n;
        if (with_grads) {
            bc.lam_off = B.arena_amps;
            B.arena_amps += 1ULL << bc.n;
        }
        int32_t pb = 0, pe = 0;
        append_plan(B.prog, W.prog, &pb, &pe);
        bc.step_begin = (int32_t)B.steps.size();
        for (BStep st : W.steps) {
            if (st.kind == BS_PROG || st.kind == BS_JOINT) {
                st.pass_begin += pb;
                st.pass_end += pb;
            }
            if (st.kind == BS_ENERGY || st.kind == BS_JOINT) st.out += B.nvals;
            B.steps.push_back(st);
        }
        bc.step_end = (int32_t)B.steps.size();
        for (BatchGradRef g : W.grads) {
            g.slot += B.nvals;
            B.grads.push_back(g);
        }
        B.energy_slot.push_back(W.energy_slot + B.nvals);
        B.nvals += W.nvals;
        B.circs.push_back(bc);
        B.kmax = std::max(B.kmax, bc.k);
        B.max_probes = std::max(B.max_probes, W.max_probes);
        B.est_bytes += W.cost;
        B.est_bytes_rev += W.cost_rev;
        W.prog = HostPlan();  // merged: release as we go
    }
    B.order.resize(ncirc);
    std::iota(B.order.begin(), B.order.end(), 0);
    std::stable_sort(B.order.begin(), B.order.end(), [&](int a, int b) { return work[a].cost > work[b].cost; });
    return QSE_OK;
}

void batch_results(const BatchPlan& B, const double* vals, const int64_t* par_off, double* energies, double* grads) {
    const int ncirc = (int)B.circs.size();
    for (int c = 0; c < ncirc; ++c) energies[c] = vals[B.energy_slot[c]];
    if (!grads || !par_off) return;
    for (int c = 0; c < ncirc; ++c)
        for (int64_t j = par_off[c]; j < par_off[c + 1]; ++j) grads[j] = 0.0;
    for (const BatchGradRef& g : B.grads) grads[par_off[g.circ] + g.param] += g.coeff * vals[g.slot];
}

}  // namespace qse

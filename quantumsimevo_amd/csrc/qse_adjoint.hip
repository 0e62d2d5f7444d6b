// qse_adjoint.hip — adjoint-method gradients and batched expectation values
// (no counterpart in the reference simulator).
//
// For E(theta) = <psi|O|psi>, psi = U_L ... U_1 |0>, O a weighted sum of
// Z-strings, the reverse sweep keeps psi_l (state after gate l) and
// lambda_l = U_{l+1}^dag ... U_L^dag O psi_L and, at every parameterised gate,
// adds 2 Re<lambda_l| dU_l/dtheta |psi_{l-1}> = generator term below.
// Generators (every parameterised kind has two eigenvalues one apart):
//   RX/RY/RZ = exp(-i theta G/2):  dE/dtheta = Im<lambda|G|psi>,  G = X, Y, Z
//   CNOT(c,t) RZ(theta,t) CNOT(c,t) (the QAOA cost gate): G = Z_c Z_t
//   PHASE = diag(1, e^{i theta}):  dE/dtheta = -2 Im<lambda|P1|psi>
//                                  = -Im<lambda|psi> + Im<lambda|Z|psi>
//   CPHASE:  -2 Im<lambda|P11|psi> = -(Im<l|psi> - Im<l|Z_c|psi> - Im<l|Z_t|psi> + Im<l|Z_cZ_t|psi>)/2
// Gates are grouped, from the end, into blocks of pairwise-commuting gates;
// every generator term of a block is evaluated at the block's output (valid
// because any gate of the block can be moved to its end), with one reduction
// pass per 16 Z-string terms (X and Y generators are turned into Z-strings by
// rotating psi and lambda, qse_adjoint_plan.hpp); blocks without parameters
// are merged into the un-application that follows (the segment schedule of
// qse_adjoint_plan.cpp), which is run on psi and lambda with the fused passes.
// Terms accumulate in FP64 whatever the state precision. With ckpt_every > 0
// the forward pass stores psi every ckpt_every segments and the reverse sweep
// restarts psi from those states, which bounds the drift of un-applying gates
// in low precision.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "qse_adjoint_plan.hpp"
#include "qse_device.hpp"
#include "qse_internal.hpp"

namespace qse {

int run_fused(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int* passes_out,
              bool dry_run, bool allow_jit);
int validate_ops(int n, const qse_op* ops, int nops, const double* params);
int reduce(qse_ctx* ctx, const qse_state* st, int mode, int q, const uint64_t* masks, const double* w, int nterms,
           double* out);

static constexpr int AD_THREADS = 256;
static constexpr int DIAG_CHUNK = 16;

struct DiagMasks {
    uint64_t m[DIAG_CHUNK];
};

// lambda_i = f(i) psi_i, f(i) = sum_k w_k (-1)^popcount(i & m_k)
template <class R>
__global__ __launch_bounds__(AD_THREADS) void k_apply_obs(cx<R>* __restrict__ lam, uint64_t N,
                                                          const uint64_t* __restrict__ masks,
                                                          const double* __restrict__ w, int nterms) {
    for (uint64_t i = gtid(); i < N; i += gstride()) {
        double f = 0.0;
        for (int t = 0; t < nterms; ++t) f += (__popcll(i & masks[t]) & 1) ? -w[t] : w[t];
        const cx<R> a = lam[i];
        lam[i] = {(R)(f * (double)a.re), (R)(f * (double)a.im)};
    }
}

// partial[b][d] = sum_i (-1)^popcount(i & m_d) Im(conj(lambda_i) psi_i)
template <class R>
__global__ __launch_bounds__(AD_THREADS) void k_grad_diag(const cx<R>* __restrict__ psi, const cx<R>* __restrict__ lam,
                                                          uint64_t N, DiagMasks M, double* __restrict__ part) {
    double acc[DIAG_CHUNK];
#pragma unroll
    for (int d = 0; d < DIAG_CHUNK; ++d) acc[d] = 0.0;
    for (uint64_t i = gtid(); i < N; i += gstride()) {
        const cx<R> p = psi[i], l = lam[i];
        const double w = (double)l.re * (double)p.im - (double)l.im * (double)p.re;
#pragma unroll
        for (int d = 0; d < DIAG_CHUNK; ++d) acc[d] += (__popcll(i & M.m[d]) & 1) ? -w : w;
    }
    __shared__ double sh[AD_THREADS / 64][DIAG_CHUNK];
#pragma unroll
    for (int d = 0; d < DIAG_CHUNK; ++d) {
        const double v = wave_sum(acc[d]);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][d] = v;
    }
    __syncthreads();
    if (threadIdx.x < DIAG_CHUNK) {
        double v = 0.0;
        for (int w = 0; w < AD_THREADS / 64; ++w) v += sh[w][threadIdx.x];
        part[(uint64_t)blockIdx.x * DIAG_CHUNK + threadIdx.x] = v;
    }
}

namespace {

struct Ws {
    qse_state* psi = nullptr;
    qse_state* lam = nullptr;
    std::vector<qse_state*> ckpt;
    qse_ctx* ctx = nullptr;
    ~Ws() {
        if (psi) qse_state_destroy(ctx, psi);
        if (lam) qse_state_destroy(ctx, lam);
        for (auto* c : ckpt)
            if (c) qse_state_destroy(ctx, c);
    }
};

template <class R>
int launch_apply_obs(qse_ctx* ctx, qse_state* lam, const uint64_t* dm, const double* dw, int nterms) {
    const uint64_t N = 1ULL << lam->n;
    k_apply_obs<R><<<grid_for(ctx, N, AD_THREADS), AD_THREADS, 0, ctx->stream>>>((cx<R>*)lam->d, N, dm, dw, nterms);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

template <class R>
int grad_diag(qse_ctx* ctx, qse_state* psi, qse_state* lam, const std::vector<uint64_t>& masks, double* out) {
    const uint64_t N = 1ULL << psi->n;
    const unsigned g = std::min<unsigned>(grid_for(ctx, N, AD_THREADS, 8), 1024u);
    void* pbuf;
    int rc = scratch(ctx, 6, sizeof(double) * DIAG_CHUNK * g, &pbuf);
    if (rc) return rc;
    std::vector<double> h((size_t)DIAG_CHUNK * g);
    for (size_t c0 = 0; c0 < masks.size(); c0 += DIAG_CHUNK) {
        DiagMasks M{};
        const size_t cnt = std::min<size_t>(DIAG_CHUNK, masks.size() - c0);
        for (size_t d = 0; d < cnt; ++d) M.m[d] = masks[c0 + d];
        TimedLaunch t;
        launch_begin(ctx, FAM_ADJOINT, &t);
        k_grad_diag<R><<<g, AD_THREADS, 0, ctx->stream>>>((const cx<R>*)psi->d, (const cx<R>*)lam->d, N, M,
                                                          (double*)pbuf);
        QSE_HIP(hipGetLastError());
        launch_end(ctx, &t, 2.0 * amp_bytes(psi->prec) * (double)N);
        QSE_HIP(hipMemcpyAsync(h.data(), pbuf, sizeof(double) * DIAG_CHUNK * g, hipMemcpyDeviceToHost, ctx->stream));
        QSE_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t d = 0; d < cnt; ++d) {
            double v = 0.0;
            for (unsigned b = 0; b < g; ++b) v += h[(size_t)b * DIAG_CHUNK + d];
            out[c0 + d] = v;
        }
    }
    return QSE_OK;
}

template <class R>
int adjoint_one(qse_ctx* ctx, Ws& ws, const qse_op* ops, int nops, const double* params,
                const std::vector<AdjSegment>& segs, const std::vector<qse_op>& prefix, const uint64_t* dm,
                const double* dw, int nterms, const uint64_t* obs_m_h, const double* obs_w_h, int ckpt_every,
                double* energy, double* grad) {
    int rc = qse_state_set_basis(ctx, ws.psi, 0);
    if (rc) return rc;
    const int S = (int)segs.size();
    // forward: whole list fused, or segment by segment storing checkpoints
    std::map<int, int> ckpt_at;  // segment -> checkpoint slot holding psi at its evaluation point
    if (ckpt_every > 0 && S > 1) {
        rc = run_fused(ctx, ws.psi, prefix.data(), (int)prefix.size(), params, nullptr, false, false);
        if (rc) return rc;
        int slot = 0;
        for (int g = S - 2; g >= 0; --g) {  // forward order; the last segment's gates are in the prefix
            rc = run_fused(ctx, ws.psi, segs[g].forward.data(), (int)segs[g].forward.size(), params, nullptr,
                           false, false);
            if (rc) return rc;
            if (g > 0 && ((S - 1 - g) % ckpt_every) == 0 && slot < (int)ws.ckpt.size()) {
                rc = qse_state_copy(ctx, ws.ckpt[slot], ws.psi);
                if (rc) return rc;
                ckpt_at[g] = slot++;
            }
        }
    } else {
        rc = run_fused(ctx, ws.psi, ops, nops, params, nullptr, false, false);
        if (rc) return rc;
    }
    rc = reduce(ctx, ws.psi, 2, 0, obs_m_h, obs_w_h, nterms, energy);
    if (rc) return rc;
    rc = qse_state_copy(ctx, ws.lam, ws.psi);
    if (rc) return rc;
    rc = (ws.lam->prec == QSE_C128) ? launch_apply_obs<double>(ctx, ws.lam, dm, dw, nterms)
                                    : launch_apply_obs<float>(ctx, ws.lam, dm, dw, nterms);
    if (rc) return rc;
    if (S > 0 && !segs[0].pre.empty()) {
        // rotate into the first evaluation point's frame (X/Y generators become Z-strings)
        for (qse_state* v : {ws.psi, ws.lam}) {
            rc = run_fused(ctx, v, segs[0].pre.data(), (int)segs[0].pre.size(), nullptr, nullptr, false, false);
            if (rc) return rc;
        }
    }
    for (int g = 0; g < S; ++g) {
        const AdjSegment& seg = segs[g];
        auto ck = ckpt_at.find(g);
        if (ck != ckpt_at.end()) {
            rc = qse_state_copy(ctx, ws.psi, ws.ckpt[ck->second]);
            if (rc) return rc;
            if (!seg.pre.empty()) {
                rc = run_fused(ctx, ws.psi, seg.pre.data(), (int)seg.pre.size(), nullptr, nullptr, false, false);
                if (rc) return rc;
            }
        }
        // generator terms at this point: one pass for up to 16 Z-strings
        std::vector<uint64_t> masks;
        std::vector<std::vector<GradTerm>> mterms;
        for (const auto& dt : seg.diag) {
            auto pos = std::find(masks.begin(), masks.end(), dt.first);
            if (pos == masks.end()) {
                masks.push_back(dt.first);
                mterms.push_back({dt.second});
            } else {
                mterms[pos - masks.begin()].push_back(dt.second);
            }
        }
        if (!masks.empty()) {
            std::vector<double> vals(masks.size());
            rc = grad_diag<R>(ctx, ws.psi, ws.lam, masks, vals.data());
            if (rc) return rc;
            for (size_t i = 0; i < masks.size(); ++i)
                for (const GradTerm& t : mterms[i]) grad[t.param] += t.coeff * vals[i];
        }
        if (g + 1 < S && !seg.unapply.empty()) {
            if (ckpt_at.find(g + 1) == ckpt_at.end()) {
                rc = run_fused(ctx, ws.psi, seg.unapply.data(), (int)seg.unapply.size(), nullptr, nullptr, false,
                               false);
                if (rc) return rc;
            }
            rc = run_fused(ctx, ws.lam, seg.unapply.data(), (int)seg.unapply.size(), nullptr, nullptr, false, false);
            if (rc) return rc;
        }
    }
    return QSE_OK;
}

int check_common(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, const double* params,
                 int nparams, int batch, const uint64_t* masks, const double* w, int nterms) {
    if (!ctx || nqubits < 1 || nqubits > 40 || (precision != QSE_C128 && precision != QSE_C64) || nops < 0 ||
        (nops > 0 && !ops) || nparams < 0 || batch < 1 || nterms < 0 || (nterms > 0 && (!masks || !w)) ||
        (nparams > 0 && !params))
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    for (int k = 0; k < nops; ++k)
        if (ops[k].param >= nparams) return set_error(QSE_ERR_INVALID_ARG, "op parameter index out of range");
    return validate_ops(nqubits, ops, nops, params);
}

}  // namespace

}  // namespace qse

using namespace qse;

extern "C" {

int qse_adjoint_expect(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, const double* params,
                       int nparams, int batch, const uint64_t* obs_masks, const double* obs_weights, int nterms,
                       int ckpt_every, double* energies, double* grads) {
    int rc = check_common(ctx, nqubits, precision, ops, nops, params, nparams, batch, obs_masks, obs_weights, nterms);
    if (rc) return rc;
    if (!energies || (nparams > 0 && !grads)) return set_error(QSE_ERR_INVALID_ARG, "null output");
    // the schedule's shape does not depend on the parameter values
    std::vector<AdjSegment> segs0;
    std::vector<qse_op> prefix;
    rc = adjoint_segments(ops, nops, params, segs0, prefix);
    if (rc) return rc;
    Ws ws;
    ws.ctx = ctx;
    rc = qse_state_create(ctx, nqubits, precision, &ws.psi);
    if (rc) return rc;
    rc = qse_state_create(ctx, nqubits, precision, &ws.lam);
    if (rc) return rc;
    if (ckpt_every > 0) {
        const int nck = std::min<int>(64, ((int)segs0.size() - 1) / ckpt_every);
        ws.ckpt.assign(std::max(nck, 0), nullptr);
        for (auto& c : ws.ckpt) {
            rc = qse_state_create(ctx, nqubits, precision, &c);
            if (rc) return rc;
        }
    }
    // observable on the device
    void* obuf;
    rc = scratch(ctx, 7, 16 * (size_t)std::max(nterms, 1) + 256, &obuf);
    if (rc) return rc;
    uint64_t* dm = (uint64_t*)obuf;
    double* dw = (double*)((char*)obuf + 8 * (((size_t)std::max(nterms, 1) + 31) / 32 * 32));
    if (nterms > 0) {
        QSE_HIP(hipMemcpyAsync(dm, obs_masks, 8 * nterms, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(dw, obs_weights, 8 * nterms, hipMemcpyHostToDevice, ctx->stream));
    }
    for (int b = 0; b < batch; ++b) {
        const double* pb = params ? params + (size_t)b * nparams : nullptr;
        double* gb = grads ? grads + (size_t)b * nparams : nullptr;
        if (gb) std::fill(gb, gb + nparams, 0.0);
        std::vector<double> gtmp(std::max(nparams, 1), 0.0);
        std::vector<AdjSegment> segs;
        rc = adjoint_segments(ops, nops, pb, segs, prefix);
        if (rc) return rc;
        rc = (precision == QSE_C128)
                 ? adjoint_one<double>(ctx, ws, ops, nops, pb, segs, prefix, dm, dw, nterms, obs_masks, obs_weights,
                                       ckpt_every, &energies[b], gtmp.data())
                 : adjoint_one<float>(ctx, ws, ops, nops, pb, segs, prefix, dm, dw, nterms, obs_masks, obs_weights,
                                      ckpt_every, &energies[b], gtmp.data());
        if (rc) return rc;
        if (gb) std::copy(gtmp.begin(), gtmp.begin() + nparams, gb);
    }
    return QSE_OK;
}

int qse_batch_expect(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, const double* params,
                     int nparams, int batch, const uint64_t* obs_masks, const double* obs_weights, int nterms,
                     double* energies) {
    int rc = check_common(ctx, nqubits, precision, ops, nops, params, nparams, batch, obs_masks, obs_weights, nterms);
    if (rc) return rc;
    if (!energies) return set_error(QSE_ERR_INVALID_ARG, "null output");
    qse_state* st = nullptr;
    rc = qse_state_create(ctx, nqubits, precision, &st);
    if (rc) return rc;
    for (int b = 0; b < batch && rc == QSE_OK; ++b) {
        const double* pb = params ? params + (size_t)b * nparams : nullptr;
        rc = qse_state_set_basis(ctx, st, 0);
        if (!rc) rc = run_fused(ctx, st, ops, nops, pb, nullptr, false, false);
        if (!rc) rc = reduce(ctx, st, 2, 0, obs_masks, obs_weights, nterms, &energies[b]);
    }
    qse_state_destroy(ctx, st);
    return rc;
}

}  // extern "C"

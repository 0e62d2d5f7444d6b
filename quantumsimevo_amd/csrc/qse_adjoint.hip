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
// pass for all Z-string terms and one per X/Y term, then the block is
// un-applied on psi and lambda with the fused passes. Terms accumulate in
// FP64 whatever the state precision. With ckpt_every > 0 the forward pass
// stores psi every ckpt_every blocks and the reverse sweep restarts psi from
// those states, which bounds the drift of un-applying gates in low precision.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "qse_device.hpp"
#include "qse_internal.hpp"

namespace qse {

int run_fused(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int* passes_out,
              bool dry_run, bool allow_jit);
int validate_ops(int n, const qse_op* ops, int nops, const double* params);
double op_angle(const qse_op& op, const double* params);
bool is_zz_triple(const qse_op* ops, int nops, int k);
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

// partial[b] = Im<lambda|X_q|psi> (type 0) or Im<lambda|Y_q|psi> (type 1) over this block's pairs
template <class R>
__global__ __launch_bounds__(AD_THREADS) void k_grad_pair(const cx<R>* __restrict__ psi, const cx<R>* __restrict__ lam,
                                                          uint64_t npairs, int q, int type, double* __restrict__ part) {
    double acc = 0.0;
    for (uint64_t k = gtid(); k < npairs; k += gstride()) {
        const uint64_t i0 = ins0(k, q), i1 = i0 | (1ULL << q);
        const cx<R> p0 = psi[i0], p1 = psi[i1], l0 = lam[i0], l1 = lam[i1];
        // conj(l) * p = (lr pr + li pi) + i (lr pi - li pr)
        if (type == 0) {
            acc += (double)l0.re * p1.im - (double)l0.im * p1.re + (double)l1.re * p0.im - (double)l1.im * p0.re;
        } else {
            // (Y psi)_i0 = -i psi_i1, (Y psi)_i1 = +i psi_i0: Im(conj(l) (+-i) p) = +-Re(conj(l) p)
            acc += -((double)l0.re * p1.re + (double)l0.im * p1.im) + ((double)l1.re * p0.re + (double)l1.im * p0.im);
        }
    }
    acc = wave_sum(acc);
    __shared__ double sh[AD_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < AD_THREADS / 64; ++w) v += sh[w];
        part[blockIdx.x] = v;
    }
}

namespace {

struct GradTerm {
    int param;
    double coeff;  // d(angle)/d(param) times the generator coefficient
};

struct AdjItem {
    int first, count;       // op range [first, first + count)
    uint64_t nd = 0, dg = 0;  // non-diagonal / diagonal qubit masks
    std::vector<std::pair<uint64_t, GradTerm>> diag;  // Z-string mask -> term
    int pair_q = -1, pair_type = 0;
    GradTerm pair{-1, 0.0};
};

int build_items(int n, const qse_op* ops, int nops, std::vector<AdjItem>& items) {
    for (int k = 0; k < nops; ++k) {
        const qse_op& o = ops[k];
        AdjItem it;
        it.first = k;
        it.count = 1;
        const uint64_t tb = 1ULL << o.target;
        const double sc = o.scale;
        if (is_zz_triple(ops, nops, k)) {
            const qse_op& rz = ops[k + 1];
            it.count = 3;
            it.dg = tb | (1ULL << o.control);
            if (rz.param >= 0) it.diag.push_back({it.dg, {rz.param, rz.scale}});
            items.push_back(it);
            k += 2;
            continue;
        }
        switch (o.kind) {
            case QSE_GATE_IDENTITY: continue;
            case QSE_GATE_H: case QSE_GATE_X: case QSE_GATE_Y: case QSE_GATE_MATRIX: it.nd = tb; break;
            case QSE_GATE_RX: it.nd = tb; break;
            case QSE_GATE_RY: it.nd = tb; break;
            case QSE_GATE_Z: case QSE_GATE_PHASE: case QSE_GATE_RZ: it.dg = tb; break;
            case QSE_GATE_CNOT: it.nd = tb; it.dg = 1ULL << o.control; break;
            case QSE_GATE_CPHASE: it.dg = tb | (1ULL << o.control); break;
            case QSE_GATE_MCPHASE: it.dg = tb | o.controls; break;
            case QSE_GATE_SWAP: it.nd = tb | (1ULL << o.control); break;
            default: return set_error(QSE_ERR_UNSUPPORTED, "adjoint: unsupported gate kind");
        }
        if (o.param >= 0) {
            switch (o.kind) {
                case QSE_GATE_RX: it.pair_q = o.target; it.pair_type = 0; it.pair = {o.param, sc}; break;
                case QSE_GATE_RY: it.pair_q = o.target; it.pair_type = 1; it.pair = {o.param, sc}; break;
                case QSE_GATE_RZ: it.diag.push_back({tb, {o.param, sc}}); break;
                case QSE_GATE_PHASE:
                    it.diag.push_back({0, {o.param, -sc}});
                    it.diag.push_back({tb, {o.param, sc}});
                    break;
                case QSE_GATE_CPHASE: {
                    if (o.control == o.target) {
                        it.diag.push_back({0, {o.param, -sc}});
                        it.diag.push_back({tb, {o.param, sc}});
                    } else {
                        const uint64_t cb = 1ULL << o.control;
                        it.diag.push_back({0, {o.param, -0.5 * sc}});
                        it.diag.push_back({cb, {o.param, 0.5 * sc}});
                        it.diag.push_back({tb, {o.param, 0.5 * sc}});
                        it.diag.push_back({cb | tb, {o.param, -0.5 * sc}});
                    }
                    break;
                }
                default:
                    return set_error(QSE_ERR_UNSUPPORTED,
                                     "adjoint: parameters are supported on RX, RY, RZ, PHASE and CPHASE only");
            }
        }
        items.push_back(it);
    }
    (void)n;
    return QSE_OK;
}

// Blocks of pairwise-commuting items, listed from the end of the circuit.
void build_blocks(const std::vector<AdjItem>& items, std::vector<std::vector<int>>& blocks) {
    uint64_t ndB = 0, dgB = 0;
    for (int i = (int)items.size() - 1; i >= 0; --i) {
        const AdjItem& it = items[i];
        const bool commutes = !blocks.empty() && ((it.nd & (ndB | dgB)) == 0) && ((it.dg & ndB) == 0);
        if (!commutes) {
            blocks.emplace_back();
            ndB = dgB = 0;
        }
        blocks.back().push_back(i);
        ndB |= it.nd;
        dgB |= it.dg;
    }
}

// U^dag of op o with its angle resolved from params.
qse_op dagger(const qse_op& o, const double* params) {
    qse_op d = o;
    d.param = -1;
    d.scale = 0.0;
    switch (o.kind) {
        case QSE_GATE_PHASE: case QSE_GATE_RX: case QSE_GATE_RY: case QSE_GATE_RZ: case QSE_GATE_CPHASE:
        case QSE_GATE_MCPHASE:
            d.theta = -op_angle(o, params);
            break;
        case QSE_GATE_MATRIX:
            // conjugate transpose
            d.m[0] = o.m[0]; d.m[1] = -o.m[1];
            d.m[2] = o.m[4]; d.m[3] = -o.m[5];
            d.m[4] = o.m[2]; d.m[5] = -o.m[3];
            d.m[6] = o.m[6]; d.m[7] = -o.m[7];
            break;
        default: break;  // H, X, Y, Z, CNOT, SWAP: self-inverse
    }
    return d;
}

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
int grad_pair(qse_ctx* ctx, qse_state* psi, qse_state* lam, int q, int type, double* out) {
    const uint64_t np = (1ULL << psi->n) / 2;
    const unsigned g = std::min<unsigned>(grid_for(ctx, np, AD_THREADS, 8), 1024u);
    void* pbuf;
    int rc = scratch(ctx, 6, sizeof(double) * g, &pbuf);
    if (rc) return rc;
    TimedLaunch t;
    launch_begin(ctx, FAM_ADJOINT, &t);
    k_grad_pair<R><<<g, AD_THREADS, 0, ctx->stream>>>((const cx<R>*)psi->d, (const cx<R>*)lam->d, np, q, type,
                                                      (double*)pbuf);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, 4.0 * amp_bytes(psi->prec) * (double)np);
    std::vector<double> h(g);
    QSE_HIP(hipMemcpyAsync(h.data(), pbuf, sizeof(double) * g, hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    double v = 0.0;
    for (unsigned b = 0; b < g; ++b) v += h[b];
    *out = v;
    return QSE_OK;
}

template <class R>
int adjoint_one(qse_ctx* ctx, Ws& ws, const qse_op* ops, int nops, const double* params, int nparams,
                const std::vector<AdjItem>& items, const std::vector<std::vector<int>>& blocks,
                const uint64_t* dm, const double* dw, int nterms, const uint64_t* obs_m_h, const double* obs_w_h,
                int ckpt_every, double* energy, double* grad) {
    (void)nparams;
    int rc = qse_state_set_basis(ctx, ws.psi, 0);
    if (rc) return rc;
    // forward: whole list fused, or block by block storing checkpoints
    std::map<int, int> ckpt_at;  // block index (reverse numbering) -> checkpoint slot
    if (ckpt_every > 0) {
        const int nb = (int)blocks.size();
        std::vector<qse_op> seg;
        int slot = 0;
        for (int b = nb - 1; b >= 0; --b) {  // forward order
            seg.clear();
            for (int ii = (int)blocks[b].size() - 1; ii >= 0; --ii) {
                const AdjItem& it = items[blocks[b][ii]];
                for (int k = 0; k < it.count; ++k) seg.push_back(ops[it.first + k]);
            }
            rc = run_fused(ctx, ws.psi, seg.data(), (int)seg.size(), params, nullptr, false, false);
            if (rc) return rc;
            // checkpoint the state after this block every ckpt_every blocks
            if (b > 0 && ((nb - b) % ckpt_every) == 0 && slot < (int)ws.ckpt.size()) {
                rc = qse_state_copy(ctx, ws.ckpt[slot], ws.psi);
                if (rc) return rc;
                ckpt_at[b] = slot++;  // psi at the output of block b: the reverse sweep's input at b
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
    std::vector<qse_op> dag;
    for (int b = 0; b < (int)blocks.size(); ++b) {
        auto ck = ckpt_at.find(b);
        if (ck != ckpt_at.end()) {
            rc = qse_state_copy(ctx, ws.psi, ws.ckpt[ck->second]);
            if (rc) return rc;
        }
        // generator terms at the block output
        std::vector<uint64_t> masks;
        std::vector<std::vector<GradTerm>> mterms;
        for (int idx : blocks[b]) {
            const AdjItem& it = items[idx];
            for (auto& dt : it.diag) {
                auto pos = std::find(masks.begin(), masks.end(), dt.first);
                if (pos == masks.end()) {
                    masks.push_back(dt.first);
                    mterms.push_back({dt.second});
                } else {
                    mterms[pos - masks.begin()].push_back(dt.second);
                }
            }
            if (it.pair_q >= 0) {
                double v = 0.0;
                rc = grad_pair<R>(ctx, ws.psi, ws.lam, it.pair_q, it.pair_type, &v);
                if (rc) return rc;
                grad[it.pair.param] += it.pair.coeff * v;
            }
        }
        if (!masks.empty()) {
            std::vector<double> vals(masks.size());
            rc = grad_diag<R>(ctx, ws.psi, ws.lam, masks, vals.data());
            if (rc) return rc;
            for (size_t i = 0; i < masks.size(); ++i)
                for (const GradTerm& t : mterms[i]) grad[t.param] += t.coeff * vals[i];
        }
        // un-apply the block (items in reverse order, each item's ops reversed)
        dag.clear();
        for (int idx : blocks[b]) {
            const AdjItem& it = items[idx];
            for (int k = it.count - 1; k >= 0; --k) dag.push_back(dagger(ops[it.first + k], params));
        }
        if (b + 1 < (int)blocks.size()) {
            if (ckpt_at.find(b + 1) == ckpt_at.end()) {
                rc = run_fused(ctx, ws.psi, dag.data(), (int)dag.size(), nullptr, nullptr, false, false);
                if (rc) return rc;
            }
            rc = run_fused(ctx, ws.lam, dag.data(), (int)dag.size(), nullptr, nullptr, false, false);
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
    std::vector<AdjItem> items;
    rc = build_items(nqubits, ops, nops, items);
    if (rc) return rc;
    std::vector<std::vector<int>> blocks;
    build_blocks(items, blocks);
    Ws ws;
    ws.ctx = ctx;
    rc = qse_state_create(ctx, nqubits, precision, &ws.psi);
    if (rc) return rc;
    rc = qse_state_create(ctx, nqubits, precision, &ws.lam);
    if (rc) return rc;
    if (ckpt_every > 0) {
        const int nck = std::min<int>(64, ((int)blocks.size() - 1) / ckpt_every);
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
        rc = (precision == QSE_C128)
                 ? adjoint_one<double>(ctx, ws, ops, nops, pb, nparams, items, blocks, dm, dw, nterms, obs_masks,
                                       obs_weights, ckpt_every, &energies[b], gtmp.data())
                 : adjoint_one<float>(ctx, ws, ops, nops, pb, nparams, items, blocks, dm, dw, nterms, obs_masks,
                                      obs_weights, ckpt_every, &energies[b], gtmp.data());
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

int qse_batch_run_small(qse_ctx*, int, int, const int*, const int64_t*, const qse_op*, const int64_t*, const double*,
                        const uint64_t*, const double*, int, double*, double*) {
    return set_error(QSE_ERR_UNSUPPORTED, "batch run not built yet");
}

}  // extern "C"

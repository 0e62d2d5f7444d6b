// qse_readout.hip — device readout for gfx950: norms, sampling, measurement,
// single-qubit collapse and diagonal (Z-string) expectation values.
//
// Reference semantics (StateVector.cpp:77-156): the cumulative array of |a_i|^2
// is summed sequentially, total = cum[N-1], and a shot with canonical uniform u
// returns lower_bound(cum, u*total). A parallel device CDF rounds differently
// from the sequential one, so sampling here is two-stage:
//   1. a blocked FP64 CDF (block sums of B amplitudes, a scan over blocks, then
//      a sequential walk inside the chosen block) resolves every shot and
//      checks it against a proven rounding band: the outcome is certain when
//      the walk's values at j-1 and j sit outside the band around the target,
//      taking into account that both the device and the sequential sums carry
//      at most (index + depth) unit roundoffs of relative error;
//   2. shots inside the band are re-resolved exactly: one workgroup per frame
//      reproduces the sequential sum (same order, same rounding) once over the
//      whole state and keeps it every 64 amplitudes; a shot then finds its
//      group among these checkpoints and replays at most 64 additions.
// When the band would catch about one shot per trajectory or more (large
// states, many shots), stage 1 is skipped and every shot is resolved from the
// checkpoints. Outcomes therefore equal the sequential reference on identical
// amplitudes.
// Pauli frames (noise trajectories) are sampled the same way over the framed
// order psi[i ^ x] without materialising the framed state.
#include "qse_device.hpp"
#include "qse_internal.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

namespace qse {

static constexpr int RB_THREADS = 256;
static constexpr double UR = 2.220446049250313e-16;  // 2^-52: twice the unit roundoff

// Block sums of framed probabilities: block b of trajectory t covers framed
// indices [b*B, (b+1)*B), i.e. amplitudes psi[i ^ x_t].
template <class R>
__global__ __launch_bounds__(RB_THREADS) void k_block_sums(const cx<R>* __restrict__ s, uint64_t B, uint64_t nb,
                                                           const uint64_t* __restrict__ xmasks,
                                                           double* __restrict__ bsum) {
    const uint64_t b = blockIdx.x;
    const int t = blockIdx.y;
    const uint64_t x = xmasks ? xmasks[t] : 0;
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < B; i += RB_THREADS) acc += prob(s[(b * B + i) ^ x]);
    acc = wave_sum(acc);
    __shared__ double part[RB_THREADS / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < RB_THREADS / 64; ++w) v += part[w];
        bsum[(uint64_t)t * nb + b] = v;
    }
}

// Inclusive-start prefix: pre[t][0] = 0, pre[t][b+1] = pre[t][b] + bsum[t][b].
// One workgroup of 1024 threads per trajectory; thread chunks summed
// sequentially, chunk totals scanned in LDS.
__global__ __launch_bounds__(1024) void k_scan_blocks(const double* __restrict__ bsum, uint64_t nb,
                                                      double* __restrict__ pre) {
    const int t = blockIdx.x;
    const double* in = bsum + (uint64_t)t * nb;
    double* out = pre + (uint64_t)t * (nb + 1);
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t lo = std::min<uint64_t>(nb, threadIdx.x * per), hi = std::min<uint64_t>(nb, lo + per);
    double tot = 0.0;
    for (uint64_t i = lo; i < hi; ++i) tot += in[i];
    __shared__ double sh[1024];
    sh[threadIdx.x] = tot;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const double v = (threadIdx.x >= (unsigned)o) ? sh[threadIdx.x - o] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    // exclusive start of this chunk
    double run = (threadIdx.x == 0) ? 0.0 : sh[threadIdx.x - 1];
    if (threadIdx.x == 0) out[0] = 0.0;
    for (uint64_t i = lo; i < hi; ++i) {
        run += in[i];
        out[i + 1] = run;
    }
}

// One thread per shot: coarse binary search over block prefixes, then a
// sequential walk in the block, then the rounding-band certainty check.
template <class R>
__global__ __launch_bounds__(RB_THREADS) void k_shots(const cx<R>* __restrict__ s, uint64_t N, uint64_t B, uint64_t nb,
                                                      const uint64_t* __restrict__ xmasks,
                                                      const double* __restrict__ pre, const double* __restrict__ u,
                                                      int64_t nshots, int ntraj, uint64_t* __restrict__ out,
                                                      uint8_t* __restrict__ flag) {
    const uint64_t total_shots = (uint64_t)nshots * ntraj;
    for (uint64_t g = gtid(); g < total_shots; g += gstride()) {
        const int t = (int)(g / (uint64_t)nshots);
        const uint64_t x = xmasks ? xmasks[t] : 0;
        const double* P = pre + (uint64_t)t * (nb + 1);
        const double T = P[nb];
        const double target = u[g] * T;
        // largest b with P[b] < target (P[0] = 0), clamped to [0, nb-1]
        uint64_t lo = 0, hi = nb;  // invariant: P[lo] < target or lo == 0
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (P[mid] < target) lo = mid;
            else hi = mid;
        }
        const uint64_t b = lo;
        double prev = P[b], cur = prev;
        uint64_t j = b * B;
        bool found = false;
        // the walk's running sum, 16 amplitudes per round: the loads of a round
        // are issued together, the additions stay in index order (same rounding)
        for (uint64_t i0 = 0; i0 < B && !found; i0 += 16) {
            double pr[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) pr[c] = (i0 + c < B) ? prob(s[(b * B + i0 + c) ^ x]) : 0.0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                if (!found && i0 + c < B) {
                    cur = prev + pr[c];
                    if (cur >= target) {
                        j = b * B + i0 + c;
                        found = true;
                    } else {
                        prev = cur;
                    }
                }
            }
        }
        const double depth = (double)(B + nb + 64);
        const double delta = ((double)N + depth) * UR;
        const double t_lo = target * (1.0 - delta - 4.0 * UR);
        const double t_hi = target * (1.0 + delta + 4.0 * UR);
        bool certain = found;
        if (certain) {
            const double e_j = ((double)j + depth) * UR * cur;
            certain = (cur - e_j >= t_hi);
            if (certain && j > 0) {
                const double e_p = ((double)j + depth) * UR * prev;
                certain = (prev + e_p < t_lo);
            }
        }
        out[g] = found ? j : (b * B + B - 1);
        flag[g] = certain ? 0 : 1;
    }
}

// ---- exact sequential cumulative sums, from checkpoints ----
// One workgroup per frame f reproduces the reference's cumulative array over
// the framed order i -> psi[i ^ x_f] (StateVector.cpp:84-90: cum[0] = p_0,
// cum[i] = cum[i-1] + p_i, FP64, rounding for rounding) and stores it at the
// end of every G-amplitude group: ck[f][g] = cum[g*G + G - 1], so the last
// entry is the total. The sum is a serial chain: wave 0 adds, every lane the
// same values (lane 0 stores), while waves 1-3 compute the probabilities of
// the next chunk into the other LDS buffer.
static constexpr uint64_t CK_G = 64;    // amplitudes per checkpoint
static constexpr int CH_CHUNK = 2048;   // probabilities per LDS buffer

template <class R>
__global__ __launch_bounds__(256) void k_chain(const cx<R>* __restrict__ s, uint64_t N, uint64_t G,
                                               const uint64_t* __restrict__ xs, double* __restrict__ ck) {
    __shared__ double buf[2][CH_CHUNK];
    const uint64_t x = xs[blockIdx.x];
    const uint64_t ng = N / G;
    double* out = ck + (uint64_t)blockIdx.x * ng;
    const uint64_t nchunks = (N + CH_CHUNK - 1) / CH_CHUNK;
    const int wave = threadIdx.x >> 6;
    {
        const uint64_t len = std::min<uint64_t>(CH_CHUNK, N);
        for (uint64_t i = threadIdx.x; i < len; i += 256) buf[0][i] = prob(s[i ^ x]);
    }
    __syncthreads();
    double run = 0.0;
    for (uint64_t c = 0; c < nchunks; ++c) {
        const int b = (int)(c & 1);
        const uint64_t base = c * CH_CHUNK;
        const uint64_t len = std::min<uint64_t>(CH_CHUNK, N - base);
        if (wave == 0) {
            if (G == CK_G && len == (uint64_t)CH_CHUNK) {
                for (int g0 = 0; g0 < CH_CHUNK; g0 += (int)CK_G) {
                    double v[CK_G];
#pragma unroll
                    for (int i = 0; i < (int)CK_G; ++i) v[i] = buf[b][g0 + i];
#pragma unroll
                    for (int i = 0; i < (int)CK_G; ++i) run += v[i];
                    if (threadIdx.x == 0) out[(base + g0) / CK_G] = run;
                }
            } else {
                for (uint64_t i = 0; i < len; ++i) {
                    run += buf[b][i];
                    if ((base + i + 1) % G == 0 && threadIdx.x == 0) out[(base + i) / G] = run;
                }
            }
        } else if (c + 1 < nchunks) {
            const uint64_t nbase = base + CH_CHUNK;
            const uint64_t nlen = std::min<uint64_t>(CH_CHUNK, N - nbase);
            for (uint64_t i = threadIdx.x - 64; i < nlen; i += 192) buf[b ^ 1][i] = prob(s[(nbase + i) ^ x]);
        }
        __syncthreads();
    }
}

// Exact outcome of a shot from its frame's checkpoints: target = u * total,
// the first group whose checkpoint reaches the target (binary search), then
// the sequential sum replayed from the previous checkpoint through that group
// (the same additions in the same order as k_chain, so the same values): the
// first index with cum >= target, i.e. std::lower_bound over the sequential
// cumulative array (StateVector.cpp:94-107). Shots are all of them
// (list == nullptr, out[g]) or the listed ones (out[k] for list[k]).
template <class R>
__global__ __launch_bounds__(RB_THREADS) void k_exact_shots(const cx<R>* __restrict__ s, uint64_t N, uint64_t G,
                                                            const uint64_t* __restrict__ xs,
                                                            const int* __restrict__ fidx,
                                                            const double* __restrict__ ck,
                                                            const double* __restrict__ u, int64_t nshots,
                                                            uint64_t nlist, const int64_t* __restrict__ list,
                                                            uint64_t* __restrict__ out) {
    const uint64_t ng = N / G;
    for (uint64_t k = gtid(); k < nlist; k += gstride()) {
        const uint64_t g = list ? (uint64_t)list[k] : k;
        const int f = fidx[g / (uint64_t)nshots];
        const uint64_t x = xs[f];
        const double* C = ck + (uint64_t)f * ng;
        const double target = u[g] * C[ng - 1];
        uint64_t lo = 0, hi = ng - 1;  // C[ng-1] = total >= u * total
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (C[mid] >= target) hi = mid;
            else lo = mid + 1;
        }
        double run = lo ? C[lo - 1] : 0.0;
        const uint64_t b0 = lo * G;
        uint64_t j = b0 + G - 1;
        bool found = false;
        for (uint64_t i0 = 0; i0 < G && !found; i0 += 16) {
            double pr[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) pr[c] = (i0 + c < G) ? prob(s[(b0 + i0 + c) ^ x]) : 0.0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                if (!found && i0 + c < G) {
                    run += pr[c];
                    if (run >= target) {
                        j = b0 + i0 + c;
                        found = true;
                    }
                }
            }
        }
        out[list ? k : g] = j;
    }
}

// Checkpoint chains for frames xs_d[0..nf) into ck, then the exact outcomes
// of the listed shots (or of all `count` shots when list is null).
template <class R>
static int exact_from_chains(qse_ctx* ctx, const qse_state* st, const uint64_t* xs_d, int nf, const int* fidx_d,
                             double* ck_d, const double* u_d, int64_t nshots, uint64_t count, const int64_t* list_d,
                             uint64_t* out_d) {
    const uint64_t N = 1ULL << st->n;
    const uint64_t G = std::min<uint64_t>(CK_G, N);
    const cx<R>* s = (const cx<R>*)st->d;
    TimedLaunch tl;
    launch_begin(ctx, FAM_READOUT, &tl);
    k_chain<R><<<nf, 256, 0, ctx->stream>>>(s, N, G, xs_d, ck_d);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &tl, (double)amp_bytes(st->prec) * N * nf);
    TimedLaunch t2;
    launch_begin(ctx, FAM_READOUT, &t2);
    k_exact_shots<R><<<grid_for(ctx, count, RB_THREADS, 8), RB_THREADS, 0, ctx->stream>>>(
        s, N, G, xs_d, fidx_d, ck_d, u_d, nshots, count, list_d, out_d);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t2, (double)amp_bytes(st->prec) * G * count);
    return QSE_OK;
}

// Sampling driver. Frames are grouped (trajectories with equal X masks share
// one framed order). When the blocked CDF would leave about one shot per
// trajectory or more inside its rounding band (shots * 2 N (N + depth) * u'
// >= 1, u' = 2^-52), every shot is resolved exactly from checkpoint chains of
// all frames at once; otherwise the blocked CDF resolves the shots and chains
// are built only for the frames of the flagged ones.
template <class R>
static int sample_frames_t(qse_ctx* ctx, qse_state* st, const uint64_t* xmask_h, int ntraj, const double* u_h,
                           int64_t nshots, uint64_t* out_h) {
    const uint64_t N = 1ULL << st->n;
    const uint64_t B = std::min<uint64_t>(1024, N);
    const uint64_t nb = N / B;
    const uint64_t G = std::min<uint64_t>(CK_G, N);
    const uint64_t ng = N / G;
    const uint64_t total = (uint64_t)nshots * ntraj;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    // distinct frames
    std::vector<uint64_t> xs;
    std::vector<int> fidx(ntraj, 0);
    if (xmask_h) {
        xs.assign(xmask_h, xmask_h + ntraj);
        std::sort(xs.begin(), xs.end());
        xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
        for (int t = 0; t < ntraj; ++t)
            fidx[t] = (int)(std::lower_bound(xs.begin(), xs.end(), xmask_h[t]) - xs.begin());
    } else {
        xs.assign(1, 0);
    }
    const int F = (int)xs.size();
    const double depth = (double)(B + nb + 64);
    const double flag_rate = 2.0 * ((double)N + depth) * UR * (double)N;
    const bool direct = (double)nshots * flag_rate >= 1.0 && (double)F * ng * 8.0 <= (double)(1ULL << 30);
    const cx<R>* s = (const cx<R>*)st->d;
    const size_t sz_u = sizeof(double) * (size_t)total;
    const size_t sz_o = sizeof(uint64_t) * (size_t)total;
    if (direct) {
        // scratch layout (slot 0): xs | fidx | checkpoints | u | out
        const size_t sz_x = sizeof(uint64_t) * F, sz_fi = sizeof(int) * ntraj, sz_ck = sizeof(double) * F * ng;
        void* base;
        int rc = scratch(ctx, 0, al(sz_x) + al(sz_fi) + al(sz_ck) + al(sz_u) + al(sz_o), &base);
        if (rc) return rc;
        char* p = (char*)base;
        uint64_t* d_x = (uint64_t*)p; p += al(sz_x);
        int* d_fi = (int*)p; p += al(sz_fi);
        double* d_ck = (double*)p; p += al(sz_ck);
        double* d_u = (double*)p; p += al(sz_u);
        uint64_t* d_o = (uint64_t*)p;
        QSE_HIP(hipMemcpyAsync(d_x, xs.data(), sz_x, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(d_fi, fidx.data(), sz_fi, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(d_u, u_h, sz_u, hipMemcpyHostToDevice, ctx->stream));
        rc = exact_from_chains<R>(ctx, st, d_x, F, d_fi, d_ck, d_u, nshots, total, nullptr, d_o);
        if (rc) return rc;
        QSE_HIP(hipMemcpyAsync(out_h, d_o, sz_o, hipMemcpyDeviceToHost, ctx->stream));
        QSE_HIP(hipStreamSynchronize(ctx->stream));
        return QSE_OK;
    }
    // blocked CDF; scratch layout (slot 0): xmasks | bsum | pre | u | out | flags
    const size_t sz_x = sizeof(uint64_t) * (size_t)ntraj;
    const size_t sz_b = sizeof(double) * (size_t)(nb * ntraj);
    const size_t sz_p = sizeof(double) * (size_t)((nb + 1) * ntraj);
    const size_t sz_f = (size_t)total;
    const size_t need = al(sz_x) + al(sz_b) + al(sz_p) + al(sz_u) + al(sz_o) + al(sz_f);
    void* base;
    int rc = scratch(ctx, 0, need, &base);
    if (rc) return rc;
    char* p = (char*)base;
    uint64_t* d_x = (uint64_t*)p; p += al(sz_x);
    double* d_b = (double*)p; p += al(sz_b);
    double* d_p = (double*)p; p += al(sz_p);
    double* d_u = (double*)p; p += al(sz_u);
    uint64_t* d_o = (uint64_t*)p; p += al(sz_o);
    uint8_t* d_f = (uint8_t*)p;
    const bool framed = (xmask_h != nullptr);
    if (framed) QSE_HIP(hipMemcpyAsync(d_x, xmask_h, sz_x, hipMemcpyHostToDevice, ctx->stream));
    QSE_HIP(hipMemcpyAsync(d_u, u_h, sz_u, hipMemcpyHostToDevice, ctx->stream));
    TimedLaunch tl;
    launch_begin(ctx, FAM_READOUT, &tl);
    k_block_sums<R><<<dim3((unsigned)nb, (unsigned)ntraj), RB_THREADS, 0, ctx->stream>>>(s, B, nb, framed ? d_x : nullptr,
                                                                                          d_b);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &tl, (double)amp_bytes(st->prec) * N * ntraj);
    k_scan_blocks<<<ntraj, 1024, 0, ctx->stream>>>(d_b, nb, d_p);
    QSE_HIP(hipGetLastError());
    if (total > 0) {
        k_shots<R><<<grid_for(ctx, total, RB_THREADS, 8), RB_THREADS, 0, ctx->stream>>>(
            s, N, B, nb, framed ? d_x : nullptr, d_p, d_u, nshots, ntraj, d_o, d_f);
        QSE_HIP(hipGetLastError());
    }
    std::vector<uint8_t> flags(total);
    if (total > 0) {
        QSE_HIP(hipMemcpyAsync(out_h, d_o, sz_o, hipMemcpyDeviceToHost, ctx->stream));
        QSE_HIP(hipMemcpyAsync(flags.data(), d_f, sz_f, hipMemcpyDeviceToHost, ctx->stream));
    }
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    // flagged shots, and the frames they need (compact frame numbering)
    std::vector<int64_t> list;
    std::vector<int> need_f(F, -1), fidx2(ntraj, 0);
    std::vector<uint64_t> xs2;
    for (uint64_t g = 0; g < total; ++g)
        if (flags[g]) {
            list.push_back((int64_t)g);
            const int f = fidx[g / (uint64_t)nshots];
            if (need_f[f] < 0) {
                need_f[f] = (int)xs2.size();
                xs2.push_back(xs[f]);
            }
        }
    if (list.empty()) return QSE_OK;
    for (int t = 0; t < ntraj; ++t) fidx2[t] = std::max(need_f[fidx[t]], 0);
    const int F2 = (int)xs2.size();
    const uint64_t nl = list.size();
    const size_t s_x = sizeof(uint64_t) * F2, s_fi = sizeof(int) * ntraj, s_ck = sizeof(double) * F2 * ng,
                 s_l = sizeof(int64_t) * nl, s_o = sizeof(uint64_t) * nl;
    void* base2;
    rc = scratch(ctx, 1, al(s_x) + al(s_fi) + al(s_ck) + al(s_l) + al(s_o), &base2);
    if (rc) return rc;
    char* q = (char*)base2;
    uint64_t* d_x2 = (uint64_t*)q; q += al(s_x);
    int* d_fi2 = (int*)q; q += al(s_fi);
    double* d_ck = (double*)q; q += al(s_ck);
    int64_t* d_l = (int64_t*)q; q += al(s_l);
    uint64_t* d_o2 = (uint64_t*)q;
    QSE_HIP(hipMemcpyAsync(d_x2, xs2.data(), s_x, hipMemcpyHostToDevice, ctx->stream));
    QSE_HIP(hipMemcpyAsync(d_fi2, fidx2.data(), s_fi, hipMemcpyHostToDevice, ctx->stream));
    QSE_HIP(hipMemcpyAsync(d_l, list.data(), s_l, hipMemcpyHostToDevice, ctx->stream));
    rc = exact_from_chains<R>(ctx, st, d_x2, F2, d_fi2, d_ck, d_u, nshots, nl, d_l, d_o2);
    if (rc) return rc;
    std::vector<uint64_t> res(nl);
    QSE_HIP(hipMemcpyAsync(res.data(), d_o2, s_o, hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t i = 0; i < nl; ++i) out_h[list[i]] = res[i];
    return QSE_OK;
}

int sample_frames(qse_ctx* ctx, qse_state* st, const uint64_t* xmask, int ntraj, const double* u, int64_t nshots,
                  uint64_t* out) {
    if (st->prec == QSE_C128) return sample_frames_t<double>(ctx, st, xmask, ntraj, u, nshots, out);
    return sample_frames_t<float>(ctx, st, xmask, ntraj, u, nshots, out);
}

// ---- reductions: sum over i of |a_i|^2 * f(i) ----
// f(i) = 1 (norm), [bit q set] (prob of qubit q), or a weighted Z-string sum.
enum RedMode { RED_NORM = 0, RED_BIT = 1, RED_ZSTR = 2 };

template <class R, int MODE>
__global__ __launch_bounds__(RB_THREADS) void k_reduce(const cx<R>* __restrict__ s, uint64_t nk, int q,
                                                       const uint64_t* __restrict__ masks,
                                                       const double* __restrict__ w, int nterms,
                                                       double* __restrict__ part) {
    double acc = 0.0;
    for (uint64_t k = gtid(); k < nk; k += gstride()) {
        if constexpr (MODE == RED_NORM) {
            acc += prob(s[k]);
        } else if constexpr (MODE == RED_BIT) {
            const uint64_t i = ins0(k, q) | (1ULL << q);
            acc += prob(s[i]);
        } else {
            double f = 0.0;
            for (int t = 0; t < nterms; ++t) f += (__popcll(k & masks[t]) & 1) ? -w[t] : w[t];
            acc += prob(s[k]) * f;
        }
    }
    acc = wave_sum(acc);
    __shared__ double sh[RB_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int i = 0; i < RB_THREADS / 64; ++i) v += sh[i];
        part[blockIdx.x] = v;
    }
}

// Deterministic final sum of partials by one thread (fixed order).
__global__ void k_sum_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double v = 0.0;
        for (int i = 0; i < n; ++i) v += part[i];
        *out = v;
    }
}

template <class R>
static int reduce_t(qse_ctx* ctx, const qse_state* st, int mode, int q, const uint64_t* masks_h, const double* w_h,
                    int nterms, double* out_h) {
    const uint64_t N = 1ULL << st->n;
    const uint64_t nk = (mode == RED_BIT) ? N / 2 : N;
    const unsigned g = std::min<unsigned>(grid_for(ctx, nk, RB_THREADS, 8), 2048u);
    const size_t szm = sizeof(uint64_t) * std::max(nterms, 1), szw = sizeof(double) * std::max(nterms, 1);
    void* base;
    int rc = scratch(ctx, 2, 256 + szm + szw + sizeof(double) * (g + 1) + 256, &base);
    if (rc) return rc;
    char* p = (char*)base;
    uint64_t* d_m = (uint64_t*)p; p += (szm + 255) & ~(size_t)255;
    double* d_w = (double*)p; p += (szw + 255) & ~(size_t)255;
    double* d_part = (double*)p;
    double* d_out = d_part + g;
    if (mode == RED_ZSTR && nterms > 0) {
        QSE_HIP(hipMemcpyAsync(d_m, masks_h, sizeof(uint64_t) * nterms, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(d_w, w_h, sizeof(double) * nterms, hipMemcpyHostToDevice, ctx->stream));
    }
    const cx<R>* s = (const cx<R>*)st->d;
    TimedLaunch t;
    launch_begin(ctx, FAM_READOUT, &t);
    if (mode == RED_NORM) k_reduce<R, RED_NORM><<<g, RB_THREADS, 0, ctx->stream>>>(s, nk, q, d_m, d_w, nterms, d_part);
    else if (mode == RED_BIT) k_reduce<R, RED_BIT><<<g, RB_THREADS, 0, ctx->stream>>>(s, nk, q, d_m, d_w, nterms, d_part);
    else k_reduce<R, RED_ZSTR><<<g, RB_THREADS, 0, ctx->stream>>>(s, nk, q, d_m, d_w, nterms, d_part);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, (double)amp_bytes(st->prec) * nk);
    k_sum_partials<<<1, 64, 0, ctx->stream>>>(d_part, (int)g, d_out);
    QSE_HIP(hipGetLastError());
    void* pin;
    rc = pinned(ctx, sizeof(double), &pin);
    if (rc) return rc;
    QSE_HIP(hipMemcpyAsync(pin, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    *out_h = *(double*)pin;
    return QSE_OK;
}

int reduce(qse_ctx* ctx, const qse_state* st, int mode, int q, const uint64_t* masks, const double* w, int nterms,
           double* out) {
    if (st->prec == QSE_C128) return reduce_t<double>(ctx, st, mode, q, masks, w, nterms, out);
    return reduce_t<float>(ctx, st, mode, q, masks, w, nterms, out);
}

}  // namespace qse

// ---- allocation-free expectation values (qse_expect_zstrings_noalloc) ----
// The scratch memory of that entry point: masks (64 x 8 B) | weights
// (64 x 8 B) | per-workgroup partial sums (at most 2048 x 8 B) | result. One
// copy of the array lives on every GPU the library runs on, until the process
// ends. In host code its name is only a stand-in, not a device address:
// noalloc_init() looks the address up once per device (with that device
// current) and host code uses that address only.
static constexpr int NOALLOC_MAX_TERMS = 64;
static constexpr unsigned NOALLOC_MAX_PARTS = 2048;
static constexpr size_t NOALLOC_BYTES = 32768;
static_assert(8 * (2 * NOALLOC_MAX_TERMS + NOALLOC_MAX_PARTS + 1) <= NOALLOC_BYTES, "scratch layout");
static __device__ __align__(256) unsigned char noalloc_scratch[NOALLOC_BYTES];

namespace qse {

static constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_noalloc_mu;

// Called by qse_ctx_create with `device` current.
int noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return set_error(QSE_ERR_LIMIT, "device index above the library's per-device table");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return QSE_OK;
    void* p = nullptr;
    QSE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    if (!p) return set_error(QSE_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_noalloc_ptr[device] = p;
    return QSE_OK;
}

// reduce(RED_ZSTR) without allocating: the same kernels, grid and summation
// order (so bit-identical results), with masks, weights, partial sums and the
// result in the device's scratch array instead of the context's grown scratch
// and pinned buffers, and without the timing hooks (which grow host lists).
int reduce_zstr_noalloc(qse_ctx* ctx, const qse_state* st, const uint64_t* masks_h, const double* w_h, int nterms,
                        double* out_h) {
    if (nterms > NOALLOC_MAX_TERMS)
        return set_error(QSE_ERR_LIMIT, "qse_expect_zstrings_noalloc: at most 64 Z-string terms (got " +
                                            std::to_string(nterms) + ")");
    void* base = (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) ? g_noalloc_ptr[ctx->device] : nullptr;
    if (!base) return set_error(QSE_ERR_HIP, "scratch array of this device was not set up at context creation");
    const uint64_t N = 1ULL << st->n;
    const unsigned g = std::min<unsigned>(grid_for(ctx, N, RB_THREADS, 8), NOALLOC_MAX_PARTS);
    uint64_t* d_m = (uint64_t*)base;
    double* d_w = (double*)(d_m + NOALLOC_MAX_TERMS);
    double* d_part = d_w + NOALLOC_MAX_TERMS;
    double* d_out = d_part + NOALLOC_MAX_PARTS;
    if (nterms > 0) {
        QSE_HIP(hipMemcpyAsync(d_m, masks_h, sizeof(uint64_t) * nterms, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(d_w, w_h, sizeof(double) * nterms, hipMemcpyHostToDevice, ctx->stream));
    }
    if (st->prec == QSE_C128)
        k_reduce<double, RED_ZSTR><<<g, RB_THREADS, 0, ctx->stream>>>((const cx<double>*)st->d, N, 0, d_m, d_w,
                                                                       nterms, d_part);
    else
        k_reduce<float, RED_ZSTR><<<g, RB_THREADS, 0, ctx->stream>>>((const cx<float>*)st->d, N, 0, d_m, d_w,
                                                                      nterms, d_part);
    QSE_HIP(hipGetLastError());
    k_sum_partials<<<1, 64, 0, ctx->stream>>>(d_part, (int)g, d_out);
    QSE_HIP(hipGetLastError());
    QSE_HIP(hipMemcpyAsync(out_h, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    return QSE_OK;
}

// sum_k w_k <Z_{m_k}> with masks and weights already on the device, written to
// d_out on the device (stream-ordered, no host wait): the same partial sums
// and final order as reduce().
int reduce_zstr_async(qse_ctx* ctx, const qse_state* st, const uint64_t* d_m, const double* d_w, int nterms,
                      double* d_out) {
    const uint64_t N = 1ULL << st->n;
    const unsigned g = std::min<unsigned>(grid_for(ctx, N, RB_THREADS, 8), 2048u);
    void* part;
    int rc = scratch(ctx, 9, sizeof(double) * g, &part);
    if (rc) return rc;
    TimedLaunch t;
    launch_begin(ctx, FAM_READOUT, &t);
    if (st->prec == QSE_C128)
        k_reduce<double, RED_ZSTR><<<g, RB_THREADS, 0, ctx->stream>>>((const cx<double>*)st->d, N, 0, d_m, d_w,
                                                                       nterms, (double*)part);
    else
        k_reduce<float, RED_ZSTR><<<g, RB_THREADS, 0, ctx->stream>>>((const cx<float>*)st->d, N, 0, d_m, d_w,
                                                                      nterms, (double*)part);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, (double)amp_bytes(st->prec) * N);
    k_sum_partials<<<1, 64, 0, ctx->stream>>>((const double*)part, (int)g, d_out);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

// FP64 renormalisation on the device: norm^2 = sum |a_i|^2 with the partial
// sums and summation order of reduce(RED_NORM), then every amplitude times
// 1/sqrt(norm^2), computed in FP64 from the device copy of norm^2 (no host
// round trip). A state whose norm^2 is zero or not finite is left unchanged.
template <class R>
__global__ __launch_bounds__(RB_THREADS) void k_scale_norm(cx<R>* __restrict__ s, uint64_t n,
                                                           const double* __restrict__ nrm2) {
    const double v = *nrm2;
    const double f = (v > 0.0 && v <= 1.7976931348623157e308) ? 1.0 / sqrt(v) : 1.0;
    for (uint64_t k = gtid(); k < n; k += gstride()) {
        const cx<R> a = s[k];
        s[k] = {(R)((double)a.re * f), (R)((double)a.im * f)};
    }
}

int renormalize_async(qse_ctx* ctx, qse_state* st, double** d_nrm2) {
    const uint64_t N = 1ULL << st->n;
    const unsigned g = std::min<unsigned>(grid_for(ctx, N, RB_THREADS, 8), 2048u);
    void* base;
    int rc = scratch(ctx, 10, sizeof(double) * (g + 1), &base);
    if (rc) return rc;
    double* d_part = (double*)base;
    double* d_out = d_part + g;
    TimedLaunch t;
    launch_begin(ctx, FAM_READOUT, &t);
    if (st->prec == QSE_C128) {
        k_reduce<double, RED_NORM><<<g, RB_THREADS, 0, ctx->stream>>>((const cx<double>*)st->d, N, 0, nullptr,
                                                                       nullptr, 0, d_part);
    } else {
        k_reduce<float, RED_NORM><<<g, RB_THREADS, 0, ctx->stream>>>((const cx<float>*)st->d, N, 0, nullptr,
                                                                      nullptr, 0, d_part);
    }
    QSE_HIP(hipGetLastError());
    k_sum_partials<<<1, 64, 0, ctx->stream>>>(d_part, (int)g, d_out);
    QSE_HIP(hipGetLastError());
    const unsigned gs = grid_for(ctx, N, RB_THREADS);
    if (st->prec == QSE_C128)
        k_scale_norm<double><<<gs, RB_THREADS, 0, ctx->stream>>>((cx<double>*)st->d, N, d_out);
    else
        k_scale_norm<float><<<gs, RB_THREADS, 0, ctx->stream>>>((cx<float>*)st->d, N, d_out);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, 3.0 * (double)amp_bytes(st->prec) * N);
    if (d_nrm2) *d_nrm2 = d_out;
    return QSE_OK;
}

// Periodic renormalisation of qse_run / qse_run_repeat results: count the
// passes a call made over the state and renormalise once the count reaches
// the context's period.
int renorm_tick(qse_ctx* ctx, qse_state* st, int64_t passes) {
    if (ctx->renorm_every <= 0) return QSE_OK;
    st->passes_since_renorm += passes;
    if (st->passes_since_renorm < ctx->renorm_every) return QSE_OK;
    st->passes_since_renorm = 0;
    return renormalize_async(ctx, st, nullptr);
}

// Single-qubit collapse (StateVector.cpp:158-221): keep the `outcome` half
// scaled by 1/nf; with reset, move it to the |0> half and zero the |1> half.
template <class R>
__global__ __launch_bounds__(RB_THREADS) void k_collapse(cx<R>* __restrict__ s, uint64_t nk, int q, int outcome,
                                                         int reset, double nf) {
    const uint64_t bit = 1ULL << q;
    for (uint64_t k = gtid(); k < nk; k += gstride()) {
        const uint64_t i0 = ins0(k, q), i1 = i0 | bit;
        const cx<R> a0 = s[i0], a1 = s[i1];
        const cx<R> z = {(R)0, (R)0};
        if (!reset) {
            if (outcome) {
                s[i0] = z;
                s[i1] = {(R)(a1.re / nf), (R)(a1.im / nf)};
            } else {
                s[i0] = {(R)(a0.re / nf), (R)(a0.im / nf)};
                s[i1] = z;
            }
        } else {
            const cx<R> keep = outcome ? a1 : a0;
            s[i0] = {(R)(keep.re / nf), (R)(keep.im / nf)};
            s[i1] = z;
        }
    }
}

int launch_collapse(qse_ctx* ctx, qse_state* st, int q, int outcome, int reset, double nf) {
    const uint64_t nk = (1ULL << st->n) / 2;
    const unsigned g = grid_for(ctx, nk, RB_THREADS);
    if (st->prec == QSE_C128)
        k_collapse<double><<<g, RB_THREADS, 0, ctx->stream>>>((cx<double>*)st->d, nk, q, outcome, reset, nf);
    else
        k_collapse<float><<<g, RB_THREADS, 0, ctx->stream>>>((cx<float>*)st->d, nk, q, outcome, reset, nf);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

}  // namespace qse

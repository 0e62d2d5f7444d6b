// qse_gates.hip — one-gate amplitude sweeps for gfx950.
//
// Every gate of the reference's backend interface (ICircuitUnitaryOperation.h:19-31)
// maps onto one index-compressed pair sweep: a compressed index k is expanded by
// inserting the gate's "fixed" bits (target bit = 0, control bits = 1), so no
// work-item idles on amplitudes the gate does not change (the reference's CUDA
// kernels launch N threads with 1/2 to 3/4 idle for Z, phase, CNOT and CPhase).
// Consecutive k map to consecutive amplitudes except across the inserted bits,
// so each wavefront reads and writes whole contiguous runs.
//
// Algorithmic traffic per gate (read + write of each touched amplitude,
// b = bytes/amplitude): 2bN for dense 1q gates and RZ, bN for X-type swaps of
// N/2 amplitudes... see launch sites in qse_capi.cpp for the exact figures.
#include "qse_device.hpp"
#include "qse_internal.hpp"

namespace qse {

template <class R>
struct PairCoef {
    cx<R> m00, m01, m10, m11;  // dense
};

template <class R, int MODE>
__global__ __launch_bounds__(256) void k_pairs(cx<R>* __restrict__ s, uint64_t nk, Deposit dep, uint64_t off0,
                                               uint64_t off1, PairCoef<R> c) {
    const uint64_t stride = gstride();
    for (uint64_t k = gtid(); k < nk; k += stride) {
        const uint64_t base = deposit(k, dep);
        if constexpr (MODE == PAIR_DIAG1) {
            const uint64_t i1 = base | off1;
            s[i1] = cmul(c.m11, s[i1]);
        } else {
            const uint64_t i0 = base | off0, i1 = base | off1;
            const cx<R> v0 = s[i0], v1 = s[i1];
            if constexpr (MODE == PAIR_DENSE) {
                s[i0] = cadd(cmul(c.m00, v0), cmul(c.m01, v1));
                s[i1] = cadd(cmul(c.m10, v0), cmul(c.m11, v1));
            } else if constexpr (MODE == PAIR_SWAP) {
                s[i0] = v1;
                s[i1] = v0;
            } else if constexpr (MODE == PAIR_DIAG) {
                s[i0] = cmul(c.m00, v0);
                s[i1] = cmul(c.m11, v1);
            } else {  // PAIR_HADAMARD: h*(v0+v1), h*(v0-v1) (CircuitUnitaryOperation.cpp:101-102)
                const R h = c.m00.re;
                s[i0] = {h * (v0.re + v1.re), h * (v0.im + v1.im)};
                s[i1] = {h * (v0.re - v1.re), h * (v0.im - v1.im)};
            }
        }
    }
}

template <class R>
static int launch_pairs_t(qse_ctx* ctx, qse_state* st, int mode, const Deposit& dep, uint64_t nk, uint64_t off0,
                          uint64_t off1, const double m[8], double algo_bytes) {
    PairCoef<R> c;
    c.m00 = {(R)m[0], (R)m[1]};
    c.m01 = {(R)m[2], (R)m[3]};
    c.m10 = {(R)m[4], (R)m[5]};
    c.m11 = {(R)m[6], (R)m[7]};
    cx<R>* s = (cx<R>*)st->d;
    const unsigned g = grid_for(ctx, nk, 256);
    TimedLaunch t;
    launch_begin(ctx, FAM_SWEEP, &t);
    switch (mode) {
        case PAIR_DENSE: k_pairs<R, PAIR_DENSE><<<g, 256, 0, ctx->stream>>>(s, nk, dep, off0, off1, c); break;
        case PAIR_SWAP: k_pairs<R, PAIR_SWAP><<<g, 256, 0, ctx->stream>>>(s, nk, dep, off0, off1, c); break;
        case PAIR_DIAG: k_pairs<R, PAIR_DIAG><<<g, 256, 0, ctx->stream>>>(s, nk, dep, off0, off1, c); break;
        case PAIR_DIAG1: k_pairs<R, PAIR_DIAG1><<<g, 256, 0, ctx->stream>>>(s, nk, dep, off0, off1, c); break;
        case PAIR_HADAMARD: k_pairs<R, PAIR_HADAMARD><<<g, 256, 0, ctx->stream>>>(s, nk, dep, off0, off1, c); break;
        default: return set_error(QSE_ERR_INVALID_ARG, "bad pair mode");
    }
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, algo_bytes);
    return QSE_OK;
}

int launch_pairs(qse_ctx* ctx, qse_state* st, int mode, const uint64_t* fixed0, int nfixed0, uint64_t ones,
                 uint64_t off0, uint64_t off1, const double m[8], double algo_bytes) {
    // Fixed bits: zero bits listed in fixed0 plus the one bits of `ones`; all of
    // them are inserted (ascending) into the compressed index.
    uint64_t fixed = ones;
    for (int i = 0; i < nfixed0; ++i) fixed |= fixed0[i];
    Deposit dep;
    dep.nb = 0;
    for (int b = 0; b < 64; ++b)
        if ((fixed >> b) & 1) dep.pos[dep.nb++] = (uint8_t)b;
    dep.ones = ones;
    if (dep.nb > st->n) return set_error(QSE_ERR_INVALID_ARG, "fixed bits exceed qubit count");
    const uint64_t nk = 1ULL << (st->n - dep.nb);
    if (st->prec == QSE_C128) return launch_pairs_t<double>(ctx, st, mode, dep, nk, off0, off1, m, algo_bytes);
    return launch_pairs_t<float>(ctx, st, mode, dep, nk, off0, off1, m, algo_bytes);
}

// Pauli frame psi'[i] = (-1)^popcount(i & z) psi[i ^ x] (Noise.cpp:34-44 as one pass).
template <class R>
__global__ __launch_bounds__(256) void k_frame(cx<R>* __restrict__ s, uint64_t nk, int h, uint64_t x, uint64_t z) {
    const uint64_t stride = gstride();
    if (x != 0) {
        for (uint64_t k = gtid(); k < nk; k += stride) {
            const uint64_t i = ins0(k, h), j = i ^ x;
            const cx<R> a = s[i], b = s[j];
            s[i] = (__popcll(i & z) & 1) ? cneg(b) : b;
            s[j] = (__popcll(j & z) & 1) ? cneg(a) : a;
        }
    } else {
        for (uint64_t k = gtid(); k < nk; k += stride)
            if (__popcll(k & z) & 1) s[k] = cneg(s[k]);
    }
}

int launch_frame(qse_ctx* ctx, qse_state* st, uint64_t xmask, uint64_t zmask) {
    const uint64_t N = 1ULL << st->n;
    int h = 63;
    while (h >= 0 && !((xmask >> h) & 1)) --h;
    const uint64_t nk = xmask ? N / 2 : N;
    const unsigned g = grid_for(ctx, nk, 256);
    const double b = (double)amp_bytes(st->prec);
    TimedLaunch t;
    launch_begin(ctx, FAM_SWEEP, &t);
    if (st->prec == QSE_C128)
        k_frame<double><<<g, 256, 0, ctx->stream>>>((cx<double>*)st->d, nk, h, xmask, zmask);
    else
        k_frame<float><<<g, 256, 0, ctx->stream>>>((cx<float>*)st->d, nk, h, xmask, zmask);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, xmask ? 2.0 * b * N : b * N);
    return QSE_OK;
}

template <class R>
__global__ void k_set_one(cx<R>* s, uint64_t idx) {
    if (threadIdx.x == 0 && blockIdx.x == 0) s[idx] = {(R)1, (R)0};
}

int launch_set_basis(qse_ctx* ctx, qse_state* st, uint64_t index) {
    QSE_HIP(hipMemsetAsync(st->d, 0, st->bytes, ctx->stream));
    if (st->prec == QSE_C128)
        k_set_one<double><<<1, 64, 0, ctx->stream>>>((cx<double>*)st->d, index);
    else
        k_set_one<float><<<1, 64, 0, ctx->stream>>>((cx<float>*)st->d, index);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

__global__ __launch_bounds__(256) void k_c128_to_c64(cx<float>* __restrict__ dst, const cx<double>* __restrict__ src,
                                                     uint64_t count) {
    for (uint64_t i = gtid(); i < count; i += gstride()) dst[i] = {(float)src[i].re, (float)src[i].im};
}
__global__ __launch_bounds__(256) void k_c64_to_c128(cx<double>* __restrict__ dst, const cx<float>* __restrict__ src,
                                                     uint64_t count) {
    for (uint64_t i = gtid(); i < count; i += gstride()) dst[i] = {(double)src[i].re, (double)src[i].im};
}

int launch_convert_upload(qse_ctx* ctx, qse_state* st, const double* dev_c128, uint64_t count, uint64_t first) {
    k_c128_to_c64<<<grid_for(ctx, count, 256), 256, 0, ctx->stream>>>((cx<float>*)st->d + first,
                                                                       (const cx<double>*)dev_c128, count);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

int launch_convert_download(qse_ctx* ctx, const qse_state* st, double* dev_c128, uint64_t count, uint64_t first) {
    k_c64_to_c128<<<grid_for(ctx, count, 256), 256, 0, ctx->stream>>>((cx<double>*)dev_c128,
                                                                       (const cx<float>*)st->d + first, count);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

template <class R>
__global__ __launch_bounds__(256) void k_random(cx<R>* __restrict__ s, uint64_t n, uint64_t seed) {
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        double re, im;
        random_amp(seed, i, re, im);
        s[i] = {(R)re, (R)im};
    }
}

int launch_random(qse_ctx* ctx, qse_state* st, uint64_t seed) {
    const uint64_t N = 1ULL << st->n;
    if (st->prec == QSE_C128)
        k_random<double><<<grid_for(ctx, N, 256), 256, 0, ctx->stream>>>((cx<double>*)st->d, N, seed);
    else
        k_random<float><<<grid_for(ctx, N, 256), 256, 0, ctx->stream>>>((cx<float>*)st->d, N, seed);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

template <class R>
__global__ __launch_bounds__(256) void k_scale(cx<R>* __restrict__ s, uint64_t n, R f) {
    for (uint64_t i = gtid(); i < n; i += gstride()) s[i] = {s[i].re * f, s[i].im * f};
}

int launch_scale(qse_ctx* ctx, qse_state* st, double f) {
    const uint64_t N = 1ULL << st->n;
    if (st->prec == QSE_C128)
        k_scale<double><<<grid_for(ctx, N, 256), 256, 0, ctx->stream>>>((cx<double>*)st->d, N, f);
    else
        k_scale<float><<<grid_for(ctx, N, 256), 256, 0, ctx->stream>>>((cx<float>*)st->d, N, (float)f);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

}  // namespace qse

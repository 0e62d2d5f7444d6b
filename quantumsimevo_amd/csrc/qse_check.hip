// qse_check.hip — closed-form checks of full-size runs, evaluated on the GPU
// without a second state vector.
//
// A 33-qubit complex128 or 34-qubit complex64 state (128 GiB) leaves no room
// on the device, and far too little on the host, to keep the input of a QFT
// for comparison. The QFT of QFT.cpp:9-22 (H on qubit 0 first, final swaps)
// is the bit-reversed unitary DFT with positive exponent:
//     out[k] = (1/sqrt(N S)) sum_j psi[j] exp(+2 pi i rev(j) rev(k) / N),
// rev = n-bit reversal, S = sum_j |psi[j]|^2, and the input psi of
// qse_state_set_random(seed) is a pure function of (seed, j) (random_amp), so
// any output amplitude can be evaluated in FP64 from the seed alone: one sweep
// over j regenerates psi[j] and accumulates every requested k at once.
#include <cmath>
#include <vector>

#include "qse_device.hpp"
#include "qse_internal.hpp"

namespace qse {

namespace {

constexpr int PROBE_K = 16;       // output amplitudes per sweep
constexpr int PROBE_CHUNK = 1024;  // consecutive y = rev(j) per thread between exact re-anchoring
constexpr int PROBE_BLOCK = 256;
constexpr int PROBE_VALS = 2 * PROBE_K + 1;  // per block: K complex sums and S

// Per block: sums over its chunks of psi[j] * exp(2 pi i y rk[c] / N) for
// y = rev(j) (the twiddle advances by a fixed factor along y and is evaluated
// exactly at each chunk start), and of |psi[j]|^2.
__global__ __launch_bounds__(PROBE_BLOCK) void k_qft_probe(int n, uint64_t seed, const uint64_t* __restrict__ rk,
                                                           int K, uint64_t chunk, double* __restrict__ partial) {
    const uint64_t N = 1ULL << n, mask = N - 1;
    const double inv_half_n = 2.0 / (double)N;  // sincospi argument per unit of (y * rk mod N)
    double wr[PROBE_K], wi[PROBE_K], ar[PROBE_K], ai[PROBE_K];
#pragma unroll
    for (int c = 0; c < PROBE_K; ++c) {
        ar[c] = ai[c] = 0.0;
        const uint64_t r = c < K ? rk[c] : 0;
        sincospi((double)(r & mask) * inv_half_n, &wi[c], &wr[c]);
    }
    double s = 0.0;
    const uint64_t nchunks = N / chunk;
    for (uint64_t ch = gtid(); ch < nchunks; ch += gstride()) {
        const uint64_t y0 = ch * chunk;
        double zr[PROBE_K], zi[PROBE_K];
#pragma unroll
        for (int c = 0; c < PROBE_K; ++c) {
            const uint64_t r = c < K ? rk[c] : 0;
            sincospi((double)((y0 * r) & mask) * inv_half_n, &zi[c], &zr[c]);
        }
        for (uint64_t t = 0; t < chunk; ++t) {
            const uint64_t y = y0 + t;
            const uint64_t j = __builtin_bitreverse64(y) >> (64 - n);
            double pr, pi;
            random_amp(seed, j, pr, pi);
            s += pr * pr + pi * pi;
#pragma unroll
            for (int c = 0; c < PROBE_K; ++c) {
                ar[c] += pr * zr[c] - pi * zi[c];
                ai[c] += pr * zi[c] + pi * zr[c];
                const double nr = zr[c] * wr[c] - zi[c] * wi[c];
                zi[c] = zr[c] * wi[c] + zi[c] * wr[c];
                zr[c] = nr;
            }
        }
    }
    __shared__ double red[PROBE_BLOCK / 64][PROBE_VALS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < PROBE_K; ++c) {
        const double a = wave_sum(ar[c]), b = wave_sum(ai[c]);
        if (lane == 0) {
            red[wave][2 * c] = a;
            red[wave][2 * c + 1] = b;
        }
    }
    const double sw = wave_sum(s);
    if (lane == 0) red[wave][2 * PROBE_K] = sw;
    __syncthreads();
    if (threadIdx.x < PROBE_VALS) {
        double v = 0.0;
        for (int w = 0; w < PROBE_BLOCK / 64; ++w) v += red[w][threadIdx.x];
        partial[(uint64_t)blockIdx.x * PROBE_VALS + threadIdx.x] = v;
    }
}

// Out-of-place copy of 2^30-element-class buffers in tiles of 4096 16-byte
// elements (64 KiB): thread t of 256 moves elements t + 256 j, j < 16, of its
// tile. Element li of tile ti lives at ((li >> R) << S | ti) << R | (li & (2^R - 1)):
// runs of 2^R contiguous elements, 2^(12-R) runs per tile spread 2^S runs apart,
// the access pattern of a fused pass with R contiguous low qubits.
__global__ __launch_bounds__(256) void k_copy_probe(const double2* __restrict__ s, double2* __restrict__ d,
                                                    uint64_t ntiles, int R, int S) {
    const uint64_t rmask = (1ULL << R) - 1;
    for (uint64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        double2 v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint64_t li = (uint64_t)threadIdx.x + 256 * j;
            v[j] = s[((((li >> R) << S) | ti) << R) | (li & rmask)];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint64_t li = (uint64_t)threadIdx.x + 256 * j;
            d[((((li >> R) << S) | ti) << R) | (li & rmask)] = v[j];
        }
    }
}

// ---- pair exchange for gates on low qubits, two ways (a measured kernel choice) ----
// Both kernels apply `reps` layers of RY(theta_q), q = 0..5 in that order
// (theta_q = 0.3 + 0.1 q), in place to a complex128 state, tile by tile (4096
// contiguous amplitudes = qubits 0..11 per tile, 256 threads x 16 amplitudes),
// with the same arithmetic per amplitude pair.
constexpr int XP_REG = 16;
typedef double v2d __attribute__((ext_vector_type(2)));  // one complex128 amplitude

__device__ __forceinline__ void ry_pair(v2d& a0, v2d& a1, double c, double s) {
    const v2d v0 = a0, v1 = a1;
    a0 = v2d{c * v0.x - s * v1.x, c * v0.y - s * v1.y};
    a1 = v2d{s * v0.x + c * v1.x, s * v0.y + c * v1.y};
}

// LDS exchange (the fused pass's): registers hold qubits 0..3 (layout A) or
// 4..7 (layout B); between them the tile goes through LDS, padded by one
// 16-byte slot per 16 (element li at li + li/16).
__global__ __launch_bounds__(256) void k_exchange_lds(v2d* __restrict__ st, uint64_t ntiles, int reps) {
    __shared__ v2d t[4096 + 256];
    const int tid = threadIdx.x;
    const int offA = 17 * tid;                                       // li = 16 tid + i
    const int liB = (tid & 15) | ((tid >> 4) << 8);                  // li = liB + 16 i
    auto pad = [](int li) { return li + (li >> 4); };
    for (uint64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        v2d* p = st + (ti << 12);
        v2d a[XP_REG];
#pragma unroll
        for (int i = 0; i < XP_REG; ++i) t[pad(tid + 256 * i)] = __builtin_nontemporal_load(p + tid + 256 * i);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < XP_REG; ++i) a[i] = t[offA + i];
        for (int r = 0; r < reps; ++r) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double sn, cs;
                sincos(0.5 * (0.3 + 0.1 * q), &sn, &cs);
#pragma unroll
                for (int i = 0; i < XP_REG; ++i)
                    if (!((i >> q) & 1)) ry_pair(a[i], a[i | (1 << q)], cs, sn);
            }
#pragma unroll
            for (int i = 0; i < XP_REG; ++i) t[offA + i] = a[i];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < XP_REG; ++i) a[i] = t[pad(liB + 16 * i)];
#pragma unroll
            for (int q = 4; q < 6; ++q) {
                double sn, cs;
                sincos(0.5 * (0.3 + 0.1 * q), &sn, &cs);
#pragma unroll
                for (int i = 0; i < XP_REG; ++i)
                    if (!((i >> (q - 4)) & 1)) ry_pair(a[i], a[i | (1 << (q - 4))], cs, sn);
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < XP_REG; ++i) t[pad(liB + 16 * i)] = a[i];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < XP_REG; ++i) a[i] = t[offA + i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < XP_REG; ++i) t[offA + i] = a[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < XP_REG; ++i) __builtin_nontemporal_store(t[pad(tid + 256 * i)], p + tid + 256 * i);
        __syncthreads();
    }
}

// The value of lane (lane ^ (1 << Q)) of a 32-bit word: DPP quad_perm (Q = 0, 1),
// two DPP mirrors (Q = 2: row_half_mirror then quad_perm [3,2,1,0]; Q = 3:
// row_mirror then row_half_mirror), ds_swizzle in bit mode (Q = 4), ds_bpermute
// (Q = 5).
template <int Q>
__device__ __forceinline__ int xlane(int v) {
    if constexpr (Q == 0) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
    else if constexpr (Q == 1) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
    else if constexpr (Q == 2) return __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
    else if constexpr (Q == 3) return __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false), 0x141, 0xF, 0xF, false);
    else if constexpr (Q == 4) return __builtin_amdgcn_ds_swizzle(v, 0x401F);
    else return __builtin_amdgcn_ds_bpermute((int)((threadIdx.x & 63) ^ 32) << 2, v);
}

template <int Q>
__device__ __forceinline__ double xlane_f64(double v) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)xlane<Q>((int)(uint32_t)u), hi = (uint32_t)xlane<Q>((int)(uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int Q>
__device__ __forceinline__ void ry_lanes(v2d* a, bool upper, double c, double s) {
    // lane with bit Q clear: c a0 - s a1; set: s a0 + c a1 = c a1 + s (partner)
    const double sg = upper ? s : -s;
#pragma unroll
    for (int i = 0; i < XP_REG; ++i) {
        const double px = xlane_f64<Q>(a[i].x), py = xlane_f64<Q>(a[i].y);
        a[i] = v2d{c * a[i].x + sg * px, c * a[i].y + sg * py};
    }
}

// Cross-lane exchange: lanes hold qubits 0..5, registers qubits 6..9, the
// four waves qubits 10, 11; no LDS.
__global__ __launch_bounds__(256) void k_exchange_dpp(v2d* __restrict__ st, uint64_t ntiles, int reps) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        v2d* p = st + (ti << 12) + (w << 10) + lane;
        v2d a[XP_REG];
#pragma unroll
        for (int i = 0; i < XP_REG; ++i) a[i] = __builtin_nontemporal_load(p + (i << 6));
        for (int r = 0; r < reps; ++r) {
            double sn, cs;
            sincos(0.5 * 0.3, &sn, &cs);
            ry_lanes<0>(a, lane & 1, cs, sn);
            sincos(0.5 * 0.4, &sn, &cs);
            ry_lanes<1>(a, (lane >> 1) & 1, cs, sn);
            sincos(0.5 * 0.5, &sn, &cs);
            ry_lanes<2>(a, (lane >> 2) & 1, cs, sn);
            sincos(0.5 * 0.6, &sn, &cs);
            ry_lanes<3>(a, (lane >> 3) & 1, cs, sn);
            sincos(0.5 * 0.7, &sn, &cs);
            ry_lanes<4>(a, (lane >> 4) & 1, cs, sn);
            sincos(0.5 * 0.8, &sn, &cs);
            ry_lanes<5>(a, (lane >> 5) & 1, cs, sn);
        }
#pragma unroll
        for (int i = 0; i < XP_REG; ++i) __builtin_nontemporal_store(a[i], p + (i << 6));
    }
}

uint64_t bitrev(uint64_t v, int n) {
    uint64_t r = 0;
    for (int b = 0; b < n; ++b)
        if ((v >> b) & 1) r |= 1ULL << (n - 1 - b);
    return r;
}

}  // namespace

}  // namespace qse

using namespace qse;

extern "C" int qse_qft_probe(qse_ctx* ctx, int nqubits, uint64_t seed, const uint64_t* indices, int count,
                             double* out) {
    if (!ctx || !indices || !out || count < 1 || count > 1024 || nqubits < 1 || nqubits > 40)
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    const uint64_t N = 1ULL << nqubits;
    for (int i = 0; i < count; ++i)
        if (indices[i] >= N) return set_error(QSE_ERR_INVALID_ARG, "index out of range");
    QSE_HIP(hipSetDevice(ctx->device));
    const uint64_t chunk = std::min<uint64_t>(PROBE_CHUNK, N);
    const uint64_t nchunks = N / chunk;
    // one chunk per thread at most, at most 8 blocks per CU (grid-stride beyond)
    const unsigned blocks = (unsigned)std::min<uint64_t>((nchunks + PROBE_BLOCK - 1) / PROBE_BLOCK,
                                                         (uint64_t)ctx->num_cus * 8);
    void* dptr = nullptr;
    int rc = scratch(ctx, 6, sizeof(uint64_t) * PROBE_K + sizeof(double) * (size_t)blocks * PROBE_VALS, &dptr);
    if (rc) return rc;
    uint64_t* d_rk = (uint64_t*)dptr;
    double* d_part = (double*)((char*)dptr + sizeof(uint64_t) * PROBE_K);
    std::vector<double> part((size_t)blocks * PROBE_VALS);
    for (int first = 0; first < count; first += PROBE_K) {
        const int K = std::min(PROBE_K, count - first);
        uint64_t rk[PROBE_K] = {0};
        for (int c = 0; c < K; ++c) rk[c] = bitrev(indices[first + c], nqubits);
        QSE_HIP(hipMemcpyAsync(d_rk, rk, sizeof(rk), hipMemcpyHostToDevice, ctx->stream));
        k_qft_probe<<<blocks, PROBE_BLOCK, 0, ctx->stream>>>(nqubits, seed, d_rk, K, chunk, d_part);
        QSE_HIP(hipGetLastError());
        QSE_HIP(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
        QSE_HIP(hipStreamSynchronize(ctx->stream));
        double S = 0.0;
        std::vector<double> acc(2 * K, 0.0);
        for (unsigned b = 0; b < blocks; ++b) {
            const double* v = &part[(size_t)b * PROBE_VALS];
            for (int c = 0; c < 2 * K; ++c) acc[c] += v[c];
            S += v[2 * PROBE_K];
        }
        // unitary DFT (1/sqrt N) of the normalised input (1/sqrt S)
        const double f = 1.0 / (std::sqrt(S) * std::sqrt((double)N));
        for (int c = 0; c < 2 * K; ++c) out[2 * first + c] = acc[c] * f;
    }
    return QSE_OK;
}

extern "C" int qse_exchange_probe(qse_ctx* ctx, qse_state* st, int method, int reps, double* ms) {
    if (!ctx || !st || !st->d || !ms || (method != 0 && method != 1) || reps < 1 || reps > 64)
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    if (st->prec != QSE_C128 || st->n < 12)
        return set_error(QSE_ERR_UNSUPPORTED, "exchange probe: complex128 states of 12 or more qubits");
    QSE_HIP(hipSetDevice(ctx->device));
    const uint64_t ntiles = 1ULL << (st->n - 12);
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 16);
    v2d* s = (v2d*)st->d;
    auto launch = [&] {
        if (method == 0) k_exchange_lds<<<grid, 256, 0, ctx->stream>>>(s, ntiles, reps);
        else k_exchange_dpp<<<grid, 256, 0, ctx->stream>>>(s, ntiles, reps);
        return hipGetLastError();
    };
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (e != hipSuccess) return hip_error(e, "exchange probe");
    *ms = t;
    return QSE_OK;
}

extern "C" int qse_copy_bandwidth(qse_ctx* ctx, int log2_elems, int run_log2, double* gbs) {
    if (!ctx || !gbs || log2_elems < 12 || log2_elems > 31 || run_log2 < 0 || run_log2 > 12)
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    QSE_HIP(hipSetDevice(ctx->device));
    const uint64_t n = 1ULL << log2_elems;  // 16-byte elements per buffer
    const uint64_t ntiles = n >> 12;
    const int S = log2_elems - 12;  // tile index bits between a tile's runs
    // largest index: (((2^(12-R) - 1) << S | (ntiles - 1)) << R) | (2^R - 1) = n - 1
    void* buf = nullptr;
    if (hipMalloc(&buf, 2 * n * 16) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(QSE_ERR_OUT_OF_MEMORY, "copy probe: allocation failed");
    }
    double2* s = (double2*)buf;
    double2* d = s + n;
    int rc = QSE_OK;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipMemsetAsync(buf, 0, 2 * n * 16, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 16);
    if (e == hipSuccess) {
        k_copy_probe<<<grid, 256, 0, ctx->stream>>>(s, d, ntiles, run_log2, S);  // warm-up
        e = hipEventRecord(a, ctx->stream);
    }
    for (int r = 0; r < 5 && e == hipSuccess; ++r) {
        k_copy_probe<<<grid, 256, 0, ctx->stream>>>(s, d, ntiles, run_log2, S);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (e != hipSuccess) rc = hip_error(e, "copy probe");
    else *gbs = 2.0 * 16.0 * (double)n / (ms / 5.0 * 1e-3) / 1e9;
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    hipStreamSynchronize(ctx->stream);
    hipFree(buf);
    return rc;
}

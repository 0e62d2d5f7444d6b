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
int run_plan(qse_ctx* ctx, qse_state* st, const HostPlan& plan, int k, int jit_kind);
int choose_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P, bool allow_oop);
int tuned_plan(const qse_ctx* ctx, int n, int prec, const qse_op* ops, int nops, const double* params, HostPlan& P,
               bool allow_oop, int* k_out, int* L_out);
int validate_ops(int n, const qse_op* ops, int nops, const double* params);
int reduce_zstr_async(qse_ctx* ctx, const qse_state* st, const uint64_t* d_m, const double* d_w, int nterms,
                      double* d_out);
bool is_zz_triple(const qse_op* ops, int nops, int k);
struct SlotVal;
int program_on_device(qse_ctx* ctx, const Serialized& S, int prec, ProgEntry** entry);
int jit_joint_for(qse_ctx* ctx, const HostPlan& P, const Serialized& S, int prec,
                  const std::vector<hipFunction_t>** fns);
int slot_table_async(qse_ctx* ctx, const char* dev, const Serialized& S, const HostPlan& P, int p, uint64_t ntiles,
                     SlotVal* table);

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

// Observable terms grouped by the top three bits of their masks, u = m >> (n - 3):
// group u holds entries [ubeg[u], ubeg[u + 1]) of the masks and weights.
struct ObsGroups {
    int ubeg[9];
};

// lambda = O psi and the energy <psi|O|psi> in one pass over psi. A thread
// takes blocks of 8 amplitudes i = i0 + S j (S = N / 8, i0 < S, j < 8):
// popcount(i & m) = popcount(i0 & m) + popcount(j & u), so with the group sums
// G[u] = sum over group u of w_t (-1)^popcount(i0 & m_t), the factor
// f(i0 + S j) = sum_u G[u] (-1)^popcount(j & u) is an 8-point Walsh-Hadamard
// transform of G: each term is evaluated once per 8 amplitudes. part[block] =
// the workgroup's sum of f(i) |psi_i|^2 in FP64 (threads by wave shuffles,
// waves in order), reduced by k_sum_columns.
template <class R>
__global__ __launch_bounds__(AD_THREADS) void k_obs_energy(const cx<R>* __restrict__ psi, cx<R>* __restrict__ lam,
                                                           uint64_t S, const uint64_t* __restrict__ masks,
                                                           const double* __restrict__ w, ObsGroups g,
                                                           double* __restrict__ part) {
    double e = 0.0;
    for (uint64_t i0 = gtid(); i0 < S; i0 += gstride()) {
        double G[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            double s = 0.0;
            for (int t = g.ubeg[u]; t < g.ubeg[u + 1]; ++t) s += (__popcll(i0 & masks[t]) & 1) ? -w[t] : w[t];
            G[u] = s;
        }
#pragma unroll
        for (int h = 1; h < 8; h <<= 1)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (!(j & h)) {
                    const double a = G[j], b = G[j + h];
                    G[j] = a + b;
                    G[j + h] = a - b;
                }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const cx<R> a = psi[i0 + S * j];
            const double re = (double)a.re, im = (double)a.im;
            lam[i0 + S * j] = {(R)(G[j] * re), (R)(G[j] * im)};
            e += G[j] * (re * re + im * im);
        }
    }
    __shared__ double sh[AD_THREADS / 64];
    const double v = wave_sum(e);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w2 = 0; w2 < AD_THREADS / 64; ++w2) t += sh[w2];
        part[blockIdx.x] = t;
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

// Up to NM Z-strings in one pass: partial[b][d] = sum over the block's
// amplitudes of (-1)^popcount(i & m_d) w_i, w_i = Im(conj(lambda_i) psi_i).
// With S threads in the grid (a power of two), thread t takes the amplitudes
// i = t + S (j + 32 k), j < 32, so the sign splits into
//   parity(t & m) ^ parity(j & (m >> log2 S)) ^ parity(32 S k & m):
// a per-thread constant, one entry of the 32-point Walsh-Hadamard transform
// H[u] = sum_j (-1)^popcount(j & u) w_j of the thread's 32 values, and a
// sign that is the same for the whole grid in iteration k. The transform is
// kept in LDS (u-major), one read per mask per 32 amplitudes. Requires
// 32 S <= N and N a multiple of 32 S.
constexpr int WHT_J = 32;
template <int NM>
struct DiagMasksN {
    uint64_t m[NM];
};

// WHT_GRAD: w_i = Im(conj(lambda_i) psi_i) (generator terms); WHT_PROB:
// w_i = |psi_i|^2 (expectation values; lam unused).
enum WhtMode { WHT_GRAD = 0, WHT_PROB = 1 };

template <class R, int NM, int MODE>
__global__ __launch_bounds__(AD_THREADS) void k_grad_diag_wht(const cx<R>* __restrict__ psi,
                                                              const cx<R>* __restrict__ lam, uint64_t N, int log2S,
                                                              DiagMasksN<NM> M, double* __restrict__ part) {
    __shared__ double hs[WHT_J][AD_THREADS];
    __shared__ double red[AD_THREADS / 64][NM];
    const uint64_t S = 1ULL << log2S;
    const uint64_t t = (uint64_t)blockIdx.x * AD_THREADS + threadIdx.x;
    uint64_t cpar = 0;  // bit d: parity(t & m_d)
#pragma unroll
    for (int d = 0; d < NM; ++d) cpar |= (uint64_t)(__popcll(t & M.m[d]) & 1) << d;
    double acc[NM];
#pragma unroll
    for (int d = 0; d < NM; ++d) acc[d] = 0.0;
    const uint64_t iters = N >> (log2S + 5);
    for (uint64_t k = 0; k < iters; ++k) {
        double w[WHT_J];
#pragma unroll
        for (int j = 0; j < WHT_J; ++j) {
            const uint64_t i = t + S * ((uint64_t)j + WHT_J * k);
            const cx<R> p = psi[i];
            if constexpr (MODE == WHT_PROB) {
                w[j] = prob(p);
            } else {
                const cx<R> l = lam[i];
                w[j] = (double)l.re * (double)p.im - (double)l.im * (double)p.re;
            }
        }
#pragma unroll
        for (int h = 1; h < WHT_J; h <<= 1)
#pragma unroll
            for (int j = 0; j < WHT_J; ++j)
                if (!(j & h)) {
                    const double a = w[j], b = w[j + h];
                    w[j] = a + b;
                    w[j + h] = a - b;
                }
#pragma unroll
        for (int u = 0; u < WHT_J; ++u) hs[u][threadIdx.x] = w[u];
        // no barrier: each thread reads back only its own column
        const uint64_t kb = (WHT_J * S) * k;
#pragma unroll
        for (int d = 0; d < NM; ++d) {
            const int u = (int)((M.m[d] >> log2S) & (WHT_J - 1));
            const double v = hs[u][threadIdx.x];
            const bool neg = (((cpar >> d) ^ (uint64_t)__popcll(kb & M.m[d])) & 1) != 0;
            acc[d] += neg ? -v : v;
        }
    }
#pragma unroll
    for (int d = 0; d < NM; ++d) {
        const double v = wave_sum(acc[d]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][d] = v;
    }
    __syncthreads();
    if (threadIdx.x < NM) {
        double v = 0.0;
        for (int w = 0; w < AD_THREADS / 64; ++w) v += red[w][threadIdx.x];
        part[(uint64_t)blockIdx.x * NM + threadIdx.x] = v;
    }
}

// out[d] = sum over rows r of part[r * stride + d], one workgroup per column
// d < cols: thread t adds rows t, t + 256, ... in order, then a fixed tree
// over the 256 threads (the same order on every run).
__global__ __launch_bounds__(AD_THREADS) void k_sum_columns(const double* __restrict__ part, int rows, int stride,
                                                            double* __restrict__ out) {
    const int d = blockIdx.x;
    double v = 0.0;
    for (int r = threadIdx.x; r < rows; r += AD_THREADS) v += part[(size_t)r * stride + d];
    __shared__ double sh[AD_THREADS];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = AD_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[d] = sh[0];
}

// X/Y generator terms without rotating psi and lambda: for a tile qubit q with
// local bit b, Im<lambda|X_q|psi> = sum over pairs (a, a|b) of
// Im(conj(l_a) p_{a|b}) + Im(conj(l_{a|b}) p_a), and Im<lambda|Y_q|psi> =
// sum of -Re(conj(l_a) p_{a|b}) + Re(conj(l_{a|b}) p_a). One read-only pass
// over both vectors covers the terms of up to 12 tile qubits: the tile's 2^12
// amplitudes of each vector sit 16 per thread in registers, 4 local bits at a
// time (groups {3,4,5,6}, {7,8,9,10}, {0,1,2,11}), and move between the
// groups through LDS, one vector at a time. Local bits 0..2 are global qubits
// 0..2 (runs of 8 contiguous amplitudes); the other 9 local bits and the
// tile index bits map to global qubits through the pass description.
constexpr int PAIR_K = 12;
constexpr int PAIR_THREADS = 256;
// Z-string terms Im<lambda|Z_m|psi> = sum_x (-1)^popcount(x & m) Im(conj(l_x) p_x)
// evaluated in the same pass, up to PAIR_Z per pass: in the last stage each
// thread takes the 16-point Walsh-Hadamard transform W of its 16 values w_r
// over the register qubits, and the term gets s_tile(m) s_thread(m) W[u_m]
// (parities of m on the tile-index and thread qubits, u_m = m on the register
// qubits).
constexpr int PAIR_Z = 12;
constexpr int PAIR_COLS = PAIR_K + PAIR_Z;

struct PairPass {
    uint32_t tq[PAIR_K];     // local bit -> global qubit (tq[0..2] = 0, 1, 2)
    uint32_t epos[64];       // tile index bit -> global qubit (ascending)
    int32_t next;            // tile index bits (n - 12)
    int32_t nstages;         // register groups to visit (terms beyond are absent)
    int32_t col[PAIR_K];     // output column of the term on each local bit, -1: none
    int32_t type[PAIR_K];    // 0: X, 1: Y
    int32_t nz;              // Z-string terms of this pass (<= PAIR_Z), columns zcol0 + z
    int32_t zcol0;
    uint32_t zu[PAIR_Z];     // mask bits on the last stage's register qubits (4-bit pattern)
    uint64_t zm[PAIR_Z];     // the masks (global qubits)
};

__device__ __forceinline__ constexpr int pair_group(int s, int j) {
    return s == 0 ? 3 + j : s == 1 ? 7 + j : (j < 3 ? j : 11);
}
// thread-id bit i of stage s -> local bit (the 8 local bits outside group s, ascending)
__device__ __forceinline__ constexpr int pair_tbit(int s, int i) {
    return s == 0 ? (i < 3 ? i : i + 4) : s == 1 ? (i < 7 ? i : 11) : 3 + i;
}
// LDS place of local index li: one element of padding after every 16, so that
// the 16 lanes of each quarter-wave hit 16 different 16-byte bank groups in all
// three layouts. The padding is additive over disjoint bits, so a stage's 16
// places are one per-thread base plus constant offsets (no address registers).
__device__ __forceinline__ constexpr uint32_t pair_pad(uint32_t li) { return li + (li >> 4); }
constexpr int PAIR_LDS = (1 << PAIR_K) + (1 << (PAIR_K - 4));
__device__ __forceinline__ uint32_t pair_tpart(int s, uint32_t t) {
    uint32_t li = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) li |= ((t >> i) & 1u) << pair_tbit(s, i);
    return li;
}
__device__ __forceinline__ constexpr uint32_t pair_rpart(int s, int r) {
    uint32_t li = 0;
    for (int j = 0; j < 4; ++j) li |= (uint32_t)((r >> j) & 1) << pair_group(s, j);
    return li;
}

template <class R, int S>
__device__ __forceinline__ void pair_terms(const cx<R> (&p)[16], const cx<R> (&l)[16], const PairPass& P,
                                           double (&acc)[PAIR_K]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int lb = pair_group(S, j);
        if (P.col[lb] < 0) continue;
        const bool y = P.type[lb] != 0;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < 16; ++a) {
            if ((a >> j) & 1) continue;
            const int b = a | (1 << j);
            const double lar = l[a].re, lai = l[a].im, lbr = l[b].re, lbi = l[b].im;
            const double par = p[a].re, pai = p[a].im, pbr = p[b].re, pbi = p[b].im;
            if (!y) v += (lar * pbi - lai * pbr) + (lbr * pai - lbi * par);
            else v += -(lar * pbr + lai * pbi) + (lbr * par + lbi * pai);
        }
        acc[lb] += v;
    }
}

template <class R, int FROM>
__device__ __forceinline__ void pair_exchange(cx<R> (&v)[16], cx<R>* lds, uint32_t t) {
    __syncthreads();
    const uint32_t wb = pair_pad(pair_tpart(FROM, t)), rb = pair_pad(pair_tpart(FROM + 1, t));
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[wb + pair_pad(pair_rpart(FROM, r))] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = lds[rb + pair_pad(pair_rpart(FROM + 1, r))];
}

template <class R>
__global__ __launch_bounds__(PAIR_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_grad_pair(const cx<R>* __restrict__ psi,
                                                            const cx<R>* __restrict__ lam, uint64_t ntiles,
                                                            PairPass P, double* __restrict__ part) {
    __shared__ cx<R> lds[PAIR_LDS];
    __shared__ double red[PAIR_THREADS / 64][PAIR_COLS];
    double* hw = reinterpret_cast<double*>(lds);  // W[u] of thread t at hw[u * 256 + t] (last stage)
    static_assert(sizeof(cx<R>) * PAIR_LDS >= sizeof(double) * 16 * PAIR_THREADS, "LDS too small for W");
    const uint32_t t = threadIdx.x;
    double acc[PAIR_K], zacc[PAIR_Z];
#pragma unroll
    for (int b = 0; b < PAIR_K; ++b) acc[b] = 0.0;
#pragma unroll
    for (int z = 0; z < PAIR_Z; ++z) zacc[z] = 0.0;
    // stage-0 offsets: thread part and the 16 register combinations
    uint64_t toff = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) toff |= (uint64_t)((t >> i) & 1u) << P.tq[pair_tbit(0, i)];
    uint64_t roff[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) roff[j] = 1ULL << P.tq[pair_group(0, j)];
    // bit z: parity of mask z on the thread's qubits in the last stage's layout
    const int zs = P.nstages - 1;
    uint64_t toffz = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        toffz |= (uint64_t)((t >> i) & 1u) << P.tq[zs == 0 ? pair_tbit(0, i) : zs == 1 ? pair_tbit(1, i) : pair_tbit(2, i)];
    uint32_t sthr = 0;
    for (int z = 0; z < P.nz; ++z) sthr |= (uint32_t)(__popcll(P.zm[z] & toffz) & 1) << z;
    for (uint64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        uint64_t ebase = 0;
        for (int e = 0; e < P.next; ++e) ebase |= ((ti >> e) & 1ULL) << P.epos[e];
        const uint64_t base = ebase | toff;
        cx<R> p[16], l[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            uint64_t g = base;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((r >> j) & 1) g |= roff[j];
            p[r] = psi[g];
            l[r] = lam[g];
        }
        pair_terms<R, 0>(p, l, P, acc);
        if (P.nstages > 1) {
            pair_exchange<R, 0>(p, lds, t);
            pair_exchange<R, 0>(l, lds, t);
            pair_terms<R, 1>(p, l, P, acc);
        }
        if (P.nstages > 2) {
            pair_exchange<R, 1>(p, lds, t);
            pair_exchange<R, 1>(l, lds, t);
            pair_terms<R, 2>(p, l, P, acc);
        }
        if (P.nz > 0) {
            // in the last stage's layout, after which p and l are no longer needed
            double w[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) w[r] = (double)l[r].re * (double)p[r].im - (double)l[r].im * (double)p[r].re;
#pragma unroll
            for (int h = 1; h < 16; h <<= 1)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    if (!(r & h)) {
                        const double a = w[r], b = w[r | h];
                        w[r] = a + b;
                        w[r | h] = a - b;
                    }
            __syncthreads();  // every thread's exchange reads before W overwrites the LDS
#pragma unroll
            for (int u = 0; u < 16; ++u) hw[u * PAIR_THREADS + t] = w[u];
            // each thread reads back only its own column: no barrier
#pragma unroll
            for (int z = 0; z < PAIR_Z; ++z) {
                if (z >= P.nz) break;
                const double v = hw[P.zu[z] * PAIR_THREADS + t];
                const bool neg = (((sthr >> z) ^ (uint32_t)__popcll(P.zm[z] & ebase)) & 1) != 0;
                zacc[z] += neg ? -v : v;
            }
        }
        __syncthreads();  // LDS reads of this tile before the next tile's writes
    }
#pragma unroll
    for (int b = 0; b < PAIR_K; ++b) {
        const double v = wave_sum(acc[b]);
        if ((t & 63) == 0) red[t >> 6][b] = v;
    }
#pragma unroll
    for (int z = 0; z < PAIR_Z; ++z) {
        const double v = wave_sum(zacc[z]);
        if ((t & 63) == 0) red[t >> 6][PAIR_K + z] = v;
    }
    __syncthreads();
    int c = -1;
    if (t < PAIR_K) c = P.col[t];
    else if (t < PAIR_K + (uint32_t)P.nz) c = P.zcol0 + (int)(t - PAIR_K);
    if (t < PAIR_COLS && c >= 0) {
        double v = 0.0;
        for (int w = 0; w < PAIR_THREADS / 64; ++w) v += red[w][t];
        part[(uint64_t)blockIdx.x * PAIR_COLS + c] = v;
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

// lambda = O psi and the energy written to *d_out, in one pass over psi
// (k_obs_energy, then its per-workgroup sums in workgroup order). dm and dw
// hold the terms sorted by group (ObsGroups); n >= 3.
template <class R>
int obs_energy_async(qse_ctx* ctx, const qse_state* psi, qse_state* lam, const uint64_t* dm, const double* dw,
                     const ObsGroups& g, double* d_out) {
    const uint64_t S = 1ULL << (psi->n - 3);
    const unsigned grid = grid_for(ctx, S, AD_THREADS);
    void* pbuf;
    int rc = scratch(ctx, 6, sizeof(double) * grid, &pbuf);
    if (rc) return rc;
    k_obs_energy<R><<<grid, AD_THREADS, 0, ctx->stream>>>((const cx<R>*)psi->d, (cx<R>*)lam->d, S, dm, dw, g,
                                                          (double*)pbuf);
    QSE_HIP(hipGetLastError());
    k_sum_columns<<<1, AD_THREADS, 0, ctx->stream>>>((const double*)pbuf, (int)grid, 1, d_out);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

template <class R, int NM, int MODE>
int launch_wht(qse_ctx* ctx, const qse_state* psi, const qse_state* lam, const uint64_t* masks, int cnt,
               unsigned grid, int log2S, double* pbuf, double* d_out) {
    DiagMasksN<NM> M{};
    for (int d = 0; d < cnt; ++d) M.m[d] = masks[d];
    const uint64_t N = 1ULL << psi->n;
    TimedLaunch t;
    launch_begin(ctx, MODE == WHT_PROB ? FAM_READOUT : FAM_ADJOINT, &t);
    k_grad_diag_wht<R, NM, MODE><<<grid, AD_THREADS, 0, ctx->stream>>>(
        (const cx<R>*)psi->d, (const cx<R>*)(lam ? lam->d : psi->d), N, log2S, M, pbuf);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, (MODE == WHT_PROB ? 1.0 : 2.0) * amp_bytes(psi->prec) * (double)N);
    k_sum_columns<<<(unsigned)cnt, AD_THREADS, 0, ctx->stream>>>(pbuf, (int)grid, NM, d_out);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

constexpr int WHT_MIN_QUBITS = 13;

// Z-string sums of `count` masks, 64 per pass, written to d_out[0 .. count) on
// the device: sum_i (-1)^popcount(i & m) w_i with w_i from MODE.
template <class R, int MODE>
int zstrings_wht(qse_ctx* ctx, const qse_state* psi, const qse_state* lam, const uint64_t* masks, int count,
                 double* d_out) {
    // grid: a power of two, at most 1024 workgroups, 32 amplitudes per thread and iteration
    int log2S = psi->n - 5;
    if (log2S > 18) log2S = 18;  // 1024 x 256 threads
    const unsigned grid = 1u << (log2S - 8);
    void* pbuf;
    int rc = scratch(ctx, 6, sizeof(double) * 64 * grid, &pbuf);
    if (rc) return rc;
    for (int c0 = 0; c0 < count; c0 += 64) {
        const int cnt = std::min(64, count - c0);
        if (cnt <= 16) rc = launch_wht<R, 16, MODE>(ctx, psi, lam, masks + c0, cnt, grid, log2S, (double*)pbuf, d_out + c0);
        else if (cnt <= 32) rc = launch_wht<R, 32, MODE>(ctx, psi, lam, masks + c0, cnt, grid, log2S, (double*)pbuf, d_out + c0);
        else if (cnt <= 48) rc = launch_wht<R, 48, MODE>(ctx, psi, lam, masks + c0, cnt, grid, log2S, (double*)pbuf, d_out + c0);
        else rc = launch_wht<R, 64, MODE>(ctx, psi, lam, masks + c0, cnt, grid, log2S, (double*)pbuf, d_out + c0);
        if (rc) return rc;
    }
    return QSE_OK;
}

// *out = sum_k w[k] s[k] in order (one thread).
__global__ void k_weighted_sum(const double* __restrict__ s, const double* __restrict__ w, int n,
                               double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double v = 0.0;
        for (int k = 0; k < n; ++k) v += w[k] * s[k];
        *out = v;
    }
}

// sum_k w_k <psi|Z_{m_k}|psi> written to *d_out on the device: from 13 qubits
// 64 Z-strings per pass (Walsh-Hadamard reduction), below one pass over all
// terms per amplitude.
template <class R>
int energy_async(qse_ctx* ctx, const qse_state* psi, const uint64_t* h_m, const uint64_t* d_m, const double* d_w,
                 int nterms, double* d_out) {
    if (psi->n < WHT_MIN_QUBITS || nterms == 0) return reduce_zstr_async(ctx, psi, d_m, d_w, nterms, d_out);
    void* sbuf;
    int rc = scratch(ctx, 9, sizeof(double) * (size_t)nterms, &sbuf);
    if (rc) return rc;
    rc = zstrings_wht<R, WHT_PROB>(ctx, psi, nullptr, h_m, nterms, (double*)sbuf);
    if (rc) return rc;
    k_weighted_sum<<<1, 64, 0, ctx->stream>>>((const double*)sbuf, d_w, nterms, d_out);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

// Im<lambda|Z_m|psi> for every mask, written to d_out[0 .. masks.size()) on
// the device (no host wait): one pass per 64 masks (per 16 below 13 qubits).
template <class R>
int grad_diag_async(qse_ctx* ctx, qse_state* psi, qse_state* lam, const std::vector<uint64_t>& masks,
                    double* d_out) {
    const uint64_t N = 1ULL << psi->n;
    int rc;
    if (psi->n >= WHT_MIN_QUBITS) return zstrings_wht<R, WHT_GRAD>(ctx, psi, lam, masks.data(), (int)masks.size(), d_out);
    const unsigned g = std::min<unsigned>(grid_for(ctx, N, AD_THREADS, 8), 1024u);
    void* pbuf;
    rc = scratch(ctx, 6, sizeof(double) * DIAG_CHUNK * g, &pbuf);
    if (rc) return rc;
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
        k_sum_columns<<<(unsigned)cnt, AD_THREADS, 0, ctx->stream>>>((const double*)pbuf, (int)g, DIAG_CHUNK,
                                                                    d_out + c0);
        QSE_HIP(hipGetLastError());
    }
    return QSE_OK;
}

// Pair terms (qubit, type) of one evaluation point, and up to PAIR_Z Z-string
// masks per pass alongside: one read-only pass over psi and lambda per tile
// of 12 qubits (terms on qubits < 12 in the tile of qubits 0..11; the others
// 9 at a time beside qubits 0..2). Values land on the device at
// d_out[0 .. terms.size() + zmasks.size()) in the order returned in `order`
// (order[k] = i for pair term i, -1 - m for mask m). Needs n >= 13 and
// zmasks.size() <= PAIR_Z x the number of passes (pair_passes()).
int pair_passes(const std::vector<std::pair<int, int>>& terms) {
    int lo = 0, hi = 0;
    for (const auto& t : terms) (t.first < PAIR_K ? lo : hi) += 1;
    return (lo > 0 ? 1 : 0) + (hi + PAIR_K - 4) / (PAIR_K - 3);
}

template <class R>
int pair_terms_async(qse_ctx* ctx, const qse_state* psi, const qse_state* lam,
                     const std::vector<std::pair<int, int>>& terms, const std::vector<uint64_t>& zmasks,
                     double* d_out, std::vector<int>& order) {
    const int n = psi->n;
    std::vector<int> lo, hi;
    for (size_t i = 0; i < terms.size(); ++i) (terms[i].first < PAIR_K ? lo : hi).push_back((int)i);
    std::sort(hi.begin(), hi.end(), [&](int a, int b) { return terms[a].first < terms[b].first; });
    std::vector<std::vector<int>> passes;
    if (!lo.empty()) passes.push_back(lo);
    for (size_t c = 0; c < hi.size(); c += PAIR_K - 3)
        passes.emplace_back(hi.begin() + c, hi.begin() + std::min(hi.size(), c + (PAIR_K - 3)));
    if (zmasks.size() > PAIR_Z * passes.size()) return set_error(QSE_ERR_LIMIT, "adjoint: too many Z-strings for the pair passes");
    const uint64_t ntiles = 1ULL << (n - PAIR_K);
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 2);
    void* pbuf;
    int rc = scratch(ctx, 6, sizeof(double) * PAIR_COLS * grid, &pbuf);
    if (rc) return rc;
    order.clear();
    size_t zdone = 0;
    for (size_t pi = 0; pi < passes.size(); ++pi) {
        const std::vector<int>& pass = passes[pi];
        PairPass P{};
        for (int b = 0; b < PAIR_K; ++b) P.col[b] = -1;
        uint64_t in_tile = 0;
        const bool low = terms[pass[0]].first < PAIR_K;
        if (low) {
            for (int b = 0; b < PAIR_K; ++b) P.tq[b] = (uint32_t)b;
        } else {
            for (int b = 0; b < 3; ++b) P.tq[b] = (uint32_t)b;
            for (size_t c = 0; c < pass.size(); ++c) P.tq[3 + c] = (uint32_t)terms[pass[c]].first;
            for (int b = 0; b < 3 + (int)pass.size(); ++b) in_tile |= 1ULL << P.tq[b];
            int fill = 3;  // pad with the lowest qubits outside the tile
            for (int b = 3 + (int)pass.size(); b < PAIR_K; ++b) {
                while ((in_tile >> fill) & 1) ++fill;
                P.tq[b] = (uint32_t)fill;
                in_tile |= 1ULL << fill;
            }
        }
        in_tile = 0;
        for (int b = 0; b < PAIR_K; ++b) in_tile |= 1ULL << P.tq[b];
        P.next = 0;
        for (int q = 0; q < n; ++q)
            if (!((in_tile >> q) & 1)) P.epos[P.next++] = (uint32_t)q;
        const size_t first = order.size();
        int last_stage = 0;
        for (size_t c = 0; c < pass.size(); ++c) {
            const int q = terms[pass[c]].first;
            int lb = -1;
            for (int b = 0; b < PAIR_K; ++b)
                if ((int)P.tq[b] == q) lb = b;
            P.col[lb] = (int)c;
            P.type[lb] = terms[pass[c]].second;
            last_stage = std::max(last_stage, (lb >= 3 && lb <= 6) ? 0 : (lb >= 7 && lb <= 10) ? 1 : 2);
            order.push_back(pass[c]);
        }
        // Z-strings: the first passes take PAIR_Z each, in their last stage
        if (zdone < zmasks.size()) last_stage = std::max(last_stage, 0);
        P.zcol0 = (int)pass.size();
        P.nz = (int)std::min<size_t>(PAIR_Z, zmasks.size() - zdone);
        for (int z = 0; z < P.nz; ++z) {
            const uint64_t m = zmasks[zdone + z];
            P.zm[z] = m;
            uint32_t u = 0;
            for (int j = 0; j < 4; ++j) {
                const int lb = last_stage == 0 ? 3 + j : last_stage == 1 ? 7 + j : (j < 3 ? j : 11);
                u |= (uint32_t)((m >> P.tq[lb]) & 1) << j;
            }
            P.zu[z] = u;
            order.push_back(-1 - (int)(zdone + z));
        }
        zdone += (size_t)P.nz;
        P.nstages = last_stage + 1;
        TimedLaunch t;
        launch_begin(ctx, FAM_ADJOINT, &t);
        k_grad_pair<R><<<grid, PAIR_THREADS, 0, ctx->stream>>>((const cx<R>*)psi->d, (const cx<R>*)lam->d, ntiles, P,
                                                               (double*)pbuf);
        QSE_HIP(hipGetLastError());
        launch_end(ctx, &t, 2.0 * amp_bytes(psi->prec) * (double)(1ULL << n));
        k_sum_columns<<<(unsigned)(order.size() - first), AD_THREADS, 0, ctx->stream>>>(
            (const double*)pbuf, (int)grid, PAIR_COLS, d_out + first);
        QSE_HIP(hipGetLastError());
    }
    return QSE_OK;
}

bool all_diagonal(const std::vector<qse_op>& ops) {
    for (int k = 0; k < (int)ops.size(); ++k) {
        if (is_zz_triple(ops.data(), (int)ops.size(), k)) {
            k += 2;
            continue;
        }
        switch (ops[k].kind) {
            case QSE_GATE_IDENTITY: case QSE_GATE_Z: case QSE_GATE_PHASE: case QSE_GATE_RZ: case QSE_GATE_CPHASE:
            case QSE_GATE_MCPHASE:
                break;
            default: return false;
        }
    }
    return true;
}

// A segment with Z-string terms only whose un-application is diagonal can
// un-apply before its evaluation (the same diagonal unitary on psi and lambda
// leaves conj(lambda_i) psi_i unchanged): its gates join the previous
// segment's un-application, which saves that program's own passes.
void merge_diagonal_unapply(std::vector<AdjSegment>& segs) {
    for (size_t g = 1; g < segs.size(); ++g) {
        AdjSegment& s = segs[g];
        if (!s.pair.empty() || s.unapply.empty() || !all_diagonal(s.unapply)) continue;
        segs[g - 1].unapply.insert(segs[g - 1].unapply.end(), s.unapply.begin(), s.unapply.end());
        s.unapply.clear();
    }
}

// Distinct (qubit, type) pair terms of a segment, in first-seen order.
std::vector<std::pair<int, int>> distinct_pairs(const AdjSegment& seg) {
    std::vector<std::pair<int, int>> v;
    for (const PairTerm& p : seg.pair)
        if (std::find(v.begin(), v.end(), std::make_pair(p.q, p.type)) == v.end()) v.push_back({p.q, p.type});
    return v;
}

// Plan ops in place with the tile shape tuned for their structure (tuned on
// first sight, then cached in the context).
int plan_tuned_shape(qse_ctx* ctx, int n, int prec, const std::vector<qse_op>& ops, const double* params,
                     HostPlan& P, int* k_out) {
    std::vector<char> key = gate_structure_key(n, prec, ops.data(), (int)ops.size());
    for (ShapeEntry& e : ctx->shape_cache)
        if (e.key == key) {
            e.stamp = ++ctx->prog_clock;
            *k_out = e.k;
            return choose_plan(n, e.k, e.L, ops.data(), (int)ops.size(), params, P, false);
        }
    int k = 0, L = 0;
    const int rc = tuned_plan(ctx, n, prec, ops.data(), (int)ops.size(), params, P, false, &k, &L);
    if (rc) return rc;
    remember_shape(ctx, std::move(key), k, L);
    *k_out = k;
    return QSE_OK;
}

// Run ops on each of the given states (same qubit count and precision) with
// one in-place plan of the tuned shape.
int run_on(qse_ctx* ctx, std::initializer_list<qse_state*> states, const std::vector<qse_op>& ops,
           const double* params) {
    if (ops.empty()) return QSE_OK;
    qse_state* s0 = *states.begin();
    int rc;
    if (s0->n < FUSE_MIN_QUBITS) {
        for (qse_state* st : states) {
            rc = run_fused(ctx, st, ops.data(), (int)ops.size(), params, nullptr, false, false);
            if (rc) return rc;
        }
        return QSE_OK;
    }
    HostPlan P;
    int k = 0;
    rc = plan_tuned_shape(ctx, s0->n, s0->prec, ops, params, P, &k);
    if (rc) return rc;
    for (qse_state* st : states) {
        rc = run_plan(ctx, st, P, k, JIT_BY_STRUCTURE);
        if (rc) return rc;
    }
    return QSE_OK;
}

// Where one parameter set's reduction values land in the call's device result
// array: the energy, and per gradient term the value it scales.
struct SetSlots {
    int energy = -1;
    std::vector<std::pair<int, GradTerm>> terms;
};

// One parameter set: forward sweep, then the reverse sweep of the segment
// schedule. Every reduction writes its value to d_res[*next ...] on the device;
// `slots` records which value is which.
template <class R>
int adjoint_one(qse_ctx* ctx, Ws& ws, const qse_op* ops, int nops, const double* params,
                const std::vector<AdjSegment>& segs, const std::vector<qse_op>& prefix, const uint64_t* hm,
                const uint64_t* dm, const double* dw, int nterms, const ObsGroups* og, int ckpt_every, double* d_res,
                int* next, SetSlots& slots) {
    int rc = qse_state_set_basis(ctx, ws.psi, 0);
    if (rc) return rc;
    const int S = (int)segs.size();
    // forward: whole list fused, or segment by segment storing checkpoints
    std::map<int, int> ckpt_at;  // segment -> checkpoint slot holding psi at its evaluation point
    if (ckpt_every > 0 && S > 1) {
        rc = run_on(ctx, {ws.psi}, prefix, params);
        if (rc) return rc;
        int slot = 0;
        for (int g = S - 2; g >= 0; --g) {  // forward order; the last segment's gates are in the prefix
            rc = run_on(ctx, {ws.psi}, segs[g].forward, params);
            if (rc) return rc;
            if (g > 0 && ((S - 1 - g) % ckpt_every) == 0 && slot < (int)ws.ckpt.size()) {
                rc = qse_state_copy(ctx, ws.ckpt[slot], ws.psi);
                if (rc) return rc;
                ckpt_at[g] = slot++;
            }
        }
    } else {
        rc = run_on(ctx, {ws.psi}, std::vector<qse_op>(ops, ops + nops), params);
        if (rc) return rc;
    }
    slots.energy = (*next)++;
    if (og) {
        rc = obs_energy_async<R>(ctx, ws.psi, ws.lam, dm, dw, *og, d_res + slots.energy);
        if (rc) return rc;
    } else {
        rc = energy_async<R>(ctx, ws.psi, hm, dm, dw, nterms, d_res + slots.energy);
        if (rc) return rc;
        rc = qse_state_copy(ctx, ws.lam, ws.psi);
        if (rc) return rc;
        rc = launch_apply_obs<R>(ctx, ws.lam, dm, dw, nterms);
        if (rc) return rc;
    }
    // rotate into the first evaluation point's frame (X/Y generators become Z-strings)
    if (S > 0) {
        rc = run_on(ctx, {ws.psi, ws.lam}, segs[0].pre, nullptr);
        if (rc) return rc;
    }
    // Z-string masks of the current evaluation point and the terms each feeds;
    // a segment that un-applies nothing shares its point with the next one,
    // whose pair passes then take the masks along
    std::vector<uint64_t> masks;
    std::vector<std::vector<GradTerm>> mterms;
    auto land = [&](const std::vector<int>& which_masks, int first_slot) {
        for (size_t k = 0; k < which_masks.size(); ++k)
            for (const GradTerm& gt : mterms[which_masks[k]]) slots.terms.push_back({first_slot + (int)k, gt});
    };
    for (int g = 0; g < S; ++g) {
        const AdjSegment& seg = segs[g];
        auto ck = ckpt_at.find(g);
        if (ck != ckpt_at.end()) {
            rc = qse_state_copy(ctx, ws.psi, ws.ckpt[ck->second]);
            if (rc) return rc;
            rc = run_on(ctx, {ws.psi}, seg.pre, nullptr);
            if (rc) return rc;
        }
        for (const auto& dt : seg.diag) {
            const size_t pos = std::find(masks.begin(), masks.end(), dt.first) - masks.begin();
            if (pos == masks.size()) {
                masks.push_back(dt.first);
                mterms.emplace_back();
            }
            mterms[pos].push_back(dt.second);
        }
        if (g + 1 < S && seg.unapply.empty() && seg.pair.empty() && !segs[g + 1].pair.empty() &&
            ckpt_at.find(g + 1) == ckpt_at.end())
            continue;
        const std::vector<std::pair<int, int>> pt = distinct_pairs(seg);
        const size_t nfused = pt.empty() ? 0 : std::min(masks.size(), (size_t)PAIR_Z * pair_passes(pt));
        if (masks.size() > nfused) {
            // Z-strings beyond what the pair passes take: their own reduction passes
            const std::vector<uint64_t> rest(masks.begin() + nfused, masks.end());
            rc = grad_diag_async<R>(ctx, ws.psi, ws.lam, rest, d_res + *next);
            if (rc) return rc;
            std::vector<int> idx;
            for (size_t k = nfused; k < masks.size(); ++k) idx.push_back((int)k);
            land(idx, *next);
            *next += (int)rest.size();
        }
        if (!pt.empty()) {
            const std::vector<uint64_t> zf(masks.begin(), masks.begin() + nfused);
            std::vector<int> order;
            rc = pair_terms_async<R>(ctx, ws.psi, ws.lam, pt, zf, d_res + *next, order);
            if (rc) return rc;
            std::vector<int> slot_of(pt.size());
            for (size_t k = 0; k < order.size(); ++k) {
                if (order[k] >= 0) slot_of[order[k]] = *next + (int)k;
                else land({-1 - order[k]}, *next + (int)k);
            }
            for (const PairTerm& p : seg.pair) {
                const size_t i = std::find(pt.begin(), pt.end(), std::make_pair(p.q, p.type)) - pt.begin();
                slots.terms.push_back({slot_of[i], p.t});
            }
            *next += (int)order.size();
        }
        masks.clear();
        mterms.clear();
        if (g + 1 < S && !seg.unapply.empty()) {
            if (ckpt_at.find(g + 1) == ckpt_at.end()) rc = run_on(ctx, {ws.psi, ws.lam}, seg.unapply, nullptr);
            else rc = run_on(ctx, {ws.lam}, seg.unapply, nullptr);
            if (rc) return rc;
        }
    }
    return QSE_OK;
}

// Tile shape of the joint reverse sweep: 12 qubits with 3 contiguous low
// qubits; psi and lambda pass through one LDS tile in turn, so two workgroups
// of 256 threads share a CU (with a tile per state, 11-qubit complex128 tiles
// left one wave per SIMD and ran at a third of the bandwidth).
void joint_shape(int n, int prec, int* k, int* L) {
    (void)prec;
    *k = std::min(n, 12);
    *L = std::max(0, std::min(3, *k - FUSE_REG));
}

// The reverse sweep of one parameter set as joint passes over psi and lambda
// (build_reverse_plan, run-time compiled by structure, jit_joint_source): the
// forward run, the energy and lambda = O psi as in adjoint_one, then each
// chunk's passes un-apply the gates from both states and sum every probe in
// the tile at its point. Each pass's per-workgroup probe sums are reduced in
// workgroup order (k_sum_columns) into d_res[*next ...]; `slots` maps each
// value to its gradient term.
template <class R>
int adjoint_joint_one(qse_ctx* ctx, Ws& ws, const qse_op* ops, int nops, const double* params,
                      const uint64_t* hm, const uint64_t* dm, const double* dw, int nterms, const ObsGroups* og,
                      double* d_res, int* next, SetSlots& slots) {
    const int n = ws.psi->n, prec = ws.psi->prec;
    int rc = qse_state_set_basis(ctx, ws.psi, 0);
    if (rc) return rc;
    rc = run_on(ctx, {ws.psi}, std::vector<qse_op>(ops, ops + nops), params);
    if (rc) return rc;
    slots.energy = (*next)++;
    if (og) {
        rc = obs_energy_async<R>(ctx, ws.psi, ws.lam, dm, dw, *og, d_res + slots.energy);
        if (rc) return rc;
    } else {
        rc = energy_async<R>(ctx, ws.psi, hm, dm, dw, nterms, d_res + slots.energy);
        if (rc) return rc;
        rc = qse_state_copy(ctx, ws.lam, ws.psi);
        if (rc) return rc;
        rc = launch_apply_obs<R>(ctx, ws.lam, dm, dw, nterms);
        if (rc) return rc;
    }
    int k, L;
    joint_shape(n, prec, &k, &L);
    std::vector<HostPlan> chunks;
    std::vector<ProbeRef> refs;
    rc = build_reverse_plan(n, k, L, ops, nops, params, chunks, refs);
    if (rc) return rc;
    const uint64_t ntiles = 1ULL << (n - k);
    // a persistent grid: two workgroups per CU is what the tile's LDS (about
    // 78 KiB at 12 qubits) lets reside; each walks its tiles in a loop
    // (QAOA-24 joint pass 479 -> 470 us against one workgroup per tile)
    constexpr uint64_t JOINT_GROUPS_PER_CU = 2;
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * JOINT_GROUPS_PER_CU);
    for (const HostPlan& C : chunks) {
        if (C.passes.empty()) continue;
        Serialized S;
        serialize(C, S);
        ProgEntry* entry = nullptr;
        rc = program_on_device(ctx, S, prec, &entry);
        if (rc) return rc;
        const std::vector<hipFunction_t>* fns = nullptr;
        rc = jit_joint_for(ctx, C, S, prec, &fns);
        if (rc) return rc;
        int max_slots = 0, max_probes = 0;
        for (const FPass& fp : C.passes) {
            max_slots = std::max(max_slots, fp.nslots);
            max_probes = std::max(max_probes, fp.nprobes);
            if (fp.k != k) return set_error(QSE_ERR_UNSUPPORTED, "adjoint: joint pass of an unexpected tile size");
        }
        void* table = nullptr;
        if (max_slots > 0) {
            rc = scratch(ctx, 5, 32 * ntiles * (uint64_t)max_slots, &table);
            if (rc) return rc;
        }
        void* part = nullptr;
        rc = scratch(ctx, 11, sizeof(double) * (size_t)grid * (size_t)std::max(max_probes, 1), &part);
        if (rc) return rc;
        const char* dev = (const char*)entry->dev;
        for (int p = 0; p < (int)C.passes.size(); ++p) {
            const FPass& fp = C.passes[p];
            rc = slot_table_async(ctx, dev, S, C, p, ntiles, (SlotVal*)table);
            if (rc) return rc;
            TimedLaunch t;
            launch_begin(ctx, FAM_ADJOINT, &t);
            void* pd = ws.psi->d;
            void* ld = ws.lam->d;
            unsigned long long nt = ntiles;
            const void* pv = dev;
            void* args[] = {&pd, &ld, &table, &nt, &pv, &part};
            QSE_HIP(hipModuleLaunchKernel((*fns)[p], grid, 1, 1, 1u << (k - FUSE_REG), 1, 1, 0, ctx->stream, args,
                                          nullptr));
            launch_end(ctx, &t, 4.0 * (double)amp_bytes(prec) * (double)(1ULL << n));
            if (fp.nprobes > 0) {
                k_sum_columns<<<fp.nprobes, AD_THREADS, 0, ctx->stream>>>((const double*)part, (int)grid, fp.nprobes,
                                                                         d_res + *next + fp.probe_base);
                QSE_HIP(hipGetLastError());
            }
        }
        for (size_t v = 0; v < C.probe_ids.size(); ++v) {
            const ProbeRef& r = refs[C.probe_ids[v]];
            slots.terms.push_back({*next + (int)v, GradTerm{r.param, r.coeff}});
        }
        *next += (int)C.probe_ids.size();
    }
    return QSE_OK;
}

// Values adjoint_joint_one writes per parameter set (the energy and every
// probe), and its state passes in units of one read and one write of a state
// (two per joint pass); false when the joint schedule does not apply.
bool joint_schedule(int n, int prec, const qse_op* ops, int nops, const double* params, int* values, double* cost) {
    int k, L;
    joint_shape(n, prec, &k, &L);
    if (k < FUSE_MIN_K) return false;
    std::vector<HostPlan> chunks;
    std::vector<ProbeRef> refs;
    if (build_reverse_plan(n, k, L, ops, nops, params, chunks, refs) != QSE_OK) return false;
    int v = 1, passes = 0;
    for (const HostPlan& C : chunks) {
        v += (int)C.probe_ids.size();
        passes += (int)C.passes.size();
    }
    *values = v;
    *cost = 2.0 * passes;
    return true;
}

// Number of reduction values adjoint_one writes for a schedule.
int values_per_set(const std::vector<AdjSegment>& segs) {
    int v = 1;
    for (const AdjSegment& seg : segs) {
        std::vector<uint64_t> masks;
        for (const auto& dt : seg.diag)
            if (std::find(masks.begin(), masks.end(), dt.first) == masks.end()) masks.push_back(dt.first);
        v += (int)masks.size() + (int)distinct_pairs(seg).size();
    }
    return v;
}

// Passes over psi and lambda of a schedule's reverse sweep, in units of one
// read and one write of a state: un-applications (twice, psi and lambda) and
// the rotation into the first frame, plus the read-only reductions (Z-string
// passes of up to 64 masks, pair passes of up to 9 or 12 qubits, each reading
// both states).
double sweep_cost(int n, const std::vector<AdjSegment>& segs) {
    double c = 0.0;
    auto passes = [&](const std::vector<qse_op>& ops) -> int {
        if (ops.empty()) return 0;
        HostPlan P;
        if (choose_plan(n, PAIR_K, 3, ops.data(), (int)ops.size(), nullptr, P, false)) return 1 << 20;
        return (int)P.passes.size();
    };
    if (!segs.empty()) c += 2.0 * passes(segs[0].pre);
    for (const AdjSegment& seg : segs) {
        c += 2.0 * passes(seg.unapply);
        std::vector<uint64_t> masks;
        for (const auto& dt : seg.diag)
            if (std::find(masks.begin(), masks.end(), dt.first) == masks.end()) masks.push_back(dt.first);
        c += (double)((masks.size() + 63) / 64);
        c += pair_passes(distinct_pairs(seg));
    }
    return c;
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
    // X/Y generators: rotated into Z-strings, or reduced as amplitude pairs
    // when that makes fewer passes (rotations that no un-application absorbs)
    auto schedule = [&](const double* pb, bool pm, std::vector<AdjSegment>& sg, std::vector<qse_op>& pf) {
        const int r = adjoint_segments(ops, nops, pb, sg, pf, pm);
        if (r == QSE_OK && pm && ckpt_every <= 0) merge_diagonal_unapply(sg);
        return r;
    };
    std::vector<AdjSegment> segs0;
    std::vector<qse_op> prefix;
    rc = schedule(params, false, segs0, prefix);
    if (rc) return rc;
    bool pair_mode = false;
    if (nqubits >= WHT_MIN_QUBITS && nqubits > PAIR_K && ctx->adjoint_terms != QSE_ADJ_TERMS_ROTATE) {
        std::vector<AdjSegment> segsP;
        std::vector<qse_op> prefixP;
        rc = schedule(params, true, segsP, prefixP);
        if (rc) return rc;
        if (ctx->adjoint_terms == QSE_ADJ_TERMS_PAIRS || sweep_cost(nqubits, segsP) < sweep_cost(nqubits, segs0)) {
            pair_mode = true;
            segs0.swap(segsP);
            prefix.swap(prefixP);
        }
    }
    // the joint schedule: the reverse sweep as run-time compiled joint passes
    // over psi and lambda, when asked for, or when it moves fewer state passes
    bool joint = false;
    int jvals = 0;
    if (ctx->adjoint_terms == QSE_ADJ_TERMS_JOINT || ctx->adjoint_terms == QSE_ADJ_TERMS_AUTO) {
        double jcost = 0.0;
        const bool ok = ckpt_every <= 0 && ctx->jit_mode != QSE_JIT_OFF && nparams > 0 &&
                        joint_schedule(nqubits, precision, ops, nops, params, &jvals, &jcost);
        if (ctx->adjoint_terms == QSE_ADJ_TERMS_JOINT && !ok)
            return set_error(QSE_ERR_UNSUPPORTED, "adjoint: the joint schedule needs parameters, run-time "
                                                  "compilation and no checkpoints");
        joint = ok && (ctx->adjoint_terms == QSE_ADJ_TERMS_JOINT ||
                       (nqubits >= WHT_MIN_QUBITS && jcost < sweep_cost(nqubits, segs0)));
    }
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
    // lambda = O psi and the energy in one pass (k_obs_energy) take the terms
    // grouped by the top three mask bits; the host copies stay alive until the
    // call's final wait
    ObsGroups og{};
    const ObsGroups* ogp = nullptr;
    std::vector<uint64_t> sm;
    std::vector<double> sw;
    const uint64_t* up_m = obs_masks;
    const double* up_w = obs_weights;
    if (nqubits >= 3) {
        for (int u = 0; u < 8; ++u) {
            og.ubeg[u] = (int)sm.size();
            for (int t = 0; t < nterms; ++t)
                if ((int)((obs_masks[t] >> (nqubits - 3)) & 7) == u) {
                    sm.push_back(obs_masks[t]);
                    sw.push_back(obs_weights[t]);
                }
        }
        og.ubeg[8] = (int)sm.size();
        ogp = &og;
        up_m = sm.data();
        up_w = sw.data();
    }
    if (nterms > 0) {
        QSE_HIP(hipMemcpyAsync(dm, up_m, 8 * nterms, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(dw, up_w, 8 * nterms, hipMemcpyHostToDevice, ctx->stream));
    }
    // every reduction value of the call lands in one device array, read once at the end
    const int vps = joint ? jvals : values_per_set(segs0);
    void* rbuf;
    rc = scratch(ctx, 8, sizeof(double) * (size_t)vps * batch, &rbuf);
    if (rc) return rc;
    double* d_res = (double*)rbuf;
    std::vector<SetSlots> slots(batch);
    int next = 0;
    for (int b = 0; b < batch; ++b) {
        const double* pb = params ? params + (size_t)b * nparams : nullptr;
        if (joint) {
            const int start = next;
            rc = (precision == QSE_C128)
                     ? adjoint_joint_one<double>(ctx, ws, ops, nops, pb, obs_masks, dm, dw, nterms, ogp, d_res, &next,
                                                 slots[b])
                     : adjoint_joint_one<float>(ctx, ws, ops, nops, pb, obs_masks, dm, dw, nterms, ogp, d_res, &next,
                                                slots[b]);
            if (rc) return rc;
            if (next - start != vps) return set_error(QSE_ERR_UNSUPPORTED, "adjoint: schedule changed between parameter sets");
            continue;
        }
        std::vector<AdjSegment> segs;
        rc = schedule(pb, pair_mode, segs, prefix);
        if (rc) return rc;
        if (values_per_set(segs) != vps)
            return set_error(QSE_ERR_UNSUPPORTED, "adjoint: schedule changed between parameter sets");
        rc = (precision == QSE_C128)
                 ? adjoint_one<double>(ctx, ws, ops, nops, pb, segs, prefix, obs_masks, dm, dw, nterms, ogp, ckpt_every, d_res, &next,
                                       slots[b])
                 : adjoint_one<float>(ctx, ws, ops, nops, pb, segs, prefix, obs_masks, dm, dw, nterms, ogp, ckpt_every, d_res, &next,
                                      slots[b]);
        if (rc) return rc;
    }
    void* pin;
    rc = pinned(ctx, sizeof(double) * (size_t)std::max(next, 1), &pin);
    if (rc) return rc;
    QSE_HIP(hipMemcpyAsync(pin, d_res, sizeof(double) * (size_t)next, hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    const double* vals = (const double*)pin;
    for (int b = 0; b < batch; ++b) {
        energies[b] = vals[slots[b].energy];
        if (!grads) continue;
        double* gb = grads + (size_t)b * nparams;
        std::fill(gb, gb + nparams, 0.0);
        for (const auto& st : slots[b].terms) gb[st.second.param] += st.second.coeff * vals[st.first];
    }
    return QSE_OK;
}

int qse_ctx_set_adjoint_terms(qse_ctx* ctx, int mode) {
    if (!ctx || mode < QSE_ADJ_TERMS_AUTO || mode > QSE_ADJ_TERMS_JOINT)
        return set_error(QSE_ERR_INVALID_ARG, "bad adjoint term mode");
    ctx->adjoint_terms = mode;
    return QSE_OK;
}

int qse_batch_expect(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, const double* params,
                     int nparams, int batch, const uint64_t* obs_masks, const double* obs_weights, int nterms,
                     double* energies) {
    int rc = check_common(ctx, nqubits, precision, ops, nops, params, nparams, batch, obs_masks, obs_weights, nterms);
    if (rc) return rc;
    if (!energies) return set_error(QSE_ERR_INVALID_ARG, "null output");
    Ws ws;
    ws.ctx = ctx;
    rc = qse_state_create(ctx, nqubits, precision, &ws.psi);
    if (rc) return rc;
    // observable and the per-set energies on the device, read once at the end
    void* obuf;
    const size_t mpad = ((size_t)std::max(nterms, 1) + 31) / 32 * 32;
    rc = scratch(ctx, 7, 16 * mpad + 8 * (size_t)batch + 256, &obuf);
    if (rc) return rc;
    uint64_t* dm = (uint64_t*)obuf;
    double* dw = (double*)((char*)obuf + 8 * mpad);
    double* d_res = dw + mpad;
    if (nterms > 0) {
        QSE_HIP(hipMemcpyAsync(dm, obs_masks, 8 * nterms, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipMemcpyAsync(dw, obs_weights, 8 * nterms, hipMemcpyHostToDevice, ctx->stream));
    }
    const std::vector<qse_op> list(ops, ops + nops);
    for (int b = 0; b < batch; ++b) {
        const double* pb = params ? params + (size_t)b * nparams : nullptr;
        rc = qse_state_set_basis(ctx, ws.psi, 0);
        if (rc) return rc;
        rc = run_on(ctx, {ws.psi}, list, pb);
        if (rc) return rc;
        rc = (precision == QSE_C128) ? energy_async<double>(ctx, ws.psi, obs_masks, dm, dw, nterms, d_res + b)
                                     : energy_async<float>(ctx, ws.psi, obs_masks, dm, dw, nterms, d_res + b);
        if (rc) return rc;
    }
    void* pin;
    rc = pinned(ctx, sizeof(double) * (size_t)batch, &pin);
    if (rc) return rc;
    QSE_HIP(hipMemcpyAsync(pin, d_res, sizeof(double) * (size_t)batch, hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(energies, pin, sizeof(double) * (size_t)batch);
    return QSE_OK;
}

}  // extern "C"

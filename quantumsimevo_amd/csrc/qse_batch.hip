// qse_batch.hip — batched small-circuit engine (qse_batch_run_small): many
// circuits of uneven qubit counts, forward run, observable and adjoint
// gradients, in one persistent launch.
//
// Circuits are taken from device work queues (in decreasing order of estimated
// work) and each circuit's whole step list (qse_batch.hpp) is executed by one
// team: fused passes over the tiles of the circuit's state in HBM (the
// LDS-tiled pass of qse_fused.hip, in place), the observable, and FP64
// reductions for the energy and the generator terms. Circuits of 16 qubits or
// more are taken by teams of 16 workgroups that split every pass's tiles and
// every reduction's indices and meet at team barriers (global counters); the
// others by single workgroups, whose own barriers are the only
// synchronisation. Reductions sum per-workgroup partials in rank order
// (deterministic). Tiles of k < KMAX qubits use the first 2^(k-4) threads; the
// others wait at the barriers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "qse_batch.hpp"
#include "qse_device.hpp"
#include "qse_fused_dev.hpp"
#include "qse_internal.hpp"

namespace qse {

namespace {

// Tile qubits of the engine's largest tile (both precisions): 32 or 64 KiB of
// LDS per workgroup, two workgroups per CU.
constexpr int BATCH_KMAX = 12;
constexpr int BATCH_THREADS = 1 << (BATCH_KMAX - 4);
// External (tile-index) qubits are deposited in groups of BATCH_DEP_BITS from
// per-pass LDS tables: 4 groups cover the 22 external qubits of a 30-qubit
// circuit with 8-qubit tiles.
constexpr int BATCH_DEP_BITS = 6;
constexpr int BATCH_DEP_LEVELS = 4;
constexpr int BATCH_DEP_HALF = BATCH_DEP_LEVELS << BATCH_DEP_BITS;
static_assert(30 - FUSE_MIN_K <= BATCH_DEP_LEVELS * BATCH_DEP_BITS,
              "the deposit tables cover the external qubits of the largest batch circuit");

struct BArgs {
    const FPass* __restrict__ passes;
    const FStage* __restrict__ stages;
    const FOp* __restrict__ ops;
    const FTerm* __restrict__ terms;
    const FSlot* __restrict__ slots;
    const FExtTerm* __restrict__ eterms;
    const double* __restrict__ alphas;
};

// The program (passes, stages, ops, terms, factors) is read-only for the whole
// launch: read through the constant address space, its uniform fields come in
// with scalar loads. (The host pass of the compiler, which never runs this
// code, knows no such address space and keeps the pointer as it is.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* kc(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}
#else
template <class T>
__device__ __forceinline__ const T* kc(const T* p) {
    return p;
}
#endif

// One in-place pass of tile size K over the 2^n amplitudes at s: tiles r, r + T, ...
// (this workgroup's share when T workgroups work on the circuit).
template <class R, int K>
__device__ void batch_pass(cx<R>* s, int n, const FPass& Pg, const BArgs& A, cx<R>* tile, SlotVal* slotv,
                           uint64_t* dep, int r, int T) {
    const auto& P = *kc(&Pg);
    constexpr int NT = 1 << (K - 4);
    constexpr int TB = K - 4;
    const int tid = threadIdx.x;
    const bool act = tid < NT;
    const uint64_t ntiles = 1ULL << (n - K);
    // Tile index -> source / destination base addresses: the external qubits in
    // groups of 6 tile-index bits, each group's 64 deposits tabulated once per
    // pass in LDS (dep[0..255]: source, dep[256..511]: destination). The
    // previous pass's readers have all passed a workgroup barrier.
    const int nlev = (P.next + BATCH_DEP_BITS - 1) / BATCH_DEP_BITS;
    for (int idx = tid; idx < nlev << BATCH_DEP_BITS; idx += blockDim.x) {
        const int lvl = idx >> BATCH_DEP_BITS, j = idx & ((1 << BATCH_DEP_BITS) - 1);
        uint64_t b = 0, o = 0;
        for (int u = 0; u < BATCH_DEP_BITS; ++u) {
            const int e = lvl * BATCH_DEP_BITS + u;
            if (e < P.next && ((j >> u) & 1)) {
                b |= 1ULL << P.epos[e];
                o |= 1ULL << P.eout[e];
            }
        }
        dep[idx] = b;
        dep[BATCH_DEP_HALF + idx] = o;
    }
    __syncthreads();
    // Complex64: the next tile's amplitudes are loaded into registers while
    // this tile's stages run (its 16 amplitudes fit the register budget of two
    // waves per SIMD; complex128's would not). A workgroup's tiles are its own
    // in the pass, so the prefetch never reads what another one writes.
    constexpr bool PF = sizeof(R) == 4;
    auto load_tile = [&](const FPass& Q, uint64_t t, int tv, cx<R>* out) {
        uint64_t b = 0, g0 = 0;
        for (int lvl = 0; lvl < nlev; ++lvl)
            b |= dep[(int)((t >> (lvl * BATCH_DEP_BITS)) & ((1 << BATCH_DEP_BITS) - 1)) + (lvl << BATCH_DEP_BITS)];
#pragma unroll
        for (int q = 0; q < TB; ++q)
            if ((tv >> q) & 1) g0 |= 1ULL << Q.tq[q];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint64_t g = b | g0;
            if (i & 1) g |= 1ULL << Q.tq[TB + 0];
            if (i & 2) g |= 1ULL << Q.tq[TB + 1];
            if (i & 4) g |= 1ULL << Q.tq[TB + 2];
            if (i & 8) g |= 1ULL << Q.tq[TB + 3];
            out[i] = s[g];
        }
    };
    cx<R> ld[16];
    if (PF && act && (uint64_t)r < ntiles) load_tile(P, (uint64_t)r, tid, ld);
    for (uint64_t ti = (uint64_t)r; ti < ntiles; ti += (uint64_t)T) {
        // The pass's fields and this thread's tile offsets are read and computed
        // again for every tile: the pass pointer and the thread id go through
        // empty asm barriers, so the compiler cannot hoist them out of the tile
        // loop, where they would hold scalar and vector registers for its whole
        // length (and spill).
        // (The pointer is uniform: readfirstlane places it in scalar registers
        // for the barrier whatever register class its producer used.)
        const uint64_t pw = (uint64_t)&Pg;
        uint32_t plo = __builtin_amdgcn_readfirstlane((uint32_t)pw), phi = __builtin_amdgcn_readfirstlane((uint32_t)(pw >> 32));
        __asm__ volatile("" : "+s"(plo), "+s"(phi));
        const FPass* Pp = (const FPass*)(((uint64_t)phi << 32) | plo);
        int tidv = tid;
        __asm__ volatile("" : "+v"(tidv));
        const auto& P = *kc(Pp);
        uint64_t gt = 0, ot = 0;
        uint32_t st_t = 0;
#pragma unroll
        for (int b = 0; b < TB; ++b)
            if ((tidv >> b) & 1) {
                gt |= 1ULL << P.tq[b];
                ot |= 1ULL << P.so_q[b];
                st_t |= 1u << P.so_l[b];
            }
        const uint32_t swz_tid = swz((uint32_t)tidv), swz_st = swz(st_t);
        uint64_t base = 0, obase = 0;
        for (int lvl = 0; lvl < nlev; ++lvl) {
            const int j = (int)((ti >> (lvl * BATCH_DEP_BITS)) & ((1 << BATCH_DEP_BITS) - 1)) + (lvl << BATCH_DEP_BITS);
            base |= dep[j];
            obase |= dep[BATCH_DEP_HALF + j];
        }
        if (!PF && act) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                uint64_t g = base | gt;
                if (i & 1) g |= 1ULL << P.tq[TB + 0];
                if (i & 2) g |= 1ULL << P.tq[TB + 1];
                if (i & 4) g |= 1ULL << P.tq[TB + 2];
                if (i & 8) g |= 1ULL << P.tq[TB + 3];
                ld[i] = s[g];
            }
        }
        // per-tile external factors: angle sum of each slot's terms whose qubits are set
        for (int sl = tid; sl < P.nslots; sl += blockDim.x) {
            const FSlot S = A.slots[P.slot_begin + sl];
            double a = 0.0;
            for (uint32_t e = S.begin; e < S.end; ++e)
                if ((base & A.eterms[e].emask) == A.eterms[e].emask) a += A.eterms[e].ang;
            double sn, cs;
            sincos(a, &sn, &cs);
            slotv[sl] = {a, cs, sn, 0.0};
        }
        if (act) {
#pragma unroll
            for (int i = 0; i < 16; ++i) tile[swz_tid ^ swz((uint32_t)(i * NT))] = ld[i];
        }
        __syncthreads();
        if (PF && act && ti + (uint64_t)T < ntiles) load_tile(P, ti + (uint64_t)T, tidv, ld);
        for (int sg = 0; sg < P.nstages; ++sg) {
            const auto& S = *kc(A.stages + P.stage_begin + sg);
            if (act) {
                uint32_t toff = 0;
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    if ((tid >> t) & 1) toff |= 1u << S.tmap[t];
                const uint32_t r0 = swz(1u << S.rmap[0]), r1 = swz(1u << S.rmap[1]), r2 = swz(1u << S.rmap[2]),
                               r3 = swz(1u << S.rmap[3]);
                const uint32_t st_off = swz(toff);
                cx<R> a[16];
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    a[j] = tile[st_off ^ ((j & 1) ? r0 : 0u) ^ ((j & 2) ? r1 : 0u) ^ ((j & 4) ? r2 : 0u) ^
                                ((j & 8) ? r3 : 0u)];
                for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
                    const FOp op = *kc(A.ops + oi);
                    if (op.type == FOP_MONO) {
                        double gr, gi;
                        mono_factor<TB>(op, kc(A.terms), kc(A.alphas), tid, toff, base, slotv, gr, gi);
                        const cx<R> g = {(R)gr, (R)gi};
                        const uint32_t rctl = op.rctl;
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            if (((uint32_t)j & rctl) == rctl) a[j] = cmul(g, a[j]);
                    } else if (op.type == FOP_H) {
                        switch (op.p) {
                            case 0: hadamard_p<R, 0>(a); break;
                            case 1: hadamard_p<R, 1>(a); break;
                            case 2: hadamard_p<R, 2>(a); break;
                            default: hadamard_p<R, 3>(a); break;
                        }
                    } else if (op.type == FOP_REAL) {
                        switch (op.p) {
                            case 0: real_p<R, 0>(a, op); break;
                            case 1: real_p<R, 1>(a, op); break;
                            case 2: real_p<R, 2>(a, op); break;
                            default: real_p<R, 3>(a, op); break;
                        }
                    } else if (op.type == FOP_RXLIKE) {
                        switch (op.p) {
                            case 0: rxlike_p<R, 0>(a, op); break;
                            case 1: rxlike_p<R, 1>(a, op); break;
                            case 2: rxlike_p<R, 2>(a, op); break;
                            default: rxlike_p<R, 3>(a, op); break;
                        }
                    } else if (op.type == FOP_XSWAP) {
                        // X / CNOT target: an exchange of registers (controls: register bits per
                        // pair, thread bits and external qubits for the whole thread)
                        if (((toff & op.tctl) == op.tctl) && ((base & op.ectl) == op.ectl)) {
                            if (op.rctl == 0) {
                                switch (op.p) {
                                    case 0: xswap_p<R, 0>(a); break;
                                    case 1: xswap_p<R, 1>(a); break;
                                    case 2: xswap_p<R, 2>(a); break;
                                    default: xswap_p<R, 3>(a); break;
                                }
                            } else {
                                switch (op.p) {
                                    case 0: xswap_ctl_p<R, 0>(a, op.rctl); break;
                                    case 1: xswap_ctl_p<R, 1>(a, op.rctl); break;
                                    case 2: xswap_ctl_p<R, 2>(a, op.rctl); break;
                                    default: xswap_ctl_p<R, 3>(a, op.rctl); break;
                                }
                            }
                        }
                    } else if (((toff & op.tctl) == op.tctl) && ((base & op.ectl) == op.ectl)) {
                        switch (op.p) {
                            case 0: dense_p<R, 0>(a, op); break;
                            case 1: dense_p<R, 1>(a, op); break;
                            case 2: dense_p<R, 2>(a, op); break;
                            default: dense_p<R, 3>(a, op); break;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    tile[st_off ^ ((j & 1) ? r0 : 0u) ^ ((j & 2) ? r1 : 0u) ^ ((j & 4) ? r2 : 0u) ^
                         ((j & 8) ? r3 : 0u)] = a[j];
            }
            __syncthreads();
        }
        if (act) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                uint64_t g = obase | ot;
                uint32_t ls = swz_st;
                if (i & 1) { g |= 1ULL << P.so_q[TB + 0]; ls ^= swz(1u << P.so_l[TB + 0]); }
                if (i & 2) { g |= 1ULL << P.so_q[TB + 1]; ls ^= swz(1u << P.so_l[TB + 1]); }
                if (i & 4) { g |= 1ULL << P.so_q[TB + 2]; ls ^= swz(1u << P.so_l[TB + 2]); }
                if (i & 8) { g |= 1ULL << P.so_q[TB + 3]; ls ^= swz(1u << P.so_l[TB + 3]); }
                s[g] = tile[ls];
            }
        }
        __syncthreads();
    }
}

// Inlined at its one call site: there the compiler sees that the state and the
// program live in global memory, and reads the program's uniform fields with
// scalar loads (an outlined call takes generic pointers and reads every field
// with a per-lane flat load).
template <class R>
__device__ __attribute__((always_inline)) void batch_pass_k(int k, cx<R>* s, int n, const FPass& P, const BArgs& A, cx<R>* tile, SlotVal* slotv, uint64_t* dep,
                             int r, int T) {
    switch (k) {
        case 8: batch_pass<R, 8>(s, n, P, A, tile, slotv, dep, r, T); break;
        case 9: batch_pass<R, 9>(s, n, P, A, tile, slotv, dep, r, T); break;
        case 10: batch_pass<R, 10>(s, n, P, A, tile, slotv, dep, r, T); break;
        case 11: batch_pass<R, 11>(s, n, P, A, tile, slotv, dep, r, T); break;
        default: batch_pass<R, 12>(s, n, P, A, tile, slotv, dep, r, T); break;
    }
}

// Sum of v over the workgroup (every thread calls; thread 0 gets the result).
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    __syncthreads();
    return t;
}

// Reductions and observables run over blocks of WB = 8 amplitudes
// i = i0 + stride j (j < 8), whose sums against Z-strings come from one
// in-register Walsh-Hadamard transform of the block.
constexpr int WB = 8, WB_LOG = 3;

// Observable terms of the running circuit, in LDS: masks and weights of the
// kept terms (their qubits lie on the circuit), grouped by the 3-bit field
// u = (mask >> log2(stride)) & 7 that selects a Walsh-Hadamard coefficient
// (group u: entries [ubeg[u], ubeg[u + 1])); built once per circuit.
struct ObsLds {
    uint64_t m[BATCH_MAX_TERMS];
    double w[BATCH_MAX_TERMS];
    int ubeg[WB + 1];
    int n;
    // Z-string step: bit d of kflip[b] = bit (b + log2(WB stride)) of mask d
    uint32_t kflip[64];
};

__device__ __forceinline__ double obs_factor(uint64_t i, const ObsLds& O) {
    double f = 0.0;
    for (int t = 0; t < O.n; ++t) f += (__popcll(i & O.m[t]) & 1) ? -O.w[t] : O.w[t];
    return f;
}

// In-register Walsh-Hadamard transform of WB values: w[u] <- sum_j w[j] (-1)^popcount(j & u).
__device__ __forceinline__ void wht(double* w) {
#pragma unroll
    for (int h = 1; h < WB; h <<= 1)
#pragma unroll
        for (int j = 0; j < WB; ++j)
            if (!(j & h)) {
                const double a = w[j], b = w[j + h];
                w[j] = a + b;
                w[j + h] = a - b;
            }
}

// G[u] = sum over the terms of group u of w_t (-1)^parity(i0 & m_t): with
// i0 = first + WB stride kbi and i = i0 + stride j, f(i) = sum_u G[u]
// (-1)^popcount(j & u) (first < stride, so the bits of first, of WB stride kbi
// and of stride j are disjoint).
__device__ __forceinline__ void obs_groups(const ObsLds& O, uint64_t i0, double* G) {
#pragma unroll
    for (int u = 0; u < WB; ++u) {
        double g = 0.0;
        for (int t = O.ubeg[u]; t < O.ubeg[u + 1]; ++t) g += (__popcll(i0 & O.m[t]) & 1) ? -O.w[t] : O.w[t];
        G[u] = g;
    }
}

// Bound on the iterations of one barrier wait (each sleeps ~128 cycles): about
// a second, far beyond any legitimate wait for a teammate's share of a step.
constexpr uint32_t TEAM_SPIN_LIMIT = 1u << 23;

// T workgroups working on one circuit (T = 1: a single workgroup).
struct Team {
    int id, r, T;  // team index, rank of this workgroup in it, team size
    int* bar;      // the team's barrier counter (global memory)
    int gen;       // barriers passed (thread 0's copy is the one used)
    int* err;      // set when a wait times out
};

// Barrier over the team's workgroups (just a workgroup barrier when T = 1):
// this workgroup's global writes are released before arriving and the
// teammates' acquired after. Returns false when the wait timed out or another
// workgroup reported a failure; the caller then leaves the kernel.
__device__ bool team_sync(Team& tm, int* s_flag) {
    __syncthreads();
    if (tm.T == 1) return true;
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(tm.bar, 1);
        const int target = (++tm.gen) * tm.T;
        int ok = 1;
        for (uint32_t spin = 0; __hip_atomic_load(tm.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
             ++spin) {
            if (spin > TEAM_SPIN_LIMIT || __hip_atomic_load(tm.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                atomicExch(tm.err, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        *s_flag = ok;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_flag) != 0;  // one LDS word for all lanes
}

struct BKernelArgs {
    cx<double>* arena_d;  // reinterpret for the precision
    const BCirc* circs;
    const int32_t* order;
    const BStep* steps;
    BArgs prog;
    const uint64_t* obs_m;
    const double* obs_w;
    int nterms;
    double* tpart;  // per team: T x BATCH_DIAG partial sums
    double* vals;
};

// One circuit's step list by a team (or a single workgroup). False: a team
// barrier failed.
template <class R>
__device__ __attribute__((always_inline)) bool run_circuit(const BCirc& C, Team& tm, const BKernelArgs& K, cx<R>* tile, SlotVal* slotv,
                            uint64_t* dep, double* red, ObsLds& O, int* s_flag) {
    const int tid = threadIdx.x, nth = blockDim.x;
    cx<R>* arena = (cx<R>*)K.arena_d;
    cx<R>* psi = arena + C.psi_off;
    cx<R>* lam = arena + C.lam_off;
    const uint64_t N = 1ULL << C.n;
    const uint64_t first = (uint64_t)tm.r * nth + tid, stride = (uint64_t)tm.T * nth;
    double* part = K.tpart + (size_t)tm.id * tm.T * BATCH_DIAG;
    // Reductions and the observable run over blocks of WB strides when the
    // state holds at least WB strides; below, element by element.
    const int log2S = __builtin_ctzll(stride);
    const bool blocked = (stride & (stride - 1)) == 0 && N >= WB * stride;
    const uint64_t nblk = blocked ? N / (WB * stride) : 0;
    // the circuit's observable terms in LDS, grouped by their Walsh-Hadamard field
    if (tid == 0) {
        int cnt[WB] = {};
        for (int t = 0; t < K.nterms; ++t)
            if ((C.obs_keep >> t) & 1) ++cnt[(int)((K.obs_m[t] >> log2S) & (WB - 1))];
        O.ubeg[0] = 0;
        for (int u = 0; u < WB; ++u) O.ubeg[u + 1] = O.ubeg[u] + cnt[u];
        for (int u = 0; u < WB; ++u) cnt[u] = O.ubeg[u];
        for (int t = 0; t < K.nterms; ++t)
            if ((C.obs_keep >> t) & 1) {
                const int u = (int)((K.obs_m[t] >> log2S) & (WB - 1));
                O.m[cnt[u]] = K.obs_m[t];
                O.w[cnt[u]++] = K.obs_w[t];
            }
        O.n = O.ubeg[WB];
    }
    __syncthreads();
    for (int si = C.step_begin; si < C.step_end; ++si) {
        const BStep& S = K.steps[si];
        const int kind = S.kind;
        if (kind == BS_BASIS) {
            for (uint64_t i = first; i < N; i += stride) psi[i] = {(R)(i == 0 ? 1 : 0), (R)0};
        } else if (kind == BS_PROG) {
            // pass p on psi (target bit 0), then on lambda (bit 1): one call site
            const int nt = S.pass_end > S.pass_begin ? 2 * (S.pass_end - S.pass_begin) : 0;
#pragma unroll 1
            for (int it = 0; it < nt; ++it) {
                const int which = it & 1;
                if (!((S.target >> which) & 1)) continue;
                const FPass& P = K.prog.passes[S.pass_begin + (it >> 1)];
                batch_pass_k<R>(C.k, which ? lam : psi, C.n, P, K.prog, tile, slotv, dep, tm.r, tm.T);
                if (!team_sync(tm, s_flag)) return false;
            }
            continue;
        } else if (kind == BS_OBS) {
            if (blocked) {
                for (uint64_t kbi = 0; kbi < nblk; ++kbi) {
                    const uint64_t i0 = first + WB * stride * kbi;
                    cx<R> a[WB];
#pragma unroll
                    for (int j = 0; j < WB; ++j) a[j] = psi[i0 + stride * j];
                    double f[WB];
                    obs_groups(O, i0, f);
                    wht(f);  // f[j] = f(i0 + stride j)
#pragma unroll
                    for (int j = 0; j < WB; ++j)
                        lam[i0 + stride * j] = {(R)(f[j] * (double)a[j].re), (R)(f[j] * (double)a[j].im)};
                }
            } else {
                for (uint64_t i = first; i < N; i += stride) {
                    const double f = obs_factor(i, O);
                    const cx<R> a = psi[i];
                    lam[i] = {(R)(f * (double)a.re), (R)(f * (double)a.im)};
                }
            }
        } else {  // BS_ENERGY, BS_GDIAG: partial sums of this workgroup, then the team's in rank order
            const int nv = (kind == BS_ENERGY) ? 1 : S.nmask;
            double acc[BATCH_DIAG];
#pragma unroll
            for (int d = 0; d < BATCH_DIAG; ++d) acc[d] = 0.0;
            if (kind == BS_ENERGY && blocked) {
                // sum_j p_j f_j = sum_u G[u] W[u], W = transform of the block's probabilities
                for (uint64_t kbi = 0; kbi < nblk; ++kbi) {
                    const uint64_t i0 = first + WB * stride * kbi;
                    double w[WB], G[WB];
#pragma unroll
                    for (int j = 0; j < WB; ++j) w[j] = prob(psi[i0 + stride * j]);
                    wht(w);
                    obs_groups(O, i0, G);
#pragma unroll
                    for (int u = 0; u < WB; ++u) acc[0] += G[u] * w[u];
                }
            } else if (kind == BS_ENERGY) {
                for (uint64_t i = first; i < N; i += stride) acc[0] += prob(psi[i]) * obs_factor(i, O);
            } else if (blocked) {
                // Z-string d over the block = (-1)^parity(i0 & m_d) W[(m_d >> log2 stride) & 7], W the
                // block's transform, looked up through this thread's column of the (idle) tile buffer.
                // parity(i0 & m_d) = parity(first & m_d) (bit d of cbits) ^ parity(WB stride kbi & m_d).
                // Blocks go in Gray-code order g = kbi ^ (kbi >> 1), which flips bit ctz(kbi) of g per
                // block, so the block part of the signs follows from one kflip word each.
                double* hs = (double*)tile;
                {
                    const int at = (tid & 63) + log2S + WB_LOG;
                    uint32_t b = 0;
#pragma unroll
                    for (int d = 0; d < BATCH_DIAG; ++d)
                        b |= (uint32_t)(at < 64 ? (S.mask[d] >> (at & 63)) & 1 : 0) << d;
                    if (tid < 64) O.kflip[tid] = b;
                }
                uint32_t flip = 0;  // bit d: parity(i0 & mask d)
                for (int d = 0; d < S.nmask; ++d) flip |= (uint32_t)(__popcll(first & S.mask[d]) & 1) << d;
                uint32_t upk[BATCH_DIAG / 8];  // field u of mask d in bits 4 (d & 7) of word d >> 3
#pragma unroll
                for (int q = 0; q < BATCH_DIAG / 8; ++q) {
                    upk[q] = 0;
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        upk[q] |= (uint32_t)((S.mask[8 * q + e] >> log2S) & (WB - 1)) << (4 * e);
                }
                __syncthreads();
                for (uint64_t kbi = 0; kbi < nblk; ++kbi) {
                    if (kbi) flip ^= O.kflip[__builtin_ctzll(kbi)];
                    const uint64_t i0 = first + WB * stride * (kbi ^ (kbi >> 1));
                    double w[WB];
#pragma unroll
                    for (int j = 0; j < WB; ++j) {
                        const cx<R> p = psi[i0 + stride * j], l = lam[i0 + stride * j];
                        w[j] = (double)l.re * (double)p.im - (double)l.im * (double)p.re;
                    }
                    wht(w);
#pragma unroll
                    for (int u = 0; u < WB; ++u) hs[u * nth + tid] = w[u];
#pragma unroll
                    for (int d = 0; d < BATCH_DIAG; ++d) {
                        const int u = (int)((upk[d >> 3] >> (4 * (d & 7))) & (WB - 1));
                        const double v = hs[u * nth + tid];
                        acc[d] += ((flip >> d) & 1) ? -v : v;
                    }
                }
            } else {
                for (uint64_t i = first; i < N; i += stride) {
                    const cx<R> p = psi[i], l = lam[i];
                    const double w = (double)l.re * (double)p.im - (double)l.im * (double)p.re;
#pragma unroll
                    for (int d = 0; d < BATCH_DIAG; ++d) acc[d] += (__popcll(i & S.mask[d]) & 1) ? -w : w;
                }
            }
            const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
            for (int d = 0; d < BATCH_DIAG; ++d) {
                const double v = wave_sum(acc[d]);
                if (lane == 0) red[wave * BATCH_DIAG + d] = v;
            }
            __syncthreads();
            if (tid < nv) {
                double v = 0.0;
                for (int w = 0; w < nth / 64; ++w) v += red[w * BATCH_DIAG + tid];
                if (tm.T == 1) K.vals[S.out + tid] = v;
                else part[tm.r * BATCH_DIAG + tid] = v;
            }
            if (tm.T > 1) {
                if (!team_sync(tm, s_flag)) return false;
                if (tm.r == 0 && tid < nv) {
                    double v = 0.0;
                    for (int q = 0; q < tm.T; ++q) v += part[q * BATCH_DIAG + tid];
                    K.vals[S.out + tid] = v;
                }
            }
        }
        // the step's writes (and, for reductions, the leader's reads of the partials) before the next step
        if (!team_sync(tm, s_flag)) return false;
    }
    return true;
}

// Persistent engine. Phase 1: teams of T workgroups take the first nbig
// circuits of `order` from queue qctr[0]; phase 2: every workgroup alone takes
// the remaining circuits from queue qctr[1]. Launched cooperatively, so every
// workgroup of the grid is resident and team barriers cannot wait on a
// workgroup that never runs; every wait is bounded besides.
template <class R>
__global__ __launch_bounds__(BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_batch(
    BKernelArgs K, int ncirc, int nbig, int T, int* qctr, int* team_bar, int* team_cur, int* err) {
    __shared__ cx<R> tile[1 << BATCH_KMAX];
    __shared__ SlotVal slotv[FUSE_MAX_SLOTS];
    __shared__ uint64_t dep[2 * BATCH_DEP_HALF];
    __shared__ double red[BATCH_THREADS / 64 * BATCH_DIAG];
    __shared__ ObsLds obs;
    __shared__ int s_next, s_flag;
    const int tid = threadIdx.x;
    // Team circuits first (queue 0), then this workgroup alone (queue 1); one
    // call site of run_circuit serves both phases.
    bool in_team = T > 1 && nbig > 0;
    Team tm{(int)blockIdx.x / T, (int)blockIdx.x % T, T, team_bar + blockIdx.x / T, 0, err};
    if (!in_team) tm = Team{0, 0, 1, nullptr, 0, err};
    const int first = (T > 1) ? nbig : 0;
    for (;;) {
        int qi;
        if (in_team) {
            if (tm.r == 0 && tid == 0) team_cur[tm.id] = atomicAdd(&qctr[0], 1);
            if (!team_sync(tm, &s_flag)) return;
            if (tid == 0) s_next = __hip_atomic_load(&team_cur[tm.id], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            qi = __builtin_amdgcn_readfirstlane(s_next);  // the same LDS word in every lane
            if (qi >= nbig) {
                in_team = false;
                tm = Team{0, 0, 1, nullptr, 0, err};
                continue;
            }
        } else {
            __syncthreads();
            if (tid == 0) s_next = first + atomicAdd(&qctr[1], 1);
            __syncthreads();
            qi = __builtin_amdgcn_readfirstlane(s_next);
            if (qi >= ncirc) break;
        }
        if (!run_circuit<R>(K.circs[K.order[qi]], tm, K, tile, slotv, dep, red, obs, &s_flag)) return;
    }
}

}  // namespace

}  // namespace qse

using namespace qse;

extern "C" int qse_batch_run_small(qse_ctx* ctx, int precision, int ncirc, const int* nq, const int64_t* op_off,
                                   const qse_op* ops, const int64_t* par_off, const double* params,
                                   const uint64_t* obs_masks, const double* obs_weights, int nterms,
                                   double* energies, double* grads) {
    if (!ctx || (precision != QSE_C128 && precision != QSE_C64) || ncirc < 1 || !nq || !op_off ||
        (op_off[ncirc] > op_off[0] && !ops) || nterms < 0 || (nterms > 0 && (!obs_masks || !obs_weights)) ||
        !energies || (grads && !par_off))
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    QSE_HIP(hipSetDevice(ctx->device));
    const int pi = precision == QSE_C128 ? 0 : 1;
    const int kmax = ctx->batch_k[pi] ? ctx->batch_k[pi] : BATCH_KMAX;
    const int low = ctx->batch_L[pi] ? ctx->batch_L[pi] : (precision == QSE_C128 ? 3 : 4);
    BatchPlan B;
    int rc = build_batch_plan(precision, ncirc, nq, op_off, ops, par_off, params, obs_masks, nterms, grads != nullptr,
                              kmax, low, B);
    if (rc) return rc;
    // circuits of BATCH_TEAM_MIN_QUBITS or more go to teams, largest first
    constexpr int BATCH_TEAM_MIN_QUBITS = 16;
    std::stable_partition(B.order.begin(), B.order.end(),
                          [&](int c) { return B.circs[c].n >= BATCH_TEAM_MIN_QUBITS; });
    int nbig = 0;
    for (int c : B.order) nbig += B.circs[c].n >= BATCH_TEAM_MIN_QUBITS ? 1 : 0;
    // grid: every workgroup resident (cooperative launch), a whole number of teams
    int per_cu = 0;
    QSE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, precision == QSE_C128 ? (const void*)k_batch<double> : (const void*)k_batch<float>, BATCH_THREADS,
        0));
    int grid = std::max(1, per_cu) * ctx->num_cus;
    int T = ctx->batch_team[pi] ? ctx->batch_team[pi] : 16;
    if (nbig == 0 || grid < 2 * T) T = 1;
    if (T > 1) grid = grid / T * T;
    grid = std::min(grid, std::max(ncirc, T > 1 ? grid : 1));
    const int nteams = (T > 1) ? grid / T : 1;
    const size_t ab = amp_bytes(precision);
    size_t free_b = 0, total_b = 0;
    QSE_HIP(hipMemGetInfo(&free_b, &total_b));
    const double avail = (double)free_b + (double)ctx->scratch[4].bytes;  // the buffer of an earlier batch is reused
    if ((double)B.arena_amps * (double)ab > 0.9 * avail)
        return set_error(QSE_ERR_OUT_OF_MEMORY, "batch: the circuits' states do not fit the free device memory");
    // one device buffer: arena | program | circuits | order | steps | observable | counters | partials | values
    Serialized S;
    serialize(B.prog, S);
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_arena = off; off = al(off + B.arena_amps * ab);
    const size_t o_prog = off; off = al(off + S.bytes.size());
    const size_t o_circ = off; off = al(off + sizeof(BCirc) * B.circs.size());
    const size_t o_order = off; off = al(off + sizeof(int32_t) * B.order.size());
    const size_t o_steps = off; off = al(off + sizeof(BStep) * B.steps.size());
    const size_t o_obsm = off; off = al(off + sizeof(uint64_t) * std::max(nterms, 1));
    const size_t o_obsw = off; off = al(off + sizeof(double) * std::max(nterms, 1));
    const size_t o_ctl = off; off = al(off + sizeof(int) * (3 + 2 * (size_t)nteams));  // qctr[2], err, bar, cur
    const size_t o_part = off; off = al(off + sizeof(double) * (size_t)nteams * T * BATCH_DIAG);
    const size_t o_vals = off; off = al(off + sizeof(double) * std::max(B.nvals, 1));
    // context scratch (kept between calls: a batch of this size is typically run again)
    void* dptr = nullptr;
    rc = scratch(ctx, 4, off, &dptr);
    if (rc) return rc;
    char* dev = (char*)dptr;
    auto fail = [&](int code) {
        hipStreamSynchronize(ctx->stream);
        return code;
    };
    auto h2d = [&](size_t o, const void* src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dev + o, src, bytes, hipMemcpyHostToDevice, ctx->stream) : hipSuccess;
    };
    hipError_t e = h2d(o_prog, S.bytes.data(), S.bytes.size());
    if (e == hipSuccess) e = h2d(o_circ, B.circs.data(), sizeof(BCirc) * B.circs.size());
    if (e == hipSuccess) e = h2d(o_order, B.order.data(), sizeof(int32_t) * B.order.size());
    if (e == hipSuccess) e = h2d(o_steps, B.steps.data(), sizeof(BStep) * B.steps.size());
    if (e == hipSuccess && nterms > 0) e = h2d(o_obsm, obs_masks, sizeof(uint64_t) * nterms);
    if (e == hipSuccess && nterms > 0) e = h2d(o_obsw, obs_weights, sizeof(double) * nterms);
    if (e == hipSuccess) e = hipMemsetAsync(dev + o_ctl, 0, sizeof(int) * (3 + 2 * (size_t)nteams), ctx->stream);
    if (e != hipSuccess) return fail(hip_error(e, "batch upload"));
    char* pg = dev + o_prog;
    BKernelArgs K;
    K.arena_d = (cx<double>*)(dev + o_arena);
    K.circs = (const BCirc*)(dev + o_circ);
    K.order = (const int32_t*)(dev + o_order);
    K.steps = (const BStep*)(dev + o_steps);
    K.prog = BArgs{(const FPass*)(pg + S.off[0]), (const FStage*)(pg + S.off[1]), (const FOp*)(pg + S.off[2]),
                   (const FTerm*)(pg + S.off[3]),  (const FSlot*)(pg + S.off[4]),  (const FExtTerm*)(pg + S.off[5]),
                   (const double*)(pg + S.off[6])};
    K.obs_m = (const uint64_t*)(dev + o_obsm);
    K.obs_w = (const double*)(dev + o_obsw);
    K.nterms = nterms;
    K.tpart = (double*)(dev + o_part);
    K.vals = (double*)(dev + o_vals);
    int* ctl = (int*)(dev + o_ctl);
    int* qctr = ctl;
    int* err = ctl + 2;
    int* team_bar = ctl + 3;
    int* team_cur = ctl + 3 + nteams;
    void* args[] = {&K, &ncirc, &nbig, &T, &qctr, &team_bar, &team_cur, &err};
    TimedLaunch t;
    launch_begin(ctx, FAM_BATCH, &t);
    e = hipLaunchCooperativeKernel(precision == QSE_C128 ? (const void*)k_batch<double> : (const void*)k_batch<float>,
                                   dim3(grid), dim3(BATCH_THREADS), args, 0, ctx->stream);
    if (e != hipSuccess) return fail(hip_error(e, "k_batch cooperative launch"));
    launch_end(ctx, &t, B.est_bytes * (double)ab);
    std::vector<double> vals(std::max(B.nvals, 1));
    int herr = 0;
    e = hipMemcpyAsync(vals.data(), dev + o_vals, sizeof(double) * vals.size(), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(hip_error(e, "batch readback"));
    if (herr) return set_error(QSE_ERR_HIP, "batch: a team barrier timed out");
    batch_results(B, vals.data(), par_off, energies, grads);
    return QSE_OK;
}

// Measured choice of the batch engine's tile shape: the given circuits (a
// sample of a larger batch, typically) run once per candidate (tiles of 12,
// 11 or 10 qubits; 2 to 5 contiguous low qubits), and the fastest candidate
// is kept in the context for later qse_batch_run_small calls of this
// precision. Candidates that fail are skipped.
extern "C" int qse_batch_autotune(qse_ctx* ctx, int precision, int ncirc, const int* nq, const int64_t* op_off,
                                  const qse_op* ops, const int64_t* par_off, const double* params,
                                  const uint64_t* obs_masks, const double* obs_weights, int nterms, int with_grads,
                                  int* k_out, int* L_out, int* team_out, double* ms_out) {
    if (!ctx || (precision != QSE_C128 && precision != QSE_C64) || ncirc < 1 || !nq || !op_off)
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    const int pi = precision == QSE_C128 ? 0 : 1;
    std::vector<double> en(ncirc);
    std::vector<double> gr(with_grads && par_off ? (size_t)std::max<int64_t>(par_off[ncirc], 1) : 1);
    double* gp = with_grads && par_off ? gr.data() : nullptr;
    const int keep_k = ctx->batch_k[pi], keep_L = ctx->batch_L[pi], keep_T = ctx->batch_team[pi];
    auto restore = [&] {
        ctx->batch_k[pi] = keep_k;
        ctx->batch_L[pi] = keep_L;
        ctx->batch_team[pi] = keep_T;
    };
    auto run = [&](int k, int L, int team, double* ms) -> int {
        ctx->batch_k[pi] = k;
        ctx->batch_L[pi] = L;
        ctx->batch_team[pi] = team;
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = qse_batch_run_small(ctx, precision, ncirc, nq, op_off, ops, par_off, params, obs_masks,
                                           obs_weights, nterms, en.data(), gp);
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    };
    double ms = 0.0;
    int rc = run(BATCH_KMAX, precision == QSE_C128 ? 3 : 4, 16, &ms);  // warm-up: device buffer and code objects
    if (rc) {
        restore();
        return rc;
    }
    // tile shape with teams of 16, then the team size with that shape
    int best_k = 0, best_L = 0, best_T = 16;
    double best = 1e300;
    for (int k = BATCH_KMAX; k >= 10; --k)
        for (int L = 2; L <= 5 && L <= k - FUSE_REG; ++L) {
            if (run(k, L, 16, &ms) != QSE_OK) continue;
            if (ms < best) {
                best = ms;
                best_k = k;
                best_L = L;
            }
        }
    if (!best_k) {
        restore();
        return set_error(QSE_ERR_HIP, "batch autotune: no candidate ran");
    }
    for (int team : {4, 8, 32}) {
        if (run(best_k, best_L, team, &ms) != QSE_OK) continue;
        if (ms < best) {
            best = ms;
            best_T = team;
        }
    }
    ctx->batch_k[pi] = best_k;
    ctx->batch_L[pi] = best_L;
    ctx->batch_team[pi] = best_T;
    if (k_out) *k_out = best_k;
    if (L_out) *L_out = best_L;
    if (team_out) *team_out = best_T;
    if (ms_out) *ms_out = best;
    return QSE_OK;
}

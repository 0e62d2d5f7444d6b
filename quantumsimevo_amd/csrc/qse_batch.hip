// qse_batch.hip — batched small-circuit engine (qse_batch_run_small): many
// circuits of uneven qubit counts, forward run, observable and adjoint
// gradients, in one persistent launch.
//
// Circuits are taken from device work queues (in decreasing order of estimated
// work) and each circuit's whole step list (qse_batch.hpp) is executed by one
// team: fused passes over the tiles of the circuit's state in HBM (the
// LDS-tiled pass of qse_fused.hip, in place), the observable, an FP64
// reduction for the energy, and the adjoint reverse sweep as joint passes that
// carry the tiles of psi and lambda together, un-apply every gate from both
// and add each parameter's generator term Im<lambda|G|psi> (FP64, per wave)
// at its point in the sweep. Circuits of 16 qubits or
// more are taken by teams of 16 workgroups that split every pass's tiles and
// every reduction's indices and meet at team barriers (global counters); the
// others by single workgroups, whose own barriers are the only
// synchronisation. Reductions sum per-workgroup partials in rank order
// (deterministic). Tiles of k < KMAX qubits use the first 2^(k-4) threads; the
// others wait at the barriers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "qse_batch.hpp"
#include "qse_device.hpp"
#include "qse_fused_dev.hpp"
#include "qse_internal.hpp"

namespace qse {

namespace {

// Tile qubits of the engine's largest tile (BATCH_KMAX, both precisions); the
// LDS tile buffer holds 64 KiB (a complex64 tile pair of 12 qubits, for the
// joint passes), two workgroups per CU.
constexpr int BATCH_THREADS = 1 << (BATCH_KMAX - 4);
constexpr int BATCH_WAVES = BATCH_THREADS / 64;
// Complex64 single-state passes read their ops from an LDS copy when they
// have at most this many (26 KiB; complex128's tile leaves no room for it).
constexpr int BATCH_LDS_OPS = 256;
constexpr int BATCH_TILE_BYTES = (2 << BATCH_KMAX) * (int)sizeof(cx<float>);
// External (tile-index) qubits are deposited in groups of BATCH_DEP_BITS from
// per-pass LDS tables: 4 groups cover the 22 external qubits of a 30-qubit
// circuit with 8-qubit tiles.
constexpr int BATCH_DEP_BITS = 6;
constexpr int BATCH_DEP_LEVELS = 4;
constexpr int BATCH_DEP_HALF = BATCH_DEP_LEVELS << BATCH_DEP_BITS;
static_assert(30 - FUSE_MIN_K <= BATCH_DEP_LEVELS * BATCH_DEP_BITS,
              "the deposit tables cover the external qubits of the largest batch circuit");
static_assert((2 << batch_joint_kmax(QSE_C128)) * sizeof(cx<double>) <= BATCH_TILE_BYTES &&
                  (2 << batch_joint_kmax(QSE_C64)) * sizeof(cx<float>) <= BATCH_TILE_BYTES,
              "a joint pass's tile pair fits the tile buffer");

struct BArgs {
    const FPass* __restrict__ passes;
    const FStage* __restrict__ stages;
    const FOp* __restrict__ ops;
    const FTerm* __restrict__ terms;
    const FSlot* __restrict__ slots;
    const FExtTerm* __restrict__ eterms;
    const double* __restrict__ alphas;
};

// Per-tile external phase factor of a slot (cos, sin of its angle sum).
struct BSlot {
    double c, s;
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

// A non-MONO, non-probe op on this thread's 16 register amplitudes.
template <class R>
__device__ __forceinline__ void apply_regs(const FOp& op, cx<R>* a, uint32_t toff, uint64_t base) {
    if (op.type == FOP_H) {
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
    } else if (((toff & op.tctl) == op.tctl) && ((base & op.ectl) == op.ectl)) {
        if (op.type == FOP_XSWAP) {
            // X / CNOT target: an exchange of registers (controls: register bits per
            // pair, thread bits and external qubits for the whole thread)
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
        } else {
            switch (op.p) {
                case 0: dense_p<R, 0>(a, op); break;
                case 1: dense_p<R, 1>(a, op); break;
                case 2: dense_p<R, 2>(a, op); break;
                default: dense_p<R, 3>(a, op); break;
            }
        }
    }
}

// Im<lambda|X|psi> (Y: Im<lambda|Y|psi>) over the register pairs on position
// P, in the storage precision (16 terms; the sums over threads and tiles are
// FP64).
template <class R, int P>
__device__ __forceinline__ R probe_pair(const cx<R>* a, const cx<R>* b, bool y) {
    R s = 0;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        const cx<R> a0 = a[j0], a1 = a[j1], b0 = b[j0], b1 = b[j1];
        // X: Im(conj(b0) a1) + Im(conj(b1) a0);  Y: Re(conj(b1) a0) - Re(conj(b0) a1)
        s += y ? (b1.re * a0.re + b1.im * a0.im) - (b0.re * a1.re + b0.im * a1.im)
               : (b0.re * a1.im - b0.im * a1.re) + (b1.re * a0.im - b1.im * a0.re);
    }
    return s;
}

// This thread's part of a probe: Im<lambda|G|psi> over its 16 register
// amplitudes (psi in a, lambda in b), in the storage precision.
template <class R>
__device__ __forceinline__ double probe_regs(const FOp& op, const cx<R>* a, const cx<R>* b, uint32_t toff,
                                             uint64_t base) {
    if (op.mode == PROBE_X || op.mode == PROBE_Y) {
        const bool y = op.mode == PROBE_Y;
        switch (op.p) {
            case 0: return (double)probe_pair<R, 0>(a, b, y);
            case 1: return (double)probe_pair<R, 1>(a, b, y);
            case 2: return (double)probe_pair<R, 2>(a, b, y);
            default: return (double)probe_pair<R, 3>(a, b, y);
        }
    }
    const uint32_t rm = op.rctl;
    R s = 0;
    if (op.mode == PROBE_PROJ) {
        if ((toff & op.tctl) != op.tctl || (base & op.ectl) != op.ectl) return 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (((uint32_t)j & rm) == rm) s += b[j].re * a[j].im - b[j].im * a[j].re;
        return (double)s;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const R w = b[j].re * a[j].im - b[j].im * a[j].re;
        s += (__popc((uint32_t)j & rm) & 1) ? -w : w;
    }
    return ((__popc(toff & op.tctl) + __popcll(base & op.ectl)) & 1) ? -(double)s : (double)s;
}

// One DPP step of a wave reduction on a double: v from the lane that dpp_ctrl
// selects, 0 in lanes outside row_mask or without a source lane.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Sum of v over the 64 lanes of a wave (all active), in lane 63: a scan over
// each row of 16 lanes (row_shr 1, 2, 4, 8), then the row sums are chained
// with row_bcast:15 and row_bcast:31. No LDS round trip.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
    v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
    v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
    v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
    v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// One in-place pass of tile size K over the 2^n amplitudes at s: tiles r, r + T, ...
// (this workgroup's share when T workgroups work on the circuit). J: a joint
// pass, the same gates on s (psi) and s2 (lambda) with the pass's probes
// summed into pacc[probe * BATCH_WAVES + wave] and, at the end, this
// workgroup's sums written to pout[P.probe_base + probe].
template <class R, int K, bool J>
__device__ __attribute__((always_inline)) void batch_pass(cx<R>* s, cx<R>* s2, int n, const FPass& Pg, const BArgs& A, cx<R>* tile, BSlot* slotv,
                           uint64_t* dep, double* pacc, double* pout, int r, int T, FOp* lops, int lops_cap) {
    const auto& P = *kc(&Pg);
    constexpr int NT = 1 << (K - 4);
    constexpr int TB = K - 4;
    cx<R>* tile2 = tile + (1 << K);
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
    if (J)
        for (int idx = tid; idx < P.nprobes * BATCH_WAVES; idx += blockDim.x) pacc[idx] = 0.0;
    // the pass's ops (contiguous over its stages), copied to LDS when they fit
    const uint32_t op0 = P.nstages ? kc(A.stages + P.stage_begin)->op_begin : 0;
    const uint32_t op1 = P.nstages ? kc(A.stages + P.stage_begin + P.nstages - 1)->op_end : 0;
    const bool lds_ops = (int)(op1 - op0) <= lops_cap;
    if (lds_ops) {
        static_assert(sizeof(FOp) % sizeof(uint64_t) == 0, "ops are copied in 8-byte words");
        constexpr int W = sizeof(FOp) / sizeof(uint64_t);
        const uint64_t* src = (const uint64_t*)(A.ops + op0);
        for (int idx = tid; idx < (int)(op1 - op0) * W; idx += blockDim.x) ((uint64_t*)lops)[idx] = src[idx];
    }
    __syncthreads();
    // Complex64 single-state passes: the next tile's amplitudes are loaded into
    // registers while this tile's stages run (its 16 amplitudes fit the
    // register budget of two waves per SIMD; complex128's or a joint pass's
    // 32 would not). A workgroup's tiles are its own in the pass, so the
    // prefetch never reads what another one writes.
    constexpr bool PF = sizeof(R) == 4 && !J;
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
        // length (and spill). (The pointer is uniform: readfirstlane places it
        // in scalar registers for the barrier whatever register class its
        // producer used.)
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
                if (J) tile2[swz_tid ^ swz((uint32_t)(i * NT))] = s2[g];
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
            slotv[sl] = {cs, sn};
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
                cx<R> a[16], b[J ? 16 : 1];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const uint32_t at = st_off ^ ((j & 1) ? r0 : 0u) ^ ((j & 2) ? r1 : 0u) ^ ((j & 4) ? r2 : 0u) ^
                                        ((j & 8) ? r3 : 0u);
                    a[j] = tile[at];
                    if (J) b[j] = tile2[at];
                }
                for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
                    FOp op;
                    if (lds_ops) op = lops[oi - op0];
                    else op = *kc(A.ops + oi);
                    if (op.type == FOP_MONO) {
                        double gr, gi;
                        mono_factor<TB>(op, kc(A.terms), kc(A.alphas), tid, toff, base, slotv, gr, gi);
                        const cx<R> g = {(R)gr, (R)gi};
                        const uint32_t rctl = op.rctl;
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            if (((uint32_t)j & rctl) == rctl) {
                                a[j] = cmul(g, a[j]);
                                if (J) b[j] = cmul(g, b[j]);
                            }
                    } else if (J && op.type == FOP_PROBE) {
                        // this wave's sum (over its active lanes: tid < NT), added to its own accumulator
                        double v = probe_regs<R>(op, a, J ? b : a, toff, base);
                        if constexpr (NT >= 64) {
                            v = wave_sum_dpp(v);
                            if ((tid & 63) == 63) pacc[op.term_begin * BATCH_WAVES + (tid >> 6)] += v;
                        } else {
#pragma unroll
                            for (int o = NT / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                            if ((tid & 63) == 0) pacc[op.term_begin * BATCH_WAVES + (tid >> 6)] += v;
                        }
                    } else {
                        apply_regs<R>(op, a, toff, base);
                        if (J) apply_regs<R>(op, b, toff, base);
                    }
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const uint32_t at = st_off ^ ((j & 1) ? r0 : 0u) ^ ((j & 2) ? r1 : 0u) ^ ((j & 4) ? r2 : 0u) ^
                                        ((j & 8) ? r3 : 0u);
                    tile[at] = a[j];
                    if (J) tile2[at] = b[j];
                }
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
                if (J) s2[g] = tile2[ls];
            }
        }
        __syncthreads();
    }
    if (J) {
        // this workgroup's probe sums, waves in order (the tile loop ended with a barrier)
        const auto& Q = *kc(&Pg);
        for (int p = tid; p < Q.nprobes; p += blockDim.x) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < BATCH_WAVES; ++w) v += pacc[p * BATCH_WAVES + w];
            pout[Q.probe_base + p] = v;
        }
    }
}

// A wave-uniform pointer in scalar registers (arguments of a called function
// arrive in vector registers, where the compiler cannot see that they are
// uniform and would read the program through them with per-lane loads).
template <class T>
__device__ __forceinline__ T* uni(T* p) {
    const uint64_t w = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)w), hi = __builtin_amdgcn_readfirstlane((uint32_t)(w >> 32));
    return (T*)(((uint64_t)hi << 32) | lo);
}

// A joint pass as a function of its own: its registers (two tiles' amplitudes
// per thread) are allocated apart from the rest of the kernel, so the other
// tile sizes and steps do not spill around it.
template <class R, int K>
__device__ __attribute__((noinline)) void batch_joint(cx<R>* s, cx<R>* s2, int n, const FPass& Pg, const BArgs& A0,
                                                      cx<R>* tile, BSlot* slotv, uint64_t* dep, double* pacc,
                                                      double* pout, int r, int T) {
    const BArgs A{uni(A0.passes), uni(A0.stages), uni(A0.ops), uni(A0.terms), uni(A0.slots), uni(A0.eterms), uni(A0.alphas)};
    batch_pass<R, K, true>(uni(s), uni(s2), __builtin_amdgcn_readfirstlane(n), *uni(&Pg), A, tile, slotv, dep, pacc,
                           uni(pout), __builtin_amdgcn_readfirstlane(r), __builtin_amdgcn_readfirstlane(T), nullptr, -1);
}

// Single-state passes are inlined at their call sites: there the compiler sees
// that the state and the program live in global memory, and reads the
// program's uniform fields with scalar loads.
template <class R, bool J>
__device__ __attribute__((always_inline)) void batch_pass_k(int k, cx<R>* s, cx<R>* s2, int n, const FPass& P,
                                                            const BArgs& A, cx<R>* tile, BSlot* slotv, uint64_t* dep,
                                                            double* pacc, double* pout, int r, int T, FOp* lops) {
    switch (k) {
        case 8:
            if constexpr (J) batch_joint<R, 8>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T);
            else batch_pass<R, 8, J>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T, lops, sizeof(R) == 4 ? BATCH_LDS_OPS : -1);
            break;
        case 9:
            if constexpr (J) batch_joint<R, 9>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T);
            else batch_pass<R, 9, J>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T, lops, sizeof(R) == 4 ? BATCH_LDS_OPS : -1);
            break;
        case 10:
            if constexpr (J) batch_joint<R, 10>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T);
            else batch_pass<R, 10, J>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T, lops, sizeof(R) == 4 ? BATCH_LDS_OPS : -1);
            break;
        case 11:
            if constexpr (J) batch_joint<R, 11>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T);
            else batch_pass<R, 11, J>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T, lops, sizeof(R) == 4 ? BATCH_LDS_OPS : -1);
            break;
        default:
            if constexpr (!J) batch_pass<R, 12, J>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T, lops, sizeof(R) == 4 ? BATCH_LDS_OPS : -1);
            else if constexpr (sizeof(R) == 4) batch_joint<R, 12>(s, s2, n, P, A, tile, slotv, dep, pacc, pout, r, T);
            break;
    }
}

// The energy and the observable run over blocks of WB = 8 amplitudes
// i = i0 + stride j (j < 8), whose sums against Z-strings come from one
// in-register Walsh-Hadamard transform of the block.
constexpr int WB = 8;

// Observable terms of the running circuit, in LDS: masks and weights of the
// kept terms (their qubits lie on the circuit), grouped by the 3-bit field
// u = (mask >> log2(stride)) & 7 that selects a Walsh-Hadamard coefficient
// (group u: entries [ubeg[u], ubeg[u + 1])); built once per circuit.
struct ObsLds {
    uint64_t m[BATCH_MAX_TERMS];
    double w[BATCH_MAX_TERMS];
    int ubeg[WB + 1];
    int n;
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
    int pstride;    // partial values per workgroup (the most of one step)
    double* tpart;  // per team: T x pstride partial sums
    double* vals;
};

// One circuit's step list by a team (or a single workgroup). False: a team
// barrier failed.
template <class R, int PH>
__device__ __attribute__((always_inline)) bool run_circuit(const BCirc& C, Team& tm, const BKernelArgs& K, cx<R>* tile,
                                                           BSlot* slotv, uint64_t* dep, double* red, double* pacc,
                                                           FOp* lops, ObsLds& O, int* s_flag) {
    const int tid = threadIdx.x, nth = blockDim.x;
    cx<R>* arena = (cx<R>*)K.arena_d;
    cx<R>* psi = arena + C.psi_off;
    cx<R>* lam = arena + C.lam_off;
    const uint64_t N = 1ULL << C.n;
    const uint64_t first = (uint64_t)tm.r * nth + tid, stride = (uint64_t)tm.T * nth;
    double* part = K.tpart + (size_t)tm.id * tm.T * K.pstride;
    // The energy and the observable run over blocks of WB strides when the
    // state holds at least WB strides; below, element by element.
    const int log2S = __builtin_ctzll(stride);
    const bool blocked = (stride & (stride - 1)) == 0 && N >= WB * stride;
    const uint64_t nblk = blocked ? N / (WB * stride) : 0;
    // the circuit's observable terms in LDS, grouped by their Walsh-Hadamard field
    if (PH == 0 && tid == 0) {
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
        if ((kind == BS_JOINT) != (PH == 1)) continue;  // a step of the other launch
        if constexpr (PH == 1) {
            // the reverse sweep's joint passes; each workgroup's probe sums go to the values
            // directly (single workgroup) or to its row of the team's partials
            double* pout = (tm.T == 1) ? K.vals + S.out : part + (size_t)tm.r * K.pstride;
#pragma unroll 1
            for (int pi = S.pass_begin; pi < S.pass_end; ++pi) {
                const FPass& P = K.prog.passes[pi];
                batch_pass_k<R, true>(P.k, psi, lam, C.n, P, K.prog, tile, slotv, dep, pacc, pout, tm.r, tm.T, nullptr);
                if (!team_sync(tm, s_flag)) return false;
            }
            if (tm.T > 1 && tm.r == 0)
                for (int v = tid; v < S.nprobes; v += nth) {
                    double a = 0.0;
                    for (int q = 0; q < tm.T; ++q) a += part[(size_t)q * K.pstride + v];
                    K.vals[S.out + v] = a;
                }
        } else if (kind == BS_BASIS) {
            for (uint64_t i = first; i < N; i += stride) psi[i] = {(R)(i == 0 ? 1 : 0), (R)0};
        } else if (kind == BS_PROG) {
            // pass p on psi (target bit 0), then on lambda (bit 1): one call site
            const int nt = S.pass_end > S.pass_begin ? 2 * (S.pass_end - S.pass_begin) : 0;
#pragma unroll 1
            for (int it = 0; it < nt; ++it) {
                const int which = it & 1;
                if (!((S.target >> which) & 1)) continue;
                const FPass& P = K.prog.passes[S.pass_begin + (it >> 1)];
                batch_pass_k<R, false>(C.k, which ? lam : psi, nullptr, C.n, P, K.prog, tile, slotv, dep, nullptr,
                                       nullptr, tm.r, tm.T, lops);
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
        } else {  // BS_ENERGY: partial sum of this workgroup, then the team's in rank order
            double acc = 0.0;
            if (blocked) {
                // sum_j p_j f_j = sum_u G[u] W[u], W = transform of the block's probabilities
                for (uint64_t kbi = 0; kbi < nblk; ++kbi) {
                    const uint64_t i0 = first + WB * stride * kbi;
                    double w[WB], G[WB];
#pragma unroll
                    for (int j = 0; j < WB; ++j) w[j] = prob(psi[i0 + stride * j]);
                    wht(w);
                    obs_groups(O, i0, G);
#pragma unroll
                    for (int u = 0; u < WB; ++u) acc += G[u] * w[u];
                }
            } else {
                for (uint64_t i = first; i < N; i += stride) acc += prob(psi[i]) * obs_factor(i, O);
            }
            const double v = wave_sum(acc);
            if ((tid & 63) == 0) red[tid >> 6] = v;
            __syncthreads();
            if (tid == 0) {
                double t = 0.0;
                for (int w = 0; w < nth / 64; ++w) t += red[w];
                if (tm.T == 1) K.vals[S.out] = t;
                else part[(size_t)tm.r * K.pstride] = t;
            }
            if (tm.T > 1) {
                if (!team_sync(tm, s_flag)) return false;
                if (tm.r == 0 && tid == 0) {
                    double t = 0.0;
                    for (int q = 0; q < tm.T; ++q) t += part[(size_t)q * K.pstride];
                    K.vals[S.out] = t;
                }
            }
        }
        // the step's writes (and, for reductions, the leader's reads of the partials) before the next step
        if (!team_sync(tm, s_flag)) return false;
    }
    return true;
}

// Persistent engine, in two launches: PH 0 runs every step of each circuit
// but its joint steps (forward program, energy, observable), PH 1 only the
// joint steps (the reverse sweep); each launch's kernel holds only its own
// passes' code and registers. In each launch, teams of T workgroups first take
// the first nbig circuits of `order` from queue qctr[0]; then every workgroup
// alone takes the remaining circuits from queue qctr[1]. Launched
// cooperatively, so every workgroup of the grid is resident and team barriers
// cannot wait on a workgroup that never runs; every wait is bounded besides.
template <class R, int PH>
__global__ __launch_bounds__(BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_batch(
    BKernelArgs K, int ncirc, int nbig, int T, int* qctr, int* team_bar, int* team_cur, int* err) {
    __shared__ cx<R> tile[PH == 1 ? BATCH_TILE_BYTES / sizeof(cx<R>) : 1 << BATCH_KMAX];
    __shared__ BSlot slotv[FUSE_MAX_SLOTS];
    __shared__ uint64_t dep[2 * BATCH_DEP_HALF];
    __shared__ double pacc[PH == 1 ? FUSE_MAX_PROBES * BATCH_WAVES : 1];
    __shared__ FOp lops[PH == 0 && sizeof(R) == 4 ? BATCH_LDS_OPS : 1];
    __shared__ double red[BATCH_WAVES];
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
        if (!run_circuit<R, PH>(K.circs[K.order[qi]], tm, K, tile, slotv, dep, red, pacc, lops, obs, &s_flag)) return;
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
    // the kernel's tile shapes: passes of 8..12 qubits, joint passes within the tile buffer
    for (const BStep& st : B.steps)
        for (int p = st.pass_begin; (st.kind == BS_PROG || st.kind == BS_JOINT) && p < st.pass_end; ++p) {
            const int pk = B.prog.passes[p].k;
            if (pk < FUSE_MIN_K || pk > (st.kind == BS_JOINT ? batch_joint_kmax(precision) : BATCH_KMAX) ||
                B.prog.passes[p].nprobes > FUSE_MAX_PROBES)
                return set_error(QSE_ERR_UNSUPPORTED, "batch: a pass's shape does not fit the kernel");
        }
    // circuits of BATCH_TEAM_MIN_QUBITS or more go to teams, largest first
    constexpr int BATCH_TEAM_MIN_QUBITS = 16;
    std::stable_partition(B.order.begin(), B.order.end(),
                          [&](int c) { return B.circs[c].n >= BATCH_TEAM_MIN_QUBITS; });
    int nbig = 0;
    for (int c : B.order) nbig += B.circs[c].n >= BATCH_TEAM_MIN_QUBITS ? 1 : 0;
    bool has_joint = false;
    for (const BStep& st : B.steps) has_joint |= st.kind == BS_JOINT;
    // per launch: grid with every workgroup resident (cooperative launch), a whole number of teams
    const void* kern[2] = {precision == QSE_C128 ? (const void*)k_batch<double, 0> : (const void*)k_batch<float, 0>,
                           precision == QSE_C128 ? (const void*)k_batch<double, 1> : (const void*)k_batch<float, 1>};
    int grid[2] = {1, 1}, T[2] = {1, 1}, nteams[2] = {1, 1};
    for (int ph = 0; ph < (has_joint ? 2 : 1); ++ph) {
        int per_cu = 0;
        QSE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern[ph], BATCH_THREADS, 0));
        int g = std::max(1, per_cu) * ctx->num_cus;
        int t = ctx->batch_team[pi] ? ctx->batch_team[pi] : 16;
        if (nbig == 0 || g < 2 * t) t = 1;
        if (t > 1) g = g / t * t;
        g = std::min(g, std::max(ncirc, t > 1 ? g : 1));
        grid[ph] = g;
        T[ph] = t;
        nteams[ph] = (t > 1) ? g / t : 1;
    }
    const int max_teams = std::max(nteams[0], nteams[1]);
    const int max_wgs = std::max(nteams[0] * T[0], nteams[1] * T[1]);
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
    const size_t o_ctl = off; off = al(off + sizeof(int) * (3 + 2 * (size_t)max_teams));  // err, qctr[2], bar, cur
    const int pstride = std::max(1, B.max_probes);
    const size_t o_part = off; off = al(off + sizeof(double) * (size_t)max_wgs * pstride);
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
    if (e == hipSuccess) e = hipMemsetAsync(dev + o_ctl, 0, sizeof(int), ctx->stream);
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
    K.pstride = pstride;
    K.tpart = (double*)(dev + o_part);
    K.vals = (double*)(dev + o_vals);
    int* ctl = (int*)(dev + o_ctl);
    int* err = ctl;
    int* qctr = ctl + 1;
    for (int ph = 0; ph < (has_joint ? 2 : 1); ++ph) {
        // fresh queues and team barriers per launch (an error flag of the first stays set)
        e = hipMemsetAsync(qctr, 0, sizeof(int) * (2 + 2 * (size_t)nteams[ph]), ctx->stream);
        if (e != hipSuccess) return fail(hip_error(e, "batch counters"));
        int* team_bar = ctl + 3;
        int* team_cur = ctl + 3 + nteams[ph];
        int nb = nbig;
        void* args[] = {&K, &ncirc, &nb, &T[ph], &qctr, &team_bar, &team_cur, &err};
        TimedLaunch t;
        launch_begin(ctx, FAM_BATCH, &t);
        e = hipLaunchCooperativeKernel(kern[ph], dim3(grid[ph]), dim3(BATCH_THREADS), args, 0, ctx->stream);
        if (e != hipSuccess) return fail(hip_error(e, "k_batch cooperative launch"));
        launch_end(ctx, &t, (ph == 0 ? B.est_bytes - B.est_bytes_rev : B.est_bytes_rev) * (double)ab);
    }
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
// 11 or 10 qubits; 2 to 5 contiguous low qubits; then teams of 4, 8, 16 or 32
// workgroups), timed by the engine's kernel events, and the fastest
// candidate is kept in the context for later qse_batch_run_small calls of
// this precision. Candidates that fail are skipped.
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
    // Candidates are compared by the engine's kernel time (HIP events of the
    // batch family; host planning and uploads are the same for every shape).
    // The caller's timing switch and batch-family totals are restored after.
    int rc0 = fold_timing(ctx);
    if (rc0) return rc0;
    const bool keep_timing = ctx->timing;
    const int64_t keep_launches = ctx->launches[FAM_BATCH];
    const double keep_ms = ctx->total_ms[FAM_BATCH], keep_bytes = ctx->total_bytes[FAM_BATCH];
    ctx->timing = true;
    auto run = [&](int k, int L, int team, double* ms) -> int {
        ctx->batch_k[pi] = k;
        ctx->batch_L[pi] = L;
        ctx->batch_team[pi] = team;
        int rc = fold_timing(ctx);
        const double before = ctx->total_ms[FAM_BATCH];
        if (!rc)
            rc = qse_batch_run_small(ctx, precision, ncirc, nq, op_off, ops, par_off, params, obs_masks, obs_weights,
                                     nterms, en.data(), gp);
        if (!rc) rc = fold_timing(ctx);
        *ms = ctx->total_ms[FAM_BATCH] - before;
        return rc;
    };
    auto restore_timing = [&] {
        fold_timing(ctx);
        ctx->timing = keep_timing;
        ctx->launches[FAM_BATCH] = keep_launches;
        ctx->total_ms[FAM_BATCH] = keep_ms;
        ctx->total_bytes[FAM_BATCH] = keep_bytes;
    };
    auto restore = [&] {
        ctx->batch_k[pi] = keep_k;
        ctx->batch_L[pi] = keep_L;
        ctx->batch_team[pi] = keep_T;
        restore_timing();
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
    restore_timing();
    ctx->batch_k[pi] = best_k;
    ctx->batch_L[pi] = best_L;
    ctx->batch_team[pi] = best_T;
    if (k_out) *k_out = best_k;
    if (L_out) *L_out = best_L;
    if (team_out) *team_out = best_T;
    if (ms_out) *ms_out = best;
    return QSE_OK;
}

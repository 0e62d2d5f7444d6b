// qse_fused.hip — the LDS-tiled fused pass for gfx950 and its host driver.
//
// One launch = one pass = one HBM read and one HBM write of every amplitude
// (algorithmic traffic 2*b*N bytes, b = 16 for complex128, 8 for complex64).
// Workgroup w processes tiles t = w, w + grid, ...: a tile is the 2^K
// amplitudes obtained by fixing the n-K external qubits of t and varying the
// K tile qubits (program format and planning rules: qse_fused.hpp,
// qse_plan.cpp). Per tile:
//   1. per-tile external factors (products of monomial phases whose
//      conditions are on external qubits), one slot per thread, into LDS;
//   2. load: thread tid reads tile positions i*NT + tid (i < 16) — the low
//      tile qubits are the state's low qubits, so a wavefront reads
//      contiguous runs — into the LDS tile (bank-swizzled);
//   3. stages: gather 16 amplitudes per thread over the stage's 4 register
//      bits, apply the stage's ops in registers, scatter back;
//   4. store with the pass's final tile-bit labelling (in-tile SWAPs).
// LDS is XOR-swizzled in 16-byte units (position ^ fold of bits 4.. into bits
// 0..3) so that both the contiguous load/store and the strided stage gathers
// spread a wavefront's accesses over the banks.
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "qse_device.hpp"
#include "qse_fused.hpp"
#include "qse_fused_dev.hpp"
#include "qse_internal.hpp"

namespace qse {

void fusion_shape(const qse_ctx* ctx, int n, int prec, int* k_out, int* L_out);
int build_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P, int flags);
int choose_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P,
                bool allow_oop);
int tuned_plan(const qse_ctx* ctx, int n, int prec, const qse_op* ops, int nops, const double* params, HostPlan& P,
               bool allow_oop, int* k_out, int* L_out);
int run_sweeps(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params);
int validate_ops(int n, const qse_op* ops, int nops, const double* params);
int jit_build(const HostPlan& P, int prec, const size_t* toff, hipModule_t* mod, std::vector<hipFunction_t>& fns);
int jit_launch(qse_ctx* ctx, hipFunction_t fn, const void* src, void* dst, const void* table, uint64_t ntiles, int K,
               const void* prog);
int jit_for_structure(qse_ctx* ctx, const HostPlan& P, const Serialized& S, int prec,
                      const std::vector<hipFunction_t>** fns);

// Per-tile external factors of one pass, for every tile at once: entry
// (tile, slot) = angle sum of the slot's external terms whose qubits are all
// set in the tile's external bits, with its cos and sin. The fused pass reads
// one 32-byte entry per slot per tile instead of summing the terms itself.
__global__ __launch_bounds__(256) void k_slot_table(const FPass* __restrict__ passes, const FSlot* __restrict__ slots,
                                                    const FExtTerm* __restrict__ eterms, int pass_idx,
                                                    uint64_t ntiles, SlotVal* __restrict__ table) {
    const FPass& P = passes[pass_idx];
    const uint64_t total = ntiles * (uint64_t)P.nslots;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t ti = g / (uint64_t)P.nslots;
        const int sl = (int)(g - ti * (uint64_t)P.nslots);
        uint64_t base = 0, t = ti;
        for (int e = 0; e < P.next; ++e) {
            base |= (t & 1ULL) << P.epos[e];
            t >>= 1;
        }
        const FSlot S = slots[P.slot_begin + sl];
        double a = 0.0;
        for (uint32_t e = S.begin; e < S.end; ++e)
            if ((base & eterms[e].emask) == eterms[e].emask) a += eterms[e].ang;
        double sn, cs;
        sincos(a, &sn, &cs);
        table[g] = {a, cs, sn, 0.0};
    }
}

// Two waves per SIMD: two workgroups share a CU (one computes while the other
// waits on memory).
template <class R, int K>
__global__ __launch_bounds__(1 << (K - 4)) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_fused(const cx<R>* src, cx<R>* dst,
                                                        const FPass* __restrict__ passes,
                                                        const FStage* __restrict__ stages,
                                                        const FOp* __restrict__ ops,
                                                        const FTerm* __restrict__ terms,
                                                        const FSlot* __restrict__ slots,
                                                        const FExtTerm* __restrict__ eterms,
                                                        const double* __restrict__ alphas,
                                                        const SlotVal* __restrict__ slot_table, int pass_idx,
                                                        uint64_t ntiles) {
    const KArgs A{passes, stages, ops, terms, slots, eterms, alphas};
    constexpr int NT = 1 << (K - 4);
    constexpr int TB = K - 4;
    __shared__ cx<R> tile[1 << K];
    __shared__ SlotVal slotv[FUSE_MAX_SLOTS];
    const FPass& P = A.passes[pass_idx];
    const int tid = threadIdx.x;
    uint64_t gt = 0, ot = 0;
    uint32_t st_t = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1) {
            gt |= 1ULL << P.tq[b];
            ot |= 1ULL << P.so_q[b];
            st_t |= 1u << P.so_l[b];
        }
    uint64_t gw[4], ow[4];
    uint32_t sw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        gw[c] = 1ULL << P.tq[TB + c];
        ow[c] = 1ULL << P.so_q[TB + c];
        sw[c] = swz(1u << P.so_l[TB + c]);
    }
    const uint32_t swz_tid = swz((uint32_t)tid), swz_st = swz(st_t);
    const int next = P.next, nstages = P.nstages, nslots = P.nslots;
    const uint32_t stage_begin = P.stage_begin;
    for (uint64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        uint64_t base = 0, obase = 0;
        {
            uint64_t t = ti;
            for (int e = 0; e < next; ++e) {
                base |= (t & 1ULL) << P.epos[e];
                obase |= (t & 1ULL) << P.eout[e];
                t >>= 1;
            }
        }
        // load (issued first so the slot sums below overlap the memory latency)
        cx<R> ld[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint64_t g = base | gt;
            if (i & 1) g |= gw[0];
            if (i & 2) g |= gw[1];
            if (i & 4) g |= gw[2];
            if (i & 8) g |= gw[3];
            ld[i] = src[g];
        }
        // per-tile external factors (precomputed by k_slot_table)
        for (int sl = tid; sl < nslots; sl += NT) slotv[sl] = slot_table[ti * (uint64_t)nslots + sl];
#pragma unroll
        for (int i = 0; i < 16; ++i) tile[swz_tid ^ swz((uint32_t)(i * NT))] = ld[i];
        __syncthreads();
        for (int sg = 0; sg < nstages; ++sg) {
            const FStage& S = A.stages[stage_begin + sg];
            uint32_t toff = 0;
#pragma unroll
            for (int t = 0; t < TB; ++t)
                if ((tid >> t) & 1) toff |= 1u << S.tmap[t];
            const uint32_t r0 = swz(1u << S.rmap[0]), r1 = swz(1u << S.rmap[1]), r2 = swz(1u << S.rmap[2]),
                           r3 = swz(1u << S.rmap[3]);  // wave-uniform
            const uint32_t st_off = swz(toff);
            cx<R> a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                a[j] = tile[st_off ^ ((j & 1) ? r0 : 0u) ^ ((j & 2) ? r1 : 0u) ^ ((j & 4) ? r2 : 0u) ^
                            ((j & 8) ? r3 : 0u)];
            const uint32_t oe = S.op_end;
            for (uint32_t oi = S.op_begin; oi < oe; ++oi) {
                const FOp& op = A.ops[oi];
                if (op.type == FOP_MONO) {
                    double gr, gi;
                    mono_factor<TB>(op, A.terms, A.alphas, tid, toff, base, slotv, gr, gi);
                    const cx<R> g = {(R)gr, (R)gi};
                    const uint32_t rctl = op.rctl;
                    if (rctl == 0) {
#pragma unroll
                        for (int j = 0; j < 16; ++j) a[j] = cmul(g, a[j]);
                    } else {
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            if (((uint32_t)j & rctl) == rctl) a[j] = cmul(g, a[j]);  // wave-uniform branch
                    }
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
                tile[st_off ^ ((j & 1) ? r0 : 0u) ^ ((j & 2) ? r1 : 0u) ^ ((j & 4) ? r2 : 0u) ^ ((j & 8) ? r3 : 0u)] =
                    a[j];
            __syncthreads();
        }
        // store
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint64_t g = obase | ot;
            uint32_t ls = swz_st;
            if (i & 1) { g |= ow[0]; ls ^= sw[0]; }
            if (i & 2) { g |= ow[1]; ls ^= sw[1]; }
            if (i & 4) { g |= ow[2]; ls ^= sw[2]; }
            if (i & 8) { g |= ow[3]; ls ^= sw[3]; }
            dst[g] = tile[ls];
        }
        __syncthreads();
    }
}

// Device copy of a serialized program, reused when the same program runs again.
namespace {
size_t ring_len(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
}  // namespace

// Copy `bytes` into a fresh place of the program ring (growing the ring when
// the program does not fit) and return its offset. Cache entries whose device
// copy that place overwrites lose it.
int ring_upload(qse_ctx* ctx, const std::vector<char>& bytes, size_t* off) {
    ProgRing& R = ctx->ring;
    const size_t need = ring_len(bytes.size());
    if (need > R.cap) {
        QSE_HIP(hipStreamSynchronize(ctx->stream));
        for (auto& c : R.inflight) R.spare.push_back(c.done);
        R.inflight.clear();
        if (R.dev) QSE_HIP(hipFree(R.dev));
        if (R.host) QSE_HIP(hipHostFree(R.host));
        R.dev = nullptr;
        R.host = nullptr;
        R.cap = 0;
        const size_t cap = std::max<size_t>(2 * need, (size_t)16 << 20);
        QSE_HIP(hipMalloc((void**)&R.dev, cap));
        QSE_HIP(hipHostMalloc((void**)&R.host, cap, hipHostMallocDefault));
        R.cap = cap;
        R.head = 0;
        for (auto& e : ctx->prog_cache) e.dev = nullptr;
    }
    if (R.head + need > R.cap) R.head = 0;
    const size_t b = R.head, e = b + need;
    // Copies complete in stream order: waiting for the last one that read
    // [b, e) retires it and every older one; then retire finished ones.
    int last = -1;
    for (int i = 0; i < (int)R.inflight.size(); ++i)
        if (R.inflight[i].begin < e && b < R.inflight[i].end) last = i;
    if (last >= 0) QSE_HIP(hipEventSynchronize(R.inflight[last].done));
    size_t done = (size_t)(last + 1);
    while (done < R.inflight.size() && hipEventQuery(R.inflight[done].done) == hipSuccess) ++done;
    for (size_t i = 0; i < done; ++i) R.spare.push_back(R.inflight[i].done);
    R.inflight.erase(R.inflight.begin(), R.inflight.begin() + done);
    for (auto& en : ctx->prog_cache)
        if (en.dev && en.ring_off < e && b < en.ring_off + ring_len(en.bytes.size())) en.dev = nullptr;
    std::memcpy(R.host + b, bytes.data(), bytes.size());
    QSE_HIP(hipMemcpyAsync(R.dev + b, R.host + b, bytes.size(), hipMemcpyHostToDevice, ctx->stream));
    hipEvent_t ev;
    if (!R.spare.empty()) {
        ev = R.spare.back();
        R.spare.pop_back();
    } else {
        QSE_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    QSE_HIP(hipEventRecord(ev, ctx->stream));
    R.inflight.push_back({b, e, ev});
    R.head = e;
    *off = b;
    return QSE_OK;
}

int program_on_device(qse_ctx* ctx, const Serialized& S, int prec, ProgEntry** entry) {
    ProgEntry* hit = nullptr;
    for (auto& e : ctx->prog_cache) {
        if (e.prec == prec && e.bytes.size() == S.bytes.size() &&
            std::memcmp(e.bytes.data(), S.bytes.data(), S.bytes.size()) == 0) {
            hit = &e;
            break;
        }
    }
    if (!hit) {
        if (ctx->prog_cache.size() >= 32) {
            size_t oldest = 0;
            for (size_t i = 1; i < ctx->prog_cache.size(); ++i)
                if (ctx->prog_cache[i].stamp < ctx->prog_cache[oldest].stamp) oldest = i;
            if (ctx->prog_cache[oldest].mod) {
                // its kernels may still be queued
                QSE_HIP(hipStreamSynchronize(ctx->stream));
                QSE_HIP(hipModuleUnload(ctx->prog_cache[oldest].mod));
            }
            ctx->prog_cache.erase(ctx->prog_cache.begin() + oldest);
        }
        ProgEntry e;
        e.bytes = S.bytes;
        e.prec = prec;
        ctx->prog_cache.push_back(std::move(e));
        hit = &ctx->prog_cache.back();
    }
    if (!hit->dev) {
        size_t off = 0;
        const int rc = ring_upload(ctx, hit->bytes, &off);
        if (rc) return rc;
        hit->dev = ctx->ring.dev + off;
        hit->ring_off = off;
    }
    hit->stamp = ++ctx->prog_clock;
    *entry = hit;
    return QSE_OK;
}

// Compiled kernels for the program of `e` when the context's policy asks for
// them (QSE_JIT_ALWAYS, or QSE_JIT_AUTO from the program's second run on).
// Under QSE_JIT_AUTO a failed compilation is remembered and the generic kernel
// runs; under QSE_JIT_ALWAYS it is the caller's error.
int jit_for(qse_ctx* ctx, ProgEntry& e, const HostPlan& P, int prec, bool* use) {
    *use = false;
    ++e.runs;
    if (ctx->jit_mode == QSE_JIT_OFF || e.jit_failed) return QSE_OK;
    if (ctx->jit_mode == QSE_JIT_AUTO && e.runs < 2) return QSE_OK;
    if (!e.mod) {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = jit_build(P, prec, nullptr, &e.mod, e.fns);
        ctx->jit_compile_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc) {
            if (ctx->jit_mode == QSE_JIT_ALWAYS) return rc;
            e.jit_failed = true;
            return QSE_OK;
        }
        ++ctx->jit_compiled;
    }
    *use = true;
    return QSE_OK;
}

template <class R, int K>
static void launch_k(qse_ctx* ctx, const cx<R>* src, cx<R>* dst, const FProgram& prog, const SlotVal* table,
                     int pass, uint64_t ntiles) {
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 16);
    k_fused<R, K><<<grid, 1 << (K - 4), 0, ctx->stream>>>(src, dst, prog.passes, prog.stages, prog.ops, prog.terms,
                                                          prog.slots, prog.eterms, prog.alphas, table, pass, ntiles);
}

template <class R>
static int launch_pass(qse_ctx* ctx, qse_state* st, int K, const FProgram& prog, const SlotVal* table, int pass,
                       bool oop) {
    const uint64_t ntiles = 1ULL << (st->n - K);
    const cx<R>* s = (const cx<R>*)st->d;
    cx<R>* d = oop ? (cx<R>*)st->shadow : (cx<R>*)st->d;
    switch (K) {
        case 8: launch_k<R, 8>(ctx, s, d, prog, table, pass, ntiles); break;
        case 9: launch_k<R, 9>(ctx, s, d, prog, table, pass, ntiles); break;
        case 10: launch_k<R, 10>(ctx, s, d, prog, table, pass, ntiles); break;
        case 11: launch_k<R, 11>(ctx, s, d, prog, table, pass, ntiles); break;
        case 12: launch_k<R, 12>(ctx, s, d, prog, table, pass, ntiles); break;
        case 13: launch_k<R, 13>(ctx, s, d, prog, table, pass, ntiles); break;
        default: return set_error(QSE_ERR_UNSUPPORTED, "tile size outside [8, 13]");
    }
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}


// Key of a qse_run input for the plan cache: everything the plan depends on.
static std::vector<char> plan_key(const qse_ctx* ctx, const qse_state* st, const qse_op* ops, int nops,
                                  const double* params) {
    int maxp = -1;
    for (int i = 0; i < nops; ++i) maxp = std::max(maxp, (int)ops[i].param);
    const int head[6] = {st->n, st->prec, ctx->tile_bits[st->prec], ctx->low_bits[st->prec], nops, maxp + 1};
    std::vector<char> key(sizeof(head) + sizeof(qse_op) * (size_t)nops + sizeof(double) * (size_t)(maxp + 1));
    char* w = key.data();
    std::memcpy(w, head, sizeof(head));
    w += sizeof(head);
    if (nops) std::memcpy(w, ops, sizeof(qse_op) * (size_t)nops);
    w += sizeof(qse_op) * (size_t)nops;
    if (maxp >= 0 && params) std::memcpy(w, params, sizeof(double) * (size_t)(maxp + 1));
    return key;
}

// Per-tile slot values of pass p of a program already on the device (`dev`,
// laid out as S) into table (ntiles x the pass's slots); nothing for a pass
// without slots.
int slot_table_async(qse_ctx* ctx, const char* dev, const Serialized& S, const HostPlan& P, int p, uint64_t ntiles,
                     SlotVal* table) {
    const uint64_t work = ntiles * (uint64_t)P.passes[p].nslots;
    if (work == 0) return QSE_OK;
    k_slot_table<<<grid_for(ctx, work, 256, 8), 256, 0, ctx->stream>>>(
        (const FPass*)(dev + S.off[0]), (const FSlot*)(dev + S.off[4]), (const FExtTerm*)(dev + S.off[5]), p, ntiles,
        table);
    QSE_HIP(hipGetLastError());
    return QSE_OK;
}

// Launch the passes of `plan` (tile qubits k) on st: program upload (ring),
// run-time specialised kernels when allowed and available (jit_kind: keyed by
// the exact program, literal coefficients; or keyed by the plan structure,
// coefficients read from the program), per-pass slot tables, and the pointer
// exchange after each out-of-place pass.
int run_plan(qse_ctx* ctx, qse_state* st, const HostPlan& plan, int k, int jit_kind) {
    int rc;
    if (plan_out_of_place(plan) && !st->shadow)
        return set_error(QSE_ERR_UNSUPPORTED, "out-of-place plan without a second buffer");
    Serialized S;
    serialize(plan, S);
    ProgEntry* entry = nullptr;
    rc = program_on_device(ctx, S, st->prec, &entry);
    if (rc) return rc;
    const std::vector<hipFunction_t>* fns = nullptr;
    if (jit_kind == JIT_BY_PROGRAM) {
        bool jit = false;
        rc = jit_for(ctx, *entry, plan, st->prec, &jit);
        if (jit) {
            fns = &entry->fns;
        } else if (rc == QSE_OK && ctx->jit_mode == QSE_JIT_AUTO && entry->runs == 1) {
            // a program seen for the first time: when its structure keeps
            // coming back with other angles (a parameter loop), the
            // structure-keyed kernels serve it
            rc = jit_for_structure(ctx, plan, S, st->prec, &fns);
        }
    } else if (jit_kind == JIT_BY_STRUCTURE) {
        rc = jit_for_structure(ctx, plan, S, st->prec, &fns);
    }
    if (rc) return rc;
    char* dev = (char*)entry->dev;
    FProgram prog;
    prog.passes = (const FPass*)(dev + S.off[0]);
    prog.stages = (const FStage*)(dev + S.off[1]);
    prog.ops = (const FOp*)(dev + S.off[2]);
    prog.terms = (const FTerm*)(dev + S.off[3]);
    prog.slots = (const FSlot*)(dev + S.off[4]);
    prog.eterms = (const FExtTerm*)(dev + S.off[5]);
    prog.alphas = (const double*)(dev + S.off[6]);
    const double bytes = 2.0 * (double)amp_bytes(st->prec) * (double)(1ULL << st->n);
    const uint64_t ntiles = 1ULL << (st->n - k);
    int max_slots = 0;
    for (const FPass& fp : plan.passes) max_slots = std::max(max_slots, fp.nslots);
    SlotVal* table = nullptr;
    if (max_slots > 0) {
        void* tb;
        rc = scratch(ctx, 5, sizeof(SlotVal) * ntiles * (uint64_t)max_slots, &tb);
        if (rc) return rc;
        table = (SlotVal*)tb;
    }
    for (int p = 0; p < (int)plan.passes.size(); ++p) {
        if (plan.passes[p].nslots > 0) {
            const uint64_t work = ntiles * (uint64_t)plan.passes[p].nslots;
            k_slot_table<<<grid_for(ctx, work, 256, 8), 256, 0, ctx->stream>>>(prog.passes, prog.slots, prog.eterms,
                                                                              p, ntiles, table);
            QSE_HIP(hipGetLastError());
        }
        TimedLaunch t;
        launch_begin(ctx, FAM_FUSED, &t);
        const bool oop = plan.passes[p].out_of_place != 0;
        if (fns) {
            rc = jit_launch(ctx, (*fns)[p], st->d, oop ? st->shadow : st->d, table, ntiles, k, entry->dev);
            ++ctx->jit_launches;
        } else {
            rc = (st->prec == QSE_C128) ? launch_pass<double>(ctx, st, k, prog, table, p, oop)
                                        : launch_pass<float>(ctx, st, k, prog, table, p, oop);
        }
        if (rc) return rc;
        launch_end(ctx, &t, bytes);
        if (oop) std::swap(st->d, st->shadow);
    }
    return QSE_OK;
}

int run_fused(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int* passes_out,
              bool dry_run, bool allow_jit) {
    if (st->n < FUSE_MIN_QUBITS) {
        if (passes_out) *passes_out = nops;
        return dry_run ? QSE_OK : run_sweeps(ctx, st, ops, nops, params);
    }
    // qse_run inputs (allow_jit) get the tuned shape and a cached plan; the
    // adjoint's internal runs (many distinct small programs) plan directly with
    // the default shape.
    int k, L, rc;
    HostPlan P;
    const HostPlan* plan = &P;
    if (allow_jit) {
        std::vector<char> key = plan_key(ctx, st, ops, nops, params);
        PlanEntry* pe = nullptr;
        for (auto& e : ctx->plan_cache)
            if (e.key == key) pe = &e;
        if (!pe) {
            PlanEntry e;
            e.key = std::move(key);
            rc = tuned_plan(ctx, st->n, st->prec, ops, nops, params, e.plan, true, &e.k, &e.L);
            if (rc) return rc;
            if (ctx->plan_cache.size() >= 32) {
                size_t oldest = 0;
                for (size_t i = 1; i < ctx->plan_cache.size(); ++i)
                    if (ctx->plan_cache[i].stamp < ctx->plan_cache[oldest].stamp) oldest = i;
                ctx->plan_cache.erase(ctx->plan_cache.begin() + oldest);
            }
            ctx->plan_cache.push_back(std::move(e));
            pe = &ctx->plan_cache.back();
        }
        pe->stamp = ++ctx->prog_clock;
        plan = &pe->plan;
        k = pe->k;
        L = pe->L;
    } else {
        fusion_shape(ctx, st->n, st->prec, &k, &L);
        rc = choose_plan(st->n, k, L, ops, nops, params, P, true);
        if (rc) return rc;
    }
    if (passes_out) *passes_out = (int)plan->passes.size();
    if (dry_run) return QSE_OK;
    const bool needs_shadow = plan_out_of_place(*plan);
    if (needs_shadow && !st->shadow) {
        if (hipMalloc(&st->shadow, st->bytes) != hipSuccess) {
            // no room for a second buffer: realise SWAPs in place instead
            (void)hipGetLastError();
            st->shadow = nullptr;
            P = HostPlan();
            rc = choose_plan(st->n, k, L, ops, nops, params, P, false);
            if (rc) return rc;
            plan = &P;
        }
    }
    return run_plan(ctx, st, *plan, k, allow_jit ? JIT_BY_PROGRAM : JIT_NONE);
}

}  // namespace qse

using namespace qse;

// Measured choice of the tile shape: the cost model's best candidates are each
// compiled and timed on st (three runs after two), the fastest is kept for the
// gate list's structure and used by later runs of it.
extern "C" int qse_autotune(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params,
                            int max_candidates, int* k_out, int* L_out, double* ms_out, int* tried_k, int* tried_L,
                            double* tried_ms, int* ntried) {
    if (!ctx || !st || !st->d) return set_error(QSE_ERR_INVALID_ARG, "null context or state");
    if (nops < 0 || (nops > 0 && !ops) || max_candidates < 1)
        return set_error(QSE_ERR_INVALID_ARG, "bad op list or candidate count");
    int rc = validate_ops(st->n, ops, nops, params);
    if (rc) return rc;
    if (ntried) *ntried = 0;
    if (st->n < FUSE_MIN_QUBITS || nops == 0) {
        if (k_out) *k_out = 0;
        if (L_out) *L_out = 0;
        if (ms_out) *ms_out = 0.0;
        return QSE_OK;
    }
    const int prec = st->prec;
    std::vector<ShapeCandidate> cand;
    rc = rank_shapes(st->n, prec, ops, nops, params, true, cand);
    if (rc) return rc;
    if (cand.empty()) return set_error(QSE_ERR_UNSUPPORTED, "autotune: no candidate tile shape");
    if ((int)cand.size() > max_candidates) cand.resize(max_candidates);
    const int saved_mode = ctx->jit_mode, saved_k = ctx->tile_bits[prec], saved_L = ctx->low_bits[prec];
    const bool saved_timing = ctx->timing;
    ctx->jit_mode = QSE_JIT_ALWAYS;
    ctx->timing = false;
    hipEvent_t a = nullptr, b = nullptr;
    int best = -1;
    double best_ms = 0.0;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e != hipSuccess) rc = hip_error(e, "hipEventCreate");
    for (int c = 0; c < (int)cand.size() && rc == QSE_OK; ++c) {
        ctx->tile_bits[prec] = cand[c].k;
        ctx->low_bits[prec] = cand[c].L;
        for (int r = 0; r < 2 && rc == QSE_OK; ++r) rc = run_fused(ctx, st, ops, nops, params, nullptr, false, true);
        if (rc) break;
        e = hipEventRecord(a, ctx->stream);
        for (int r = 0; r < 3 && rc == QSE_OK && e == hipSuccess; ++r)
            rc = run_fused(ctx, st, ops, nops, params, nullptr, false, true);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e != hipSuccess) {
            rc = hip_error(e, "autotune timing");
            break;
        }
        const double per = ms / 3.0;
        if (tried_k) tried_k[c] = cand[c].k;
        if (tried_L) tried_L[c] = cand[c].L;
        if (tried_ms) tried_ms[c] = per;
        if (ntried) *ntried = c + 1;
        if (best < 0 || per < best_ms) {
            best = c;
            best_ms = per;
        }
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    ctx->jit_mode = saved_mode;
    ctx->tile_bits[prec] = saved_k;
    ctx->low_bits[prec] = saved_L;
    ctx->timing = saved_timing;
    if (rc) return rc;
    remember_shape(ctx, gate_structure_key(st->n, prec, ops, nops), cand[best].k, cand[best].L);
    ctx->plan_cache.clear();  // plans cached before the measurement may have another shape
    if (k_out) *k_out = cand[best].k;
    if (L_out) *L_out = cand[best].L;
    if (ms_out) *ms_out = best_ms;
    return QSE_OK;
}

extern "C" int qse_plan_info(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, int* passes,
                             int* tile_bits, int* low_bits, int* stages) {
    if (nqubits < 1 || nqubits > 40 || (precision != QSE_C128 && precision != QSE_C64) || nops < 0 ||
        (nops > 0 && !ops))
        return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    if (nqubits < FUSE_MIN_QUBITS) {
        // one sweep per gate
        if (passes) *passes = nops;
        if (tile_bits) *tile_bits = 0;
        if (low_bits) *low_bits = 0;
        if (stages) *stages = nops;
        return QSE_OK;
    }
    // parameters do not change the plan's shape: plan with a zero parameter vector
    int maxp = -1;
    for (int i = 0; i < nops; ++i) maxp = std::max(maxp, (int)ops[i].param);
    std::vector<double> zeros(maxp + 1, 0.0);
    int k, L;
    HostPlan P;
    const int rc = tuned_plan(ctx, nqubits, precision, ops, nops, zeros.data(), P, true, &k, &L);
    if (rc) return rc;
    if (passes) *passes = (int)P.passes.size();
    if (tile_bits) *tile_bits = k;
    if (low_bits) *low_bits = L;
    if (stages) *stages = (int)P.stages.size();
    return QSE_OK;
}

extern "C" int qse_plan_passes(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, int* passes) {
    if (!passes) return set_error(QSE_ERR_INVALID_ARG, "null passes");
    return qse_plan_info(ctx, nqubits, precision, ops, nops, passes, nullptr, nullptr, nullptr);
}

extern "C" int qse_ctx_set_fusion(qse_ctx* ctx, int precision, int tile_bits, int low_bits) {
    if (!ctx || precision < 0 || precision > 1) return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    if (tile_bits != 0 && (tile_bits < FUSE_MIN_K || tile_bits > FUSE_MAX_K))
        return set_error(QSE_ERR_INVALID_ARG, "tile_bits must be 0 or in [8, 13]");
    if (low_bits < 0 || low_bits > 9) return set_error(QSE_ERR_INVALID_ARG, "low_bits must be in [0, 9]");
    ctx->tile_bits[precision] = tile_bits;
    ctx->low_bits[precision] = low_bits;
    return QSE_OK;
}

extern "C" int qse_ctx_get_fusion(qse_ctx* ctx, int precision, int* tile_bits, int* low_bits) {
    if (!ctx || precision < 0 || precision > 1) return set_error(QSE_ERR_INVALID_ARG, "bad argument");
    if (tile_bits) *tile_bits = ctx->tile_bits[precision];
    if (low_bits) *low_bits = ctx->low_bits[precision];
    return QSE_OK;
}

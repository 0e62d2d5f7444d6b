// fused_emu.cpp — CPU emulation of the fused-pass kernel (test infrastructure).
//
// Builds the same plan as the library (qse_plan.cpp, compiled in) and executes
// it with the kernel's addressing and op semantics on the host, thread by
// thread, so the planner (pass splitting, register/thread bit maps, in-tile
// SWAP relabelling, monomial flushing, external slots) can be checked against
// the CPU oracle without a GPU. Mirrors k_fused in qse_fused.hip; it is not
// linked into the product library.
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "qse_batch.hpp"
#include "qse_fused.hpp"
#include "qse_hip.h"

namespace qse {
// The planner reports errors through this hook (libqse_hip keeps a message).
int set_error(int status, const std::string&) { return status; }
int build_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P, int flags);
int choose_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P,
                bool allow_oop);
}  // namespace qse

using namespace qse;
using cd = std::complex<double>;

// Plan for the emulator entry points: flags >= 0 selects one planner variant
// (PLAN_RELABEL / PLAN_DAG bits; a relabelling plan that is not realisable
// falls back to in place), flags < 0 the library's own choice.
static int plan_for(int n, int k, int L, const qse_op* ops, int nops, const double* params, int flags, HostPlan& P) {
    if (flags < 0) return choose_plan(n, k, L, ops, nops, params, P, true);
    int rc = build_plan(n, k, L, ops, nops, params, P, flags);
    if (rc && (flags & PLAN_RELABEL)) {
        P = HostPlan();
        rc = build_plan(n, k, L, ops, nops, params, P, flags & ~PLAN_RELABEL);
    }
    return rc;
}

static uint32_t swz(uint32_t li) { return li ^ (((li >> 4) ^ (li >> 8) ^ (li >> 12)) & 15u); }

static void apply_op(const HostPlan& P, const FOp& op, cd* a, int tid, int TB, uint32_t toff, uint64_t base,
                     const cd* slotv) {
    if (op.type == FOP_MONO) {
        cd g(op.m[0], op.m[1]);
        if (op.alpha_begin != FUSE_NO_ALPHA)
            for (int t = 0; t < TB; ++t)
                if (((op.tctl >> t) & 1) && ((tid >> t) & 1))
                    g *= cd(P.alphas[op.alpha_begin + 2 * t], P.alphas[op.alpha_begin + 2 * t + 1]);
        for (uint32_t t = op.term_begin; t < op.term_end; ++t) {
            const FTerm& T = P.terms[t];
            if ((toff & T.tmask) != T.tmask) continue;
            if (T.slot >= 0) g *= std::polar(1.0, slotv[T.slot].real());
            else if ((base & T.emask) == T.emask) g *= cd(T.d[0], T.d[1]);
        }
        for (int j = 0; j < 16; ++j)
            if (((uint32_t)j & op.rctl) == op.rctl) a[j] *= g;
        return;
    }
    if ((toff & op.tctl) != op.tctl || (base & op.ectl) != op.ectl) return;
    const int p = op.p;
    const cd m00(op.m[0], op.m[1]), m01(op.m[2], op.m[3]), m10(op.m[4], op.m[5]), m11(op.m[6], op.m[7]);
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> p) << (p + 1)) | (jj & ((1 << p) - 1)), j1 = j0 | (1 << p);
        if (((uint32_t)j0 & op.rctl) != op.rctl) continue;
        const cd v0 = a[j0], v1 = a[j1];
        if (op.type == FOP_H && op.rctl == 0) {
            const double h = 0.7071067811865475244;
            a[j0] = h * (v0 + v1);
            a[j1] = h * (v0 - v1);
        } else if (op.type == FOP_XSWAP) {
            a[j0] = v1;
            a[j1] = v0;
        } else {
            a[j0] = m00 * v0 + m01 * v1;
            a[j1] = m10 * v0 + m11 * v1;
        }
    }
}

// One pass of plan P over the 2^n amplitudes of s, written to dst (dst may be
// s: each tile writes the addresses it read).
static void emu_pass(const HostPlan& P, const FPass& fp, int n, const cd* s, cd* dst) {
    const int k = fp.k;
    const int NT = 1 << (k - 4), TB = k - 4;
    std::vector<cd> tile(1u << k), slotv(FUSE_MAX_SLOTS);
    const uint64_t ntiles = 1ULL << (n - k);
    for (uint64_t ti = 0; ti < ntiles; ++ti) {
        uint64_t base = 0, obase = 0, t = ti;
        for (int e = 0; e < fp.next; ++e) {
            base |= (t & 1ULL) << fp.epos[e];
            obase |= (t & 1ULL) << fp.eout[e];
            t >>= 1;
        }
        for (int sl = 0; sl < fp.nslots; ++sl) {
            const FSlot& S = P.slots[fp.slot_begin + sl];
            double a = 0.0;
            for (uint32_t e = S.begin; e < S.end; ++e)
                if ((base & P.eterms[e].emask) == P.eterms[e].emask) a += P.eterms[e].ang;
            slotv[sl] = cd(a, 0.0);  // the slot's angle, carried in the real part
        }
        for (int tid = 0; tid < NT; ++tid)
            for (int i = 0; i < 16; ++i) {
                const uint32_t li = (uint32_t)(i * NT + tid);
                uint64_t g = base;
                for (int b = 0; b < k; ++b)
                    if ((li >> b) & 1) g |= 1ULL << fp.tq[b];
                tile[swz(li)] = s[g];
            }
        for (int sg = 0; sg < fp.nstages; ++sg) {
            const FStage& S = P.stages[fp.stage_begin + sg];
            for (int tid = 0; tid < NT; ++tid) {
                uint32_t toff = 0;
                for (int tb = 0; tb < TB; ++tb)
                    if ((tid >> tb) & 1) toff |= 1u << S.tmap[tb];
                uint32_t li[16];
                cd a[16];
                for (int j = 0; j < 16; ++j) {
                    uint32_t l = toff;
                    for (int p = 0; p < 4; ++p)
                        if ((j >> p) & 1) l |= 1u << S.rmap[p];
                    li[j] = swz(l);
                    a[j] = tile[li[j]];
                }
                for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi)
                    apply_op(P, P.ops[oi], a, tid, TB, toff, base, slotv.data());
                for (int j = 0; j < 16; ++j) tile[li[j]] = a[j];
            }
        }
        for (int tid = 0; tid < NT; ++tid)
            for (int i = 0; i < 16; ++i) {
                const uint32_t go = (uint32_t)(i * NT + tid);
                uint64_t g = obase;
                uint32_t ls = 0;
                for (int b = 0; b < k; ++b)
                    if ((go >> b) & 1) {
                        g |= 1ULL << fp.so_q[b];
                        ls |= 1u << fp.so_l[b];
                    }
                dst[g] = tile[swz(ls)];
            }
    }
}

// This thread's part of a probe (k_batch's probe_regs): Im<lambda|G|psi> over
// the 16 register amplitudes, psi in a and lambda in b.
static double probe_value(const FOp& op, const cd* a, const cd* b, uint32_t toff, uint64_t base) {
    auto im = [](cd l, cd p) { return (std::conj(l) * p).imag(); };
    double s = 0.0;
    if (op.mode == PROBE_X || op.mode == PROBE_Y) {
        const int p = op.p;
        for (int jj = 0; jj < 8; ++jj) {
            const int j0 = ((jj >> p) << (p + 1)) | (jj & ((1 << p) - 1)), j1 = j0 | (1 << p);
            if (op.mode == PROBE_X) s += im(b[j0], a[j1]) + im(b[j1], a[j0]);
            else s += (std::conj(b[j1]) * a[j0]).real() - (std::conj(b[j0]) * a[j1]).real();
        }
        return s;
    }
    if (op.mode == PROBE_PROJ) {
        if ((toff & op.tctl) != op.tctl || (base & op.ectl) != op.ectl) return 0.0;
        for (int j = 0; j < 16; ++j)
            if (((uint32_t)j & op.rctl) == op.rctl) s += im(b[j], a[j]);
        return s;
    }
    for (int j = 0; j < 16; ++j) s += (__builtin_popcount((uint32_t)j & op.rctl) & 1) ? -im(b[j], a[j]) : im(b[j], a[j]);
    return ((__builtin_popcount(toff & op.tctl) + __builtin_popcountll(base & op.ectl)) & 1) ? -s : s;
}

// One joint pass (k_batch's batch_pass with J): the pass's gates on psi and
// lambda, in place, and its probes summed into vals[fp.probe_base + probe].
static void emu_joint_pass(const HostPlan& P, const FPass& fp, int n, cd* psi, cd* lam, double* vals) {
    const int k = fp.k;
    const int NT = 1 << (k - 4), TB = k - 4;
    std::vector<cd> tile(1u << k), tile2(1u << k), slotv(FUSE_MAX_SLOTS);
    std::vector<double> acc(std::max(fp.nprobes, 1), 0.0);
    const uint64_t ntiles = 1ULL << (n - k);
    for (uint64_t ti = 0; ti < ntiles; ++ti) {
        uint64_t base = 0, obase = 0, t = ti;
        for (int e = 0; e < fp.next; ++e) {
            base |= (t & 1ULL) << fp.epos[e];
            obase |= (t & 1ULL) << fp.eout[e];
            t >>= 1;
        }
        for (int sl = 0; sl < fp.nslots; ++sl) {
            const FSlot& S = P.slots[fp.slot_begin + sl];
            double a = 0.0;
            for (uint32_t e = S.begin; e < S.end; ++e)
                if ((base & P.eterms[e].emask) == P.eterms[e].emask) a += P.eterms[e].ang;
            slotv[sl] = cd(a, 0.0);
        }
        for (int tid = 0; tid < NT; ++tid)
            for (int i = 0; i < 16; ++i) {
                const uint32_t li = (uint32_t)(i * NT + tid);
                uint64_t g = base;
                for (int b = 0; b < k; ++b)
                    if ((li >> b) & 1) g |= 1ULL << fp.tq[b];
                tile[swz(li)] = psi[g];
                tile2[swz(li)] = lam[g];
            }
        for (int sg = 0; sg < fp.nstages; ++sg) {
            const FStage& S = P.stages[fp.stage_begin + sg];
            for (int tid = 0; tid < NT; ++tid) {
                uint32_t toff = 0;
                for (int tb = 0; tb < TB; ++tb)
                    if ((tid >> tb) & 1) toff |= 1u << S.tmap[tb];
                uint32_t li[16];
                cd a[16], b[16];
                for (int j = 0; j < 16; ++j) {
                    uint32_t l = toff;
                    for (int p = 0; p < 4; ++p)
                        if ((j >> p) & 1) l |= 1u << S.rmap[p];
                    li[j] = swz(l);
                    a[j] = tile[li[j]];
                    b[j] = tile2[li[j]];
                }
                for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
                    const FOp& op = P.ops[oi];
                    if (op.type == FOP_PROBE) {
                        acc[op.term_begin] += probe_value(op, a, b, toff, base);
                    } else {
                        apply_op(P, op, a, tid, TB, toff, base, slotv.data());
                        apply_op(P, op, b, tid, TB, toff, base, slotv.data());
                    }
                }
                for (int j = 0; j < 16; ++j) {
                    tile[li[j]] = a[j];
                    tile2[li[j]] = b[j];
                }
            }
        }
        for (int tid = 0; tid < NT; ++tid)
            for (int i = 0; i < 16; ++i) {
                const uint32_t go = (uint32_t)(i * NT + tid);
                uint64_t g = obase;
                uint32_t ls = 0;
                for (int b = 0; b < k; ++b)
                    if ((go >> b) & 1) {
                        g |= 1ULL << fp.so_q[b];
                        ls |= 1u << fp.so_l[b];
                    }
                psi[g] = tile[swz(ls)];
                lam[g] = tile2[swz(ls)];
            }
    }
    for (int p = 0; p < fp.nprobes; ++p) vals[fp.probe_base + p] = acc[p];
}

extern "C" int emu_run_fused(int n, int k, int L, const qse_op* ops, int nops, const double* params, double* psi,
                             int* npasses, int relabel) {
    HostPlan P;
    int rc = plan_for(n, k, L, ops, nops, params, relabel, P);
    if (rc) return rc;
    *npasses = (int)P.passes.size();
    const uint64_t N = 1ULL << n;
    std::vector<cd> bufA(reinterpret_cast<cd*>(psi), reinterpret_cast<cd*>(psi) + N), bufB(N);
    cd* s = bufA.data();
    cd* shadow = bufB.data();
    for (const FPass& fp : P.passes) {
        emu_pass(P, fp, n, s, fp.out_of_place ? shadow : s);
        if (fp.out_of_place) std::swap(s, shadow);
    }
    std::memcpy(psi, s, sizeof(cd) * N);
    return 0;
}

// Structural checks of a plan: stage maps partition the tile bits, the store
// permutation is a permutation, passes respect the tile size.
extern "C" int emu_check_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, int relabel) {
    HostPlan P;
    int rc = plan_for(n, k, L, ops, nops, params, relabel, P);
    if (rc) return -1;
    for (const FPass& fp : P.passes) {
        if (fp.k != k || fp.next != n - k) return 1;
        uint64_t seen = 0, dq = 0, sq = 0;
        for (int b = 0; b < k; ++b) {
            seen |= 1ULL << fp.so_l[b];
            dq |= 1ULL << fp.so_q[b];
            sq |= 1ULL << fp.tq[b];
            if (b > 0 && fp.so_q[b] <= fp.so_q[b - 1]) return 6;
        }
        if (seen != ((1ULL << k) - 1)) return 2;
        for (int e = 0; e < fp.next; ++e) dq |= 1ULL << fp.eout[e];
        if (dq != ((1ULL << n) - 1)) return 7;
        if (!fp.out_of_place && sq != (dq & sq)) return 8;
        for (int b = 0; b < L; ++b)
            if (fp.so_q[b] != b) return 9;
        for (int b = 0; b < L; ++b)
            if (fp.tq[b] != b) return 3;
        for (int sg = 0; sg < fp.nstages; ++sg) {
            const FStage& S = P.stages[fp.stage_begin + sg];
            uint64_t cov = 0;
            for (int p = 0; p < 4; ++p) cov |= 1ULL << S.rmap[p];
            for (int t = 0; t < k - 4; ++t) cov |= 1ULL << S.tmap[t];
            if (cov != ((1ULL << k) - 1)) return 4;
        }
        if (fp.nslots > FUSE_MAX_SLOTS) return 5;
        // every index the kernels follow stays inside the program's arrays
        if ((size_t)fp.stage_begin + fp.nstages > P.stages.size()) return 10;
        if ((size_t)fp.slot_begin + fp.nslots > P.slots.size()) return 11;
        for (int sl = 0; sl < fp.nslots; ++sl) {
            const FSlot& S = P.slots[fp.slot_begin + sl];
            if (S.begin > S.end || S.end > P.eterms.size()) return 12;
        }
        for (int sg = 0; sg < fp.nstages; ++sg) {
            const FStage& S = P.stages[fp.stage_begin + sg];
            if (S.op_begin > S.op_end || S.op_end > P.ops.size()) return 13;
            for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
                const FOp& op = P.ops[oi];
                if (op.rctl >= 16) return 14;
                if (op.type == FOP_MONO) {
                    if (op.term_begin > op.term_end || op.term_end > P.terms.size()) return 15;
                    if (op.tctl >> (k - 4)) return 16;
                    if (op.tctl && (op.alpha_begin == FUSE_NO_ALPHA ||
                                    (size_t)op.alpha_begin + 2 * (k - 4) > P.alphas.size()))
                        return 17;
                    for (uint32_t t = op.term_begin; t < op.term_end; ++t) {
                        const FTerm& T = P.terms[t];
                        if (T.slot >= fp.nslots) return 18;
                        if (T.tmask >> k) return 19;
                    }
                } else {
                    if (op.p >= 4) return 20;
                    if (op.tctl >> k) return 21;
                    if (op.type != FOP_DENSE && op.type != FOP_XSWAP && (op.rctl || op.tctl || op.ectl)) return 22;
                }
            }
        }
    }
    return 0;
}

// Plan summary for diagnostics: per pass, stages, dense ops, MONO ops, terms,
// alpha vectors and slots.
extern "C" int emu_plan_stats(int n, int k, int L, const qse_op* ops, int nops, const double* params, int* out,
                              int max_passes, int flags) {
    HostPlan P;
    if (plan_for(n, k, L, ops, nops, params, flags, P)) return -1;
    int np = 0;
    for (const FPass& fp : P.passes) {
        if (np >= max_passes) break;
        int dense = 0, mono = 0, terms = 0, alpha = 0;
        for (int sg = 0; sg < fp.nstages; ++sg) {
            const FStage& S = P.stages[fp.stage_begin + sg];
            for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
                const FOp& op = P.ops[oi];
                if (op.type == FOP_MONO) {
                    ++mono;
                    terms += (int)(op.term_end - op.term_begin);
                    alpha += __builtin_popcount(op.tctl);
                } else {
                    ++dense;
                }
            }
        }
        int* o = out + 6 * np;
        o[0] = fp.nstages; o[1] = dense; o[2] = mono; o[3] = terms; o[4] = alpha; o[5] = fp.nslots;
        ++np;
    }
    return np;
}

// ---- run-time specialised passes (qse_jit_gen.cpp) ----
// emu_jit_prepare plans like the library (relabelling plan, in-place plan when
// it is not realisable), keeps the plan for the calls below and returns the
// length of the generated HIP source, copied into `out` when `cap` allows.
static HostPlan g_jit_plan;
static bool g_jit_paired = false;

// Generate complex128 passes with paired tiles (the library pairs passes of
// 12 tile-index bits or more whose paired kernel does not spill registers).
extern "C" void emu_jit_set_pairing(int on) { g_jit_paired = on != 0; }
static Serialized g_jit_prog;
static int g_jit_n = 0;

// table != 0: coefficients read from the serialized program (emu_jit_program)
// instead of literals.
extern "C" long emu_jit_prepare2(int n, int k, int L, const qse_op* ops, int nops, const double* params, int relabel,
                                 int prec, int table, char* out, long cap) {
    g_jit_plan = HostPlan();
    int rc = plan_for(n, k, L, ops, nops, params, relabel, g_jit_plan);
    if (rc) return -rc;
    g_jit_n = n;
    serialize(g_jit_plan, g_jit_prog);
    const std::vector<bool> paired(g_jit_plan.passes.size(), g_jit_paired);
    const std::string src = jit_source(g_jit_plan, prec, table ? g_jit_prog.off : nullptr, &paired);
    if (out && cap > (long)src.size()) std::memcpy(out, src.c_str(), src.size() + 1);
    return (long)src.size();
}

extern "C" long emu_jit_prepare(int n, int k, int L, const qse_op* ops, int nops, const double* params, int relabel,
                                int prec, char* out, long cap) {
    return emu_jit_prepare2(n, k, L, ops, nops, params, relabel, prec, 0, out, cap);
}

// The prepared plan's serialized program (what the device program buffer
// holds), its size in bytes when out is null.
extern "C" long emu_jit_program(char* out) {
    if (out) std::memcpy(out, g_jit_prog.bytes.data(), g_jit_prog.bytes.size());
    return (long)g_jit_prog.bytes.size();
}

// Structure keys of two plans compare equal (1) or not (0).
extern "C" int emu_same_structure(int n, int k, int L, const qse_op* ops_a, const double* params_a,
                                  const qse_op* ops_b, const double* params_b, int nops) {
    HostPlan A, B;
    if (plan_for(n, k, L, ops_a, nops, params_a, -1, A) || plan_for(n, k, L, ops_b, nops, params_b, -1, B)) return -1;
    return structure_key(A) == structure_key(B) ? 1 : 0;
}

// Pass count of the prepared plan, and per pass: tile qubits, out-of-place flag, slots.
extern "C" int emu_jit_pass_info(int pass, int* k, int* out_of_place, int* nslots) {
    if (pass < 0) return (int)g_jit_plan.passes.size();
    const FPass& fp = g_jit_plan.passes[pass];
    *k = fp.k;
    *out_of_place = fp.out_of_place;
    *nslots = fp.nslots;
    return 0;
}

// Slot table of `pass` as k_slot_table writes it: per tile, per slot
// {angle, cos, sin, 0}.
extern "C" int emu_jit_slot_table(int pass, double* table) {
    const FPass& fp = g_jit_plan.passes[pass];
    const uint64_t ntiles = 1ULL << (g_jit_n - fp.k);
    for (uint64_t ti = 0; ti < ntiles; ++ti) {
        uint64_t base = 0, t = ti;
        for (int e = 0; e < fp.next; ++e) {
            base |= (t & 1ULL) << fp.epos[e];
            t >>= 1;
        }
        for (int sl = 0; sl < fp.nslots; ++sl) {
            const FSlot& S = g_jit_plan.slots[fp.slot_begin + sl];
            double a = 0.0;
            for (uint32_t e = S.begin; e < S.end; ++e)
                if ((base & g_jit_plan.eterms[e].emask) == g_jit_plan.eterms[e].emask) a += g_jit_plan.eterms[e].ang;
            double* row = table + 4 * (ti * (uint64_t)fp.nslots + sl);
            row[0] = a;
            row[1] = std::cos(a);
            row[2] = std::sin(a);
            row[3] = 0.0;
        }
    }
    return 0;
}

// Tile qubits of pass `pass` of the prepared plan (load order: tq[0..k-1]).
extern "C" int emu_jit_pass_qubits(int pass, uint32_t* tq) {
    const FPass& fp = g_jit_plan.passes[pass];
    for (int b = 0; b < fp.k; ++b) tq[b] = fp.tq[b];
    return fp.k;
}

// ---- joint reverse-sweep passes (jit_joint_source) ----
// emu_joint_prepare builds the reverse sweep of a parameterised list
// (build_reverse_plan), keeps its chunks and probe references, and returns the
// generated HIP source of chunk `chunk` (its length; copied into `out` when
// `cap` allows). The calls below read the kept chunk.
static std::vector<HostPlan> g_joint_chunks;
static std::vector<ProbeRef> g_joint_refs;
static Serialized g_joint_prog;
static int g_joint_n = 0, g_joint_cur = 0;

extern "C" long emu_joint_prepare(int n, int k, int L, const qse_op* ops, int nops, const double* params, int chunk,
                                  char* out, long cap) {
    g_joint_chunks.clear();
    g_joint_refs.clear();
    const int rc = build_reverse_plan(n, k, L, ops, nops, params, g_joint_chunks, g_joint_refs);
    if (rc) return -rc;
    if (chunk < 0) return (long)g_joint_chunks.size();
    if (chunk >= (int)g_joint_chunks.size()) return -1;
    g_joint_n = n;
    g_joint_cur = chunk;
    serialize(g_joint_chunks[chunk], g_joint_prog);
    const std::string src = jit_joint_source(g_joint_chunks[chunk], QSE_C128, g_joint_prog.off);
    if (out && cap > (long)src.size()) std::memcpy(out, src.c_str(), src.size() + 1);
    return (long)src.size();
}

extern "C" long emu_joint_program(char* out) {
    if (out) std::memcpy(out, g_joint_prog.bytes.data(), g_joint_prog.bytes.size());
    return (long)g_joint_prog.bytes.size();
}

// Pass count of the kept chunk (pass < 0), or pass `pass`'s tile qubits,
// slots, probes and first probe value.
extern "C" int emu_joint_pass_info(int pass, int* k, int* nslots, int* nprobes, int* probe_base) {
    const HostPlan& P = g_joint_chunks[g_joint_cur];
    if (pass < 0) return (int)P.passes.size();
    const FPass& fp = P.passes[pass];
    *k = fp.k;
    *nslots = fp.nslots;
    *nprobes = fp.nprobes;
    *probe_base = fp.probe_base;
    return 0;
}

// Gradient term of probe value v of the kept chunk: d params[*param] +=
// *coeff * value.
extern "C" int emu_joint_probe_ref(int v, int* param, double* coeff) {
    const HostPlan& P = g_joint_chunks[g_joint_cur];
    const ProbeRef& r = g_joint_refs[P.probe_ids[v]];
    *param = r.param;
    *coeff = r.coeff;
    return 0;
}

extern "C" int emu_joint_slot_table(int pass, double* table) {
    const HostPlan& P = g_joint_chunks[g_joint_cur];
    const FPass& fp = P.passes[pass];
    const uint64_t ntiles = 1ULL << (g_joint_n - fp.k);
    for (uint64_t ti = 0; ti < ntiles; ++ti) {
        uint64_t base = 0, t = ti;
        for (int e = 0; e < fp.next; ++e) {
            base |= (t & 1ULL) << fp.epos[e];
            t >>= 1;
        }
        for (int sl = 0; sl < fp.nslots; ++sl) {
            const FSlot& S = P.slots[fp.slot_begin + sl];
            double a = 0.0;
            for (uint32_t e = S.begin; e < S.end; ++e)
                if ((base & P.eterms[e].emask) == P.eterms[e].emask) a += P.eterms[e].ang;
            double* row = table + 4 * (ti * (uint64_t)fp.nslots + sl);
            row[0] = a;
            row[1] = std::cos(a);
            row[2] = std::sin(a);
            row[3] = 0.0;
        }
    }
    return 0;
}

// ---- batched small-circuit engine (qse_batch_plan.cpp) ----
// Executes a batch plan's step lists in FP64 on the host with the kernel's
// step semantics (qse_batch.hip) and returns energies and gradients through
// the library's own result mapping. Returns the plan's status.
extern "C" int emu_batch(int ncirc, const int* nq, const int64_t* op_off, const qse_op* ops, const int64_t* par_off,
                         const double* params, const uint64_t* obs_m, const double* obs_w, int nterms, int with_grads,
                         int kmax, int L, double* energies, double* grads) {
    BatchPlan B;
    const int rc = build_batch_plan(0, ncirc, nq, op_off, ops, par_off, params, obs_m, nterms, with_grads != 0, kmax,
                                    L, B);
    if (rc) return rc;
    std::vector<double> vals(std::max(B.nvals, 1), 0.0);
    for (int c = 0; c < ncirc; ++c) {
        const BCirc& C = B.circs[c];
        const uint64_t N = 1ULL << C.n;
        std::vector<cd> psi(N), lam(N);
        auto f = [&](uint64_t i) {
            double v = 0.0;
            for (int t = 0; t < nterms; ++t)
                if ((C.obs_keep >> t) & 1) v += (__builtin_popcountll(i & obs_m[t]) & 1) ? -obs_w[t] : obs_w[t];
            return v;
        };
        for (int si = C.step_begin; si < C.step_end; ++si) {
            const BStep& S = B.steps[si];
            switch (S.kind) {
                case BS_BASIS:
                    std::fill(psi.begin(), psi.end(), cd(0.0, 0.0));
                    psi[0] = 1.0;
                    break;
                case BS_PROG:
                    for (int p = S.pass_begin; p < S.pass_end; ++p) {
                        if (B.prog.passes[p].k != C.k) return 101;
                        if (S.target & 1) emu_pass(B.prog, B.prog.passes[p], C.n, psi.data(), psi.data());
                        if (S.target & 2) emu_pass(B.prog, B.prog.passes[p], C.n, lam.data(), lam.data());
                    }
                    break;
                case BS_OBS:
                    for (uint64_t i = 0; i < N; ++i) lam[i] = f(i) * psi[i];
                    break;
                case BS_ENERGY: {
                    double acc = 0.0;
                    for (uint64_t i = 0; i < N; ++i) acc += std::norm(psi[i]) * f(i);
                    vals[S.out] = acc;
                    break;
                }
                case BS_JOINT:
                    for (int p = S.pass_begin; p < S.pass_end; ++p) {
                        const FPass& fp = B.prog.passes[p];
                        if (fp.k > C.k || fp.k > BATCH_KMAX) return 103;
                        emu_joint_pass(B.prog, fp, C.n, psi.data(), lam.data(), vals.data() + S.out);
                    }
                    break;
                default: return 102;
            }
        }
    }
    batch_results(B, vals.data(), par_off, energies, with_grads ? grads : nullptr);
    return 0;
}

// Batch plan summary: out[0] single-state passes (a joint pass counts twice),
// out[1] steps, out[2] energy steps, out[3] program and joint steps, out[4]
// value slots; returns the plan's status.
extern "C" int emu_batch_stats(int ncirc, const int* nq, const int64_t* op_off, const qse_op* ops,
                               const int64_t* par_off, const double* params, const uint64_t* obs_m, int nterms,
                               int kmax, int L, int64_t* out) {
    BatchPlan B;
    const int rc = build_batch_plan(0, ncirc, nq, op_off, ops, par_off, params, obs_m, nterms, true, kmax, L, B);
    if (rc) return rc;
    int64_t red = 0, progs = 0, passes = 0;
    for (const BStep& s : B.steps) {
        if (s.kind == BS_ENERGY) ++red;
        if (s.kind == BS_PROG || s.kind == BS_JOINT) {
            ++progs;
            passes += (int64_t)(s.pass_end - s.pass_begin) * ((s.target == 3 || s.kind == BS_JOINT) ? 2 : 1);
        }
    }
    out[0] = passes;
    out[1] = (int64_t)B.steps.size();
    out[2] = red;
    out[3] = progs;
    out[4] = B.nvals;
    return 0;
}

// qse_jit_gen.cpp — HIP source of a fused-pass program with every field of
// the program as a literal (the code generator of the run-time specialised
// passes; compiled and launched by qse_jit.cpp).
//
// Each pass becomes one kernel qse_jit_pass<i>(src, dst, slot_table, ntiles)
// with the generic kernel's (qse_fused.hip) tile, stage, LDS-swizzle and
// slot-table layout. What the generic kernel decides per op at run time is
// decided here: register pairs whose register controls fail and thread-bit
// factors and terms are emitted only where they apply, coefficients are
// literals, and products are formed in the generic kernel's order, so the two
// agree to rounding (floating-point contraction and the deferred scale of
// H and rotation ops, emit_stages, may differ by a few units in the last
// place per op). This file uses no HIP runtime call: the CPU tests compile its
// output with a host shim.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <functional>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "qse_fused.hpp"
#include "qse_hip.h"

namespace qse {

namespace {

constexpr uint32_t swz_h(uint32_t li) { return li + (li >> 4); }

// Exact decimal form of a double (17 significant digits round-trip).
std::string lit(double v) {
    if (std::isnan(v)) return "__builtin_nan(\"\")";
    if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    std::string s(b);
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    return "(" + s + ")";
}

std::string u64(uint64_t v) { return std::to_string(v) + "ull"; }

// Sum of c*x over the given terms with literal coefficients; terms with a zero
// coefficient are dropped and +-1 needs no product (both exact), so only
// products that can round are emitted.
std::string lin(std::initializer_list<std::pair<double, const char*>> terms) {
    std::string e;
    for (const auto& t : terms) {
        if (t.first == 0.0) continue;
        std::string term;
        if (t.first == 1.0) term = t.second;
        else if (t.first == -1.0) term = std::string("-") + t.second;
        else term = "(R)" + lit(t.first) + " * " + t.second;
        if (e.empty()) e = term;
        else if (term[0] == '-') e += " - " + term.substr(1);
        else e += " + " + term;
    }
    return e.empty() ? "(R)0" : e;
}

std::string sum2(const std::string& a, const std::string& b) {
    if (a == "(R)0") return b;
    if (b == "(R)0") return a;
    return "(" + a + ") + (" + b + ")";
}

void emit_mul_g(std::ostringstream& o, const std::string& dr, const std::string& di) {
    o << "      { const double nr = gr * " << dr << " - gi * " << di << "; gi = gr * " << di << " + gi * " << dr
      << "; gr = nr; }\n";
}

// A positive scalar a common to both rows of an uncontrolled 2x2 op whose rows
// have coefficient magnitudes (a0, a1) and (b0, b1): a equals one magnitude of
// each row, so with the op divided by a every output is x + c*y, one fused
// multiply-add per real component. The largest such a is taken (quotients at
// most 1 in magnitude); false when the rows share no magnitude. A rotation
// [[c, -s], [s, c]] gives a = max(|c|, |s|).
bool common_row_scale(double a0, double a1, double b0, double b1, double& a) {
    a = 0.0;
    for (double v : {std::fabs(a0), std::fabs(a1)})
        if (v > a && std::isfinite(v) && (v == std::fabs(b0) || v == std::fabs(b1))) a = v;
    return a > 0.0;
}

// One pass as kernel `name`.
// With `toff` (the serialized program's array offsets), every coefficient is
// read from the program on the device (pv, its base) instead of being a
// literal, so one compiled module serves all plans of the same structure.
void emit_pass(std::ostringstream& o, const HostPlan& P, int pi, const std::string& name, bool prefetch,
               const size_t* toff, bool paired, bool direct) {
    const FPass& fp = P.passes[pi];
    // coefficient expressions (double): literal, or a load from the device program
    auto opv = [&](uint32_t oi, int j) -> std::string {
        if (!toff) return lit(P.ops[oi].m[j]);
        return "pv[" + std::to_string((toff[2] + oi * sizeof(FOp) + offsetof(FOp, m)) / 8 + j) + "]";
    };
    auto termv = [&](uint32_t ti, int c) -> std::string {
        if (!toff) return lit(P.terms[ti].d[c]);
        return "pv[" + std::to_string((toff[3] + ti * sizeof(FTerm) + offsetof(FTerm, d)) / 8 + c) + "]";
    };
    auto alphav = [&](uint32_t ai) -> std::string {
        if (!toff) return lit(P.alphas[ai]);
        return "pv[" + std::to_string(toff[6] / 8 + ai) + "]";
    };
    const int K = fp.k, TB = K - 4, NT = 1 << TB;
    const int nslots = fp.nslots;
    o << "extern \"C\" __global__ __launch_bounds__(" << NT << ") __attribute__((amdgpu_waves_per_eu(2, 8)))\n"
      << "void " << name << "(const cx* src, cx* dst, const SlotVal* __restrict__ slot_table, unsigned long long ntiles, "
      << "const double* __restrict__ pv) {\n"
      << "  __shared__ cx tile[" << (1 << K) + (1 << (K - 4)) << "];\n"
      << "  __shared__ SlotVal slotv[" << (nslots > 0 ? nslots : 1) << "];\n"
      << "  const unsigned tid = threadIdx.x;\n"
      << "  unsigned long long gt = 0, ot = 0; unsigned st_t = 0;\n";
    for (int b = 0; b < TB; ++b)
        o << "  if (tid & " << (1u << b) << "u) { gt |= " << u64(1ULL << fp.tq[b]) << "; ot |= "
          << u64(1ULL << fp.so_q[b]) << "; st_t |= " << (1u << fp.so_l[b]) << "u; }\n";
    o << "  const unsigned swz_tid = swz(tid), swz_st = swz(st_t);\n";
    // Direct load and store. The first stage can take its 16 amplitudes from
    // HBM into registers, and the last stage can write them to HBM, without the
    // LDS round trip of the load and store mapping, when the stage's thread
    // bits keep the accesses coalesced: the lowest lane bits must select the
    // tile's contiguous low qubits, in order, as the load (store) mapping's
    // lane bits do. This removes two of the tile's LDS round trips (all of
    // them in a one-stage pass).
    const FStage& S0 = P.stages[fp.stage_begin];
    const FStage& SL = P.stages[fp.stage_begin + fp.nstages - 1];
    uint32_t dq[16];  // local bit -> destination qubit
    for (int c = 0; c < K; ++c) dq[fp.so_l[c]] = fp.so_q[c];
    int lsrc = 0, ldst = 0;
    while (lsrc < K && fp.tq[lsrc] == (uint32_t)lsrc) ++lsrc;
    while (ldst < K && fp.so_q[ldst] == (uint32_t)ldst) ++ldst;
    const int lane_bits = TB < 6 ? TB : 6;
    bool dl = direct, ds = direct;
    for (int b = 0; b < (lsrc < lane_bits ? lsrc : lane_bits); ++b) dl = dl && S0.tmap[b] == (uint32_t)b;
    for (int b = 0; b < (ldst < lane_bits ? ldst : lane_bits); ++b) ds = ds && dq[SL.tmap[b]] == (uint32_t)b;
    // register j's offset from the thread's offset, for the load and the store mapping
    uint64_t ld_off[16], st_off[16];
    for (int j = 0; j < 16; ++j) {
        ld_off[j] = st_off[j] = 0;
        for (int c = 0; c < 4; ++c)
            if ((j >> c) & 1) {
                ld_off[j] |= 1ULL << (dl ? fp.tq[S0.rmap[c]] : fp.tq[TB + c]);
                st_off[j] |= 1ULL << dq[SL.rmap[c]];
            }
    }
    o << "  // direct load " << (dl ? "yes" : "no") << ", direct store " << (ds ? "yes" : "no") << "\n";
    if (dl) {
        o << "  unsigned long long gd = 0;\n";
        for (int b = 0; b < TB; ++b)
            o << "  if (tid & " << (1u << b) << "u) gd |= " << u64(1ULL << fp.tq[S0.tmap[b]]) << ";\n";
    }
    if (ds) {
        o << "  unsigned long long sd = 0;\n";
        for (int b = 0; b < TB; ++b)
            o << "  if (tid & " << (1u << b) << "u) sd |= " << u64(1ULL << dq[SL.tmap[b]]) << ";\n";
    }
    const char* gthr = dl ? "gd" : "gt";
    // With `prefetch`, the next tile's amplitudes are loaded into registers
    // while this tile's stages run (tiles belong to one workgroup each, so the
    // prefetch never reads what another workgroup writes). It is used for
    // complex64 only: for complex128 the second set of 16 amplitudes exceeds
    // the register budget of two waves per SIMD and spills, and two
    // workgroups per CU already overlap one's memory traffic with the other's
    // stages.
    auto emit_loads = [&](const char* tile, const char* b) {
        o << "      unsigned long long " << b << " = 0;\n";
        for (int e = 0; e < fp.next; ++e)
            o << "      " << b << " |= ((" << tile << " >> " << e << ") & 1ull) << " << fp.epos[e] << ";\n";
        for (int i = 0; i < 16; ++i)
            o << "      ld" << i << " = ldnt(src + (" << b << " | " << gthr << " | " << u64(ld_off[i]) << "));\n";
    };
    o << "  cx ld0, ld1, ld2, ld3, ld4, ld5, ld6, ld7, ld8, ld9, ld10, ld11, ld12, ld13, ld14, ld15;\n";
    // The tile's stages: amplitudes in `tile`; `base` (external qubits of the
    // tile), `tidv` and `slotv` in scope.
    // ldv: the first stage's amplitudes are registers ldv<j> (direct load);
    // obase: the last stage stores to HBM at obase (direct store); after_read0
    // runs once the first stage has read its amplitudes.
    auto emit_stages = [&](const char* ldv, const char* obase, const std::function<void()>& after_read0) {
        // Stage code works on scalar components (ar<v>, ai<v>) of 16 amplitude
        // variables. nm[j] names the variable holding register j: an X/CNOT whose
        // condition is known at generation time is a renaming, not code. (Whole
        // struct copies for register exchanges made the compiler keep the
        // amplitudes in scratch memory.)
        //
        // Deferred scale (literal coefficients only). An uncontrolled H, real or
        // RX-like op whose rows share a coefficient magnitude a is emitted divided
        // by a: H as x0 +- x1, a rotation as x0 - t*x1 (one fused multiply-add per
        // component instead of a product and a multiply-add). Every amplitude of
        // the tile passes through every such op, so the factors a form one real
        // scalar for the whole tile. It is applied to the thread's 16 amplitudes
        // before the pass stores them, and earlier whenever it leaves
        // [2^-40, 2^40], so the unscaled values stay far from the complex64 range
        // limits. Rounding differs from the undivided form by a few units in the
        // last place per op.
        double dscale = 1.0;
        auto emit_scale = [&](const int* nmv) {
            if (dscale == 1.0) return;
            const std::string f = "(R)" + lit(dscale);
            for (int j = 0; j < 16; ++j)
                o << "      ar" << nmv[j] << " *= " << f << "; ai" << nmv[j] << " *= " << f << ";\n";
            dscale = 1.0;
        };
        for (int sg = 0; sg < fp.nstages; ++sg) {
            const FStage& S = P.stages[fp.stage_begin + sg];
            o << "    {\n      unsigned toff = 0;\n";
            for (int t = 0; t < TB; ++t) o << "      if (tidv & " << (1u << t) << "u) toff |= " << (1u << S.tmap[t]) << "u;\n";
            o << "      const unsigned so = swz(toff);\n      R";
            for (int v = 0; v < 16; ++v) o << (v ? ", " : " ") << "ar" << v << ", ai" << v;
            o << ";\n";
            uint32_t rj[16];
            int nm[16];
            for (int j = 0; j < 16; ++j) {
                uint32_t r = 0;
                for (int q = 0; q < FUSE_REG; ++q)
                    if ((j >> q) & 1) r |= 1u << S.rmap[q];
                rj[j] = swz_h(r);
                nm[j] = j;
                if (sg == 0 && ldv)
                    o << "      ar" << j << " = " << ldv << j << ".re; ai" << j << " = " << ldv << j << ".im;\n";
                else
                    o << "      { const cx t = tile[so + " << rj[j] << "u]; ar" << j << " = t.re; ai" << j << " = t.im; }\n";
            }
            if (sg == 0 && after_read0) after_read0();
            for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
                const FOp& op = P.ops[oi];
                if (op.type == FOP_MONO) {
                    // constant x thread-bit factors (ascending bit) x terms, as in the generic kernel
                    o << "      {\n      double gr = " << opv(oi, 0) << ", gi = " << opv(oi, 1) << ";\n";
                    for (int t = 0; t < TB; ++t)
                        if ((op.tctl >> t) & 1) {
                            o << "      if (tidv & " << (1u << t) << "u)\n";
                            emit_mul_g(o, alphav(op.alpha_begin + 2 * t), alphav(op.alpha_begin + 2 * t + 1));
                        }
                    for (uint32_t ti = op.term_begin; ti < op.term_end; ++ti) {
                        const FTerm& T = P.terms[ti];
                        std::string cond = "((toff & " + std::to_string(T.tmask) + "u) == " + std::to_string(T.tmask) + "u)";
                        if (T.slot < 0 && T.emask) cond += " && ((base & " + u64(T.emask) + ") == " + u64(T.emask) + ")";
                        o << "      if (" << cond << ")\n";
                        if (T.slot >= 0) {
                            const std::string sv = "slotv[" + std::to_string(T.slot) + "]";
                            emit_mul_g(o, sv + ".c", sv + ".s");
                        } else {
                            emit_mul_g(o, termv(ti, 0), termv(ti, 1));
                        }
                    }
                    o << "      const R g_r = (R)gr, g_i = (R)gi;\n";
                    for (int j = 0; j < 16; ++j)
                        if (((uint32_t)j & op.rctl) == op.rctl) {
                            const std::string v = std::to_string(nm[j]);
                            o << "      { const R x = ar" << v << ", y = ai" << v << "; ar" << v << " = g_r * x - g_i * y; ai"
                              << v << " = g_r * y + g_i * x; }\n";
                        }
                    o << "      }\n";
                    continue;
                }
                const bool guard = op.tctl != 0 || op.ectl != 0;
                const std::string cond = "((toff & " + std::to_string(op.tctl) + "u) == " + std::to_string(op.tctl) +
                                         "u) && ((base & " + u64(op.ectl) + ") == " + u64(op.ectl) + ")";
                const int pp = (int)op.p;
                std::vector<std::pair<int, int>> prs;
                for (int jj = 0; jj < 8; ++jj) {
                    const int j0 = ((jj >> pp) << (pp + 1)) | (jj & ((1 << pp) - 1));
                    // REAL, RXLIKE and H are uncontrolled (rctl = 0); a failed register
                    // control of DENSE/XSWAP is the identity, i.e. nothing to emit.
                    if (((uint32_t)j0 & op.rctl) == op.rctl) prs.push_back({j0, j0 | (1 << pp)});
                }
                if (op.type == FOP_XSWAP) {
                    if (!guard) {
                        for (auto& pr : prs) std::swap(nm[pr.first], nm[pr.second]);
                        continue;
                    }
                    o << "      { const bool c = " << cond << ";\n";
                    for (auto& pr : prs) {
                        const std::string a = std::to_string(nm[pr.first]), b = std::to_string(nm[pr.second]);
                        o << "      { const R x0 = ar" << a << ", y0 = ai" << a << ", x1 = ar" << b << ", y1 = ai" << b << ";\n"
                          << "        ar" << a << " = c ? x1 : x0; ai" << a << " = c ? y1 : y0; ar" << b << " = c ? x0 : x1; ai"
                          << b << " = c ? y0 : y1; }\n";
                    }
                    o << "      }\n";
                    continue;
                }
                // the op divided by a common row scale (deferred scale above), when it has one
                double fa = 0.0;
                const bool fac = !guard && op.rctl == 0 &&
                                 ((op.type == FOP_H) ||
                                  (!toff && op.type == FOP_REAL && common_row_scale(op.m[0], op.m[2], op.m[4], op.m[6], fa)) ||
                                  (!toff && op.type == FOP_RXLIKE && common_row_scale(op.m[0], op.m[3], op.m[5], op.m[6], fa)));
                if (op.type == FOP_H) fa = 0.7071067811865475244;
                o << (guard ? "      if (" + cond + ") {\n" : std::string("      {\n"));
                if (toff && op.type != FOP_H) {
                    o << "      const R";
                    for (int j = 0; j < 8; ++j) o << (j ? ", " : " ") << "c" << j << " = (R)" << opv(oi, j);
                    o << ";\n";
                }
                for (auto& pr : prs) {
                    if (toff && op.type != FOP_H) {
                        // the same products and sums as below, with every coefficient from the table
                        const std::string a = std::to_string(nm[pr.first]), b = std::to_string(nm[pr.second]);
                        o << "      { const R x0 = ar" << a << ", y0 = ai" << a << ", x1 = ar" << b << ", y1 = ai" << b << ";\n";
                        if (op.type == FOP_REAL)
                            o << "        ar" << a << " = c0 * x0 + c2 * x1; ai" << a << " = c0 * y0 + c2 * y1;\n        ar"
                              << b << " = c4 * x0 + c6 * x1; ai" << b << " = c4 * y0 + c6 * y1; }\n";
                        else if (op.type == FOP_RXLIKE)
                            o << "        ar" << a << " = c0 * x0 - c3 * y1; ai" << a << " = c0 * y0 + c3 * x1;\n        ar"
                              << b << " = c6 * x1 - c5 * y0; ai" << b << " = c6 * y1 + c5 * x0; }\n";
                        else
                            o << "        ar" << a << " = (c0 * x0 - c1 * y0) + (c2 * x1 - c3 * y1); ai" << a
                              << " = (c0 * y0 + c1 * x0) + (c2 * y1 + c3 * x1);\n        ar" << b
                              << " = (c4 * x0 - c5 * y0) + (c6 * x1 - c7 * y1); ai" << b
                              << " = (c4 * y0 + c5 * x0) + (c6 * y1 + c7 * x1); }\n";
                        continue;
                    }
                    const std::string a = std::to_string(nm[pr.first]), b = std::to_string(nm[pr.second]);
                    o << "      { const R x0 = ar" << a << ", y0 = ai" << a << ", x1 = ar" << b << ", y1 = ai" << b << ";\n";
                    // coefficients of the emitted form: the op's, or the op's over fa
                    double m[8];
                    for (int j = 0; j < 8; ++j) m[j] = fac ? op.m[j] / fa : op.m[j];
                    switch (op.type) {
                        case FOP_H:
                            if (fac)
                                o << "        ar" << a << " = x0 + x1; ai" << a << " = y0 + y1; ar" << b << " = x0 - x1; ai"
                                  << b << " = y0 - y1; }\n";
                            else
                                o << "        ar" << a << " = H * (x0 + x1); ai" << a << " = H * (y0 + y1); ar" << b
                                  << " = H * (x0 - x1); ai" << b << " = H * (y0 - y1); }\n";
                            break;
                        case FOP_REAL:
                            o << "        ar" << a << " = " << lin({{m[0], "x0"}, {m[2], "x1"}}) << "; ai" << a << " = "
                              << lin({{m[0], "y0"}, {m[2], "y1"}}) << ";\n        ar" << b << " = "
                              << lin({{m[4], "x0"}, {m[6], "x1"}}) << "; ai" << b << " = "
                              << lin({{m[4], "y0"}, {m[6], "y1"}}) << "; }\n";
                            break;
                        case FOP_RXLIKE:
                            o << "        ar" << a << " = " << lin({{m[0], "x0"}, {-m[3], "y1"}}) << "; ai" << a << " = "
                              << lin({{m[0], "y0"}, {m[3], "x1"}}) << ";\n        ar" << b << " = "
                              << lin({{m[6], "x1"}, {-m[5], "y0"}}) << "; ai" << b << " = "
                              << lin({{m[6], "y1"}, {m[5], "x0"}}) << "; }\n";
                            break;
                        default:  // FOP_DENSE: (m00 v0) + (m01 v1), (m10 v0) + (m11 v1)
                            o << "        ar" << a << " = " << sum2(lin({{op.m[0], "x0"}, {-op.m[1], "y0"}}), lin({{op.m[2], "x1"}, {-op.m[3], "y1"}}))
                              << "; ai" << a << " = " << sum2(lin({{op.m[0], "y0"}, {op.m[1], "x0"}}), lin({{op.m[2], "y1"}, {op.m[3], "x1"}}))
                              << ";\n        ar" << b << " = " << sum2(lin({{op.m[4], "x0"}, {-op.m[5], "y0"}}), lin({{op.m[6], "x1"}, {-op.m[7], "y1"}}))
                              << "; ai" << b << " = " << sum2(lin({{op.m[4], "y0"}, {op.m[5], "x0"}}), lin({{op.m[6], "y1"}, {op.m[7], "x1"}}))
                              << "; }\n";
                            break;
                    }
                }
                o << "      }\n";
                if (fac) {
                    dscale *= fa;
                    if (dscale < 0x1p-40 || dscale > 0x1p40) emit_scale(nm);
                }
            }
            if (sg == fp.nstages - 1) emit_scale(nm);
            if (sg == fp.nstages - 1 && obase) {
                for (int j = 0; j < 16; ++j)
                    o << "      { cx t; t.re = ar" << nm[j] << "; t.im = ai" << nm[j] << "; stnt(dst + (" << obase
                      << " | sd | " << u64(st_off[j]) << "), t); }\n";
                o << "    }\n";
                continue;
            }
            for (int j = 0; j < 16; ++j)
                o << "      { cx t; t.re = ar" << nm[j] << "; t.im = ai" << nm[j] << "; tile[so + " << rj[j] << "u] = t; }\n";
            o << "      __syncthreads();\n    }\n";
        }
    };
    auto emit_tile_in = [&](const char* ld) {
        if (nslots > 0)
            o << "    for (int sl = tid; sl < " << nslots << "; sl += " << NT << ") slotv[sl] = slot_table[ti * "
              << nslots << "ull + sl];\n";
        if (!dl)
            for (int i = 0; i < 16; ++i)
                o << "    tile[swz_tid + " << swz_h((uint32_t)(i * NT)) << "u] = " << ld << i << ";\n";
        if (!dl || nslots > 0) o << "    __syncthreads();\n";
    };
    auto emit_store = [&](const char* obase) {
        if (ds) return;  // the last stage stored the tile
        for (int i = 0; i < 16; ++i) {
            uint64_t g = 0;
            uint32_t sl = 0;
            for (int c = 0; c < 4; ++c)
                if ((i >> c) & 1) {
                    g |= 1ULL << fp.so_q[TB + c];
                    sl |= 1u << fp.so_l[TB + c];
                }
            o << "    stnt(dst + (" << obase << " | ot | " << u64(g) << "), tile[swz_st + " << swz_h(sl) << "u]);\n";
        }
    };
    auto emit_bases = [&](const char* ti, const char* base, const char* obase) {
        o << "    unsigned long long " << base << " = 0, " << obase << " = 0;\n";
        for (int e = 0; e < fp.next; ++e)
            o << "    " << base << " |= ((" << ti << " >> " << e << ") & 1ull) << " << fp.epos[e] << "; " << obase
              << " |= ((" << ti << " >> " << e << ") & 1ull) << " << fp.eout[e] << ";\n";
    };
    // the thread id as a value the compiler cannot carry across iterations:
    // thread-dependent phase factors are recomputed per tile instead of being
    // hoisted out of the loop, where they would hold registers for its length
    const char* opaque_tid = "    unsigned tidv = tid;\n    QSE_OPAQUE(tidv);\n";
    // Paired tiles: a workgroup takes tiles 2p and 2p + 1, which differ in the
    // lowest external qubit, and issues the loads of both at once. When that
    // qubit sits right above the tile's contiguous low qubits, the two tiles'
    // runs are adjacent, and HBM serves the pair like runs of twice the length
    // (128-byte runs: 4.75 against 4.08 TB/s for a copy of that pattern). The
    // second tile waits in registers while the first one's stages run; each
    // tile is stored when its stages are done.
    if (paired && fp.next >= 1 && !prefetch) {
        o << "  cx lb0, lb1, lb2, lb3, lb4, lb5, lb6, lb7, lb8, lb9, lb10, lb11, lb12, lb13, lb14, lb15;\n";
        o << "  for (unsigned long long tp = blockIdx.x; tp < ntiles / 2; tp += gridDim.x) {\n";
        o << opaque_tid;
        o << "    const unsigned long long ta = 2 * tp;\n";
        emit_bases("ta", "base_a", "obase_a");
        o << "    const unsigned long long base_b = base_a | " << u64(1ULL << fp.epos[0]) << ", obase_b = obase_a | "
          << u64(1ULL << fp.eout[0]) << ";\n";
        for (int i = 0; i < 16; ++i)
            o << "    ld" << i << " = ldnt(src + (base_a | " << gthr << " | " << u64(ld_off[i]) << ")); lb" << i
              << " = ldnt(src + (base_b | " << gthr << " | " << u64(ld_off[i]) << "));\n";
        o << "    {\n    const unsigned long long ti = ta, base = base_a;\n";
        emit_tile_in("ld");
        emit_stages(dl ? "ld" : nullptr, ds ? "obase_a" : nullptr, nullptr);
        emit_store("obase_a");
        o << "    __syncthreads();\n    }\n";
        o << "    {\n    const unsigned long long ti = ta + 1, base = base_b;\n";
        emit_tile_in("lb");
        emit_stages(dl ? "lb" : nullptr, ds ? "obase_b" : nullptr, nullptr);
        emit_store("obase_b");
        o << "    __syncthreads();\n    }\n  }\n}\n\n";
        return;
    }
    if (prefetch) {
        o << "  if ((unsigned long long)blockIdx.x < ntiles) {\n    const unsigned long long t0 = blockIdx.x;\n";
        emit_loads("t0", "b0");
        o << "  }\n";
    }
    o << "  for (unsigned long long ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {\n";
    o << opaque_tid;
    emit_bases("ti", "base", "obase");
    if (!prefetch) {
        o << "    {\n";
        emit_loads("ti", "bt");
        o << "    }\n";
    }
    emit_tile_in("ld");
    // prefetch: the next tile's loads go out once this tile's amplitudes have
    // left the load registers (into LDS, or into the first stage's variables)
    auto emit_prefetch = [&]() {
        o << "    if (ti + gridDim.x < ntiles) {\n      const unsigned long long tn = ti + gridDim.x;\n";
        emit_loads("tn", "bn");
        o << "    }\n";
    };
    if (prefetch && !dl) emit_prefetch();
    emit_stages(dl ? "ld" : nullptr, ds ? "obase" : nullptr,
                prefetch && dl ? std::function<void()>(emit_prefetch) : std::function<void()>());
    emit_store("obase");
    o << "    __syncthreads();\n  }\n}\n\n";
}

// One joint pass of an adjoint reverse sweep as kernel `name`: the same ops on
// the tiles of psi and lambda (both in place; the program's coefficients read
// from pv at `toff`), and each FOP_PROBE op's Im<lambda|G|psi> summed over the
// tile at its point: per thread over its 16 register amplitudes (in the
// storage precision), over the wave by shuffles, into the wave's LDS
// accumulator; at the end each workgroup writes its sums (waves in order) to
// part[blockIdx.x * nprobes + probe]. The probes follow the batch engine's
// definitions (qse_batch.hip, probe_regs): X/Y on a register position, Z-string
// and projector masks over register, thread and external bits.
void emit_joint_pass(std::ostringstream& o, const HostPlan& P, int pi, const std::string& name, const size_t* toff) {
    const FPass& fp = P.passes[pi];
    auto opv = [&](uint32_t oi, int j) {
        return "pv[" + std::to_string((toff[2] + oi * sizeof(FOp) + offsetof(FOp, m)) / 8 + j) + "]";
    };
    auto termv = [&](uint32_t ti, int c) {
        return "pv[" + std::to_string((toff[3] + ti * sizeof(FTerm) + offsetof(FTerm, d)) / 8 + c) + "]";
    };
    auto alphav = [&](uint32_t ai) { return "pv[" + std::to_string(toff[6] / 8 + ai) + "]"; };
    const int K = fp.k, TB = K - 4, NT = 1 << TB;
    const int nslots = fp.nslots, np = fp.nprobes;
    const int NW = NT >= 64 ? NT / 64 : 1;
    const int LDS = (1 << K) + (1 << (K - 4));
    o << "extern \"C\" __global__ __launch_bounds__(" << NT << ") __attribute__((amdgpu_waves_per_eu(2, 8)))\n"
      << "void " << name << "(cx* __restrict__ psi, cx* __restrict__ lam, const SlotVal* __restrict__ slot_table, "
      << "unsigned long long ntiles, const double* __restrict__ pv, double* __restrict__ part) {\n"
      << "  __shared__ cx tile[" << LDS << "];\n"
      << "  __shared__ SlotVal slotv[" << (nslots > 0 ? nslots : 1) << "];\n"
      << "  __shared__ double pacc[" << (np > 0 ? np * NW : 1) << "];\n"
      << "  const unsigned tid = threadIdx.x;\n"
      << "  for (unsigned i = tid; i < " << np * NW << "u; i += " << NT << "u) pacc[i] = 0.0;\n"
      << "  unsigned long long gt = 0, ot = 0; unsigned st_t = 0;\n";
    for (int b = 0; b < TB; ++b)
        o << "  if (tid & " << (1u << b) << "u) { gt |= " << u64(1ULL << fp.tq[b]) << "; ot |= "
          << u64(1ULL << fp.so_q[b]) << "; st_t |= " << (1u << fp.so_l[b]) << "u; }\n";
    o << "  const unsigned swz_tid = swz(tid), swz_st = swz(st_t);\n";
    o << "  for (unsigned long long ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {\n"
      << "    unsigned tidv = tid;\n    QSE_OPAQUE(tidv);\n"
      << "    unsigned long long base = 0, obase = 0;\n";
    for (int e = 0; e < fp.next; ++e)
        o << "    base |= ((ti >> " << e << ") & 1ull) << " << fp.epos[e] << "; obase |= ((ti >> " << e << ") & 1ull) << "
          << fp.eout[e] << ";\n";
    // both tiles' loads at once; psi's amplitudes go through the LDS tile first,
    // lambda's wait in registers (one tile buffer: two workgroups per CU)
    o << "    cx ld0, ld1, ld2, ld3, ld4, ld5, ld6, ld7, ld8, ld9, ld10, ld11, ld12, ld13, ld14, ld15;\n"
      << "    cx lb0, lb1, lb2, lb3, lb4, lb5, lb6, lb7, lb8, lb9, lb10, lb11, lb12, lb13, lb14, lb15;\n";
    for (int i = 0; i < 16; ++i) {
        uint64_t off = 0;
        for (int c = 0; c < 4; ++c)
            if ((i >> c) & 1) off |= 1ULL << fp.tq[TB + c];
        o << "    { const unsigned long long g = base | gt | " << u64(off) << "; ld" << i << " = psi[g]; lb" << i
          << " = lam[g]; }\n";
    }
    for (int i = 0; i < 16; ++i) o << "    tile[swz_tid + " << swz_h((uint32_t)(i * NT)) << "u] = ld" << i << ";\n";
    if (nslots > 0)
        o << "    for (int sl = tid; sl < " << nslots << "; sl += " << NT << ") slotv[sl] = slot_table[ti * " << nslots
          << "ull + sl];\n";
    o << "    __syncthreads();\n";
    o << "    R";
    for (int v = 0; v < 16; ++v) o << (v ? ", " : " ") << "ar" << v << ", ai" << v << ", br" << v << ", bi" << v;
    o << ";\n";
    for (int sg = 0; sg < fp.nstages; ++sg) {
        const FStage& S = P.stages[fp.stage_begin + sg];
        // register j of this stage: LDS offset rj[j] from the thread's offset so
        auto stage_offsets = [&](const FStage& St, uint32_t* r16) {
            for (int j = 0; j < 16; ++j) {
                uint32_t r = 0;
                for (int q = 0; q < FUSE_REG; ++q)
                    if ((j >> q) & 1) r |= 1u << St.rmap[q];
                r16[j] = swz_h(r);
            }
        };
        auto stage_so = [&](const FStage& St, const char* var) {
            o << "      unsigned " << var << " = 0;\n";
            for (int t = 0; t < TB; ++t) o << "      if (tidv & " << (1u << t) << "u) " << var << " |= " << (1u << St.tmap[t]) << "u;\n";
        };
        o << "    {\n";
        stage_so(S, "toff");
        o << "      const unsigned so = swz(toff);\n";
        uint32_t rj[16];
        int nm[16];
        stage_offsets(S, rj);
        for (int j = 0; j < 16; ++j) nm[j] = j;
        if (sg == 0) {
            // psi from the tile; then lambda through the same tile
            for (int j = 0; j < 16; ++j)
                o << "      { const cx t = tile[so + " << rj[j] << "u]; ar" << j << " = t.re; ai" << j << " = t.im; }\n";
            o << "      __syncthreads();\n";
            for (int i = 0; i < 16; ++i) o << "      tile[swz_tid + " << swz_h((uint32_t)(i * NT)) << "u] = lb" << i << ";\n";
            o << "      __syncthreads();\n";
            for (int j = 0; j < 16; ++j)
                o << "      { const cx t = tile[so + " << rj[j] << "u]; br" << j << " = t.re; bi" << j << " = t.im; }\n";
        }
        for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
            const FOp& op = P.ops[oi];
            if (op.type == FOP_MONO) {
                o << "      {\n      double gr = " << opv(oi, 0) << ", gi = " << opv(oi, 1) << ";\n";
                for (int t = 0; t < TB; ++t)
                    if ((op.tctl >> t) & 1) {
                        o << "      if (tidv & " << (1u << t) << "u)\n";
                        emit_mul_g(o, alphav(op.alpha_begin + 2 * t), alphav(op.alpha_begin + 2 * t + 1));
                    }
                for (uint32_t ti = op.term_begin; ti < op.term_end; ++ti) {
                    const FTerm& T = P.terms[ti];
                    std::string cond = "((toff & " + std::to_string(T.tmask) + "u) == " + std::to_string(T.tmask) + "u)";
                    if (T.slot < 0 && T.emask) cond += " && ((base & " + u64(T.emask) + ") == " + u64(T.emask) + ")";
                    o << "      if (" << cond << ")\n";
                    if (T.slot >= 0) {
                        const std::string sv = "slotv[" + std::to_string(T.slot) + "]";
                        emit_mul_g(o, sv + ".c", sv + ".s");
                    } else {
                        emit_mul_g(o, termv(ti, 0), termv(ti, 1));
                    }
                }
                o << "      const R g_r = (R)gr, g_i = (R)gi;\n";
                for (int j = 0; j < 16; ++j)
                    if (((uint32_t)j & op.rctl) == op.rctl) {
                        const std::string v = std::to_string(nm[j]);
                        for (const char* s2 : {"a", "b"})
                            o << "      { const R x = " << s2 << "r" << v << ", y = " << s2 << "i" << v << "; " << s2 << "r"
                              << v << " = g_r * x - g_i * y; " << s2 << "i" << v << " = g_r * y + g_i * x; }\n";
                    }
                o << "      }\n";
                continue;
            }
            if (op.type == FOP_PROBE) {
                o << "      { R v = (R)0;\n";
                auto w = [&](int j) {
                    const std::string c = std::to_string(nm[j]);
                    return "(br" + c + " * ai" + c + " - bi" + c + " * ar" + c + ")";
                };
                if (op.mode == PROBE_X || op.mode == PROBE_Y) {
                    const int pp = (int)op.p;
                    for (int jj = 0; jj < 8; ++jj) {
                        const int j0 = ((jj >> pp) << (pp + 1)) | (jj & ((1 << pp) - 1)), j1 = j0 | (1 << pp);
                        const std::string a0 = std::to_string(nm[j0]), a1 = std::to_string(nm[j1]);
                        if (op.mode == PROBE_Y)
                            o << "        v += (br" << a1 << " * ar" << a0 << " + bi" << a1 << " * ai" << a0 << ") - (br" << a0
                              << " * ar" << a1 << " + bi" << a0 << " * ai" << a1 << ");\n";
                        else
                            o << "        v += (br" << a0 << " * ai" << a1 << " - bi" << a0 << " * ar" << a1 << ") + (br" << a1
                              << " * ai" << a0 << " - bi" << a1 << " * ar" << a0 << ");\n";
                    }
                } else if (op.mode == PROBE_PROJ) {
                    o << "        if (((toff & " << op.tctl << "u) == " << op.tctl << "u) && ((base & " << u64(op.ectl)
                      << ") == " << u64(op.ectl) << ")) {\n";
                    for (int j = 0; j < 16; ++j)
                        if (((uint32_t)j & op.rctl) == op.rctl) o << "          v += " << w(j) << ";\n";
                    o << "        }\n";
                } else {
                    for (int j = 0; j < 16; ++j)
                        o << "        v " << ((__builtin_popcount((uint32_t)j & op.rctl) & 1) ? "-=" : "+=") << " " << w(j)
                          << ";\n";
                    o << "        if ((__popc(toff & " << op.tctl << "u) + __popcll(base & " << u64(op.ectl)
                      << ")) & 1) v = -v;\n";
                }
                o << "        double s = (double)v;\n";
                if (NT >= 64) {
                    // DPP row scans and row broadcasts: the wave's sum in lane 63
                    o << "        s = qse_wsum64(s);\n"
                      << "        if ((tid & 63u) == 63u) pacc[" << op.term_begin * NW << "u + (tid >> 6)] += s;\n      }\n";
                } else {
                    for (int h = NT / 2; h > 0; h >>= 1) o << "        s += __shfl_xor(s, " << h << ", 64);\n";
                    o << "        if ((tid & 63u) == 0u) pacc[" << op.term_begin * NW << "u + (tid >> 6)] += s;\n      }\n";
                }
                continue;
            }
            const bool guard = op.tctl != 0 || op.ectl != 0;
            const std::string cond = "((toff & " + std::to_string(op.tctl) + "u) == " + std::to_string(op.tctl) +
                                     "u) && ((base & " + u64(op.ectl) + ") == " + u64(op.ectl) + ")";
            const int pp = (int)op.p;
            std::vector<std::pair<int, int>> prs;
            for (int jj = 0; jj < 8; ++jj) {
                const int j0 = ((jj >> pp) << (pp + 1)) | (jj & ((1 << pp) - 1));
                if (((uint32_t)j0 & op.rctl) == op.rctl) prs.push_back({j0, j0 | (1 << pp)});
            }
            if (op.type == FOP_XSWAP) {
                if (!guard) {
                    for (auto& pr : prs) std::swap(nm[pr.first], nm[pr.second]);
                    continue;
                }
                o << "      { const bool c = " << cond << ";\n";
                for (auto& pr : prs)
                    for (const char* s2 : {"a", "b"}) {
                        const std::string a = std::to_string(nm[pr.first]), b = std::to_string(nm[pr.second]);
                        const std::string xr = std::string(s2) + "r", xi = std::string(s2) + "i";
                        o << "      { const R x0 = " << xr << a << ", y0 = " << xi << a << ", x1 = " << xr << b << ", y1 = " << xi
                          << b << ";\n        " << xr << a << " = c ? x1 : x0; " << xi << a << " = c ? y1 : y0; " << xr << b
                          << " = c ? x0 : x1; " << xi << b << " = c ? y0 : y1; }\n";
                    }
                o << "      }\n";
                continue;
            }
            o << (guard ? "      if (" + cond + ") {\n" : std::string("      {\n"));
            if (op.type != FOP_H) {
                o << "      const R";
                for (int j = 0; j < 8; ++j) o << (j ? ", " : " ") << "c" << j << " = (R)" << opv(oi, j);
                o << ";\n";
            }
            for (auto& pr : prs)
                for (const char* s2 : {"a", "b"}) {
                    const std::string a = std::to_string(nm[pr.first]), b = std::to_string(nm[pr.second]);
                    const std::string xr = std::string(s2) + "r", xi = std::string(s2) + "i";
                    o << "      { const R x0 = " << xr << a << ", y0 = " << xi << a << ", x1 = " << xr << b << ", y1 = " << xi << b
                      << ";\n";
                    if (op.type == FOP_H)
                        o << "        " << xr << a << " = H * (x0 + x1); " << xi << a << " = H * (y0 + y1); " << xr << b
                          << " = H * (x0 - x1); " << xi << b << " = H * (y0 - y1); }\n";
                    else if (op.type == FOP_REAL)
                        o << "        " << xr << a << " = c0 * x0 + c2 * x1; " << xi << a << " = c0 * y0 + c2 * y1;\n        " << xr
                          << b << " = c4 * x0 + c6 * x1; " << xi << b << " = c4 * y0 + c6 * y1; }\n";
                    else if (op.type == FOP_RXLIKE)
                        o << "        " << xr << a << " = c0 * x0 - c3 * y1; " << xi << a << " = c0 * y0 + c3 * x1;\n        " << xr
                          << b << " = c6 * x1 - c5 * y0; " << xi << b << " = c6 * y1 + c5 * x0; }\n";
                    else
                        o << "        " << xr << a << " = (c0 * x0 - c1 * y0) + (c2 * x1 - c3 * y1); " << xi << a
                          << " = (c0 * y0 + c1 * x0) + (c2 * y1 + c3 * x1);\n        " << xr << b
                          << " = (c4 * x0 - c5 * y0) + (c6 * x1 - c7 * y1); " << xi << b
                          << " = (c4 * y0 + c5 * x0) + (c6 * y1 + c7 * x1); }\n";
                }
            o << "      }\n";
        }
        // each state through the tile in turn: into the next stage's registers,
        // or out to HBM after the last stage
        auto put = [&](const char* s2) {
            for (int j = 0; j < 16; ++j)
                o << "      { cx t; t.re = " << s2 << "r" << nm[j] << "; t.im = " << s2 << "i" << nm[j] << "; tile[so + " << rj[j]
                  << "u] = t; }\n";
        };
        if (sg + 1 < fp.nstages) {
            const FStage& N = P.stages[fp.stage_begin + sg + 1];
            uint32_t rn[16];
            stage_offsets(N, rn);
            stage_so(N, "tnx");
            o << "      const unsigned sn = swz(tnx);\n";
            for (const char* s2 : {"a", "b"}) {
                o << "      __syncthreads();\n";
                put(s2);
                o << "      __syncthreads();\n";
                for (int j = 0; j < 16; ++j)
                    o << "      { const cx t = tile[sn + " << rn[j] << "u]; " << s2 << "r" << j << " = t.re; " << s2 << "i" << j
                      << " = t.im; }\n";
            }
        } else {
            for (const char* s2 : {"a", "b"}) {
                o << "      __syncthreads();\n";
                put(s2);
                o << "      __syncthreads();\n";
                const char* st = s2[0] == 'a' ? "psi" : "lam";
                // store addresses from an opaque copy of the thread's base: equal
                // to the load addresses in place, which the compiler would
                // otherwise keep live (32 64-bit addresses) across the whole tile
                o << "      { unsigned long long ov = obase | ot;\n      QSE_OPAQUE(ov);\n";
                for (int i = 0; i < 16; ++i) {
                    uint64_t g = 0;
                    uint32_t sl = 0;
                    for (int c = 0; c < 4; ++c)
                        if ((i >> c) & 1) {
                            g |= 1ULL << fp.so_q[TB + c];
                            sl |= 1u << fp.so_l[TB + c];
                        }
                    o << "      " << st << "[ov | " << u64(g) << "] = tile[swz_st + " << swz_h(sl) << "u];\n";
                }
                o << "      }\n";
            }
            o << "      __syncthreads();\n";
        }
        o << "    }\n";
    }
    o << "  }\n";
    // this workgroup's probe sums, waves in order (the tile loop ended with a barrier)
    if (np > 0) {
        o << "  for (int p = tid; p < " << np << "; p += " << NT << ") {\n    double v = 0.0;\n";
        for (int w2 = 0; w2 < NW; ++w2) o << "    v += pacc[p * " << NW << " + " << w2 << "];\n";
        o << "    part[(unsigned long long)blockIdx.x * " << np << "ull + p] = v;\n  }\n";
    }
    o << "}\n\n";
}

}  // namespace


// Kernel name of pass `pass`: qse_jit_pass<i> (literal coefficients) or
// qse_jit_tpass<i> (coefficients from the program buffer), so profiles tell
// the two apart.
std::string jit_kernel_name(int pass, bool table) {
    return (table ? "qse_jit_tpass" : "qse_jit_pass") + std::to_string(pass);
}

// HIP source of every pass of plan P (kernel qse_jit_pass<i> for pass i).
// With table_off (the offsets of P's serialized program) the coefficients are
// read from the program buffer passed as the kernels' last argument. Complex128
// passes p with (*paired)[p] set process tiles in adjacent pairs.
std::string jit_source(const HostPlan& P, int prec, const size_t* table_off, const std::vector<bool>* paired) {
    std::ostringstream o;
    o << "typedef " << (prec == QSE_C128 ? "double" : "float") << " R;\n"
      << "struct __attribute__((aligned(2 * sizeof(R)))) cx { R re, im; };\n"
         "struct SlotVal { double ang, c, s, pad; };\n"
         "#define H ((R)0.7071067811865475244)\n"
         "__device__ __forceinline__ cx cmul(cx a, cx b) { cx r; r.re = a.re * b.re - a.im * b.im; "
         "r.im = a.re * b.im + a.im * b.re; return r; }\n"
         "__device__ __forceinline__ cx cadd(cx a, cx b) { cx r; r.re = a.re + b.re; r.im = a.im + b.im; return r; }\n"
         "__device__ __forceinline__ unsigned swz(unsigned li) { return li + (li >> 4); }\n"
         "#ifndef QSE_OPAQUE\n#define QSE_OPAQUE(v) __asm__ volatile(\"\" : \"+v\"(v))\n#endif\n";
    // complex128 state loads and stores are non-temporal (streamed past the caches): every amplitude is
    // read and written once per pass, and the hint made 30-qubit QFT passes 2-5% faster; for complex64
    // (whose passes prefetch the next tile) it made them 4% slower, so they stay plain. One vector access
    // per amplitude keeps each a single dwordx4.
    if (prec == QSE_C128)
        o << "typedef R R2 __attribute__((ext_vector_type(2)));\n"
             "__device__ __forceinline__ cx ldnt(const cx* p) { const R2 v = __builtin_nontemporal_load((const R2*)p); "
             "cx r; r.re = v[0]; r.im = v[1]; return r; }\n"
             "__device__ __forceinline__ void stnt(cx* p, cx a) { R2 v; v[0] = a.re; v[1] = a.im; "
             "__builtin_nontemporal_store(v, (R2*)p); }\n\n";
    else
        o << "__device__ __forceinline__ cx ldnt(const cx* p) { return *p; }\n"
             "__device__ __forceinline__ void stnt(cx* p, cx a) { *p = a; }\n\n";
    for (int p = 0; p < (int)P.passes.size(); ++p)
        emit_pass(o, P, p, jit_kernel_name(p, table_off != nullptr), prec == QSE_C64, table_off,
                  prec == QSE_C128 && paired && p < (int)paired->size() && (*paired)[p],
                  P.passes[p].nstages >= JIT_DIRECT_MIN_STAGES);
    return o.str();
}

// Kernel name of joint pass `pass` (adjoint reverse sweep over psi and lambda).
std::string jit_joint_kernel_name(int pass) { return "qse_jit_jpass" + std::to_string(pass); }

// HIP source of every pass of the joint reverse-sweep plan P (kernel
// qse_jit_jpass<i> for pass i; coefficients read from the program buffer at
// table_off).
std::string jit_joint_source(const HostPlan& P, int prec, const size_t* table_off) {
    std::ostringstream o;
    o << "typedef " << (prec == QSE_C128 ? "double" : "float") << " R;\n"
      << "struct __attribute__((aligned(2 * sizeof(R)))) cx { R re, im; };\n"
         "struct SlotVal { double ang, c, s, pad; };\n"
         "#define H ((R)0.7071067811865475244)\n"
         "__device__ __forceinline__ unsigned swz(unsigned li) { return li + (li >> 4); }\n"
         "#ifndef QSE_OPAQUE\n#define QSE_OPAQUE(v) __asm__ volatile(\"\" : \"+v\"(v))\n#endif\n"
         // Sum of v over the 64 lanes of a wave, in lane 63: a scan over each row
         // of 16 lanes (row_shr 1, 2, 4, 8), then the row sums chained with
         // row_bcast:15 and row_bcast:31 (DPP, no LDS round trip). The host shim
         // of the tests brings its own.
         "#ifndef QSE_HOST_SHIM\n"
         "template <int CTRL, int ROWS> __device__ __forceinline__ double qse_dpp(double v) {\n"
         "  union { double d; unsigned long long u; } x; x.d = v;\n"
         "  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x.u, CTRL, ROWS, 0xf, false);\n"
         "  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x.u >> 32), CTRL, ROWS, 0xf, false);\n"
         "  x.u = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo; return x.d;\n}\n"
         "__device__ __forceinline__ double qse_wsum64(double v) {\n"
         "  v += qse_dpp<0x111, 0xf>(v); v += qse_dpp<0x112, 0xf>(v); v += qse_dpp<0x114, 0xf>(v);\n"
         "  v += qse_dpp<0x118, 0xf>(v); v += qse_dpp<0x142, 0xa>(v); v += qse_dpp<0x143, 0xc>(v);\n"
         "  return v;\n}\n#endif\n\n";
    for (int p = 0; p < (int)P.passes.size(); ++p) emit_joint_pass(o, P, p, jit_joint_kernel_name(p), table_off);
    return o.str();
}

}  // namespace qse

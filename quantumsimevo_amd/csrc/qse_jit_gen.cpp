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
// agree to rounding (only floating-point contraction may differ). This file
// uses no HIP runtime call: the CPU tests compile its output with a host shim.
#include <cmath>
#include <cstdio>
#include <sstream>
#include <string>

#include "qse_fused.hpp"
#include "qse_hip.h"

namespace qse {

namespace {

constexpr uint32_t swz_h(uint32_t li) { return li ^ (((li >> 4) ^ (li >> 8) ^ (li >> 12)) & 15u); }

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

void emit_mul_g(std::ostringstream& o, const std::string& dr, const std::string& di) {
    o << "      { const double nr = gr * " << dr << " - gi * " << di << "; gi = gr * " << di << " + gi * " << dr
      << "; gr = nr; }\n";
}

// One pass as kernel `name`.
void emit_pass(std::ostringstream& o, const HostPlan& P, int pi, const std::string& name) {
    const FPass& fp = P.passes[pi];
    const int K = fp.k, TB = K - 4, NT = 1 << TB;
    const int nslots = fp.nslots;
    o << "extern \"C\" __global__ __launch_bounds__(" << NT << ") __attribute__((amdgpu_waves_per_eu(2, 8)))\n"
      << "void " << name << "(const cx* src, cx* dst, const SlotVal* __restrict__ slot_table, unsigned long long ntiles) {\n"
      << "  __shared__ cx tile[" << (1 << K) << "];\n"
      << "  __shared__ SlotVal slotv[" << (nslots > 0 ? nslots : 1) << "];\n"
      << "  const unsigned tid = threadIdx.x;\n"
      << "  unsigned long long gt = 0, ot = 0; unsigned st_t = 0;\n";
    for (int b = 0; b < TB; ++b)
        o << "  if (tid & " << (1u << b) << "u) { gt |= " << u64(1ULL << fp.tq[b]) << "; ot |= "
          << u64(1ULL << fp.so_q[b]) << "; st_t |= " << (1u << fp.so_l[b]) << "u; }\n";
    o << "  const unsigned swz_tid = swz(tid), swz_st = swz(st_t);\n";
    o << "  for (unsigned long long ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {\n";
    o << "    unsigned long long base = 0, obase = 0;\n";
    for (int e = 0; e < fp.next; ++e)
        o << "    base |= ((ti >> " << e << ") & 1ull) << " << fp.epos[e] << "; obase |= ((ti >> " << e
          << ") & 1ull) << " << fp.eout[e] << ";\n";
    o << "    cx ld[16];\n";
    for (int i = 0; i < 16; ++i) {
        uint64_t g = 0;
        for (int c = 0; c < 4; ++c)
            if ((i >> c) & 1) g |= 1ULL << fp.tq[TB + c];
        o << "    ld[" << i << "] = src[base | gt | " << u64(g) << "];\n";
    }
    if (nslots > 0)
        o << "    for (int sl = tid; sl < " << nslots << "; sl += " << NT << ") slotv[sl] = slot_table[ti * "
          << nslots << "ull + sl];\n";
    for (int i = 0; i < 16; ++i) o << "    tile[swz_tid ^ " << swz_h((uint32_t)(i * NT)) << "u] = ld[" << i << "];\n";
    o << "    __syncthreads();\n";
    for (int sg = 0; sg < fp.nstages; ++sg) {
        const FStage& S = P.stages[fp.stage_begin + sg];
        o << "    {\n      unsigned toff = 0;\n";
        for (int t = 0; t < TB; ++t) o << "      if (tid & " << (1u << t) << "u) toff |= " << (1u << S.tmap[t]) << "u;\n";
        o << "      const unsigned so = swz(toff);\n      cx a[16];\n";
        uint32_t rj[16];
        for (int j = 0; j < 16; ++j) {
            uint32_t r = 0;
            for (int p = 0; p < FUSE_REG; ++p)
                if ((j >> p) & 1) r |= 1u << S.rmap[p];
            rj[j] = swz_h(r);
            o << "      a[" << j << "] = tile[so ^ " << rj[j] << "u];\n";
        }
        for (uint32_t oi = S.op_begin; oi < S.op_end; ++oi) {
            const FOp& op = P.ops[oi];
            if (op.type == FOP_MONO) {
                // constant x thread-bit factors (ascending bit) x terms, as in the generic kernel
                o << "      {\n      double gr = " << lit(op.m[0]) << ", gi = " << lit(op.m[1]) << ";\n";
                for (int t = 0; t < TB; ++t)
                    if ((op.tctl >> t) & 1) {
                        o << "      if (tid & " << (1u << t) << "u)\n";
                        emit_mul_g(o, lit(P.alphas[op.alpha_begin + 2 * t]), lit(P.alphas[op.alpha_begin + 2 * t + 1]));
                    }
                for (uint32_t ti = op.term_begin; ti < op.term_end; ++ti) {
                    const FTerm& T = P.terms[ti];
                    std::string cond = "((toff & " + std::to_string(T.tmask) + "u) == " + std::to_string(T.tmask) + "u)";
                    if (T.slot < 0 && T.emask) cond += " && ((base & " + u64(T.emask) + ") == " + u64(T.emask) + ")";
                    o << "      if (" << cond << ")\n";
                    if (T.slot >= 0) {
                        const std::string s = "slotv[" + std::to_string(T.slot) + "]";
                        emit_mul_g(o, s + ".c", s + ".s");
                    } else {
                        emit_mul_g(o, lit(T.d[0]), lit(T.d[1]));
                    }
                }
                o << "      cx g; g.re = (R)gr; g.im = (R)gi;\n";
                for (int j = 0; j < 16; ++j)
                    if (((uint32_t)j & op.rctl) == op.rctl) o << "      a[" << j << "] = cmul(g, a[" << j << "]);\n";
                o << "      }\n";
                continue;
            }
            const bool guard = op.tctl != 0 || op.ectl != 0;
            if (guard)
                o << "      if (((toff & " << op.tctl << "u) == " << op.tctl << "u) && ((base & " << u64(op.ectl)
                  << ") == " << u64(op.ectl) << ")) {\n";
            else
                o << "      {\n";
            if (op.type == FOP_DENSE)
                o << "      const cx m00 = {(R)" << lit(op.m[0]) << ", (R)" << lit(op.m[1]) << "}, m01 = {(R)"
                  << lit(op.m[2]) << ", (R)" << lit(op.m[3]) << "};\n"
                  << "      const cx m10 = {(R)" << lit(op.m[4]) << ", (R)" << lit(op.m[5]) << "}, m11 = {(R)"
                  << lit(op.m[6]) << ", (R)" << lit(op.m[7]) << "};\n";
            const int p = (int)op.p;
            for (int jj = 0; jj < 8; ++jj) {
                const int j0 = ((jj >> p) << (p + 1)) | (jj & ((1 << p) - 1));
                const int j1 = j0 | (1 << p);
                // REAL, RXLIKE and H are uncontrolled (rctl = 0); a failed register
                // control of DENSE/XSWAP is the identity, i.e. nothing to emit.
                if (((uint32_t)j0 & op.rctl) != op.rctl) continue;
                const std::string A0 = "a[" + std::to_string(j0) + "]", A1 = "a[" + std::to_string(j1) + "]";
                o << "      { const cx v0 = " << A0 << ", v1 = " << A1 << ";\n";
                switch (op.type) {
                    case FOP_H:
                        o << "        " << A0 << ".re = H * (v0.re + v1.re); " << A0 << ".im = H * (v0.im + v1.im);\n"
                          << "        " << A1 << ".re = H * (v0.re - v1.re); " << A1 << ".im = H * (v0.im - v1.im); }\n";
                        break;
                    case FOP_XSWAP:
                        o << "        " << A0 << " = v1; " << A1 << " = v0; }\n";
                        break;
                    case FOP_REAL:
                        o << "        " << A0 << ".re = (R)" << lit(op.m[0]) << " * v0.re + (R)" << lit(op.m[2])
                          << " * v1.re; " << A0 << ".im = (R)" << lit(op.m[0]) << " * v0.im + (R)" << lit(op.m[2])
                          << " * v1.im;\n        " << A1 << ".re = (R)" << lit(op.m[4]) << " * v0.re + (R)"
                          << lit(op.m[6]) << " * v1.re; " << A1 << ".im = (R)" << lit(op.m[4]) << " * v0.im + (R)"
                          << lit(op.m[6]) << " * v1.im; }\n";
                        break;
                    case FOP_RXLIKE:
                        o << "        " << A0 << ".re = (R)" << lit(op.m[0]) << " * v0.re - (R)" << lit(op.m[3])
                          << " * v1.im; " << A0 << ".im = (R)" << lit(op.m[0]) << " * v0.im + (R)" << lit(op.m[3])
                          << " * v1.re;\n        " << A1 << ".re = (R)" << lit(op.m[6]) << " * v1.re - (R)"
                          << lit(op.m[5]) << " * v0.im; " << A1 << ".im = (R)" << lit(op.m[6]) << " * v1.im + (R)"
                          << lit(op.m[5]) << " * v0.re; }\n";
                        break;
                    default:  // FOP_DENSE
                        o << "        " << A0 << " = cadd(cmul(m00, v0), cmul(m01, v1)); " << A1
                          << " = cadd(cmul(m10, v0), cmul(m11, v1)); }\n";
                        break;
                }
            }
            o << "      }\n";
        }
        for (int j = 0; j < 16; ++j) o << "      tile[so ^ " << rj[j] << "u] = a[" << j << "];\n";
        o << "      __syncthreads();\n    }\n";
    }
    for (int i = 0; i < 16; ++i) {
        uint64_t g = 0;
        uint32_t s = 0;
        for (int c = 0; c < 4; ++c)
            if ((i >> c) & 1) {
                g |= 1ULL << fp.so_q[TB + c];
                s |= 1u << fp.so_l[TB + c];
            }
        o << "    dst[obase | ot | " << u64(g) << "] = tile[swz_st ^ " << swz_h(s) << "u];\n";
    }
    o << "    __syncthreads();\n  }\n}\n\n";
}

}  // namespace

std::string jit_kernel_name(int pass) { return "qse_jit_pass" + std::to_string(pass); }

// HIP source of every pass of plan P (kernel qse_jit_pass<i> for pass i).
std::string jit_source(const HostPlan& P, int prec) {
    std::ostringstream o;
    o << "typedef " << (prec == QSE_C128 ? "double" : "float") << " R;\n"
      << "struct __attribute__((aligned(2 * sizeof(R)))) cx { R re, im; };\n"
         "struct SlotVal { double ang, c, s, pad; };\n"
         "#define H ((R)0.7071067811865475244)\n"
         "__device__ __forceinline__ cx cmul(cx a, cx b) { cx r; r.re = a.re * b.re - a.im * b.im; "
         "r.im = a.re * b.im + a.im * b.re; return r; }\n"
         "__device__ __forceinline__ cx cadd(cx a, cx b) { cx r; r.re = a.re + b.re; r.im = a.im + b.im; return r; }\n"
         "__device__ __forceinline__ unsigned swz(unsigned li) { return li ^ (((li >> 4) ^ (li >> 8) ^ (li >> 12)) & "
         "15u); }\n\n";
    for (int p = 0; p < (int)P.passes.size(); ++p) emit_pass(o, P, p, jit_kernel_name(p));
    return o.str();
}

}  // namespace qse

// qse_fused_dev.hpp — device-side pieces of the LDS-tiled pass shared by the
// generic pass kernel (qse_fused.hip) and the batched small-circuit engine
// (qse_batch.hip): the LDS swizzle, the register-pair ops and the per-thread
// factor of a MONO op (program format: qse_fused.hpp).
#pragma once
#include "qse_device.hpp"
#include "qse_fused.hpp"

namespace qse {

// The swizzle is linear over GF(2): swz(a ^ b) = swz(a) ^ swz(b), so the
// address of (thread part | register/iteration part) is one XOR of
// precomputed halves.
__host__ __device__ constexpr uint32_t swz(uint32_t li) { return li ^ (((li >> 4) ^ (li >> 8) ^ (li >> 12)) & 15u); }

// Uncontrolled op with real entries (RY, real 2x2): 4 FP ops per amplitude.
template <class R, int P>
__device__ __forceinline__ void real_p(cx<R>* a, const FOp& op) {
    const R m00 = (R)op.m[0], m01 = (R)op.m[2], m10 = (R)op.m[4], m11 = (R)op.m[6];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        const cx<R> v0 = a[j0], v1 = a[j1];
        a[j0] = {m00 * v0.re + m01 * v1.re, m00 * v0.im + m01 * v1.im};
        a[j1] = {m10 * v0.re + m11 * v1.re, m10 * v0.im + m11 * v1.im};
    }
}

// Uncontrolled op [c0, i s0; i s1, c1] with real c, s (RX): 4 FP ops per amplitude.
template <class R, int P>
__device__ __forceinline__ void rxlike_p(cx<R>* a, const FOp& op) {
    const R c0 = (R)op.m[0], s0 = (R)op.m[3], s1 = (R)op.m[5], c1 = (R)op.m[6];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        const cx<R> v0 = a[j0], v1 = a[j1];
        a[j0] = {c0 * v0.re - s0 * v1.im, c0 * v0.im + s0 * v1.re};
        a[j1] = {c1 * v1.re - s1 * v0.im, c1 * v1.im + s1 * v0.re};
    }
}

// Uncontrolled X: exchange registers.
template <class R, int P>
__device__ __forceinline__ void xswap_p(cx<R>* a) {
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        const cx<R> v0 = a[j0];
        a[j0] = a[j1];
        a[j1] = v0;
    }
}

// X on register position P where the register controls hold: a wave-uniform
// branch per register pair (the pair's index is known at compile time, the
// control mask at run time), then a register exchange; no arithmetic.
template <class R, int P>
__device__ __forceinline__ void xswap_ctl_p(cx<R>* a, uint32_t rctl) {
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        if (((uint32_t)j0 & rctl) == rctl) {
            const cx<R> v0 = a[j0];
            a[j0] = a[j1];
            a[j1] = v0;
        }
    }
}

// One 2x2 op on register position P over the 8 register pairs (j0, j0|2^P).
// Pairs whose register controls fail get the identity (a wave-uniform select
// of coefficients, not a branch). H and X use the same formula: X = [0,1;1,0]
// is exact, H differs from s*(v0 +- v1) by at most 1 ulp.
template <class R, int P>
__device__ __forceinline__ void dense_p(cx<R>* a, const FOp& op) {
    const uint32_t rctl = op.rctl;
    const R m00r = (R)op.m[0], m00i = (R)op.m[1], m01r = (R)op.m[2], m01i = (R)op.m[3];
    const R m10r = (R)op.m[4], m10i = (R)op.m[5], m11r = (R)op.m[6], m11i = (R)op.m[7];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        const bool act = ((uint32_t)j0 & rctl) == rctl;  // wave-uniform
        const cx<R> c00 = {act ? m00r : (R)1, act ? m00i : (R)0}, c01 = {act ? m01r : (R)0, act ? m01i : (R)0};
        const cx<R> c10 = {act ? m10r : (R)0, act ? m10i : (R)0}, c11 = {act ? m11r : (R)1, act ? m11i : (R)0};
        const cx<R> v0 = a[j0], v1 = a[j1];
        a[j0] = cadd(cmul(c00, v0), cmul(c01, v1));
        a[j1] = cadd(cmul(c10, v0), cmul(c11, v1));
    }
}

// Uncontrolled Hadamard on register position P: s(v0+v1), s(v0-v1), the
// reference's rounding (CircuitUnitaryOperation.cpp:101-102).
template <class R, int P>
__device__ __forceinline__ void hadamard_p(cx<R>* a) {
    const R h = (R)0.7071067811865475244;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const int j1 = j0 | (1 << P);
        const cx<R> v0 = a[j0], v1 = a[j1];
        a[j0] = {h * (v0.re + v1.re), h * (v0.im + v1.im)};
        a[j1] = {h * (v0.re - v1.re), h * (v0.im - v1.im)};
    }
}

// Program arrays are kernel arguments of their own, const and __restrict__:
// the compiler then reads ops and terms (wave-uniform addresses) with scalar
// loads through the constant cache instead of vector loads.
struct KArgs {
    const FPass* __restrict__ passes;
    const FStage* __restrict__ stages;
    const FOp* __restrict__ ops;
    const FTerm* __restrict__ terms;
    const FSlot* __restrict__ slots;
    const FExtTerm* __restrict__ eterms;
    const double* __restrict__ alphas;
};

struct SlotVal {
    double ang, c, s, pad;
};

// Factor of a MONO op for this thread: constant x product over the thread-id
// bits set in both tid and op.tctl of their factor x product of the terms whose
// thread/external conditions hold (external ones from the per-tile slots).
// TermPtr / AlphaPtr: generic pointers, or pointers into the constant address
// space (uniform reads then become scalar loads); Slot: any type with the
// slot's factor in fields c, s.
template <int TB, class TermPtr, class AlphaPtr, class Slot>
__device__ __forceinline__ void mono_factor(const FOp& op, TermPtr __restrict__ terms, AlphaPtr __restrict__ alphas,
                                            int tid, uint32_t toff, uint64_t base, const Slot* slotv, double& gr,
                                            double& gi) {
    gr = op.m[0];
    gi = op.m[1];
    uint32_t am = op.tctl;
    if (am) {
        const auto al = alphas + op.alpha_begin;
        while (am) {
            const int t = __builtin_ctz(am);
            am &= am - 1;
            const bool on = (tid >> t) & 1;
            const double dr = on ? al[2 * t] : 1.0, di = on ? al[2 * t + 1] : 0.0;
            const double nr = gr * dr - gi * di;
            gi = gr * di + gi * dr;
            gr = nr;
        }
    }
    for (uint32_t t = op.term_begin; t < op.term_end; ++t) {
        const auto& T = terms[t];
        double dr, di;
        bool on = (toff & T.tmask) == T.tmask;
        if (T.slot >= 0) {
            dr = slotv[T.slot].c;
            di = slotv[T.slot].s;
        } else {
            on = on && ((base & T.emask) == T.emask);
            dr = T.d[0];
            di = T.d[1];
        }
        dr = on ? dr : 1.0;
        di = on ? di : 0.0;
        const double nr = gr * dr - gi * di;
        gi = gr * di + gi * dr;
        gr = nr;
    }
}

}  // namespace qse

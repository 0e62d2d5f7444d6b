// qse_device.hpp — device-side types and helpers shared by the gfx950 kernels.
//
// Amplitudes are interleaved (re, im) pairs: cx<double> (16 B, complex128) or
// cx<float> (8 B, complex64). Indices are 64-bit throughout: a 34-qubit state
// has 2^34 amplitudes, so every sweep is a grid-stride loop (a dispatch's total
// work-items per dimension must fit in 32 bits).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qse {

template <class R>
struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

template <class R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
template <class R>
__device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <class R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class R>
__device__ __forceinline__ cx<R> cneg(cx<R> a) {
    return {-a.re, -a.im};
}

// |a|^2 rounded as re*re, im*im, then their sum (no fused multiply-add), in
// FP64 whatever the storage precision: the same rounding as the CPU oracle's
// std::norm, so cumulative-probability comparisons agree bit for bit on equal
// amplitudes.
template <class R>
__device__ __forceinline__ double prob(cx<R> a) {
#pragma clang fp contract(off)
    const double re = (double)a.re, im = (double)a.im;
    const double r2 = re * re;
    const double i2 = im * im;
    return r2 + i2;
}

// Insert a zero bit at position b.
__device__ __forceinline__ uint64_t ins0(uint64_t i, int b) {
    const uint64_t lo = i & ((1ULL << b) - 1);
    return ((i >> b) << (b + 1)) | lo;
}

// Bit-position list used to expand a compressed index: zero bits are inserted
// at pos[0..nb) (ascending), then `ones` is OR-ed in.
struct Deposit {
    int nb;
    uint8_t pos[64];
    uint64_t ones;
};

__device__ __forceinline__ uint64_t deposit(uint64_t k, const Deposit& d) {
    uint64_t i = k;
    for (int t = 0; t < d.nb; ++t) i = ins0(i, d.pos[t]);
    return i | d.ones;
}

// Grid-stride helpers.
__device__ __forceinline__ uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// Seeded complex Gaussian amplitude of index i (before normalisation): two
// splitmix64 draws from (seed, i), Box-Muller in FP64. A pure function of
// (seed, i), so a kernel can regenerate any amplitude of a seeded random state
// without reading it.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ void random_amp(uint64_t seed, uint64_t i, double& re, double& im) {
    const double two_pi = 6.283185307179586476925;
    const uint64_t h1 = splitmix64(seed ^ ((2 * i + 1) * 0x9E3779B97F4A7C15ULL));
    const uint64_t h2 = splitmix64(h1);
    const double u1 = (double)((h1 >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(two_pi * u2, &sn, &cs);
    re = r * cs;
    im = r * sn;
}

// Wave64 sum of a double with cross-lane shuffles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace qse

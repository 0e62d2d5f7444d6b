// qse_dense.hip — a dense 2^m x 2^m complex block on the m lowest qubits of a
// complex128 state (m = 4, 5, 6), applied with the FP64 matrix cores, and the
// host composition of a gate list on m qubits into that block.
//
// The block is Y = U X where column c of X is the run of 2^m amplitudes
// [c 2^m, (c+1) 2^m): the m low qubits vary inside a column, the n-m others
// select it. One wavefront takes 16 columns at a time and computes the
// 16-row blocks of Y with v_mfma_f64_16x16x4f64 (D[16x16] += A[16x4] B[4x16];
// lane j + 16k holds A[j][k] and B[k][j], and D[i][j] is element i / 4 of
// lane j + 16 (i mod 4)):
//   * MFMA kc (< 4) of block (ib, kb) pairs kernel index k with row
//     kb 16 + 4 kc + k, so lane l = j + 16q (j < 16 column, q < 4) loads
//     rows kb 16 + 4 kc + q of its column as its B operands, and the four
//     lanes of a column read 64 contiguous bytes per load instruction; its A
//     operand is U[ib 16 + j][kb 16 + 4 kc + q] (uploaded in that lane order
//     and staged in LDS);
//   * element r of lane l is then row ib 16 + 4r + q of column j, one of the
//     rows the lane loaded: each lane stores back in place of its own loads.
// Complex products take four real ones: Re += Ur Xr - Ui Xi, Im += Ur Xi + Ui Xr.
// FP64 matrix-core and vector peaks are equal on gfx950 (78.6 TFLOP/s), so the
// block costs 2^m complex multiply-adds per amplitude on either unit; what the
// matrix cores remove is the operand traffic of a vector matrix product.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "qse_device.hpp"
#include "qse_internal.hpp"

namespace qse {

int validate_ops(int n, const qse_op* ops, int nops, const double* params);
double op_angle(const qse_op& op, const double* params);

namespace {

constexpr int DENSE_BLOCK = 256;  // 4 wavefronts, each on its own 16 columns

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Order this wavefront's LDS accesses before and after (LDS serves a
// wavefront's instructions in issue order; the fences keep the compiler from
// moving accesses across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MB: 16-row blocks per column (2^m = 16 MB). STAGE: the wavefront moves its
// 16 columns (16 x 16 MB contiguous amplitudes) between HBM and a private LDS
// area with 1 KiB-contiguous load and store instructions, and the MFMA
// operands go through LDS; otherwise each lane loads and stores its operand
// rows directly (64 contiguous bytes per column and instruction).
template <int MB, bool STAGE>
__global__ __launch_bounds__(DENSE_BLOCK) void k_dense_mfma(v2d* __restrict__ psi, const v2d* __restrict__ ulane,
                                                           uint64_t ngroups) {
    constexpr int NU = MB * MB * 4 * 64;  // (ib, kb, kc, lane) -> (Ur, Ui)
    constexpr int CR = 16 * MB;           // rows per column
    constexpr int P = CR + 1;             // LDS column stride (one amplitude of padding)
    constexpr int NA = 16 * CR;           // amplitudes per wavefront and group
    __shared__ v2d su[NU];
    __shared__ v2d stage[STAGE ? DENSE_BLOCK / 64 : 1][STAGE ? 16 * P : 1];
    for (int i = threadIdx.x; i < NU; i += DENSE_BLOCK) su[i] = ulane[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, q = lane >> 4;
    v2d* sw = stage[STAGE ? threadIdx.x >> 6 : 0];
    const uint64_t nwaves = (uint64_t)gridDim.x * (DENSE_BLOCK / 64);
    for (uint64_t g = (uint64_t)blockIdx.x * (DENSE_BLOCK / 64) + (threadIdx.x >> 6); g < ngroups; g += nwaves) {
        v2d* grp = psi + g * NA;              // the wavefront's 16 columns
        v2d* col = grp + (uint64_t)j * CR + q;  // its operand rows (direct)
        v2d x[MB][4];
        if constexpr (STAGE) {
            v2d t[NA / 64];
#pragma unroll
            for (int i = 0; i < NA / 64; ++i) t[i] = grp[lane + 64 * i];
#pragma unroll
            for (int i = 0; i < NA / 64; ++i) {
                const int a = lane + 64 * i;
                sw[(a / CR) * P + a % CR] = t[i];
            }
            wave_sync();
#pragma unroll
            for (int kb = 0; kb < MB; ++kb)
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) x[kb][kc] = sw[j * P + kb * 16 + 4 * kc + q];
        } else {
#pragma unroll
            for (int kb = 0; kb < MB; ++kb)
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) x[kb][kc] = col[kb * 16 + 4 * kc];
        }
        // one output block at a time keeps the accumulators and the block's
        // U operands in registers without raising the register count with MB
#pragma unroll 1
        for (int ib = 0; ib < MB; ++ib) {
            v4d re = {0.0, 0.0, 0.0, 0.0}, im = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kb = 0; kb < MB; ++kb)
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const v2d a = su[((ib * MB + kb) * 4 + kc) * 64 + lane];
                    re = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], x[kb][kc][0], re, 0, 0, 0);
                    re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[1], x[kb][kc][1], re, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], x[kb][kc][1], im, 0, 0, 0);
                    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], x[kb][kc][0], im, 0, 0, 0);
                }
            // rows ib 16 + 4r + q are the ones this lane read, and every MFMA
            // of block 0 consumed all of the lane's reads before the first write
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const v2d y = {re[r], im[r]};
                if constexpr (STAGE) sw[j * P + ib * 16 + 4 * r + q] = y;
                else col[ib * 16 + 4 * r] = y;
            }
        }
        if constexpr (STAGE) {
            wave_sync();
            v2d t[NA / 64];
#pragma unroll
            for (int i = 0; i < NA / 64; ++i) {
                const int a = lane + 64 * i;
                t[i] = sw[(a / CR) * P + a % CR];
            }
#pragma unroll
            for (int i = 0; i < NA / 64; ++i) grp[lane + 64 * i] = t[i];
            wave_sync();  // the next group's LDS writes come after these reads
        }
    }
}

using cd = std::complex<double>;

// Apply one gate to a 2^m-amplitude host vector (the semantics of the
// reference's scalar backend, CircuitUnitaryOperation.cpp:10-338).
void host_gate(const qse_op& o, const double* params, std::vector<cd>& v) {
    const size_t N = v.size();
    const double th = op_angle(o, params);
    const size_t tb = (size_t)1 << o.target;
    auto pairs = [&](auto&& f) {
        for (size_t i = 0; i < N; ++i)
            if (!(i & tb)) f(v[i], v[i | tb], i);
    };
    switch (o.kind) {
        case QSE_GATE_IDENTITY: break;
        case QSE_GATE_H: {
            const double s = 0.7071067811865475244;
            pairs([&](cd& a, cd& b, size_t) {
                const cd x = a, y = b;
                a = s * (x + y);
                b = s * (x - y);
            });
            break;
        }
        case QSE_GATE_X: pairs([&](cd& a, cd& b, size_t) { std::swap(a, b); }); break;
        case QSE_GATE_Y:
            pairs([&](cd& a, cd& b, size_t) {
                const cd x = a, y = b;
                a = cd(0, -1) * y;
                b = cd(0, 1) * x;
            });
            break;
        case QSE_GATE_Z: pairs([&](cd&, cd& b, size_t) { b = -b; }); break;
        case QSE_GATE_PHASE: {
            const cd e(std::cos(th), std::sin(th));
            pairs([&](cd&, cd& b, size_t) { b *= e; });
            break;
        }
        case QSE_GATE_RX: case QSE_GATE_RY: case QSE_GATE_MATRIX: {
            cd m[4];
            const double c = std::cos(th / 2), s = std::sin(th / 2);
            if (o.kind == QSE_GATE_RX) m[0] = c, m[1] = cd(0, -s), m[2] = cd(0, -s), m[3] = c;
            else if (o.kind == QSE_GATE_RY) m[0] = c, m[1] = -s, m[2] = s, m[3] = c;
            else
                for (int e = 0; e < 4; ++e) m[e] = cd(o.m[2 * e], o.m[2 * e + 1]);
            pairs([&](cd& a, cd& b, size_t) {
                const cd x = a, y = b;
                a = m[0] * x + m[1] * y;
                b = m[2] * x + m[3] * y;
            });
            break;
        }
        case QSE_GATE_RZ: {
            const cd e0(std::cos(th / 2), -std::sin(th / 2)), e1(std::cos(th / 2), std::sin(th / 2));
            pairs([&](cd& a, cd& b, size_t) {
                a *= e0;
                b *= e1;
            });
            break;
        }
        case QSE_GATE_CNOT:
            if (o.control == o.target) break;
            pairs([&](cd& a, cd& b, size_t i) {
                if ((i >> o.control) & 1) std::swap(a, b);
            });
            break;
        case QSE_GATE_CPHASE: case QSE_GATE_MCPHASE: {
            const size_t mask = tb | (o.kind == QSE_GATE_CPHASE ? (size_t)1 << o.control : (size_t)o.controls);
            const cd e(std::cos(th), std::sin(th));
            for (size_t i = 0; i < N; ++i)
                if ((i & mask) == mask) v[i] *= e;
            break;
        }
        case QSE_GATE_SWAP: {
            if (o.control == o.target) break;
            const size_t b2 = (size_t)1 << o.control;
            for (size_t i = 0; i < N; ++i)
                if ((i & tb) && !(i & b2)) std::swap(v[i], v[(i ^ tb) | b2]);
            break;
        }
        default: break;  // validated by the caller
    }
}

}  // namespace

}  // namespace qse

using namespace qse;

extern "C" int qse_compose_dense(int m, const qse_op* ops, int nops, const double* params, double* u) {
    if (m < 1 || m > 10 || !u || nops < 0 || (nops > 0 && !ops))
        return set_error(QSE_ERR_INVALID_ARG, "compose: m must be in [1, 10] and pointers non-null");
    const int rc = validate_ops(m, ops, nops, params);
    if (rc) return rc;
    const size_t D = (size_t)1 << m;
    std::vector<cd> v(D);
    for (size_t c = 0; c < D; ++c) {
        std::fill(v.begin(), v.end(), cd(0, 0));
        v[c] = 1.0;
        for (int i = 0; i < nops; ++i) host_gate(ops[i], params, v);
        for (size_t r = 0; r < D; ++r) {
            u[2 * (r * D + c)] = v[r].real();
            u[2 * (r * D + c) + 1] = v[r].imag();
        }
    }
    return QSE_OK;
}

extern "C" int qse_apply_dense(qse_ctx* ctx, qse_state* st, int m, const double* u) {
    if (!ctx || !st || !u) return set_error(QSE_ERR_INVALID_ARG, "null argument");
    if (m < 4 || m > 6) return set_error(QSE_ERR_UNSUPPORTED, "dense block: m must be 4, 5 or 6");
    if (st->prec != QSE_C128)
        return set_error(QSE_ERR_UNSUPPORTED, "dense block: complex128 states only (FP64 matrix cores)");
    if (st->n < m + 4) return set_error(QSE_ERR_UNSUPPORTED, "dense block: the state needs at least m + 4 qubits");
    QSE_HIP(hipSetDevice(ctx->device));
    const int D = 1 << m, MB = D / 16;
    // lane order: entry ((ib MB + kb) 4 + kc) 64 + l = U[ib 16 + (l & 15)][kb 16 + 4 kc + (l >> 4)]
    std::vector<char> bytes(sizeof(double) * 2 * (size_t)MB * MB * 4 * 64);
    double* lu = (double*)bytes.data();
    for (int ib = 0; ib < MB; ++ib)
        for (int kb = 0; kb < MB; ++kb)
            for (int kc = 0; kc < 4; ++kc)
                for (int l = 0; l < 64; ++l) {
                    const size_t row = (size_t)ib * 16 + (l & 15);
                    const size_t colk = (size_t)kb * 16 + 4 * kc + (l >> 4);
                    const size_t e = (((size_t)ib * MB + kb) * 4 + kc) * 64 + l;
                    lu[2 * e] = u[2 * (row * D + colk)];
                    lu[2 * e + 1] = u[2 * (row * D + colk) + 1];
                }
    size_t off = 0;
    int rc = ring_upload(ctx, bytes, &off);
    if (rc) return rc;
    const v2d* dev_u = (const v2d*)(ctx->ring.dev + off);
    const uint64_t ngroups = ((uint64_t)1 << (st->n - m)) / 16;
    const unsigned grid = (unsigned)std::min<uint64_t>((ngroups + 3) / 4, (uint64_t)ctx->num_cus * 8);
    TimedLaunch t;
    launch_begin(ctx, FAM_DENSE, &t);
    v2d* psi = (v2d*)st->d;
    if (MB == 1) k_dense_mfma<1, true><<<grid, DENSE_BLOCK, 0, ctx->stream>>>(psi, dev_u, ngroups);
    else if (MB == 2) k_dense_mfma<2, true><<<grid, DENSE_BLOCK, 0, ctx->stream>>>(psi, dev_u, ngroups);
    else k_dense_mfma<4, false><<<grid, DENSE_BLOCK, 0, ctx->stream>>>(psi, dev_u, ngroups);
    QSE_HIP(hipGetLastError());
    launch_end(ctx, &t, 2.0 * (double)st->bytes);
    return QSE_OK;
}

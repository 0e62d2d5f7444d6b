/*
 * qse_oracle.cpp — CPU oracle for the state-vector gate-sweep path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle, not the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it
 * (liboracle.so via ctypes). The product path (libqse_hip.so and the C++/Python
 * layers above it) never links, loads or calls it.
 *
 * It restates, in scalar FP64 C++ with std::complex<double>, the reference
 * simulator's CPU semantics (reference files cited per function):
 *   gates     QuantumSimEvo/CircuitUnitaryOperation.cpp:10-338 (scalar backend)
 *   sampling  QuantumSimEvo/StateVector.cpp:77-116   (sequential cumulative sum)
 *   measure   QuantumSimEvo/StateVector.cpp:118-156
 *   mid-circuit collapse/reset  StateVector.cpp:158-221
 *   noise     QuantumSimEvo/Noise.cpp:3-44, draw order Circuit.cpp:64-120
 *   dispatch  QuantumSimEvo/Circuit.cpp:73-112
 *   adjoint gradients (no reference counterpart): restated from the mathematics
 * The reference draws its readout randomness from std::random_device
 * (StateVector.cpp:98-99,127-128); here the canonical uniforms are injected so
 * outcomes are reproducible: target = u * total, exactly what
 * std::uniform_real_distribution<double>(0, total) computes from
 * generate_canonical.
 *
 * Norms |a|^2 are computed as re*re + im*im with separate roundings (the build
 * uses -ffp-contract=off) so a device kernel that rounds the same way sees the
 * same probabilities bit for bit.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   every gate  : dense Kronecker-product matrices (numpy) at n <= 8, the
 *                 reference's known-answer states QSETests.cpp:76-170, golden
 *                 vectors written by the compiled reference (oracle/_ref);
 *   QFT         : unitary DFT closed form (bit-reversed), n <= 10;
 *   sample      : brute-force lower_bound over numpy cumsum (sequential order),
 *                 basis-state and GHZ invariants;
 *   measure     : first-index rule incl. the fall-back-to-0 case, basis states;
 *   noise       : std::mt19937 standard known answer (10000th output of the
 *                 default-seeded engine = 4123659995) and flip patterns written
 *                 by the reference's own Noise class (oracle/_ref);
 *   expectation : closed forms on basis/product states, numpy brute force;
 *   adjoint     : the oracle's own parameter-shift gradients (built on the
 *                 pinned gate sweeps and expectation) for every parameterised
 *                 kind, and central finite differences.
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#include "../include/qse_hip.h"

using cplx = std::complex<double>;

namespace {

inline cplx* amps(double* psi) { return reinterpret_cast<cplx*>(psi); }
inline const cplx* amps(const double* psi) { return reinterpret_cast<const cplx*>(psi); }

inline double norm2(const cplx& a) {
    const double re = a.real(), im = a.imag();
    const double r2 = re * re;
    const double i2 = im * im;
    return r2 + i2;
}

}  // namespace

extern "C" {

/* CircuitUnitaryOperation.cpp:10-36 — generic 2x2 on pairs (i0, i0|2^q). */
void orc_apply_matrix(double* psi, int n, int q, const double* m) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, sec = 1ULL << q;
    const cplx g00(m[0], m[1]), g01(m[2], m[3]), g10(m[4], m[5]), g11(m[6], m[7]);
    for (uint64_t i = 0; i < (N >> 1); ++i) {
        const uint64_t i0 = ((i >> q) << (q + 1)) | (i & (sec - 1));
        const uint64_t i1 = i0 | sec;
        const cplx lo = s[i0], hi = s[i1];
        s[i0] = g00 * lo + g01 * hi;
        s[i1] = g10 * lo + g11 * hi;
    }
}

/* CircuitUnitaryOperation.cpp:38-59 — swap (i, i^t) when the control bit is set. */
void orc_cnot(double* psi, int n, int control, int target) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, cm = 1ULL << control, tm = 1ULL << target;
    for (uint64_t i = 0; i < N; ++i)
        if ((i & cm) && (i < (i ^ tm))) std::swap(s[i], s[i ^ tm]);
}

/* CircuitUnitaryOperation.cpp:61-79 */
void orc_pauli_x(double* psi, int n, int q) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    for (uint64_t i = 0; i < N; ++i)
        if (!(i & m)) std::swap(s[i], s[i | m]);
}

/* CircuitUnitaryOperation.cpp:81-110 — s*(v0+v1), s*(v0-v1), s = 0.7071067811865475244. */
void orc_hadamard(double* psi, int n, int q) {
    cplx* s = amps(psi);
    const double h = 0.7071067811865475244;
    const uint64_t N = 1ULL << n, step = 1ULL << q;
    for (uint64_t i = 0; i < N; i += 2 * step)
        for (uint64_t j = i; j < i + step; ++j) {
            const cplx v0 = s[j], v1 = s[j + step];
            s[j] = h * (v0 + v1);
            s[j + step] = h * (v0 - v1);
        }
}

/* CircuitUnitaryOperation.cpp:112-132 — multiply by exp(i theta) where bit q = 1. */
void orc_phase(double* psi, int n, int q, double theta) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    const cplx ph = std::exp(cplx(0, theta));
    for (uint64_t i = 0; i < N; ++i)
        if (i & m) s[i] *= ph;
}

/* CircuitUnitaryOperation.cpp:134-160 — (v0, v1) -> (-i v1, i v0). */
void orc_pauli_y(double* psi, int n, int q) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    const cplx I(0.0, 1.0);
    for (uint64_t i = 0; i < N; ++i)
        if (!(i & m)) {
            const cplx v0 = s[i], v1 = s[i | m];
            s[i] = -I * v1;
            s[i | m] = I * v0;
        }
}

/* CircuitUnitaryOperation.cpp:162-179 — negate where bit q = 1. */
void orc_pauli_z(double* psi, int n, int q) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    for (uint64_t i = 0; i < N; ++i)
        if (i & m) s[i] = -s[i];
}

/* CircuitUnitaryOperation.cpp:181-210 — [c, -is; -is, c]. */
void orc_rx(double* psi, int n, int q, double theta) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    const double c = std::cos(theta / 2.0), sn = std::sin(theta / 2.0);
    const cplx miS(0, -sn);
    for (uint64_t i = 0; i < N; ++i)
        if (!(i & m)) {
            const cplx v0 = s[i], v1 = s[i | m];
            s[i] = c * v0 + miS * v1;
            s[i | m] = miS * v0 + c * v1;
        }
}

/* CircuitUnitaryOperation.cpp:212-240 — [c, -s; s, c]. */
void orc_ry(double* psi, int n, int q, double theta) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    const double c = std::cos(theta / 2.0), sn = std::sin(theta / 2.0);
    for (uint64_t i = 0; i < N; ++i)
        if (!(i & m)) {
            const cplx v0 = s[i], v1 = s[i | m];
            s[i] = c * v0 - sn * v1;
            s[i | m] = sn * v0 + c * v1;
        }
}

/* CircuitUnitaryOperation.cpp:242-265 — exp(-i theta/2) on bit 0, exp(+i theta/2) on bit 1. */
void orc_rz(double* psi, int n, int q, double theta) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m = 1ULL << q;
    const cplx p0 = std::exp(cplx(0, -theta / 2.0)), p1 = std::exp(cplx(0, theta / 2.0));
    for (uint64_t i = 0; i < N; ++i) s[i] *= (i & m) ? p1 : p0;
}

/* CircuitUnitaryOperation.cpp:267-289 — swap (idx|2^a, idx|2^b) over the quarter. */
void orc_swap(double* psi, int n, int q0, int q1) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, m0 = 1ULL << q0, m1 = 1ULL << q1;
    const int lo = std::min(q0, q1), hi = std::max(q0, q1);
    const uint64_t lmask = (1ULL << lo) - 1;
    for (uint64_t k = 0; k < N / 4; ++k) {
        const uint64_t i = ((k >> lo) << (lo + 1)) | (k & lmask);
        const uint64_t idx = ((i >> hi) << (hi + 1)) | (i & ((1ULL << hi) - 1));
        std::swap(s[idx | m0], s[idx | m1]);
    }
}

/* CircuitUnitaryOperation.cpp:291-312 — exp(i theta) where all of controls+target are 1. */
void orc_mcphase(double* psi, int n, uint64_t controls, int target, double theta) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, all = (1ULL << target) | controls;
    const cplx ph = std::exp(cplx(0, theta));
    for (uint64_t i = 0; i < N; ++i)
        if ((i & all) == all) s[i] *= ph;
}

/* CircuitUnitaryOperation.cpp:314-338 — (re c - im s, re s + im c) where both bits are 1. */
void orc_cphase(double* psi, int n, int control, int target, double theta) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, cm = 1ULL << control, tm = 1ULL << target;
    const double c = std::cos(theta), sn = std::sin(theta);
    for (uint64_t i = 0; i < N; ++i)
        if ((i & cm) && (i & tm)) {
            const double re = s[i].real(), im = s[i].imag();
            s[i] = cplx(re * c - im * sn, re * sn + im * c);
        }
}

/* Angle of an op: theta, or theta + scale * params[param] for differentiable ops. */
static double op_angle(const qse_op& op, const double* params) {
    if (op.param >= 0 && params) return op.theta + op.scale * params[op.param];
    return op.theta;
}

/* Circuit.cpp:73-112 — replay a gate list through the scalar backend.
 * Returns 0, or -1 for an unsupported kind (the reference prints an error). */
int orc_run_ops(double* psi, int n, const qse_op* ops, int nops, const double* params) {
    for (int k = 0; k < nops; ++k) {
        const qse_op& op = ops[k];
        const double th = op_angle(op, params);
        switch (op.kind) {
            case QSE_GATE_IDENTITY: break;
            case QSE_GATE_H: orc_hadamard(psi, n, op.target); break;
            case QSE_GATE_X: orc_pauli_x(psi, n, op.target); break;
            case QSE_GATE_Y: orc_pauli_y(psi, n, op.target); break;
            case QSE_GATE_Z: orc_pauli_z(psi, n, op.target); break;
            case QSE_GATE_PHASE: orc_phase(psi, n, op.target, th); break;
            case QSE_GATE_RX: orc_rx(psi, n, op.target, th); break;
            case QSE_GATE_RY: orc_ry(psi, n, op.target, th); break;
            case QSE_GATE_RZ: orc_rz(psi, n, op.target, th); break;
            case QSE_GATE_CNOT: orc_cnot(psi, n, op.control, op.target); break;
            case QSE_GATE_CPHASE: orc_cphase(psi, n, op.control, op.target, th); break;
            case QSE_GATE_MCPHASE: orc_mcphase(psi, n, op.controls, op.target, th); break;
            case QSE_GATE_SWAP: orc_swap(psi, n, op.target, op.control); break;
            case QSE_GATE_MATRIX: orc_apply_matrix(psi, n, op.target, op.m); break;
            default: return -1;
        }
    }
    return 0;
}

/* StateVector.cpp:77-116 — cumulative |a|^2 summed sequentially, then per shot
 * lower_bound(cumulative, u*total). */
void orc_sample(const double* psi, int n, const double* u, int64_t nshots, uint64_t* out) {
    const cplx* s = amps(psi);
    const uint64_t N = 1ULL << n;
    std::vector<double> cum(N);
    for (uint64_t i = 0; i < N; ++i) cum[i] = norm2(s[i]);
    for (uint64_t i = 1; i < N; ++i) cum[i] += cum[i - 1];
    const double total = cum[N - 1];
    for (int64_t k = 0; k < nshots; ++k) {
        const double target = u[k] * total;
        out[k] = (uint64_t)(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
    }
}

/* StateVector.cpp:118-156 — total, target = u*total, first running sum >= target
 * (else 0), collapse to |idx>. The reference's total is an OpenMP reduction;
 * restated here as the single-thread sequential sum. */
uint64_t orc_measure(double* psi, int n, double u) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n;
    double total = 0.0;
    for (uint64_t i = 0; i < N; ++i) total += norm2(s[i]);
    const double target = u * total;
    uint64_t idx = 0;
    double run = 0.0;
    for (uint64_t i = 0; i < N; ++i) {
        run += norm2(s[i]);
        if (run >= target) { idx = i; break; }
    }
    for (uint64_t i = 0; i < N; ++i) s[i] = 0.0;
    s[idx] = 1.0;
    return idx;
}

/* StateVector.cpp:158-185 (measureQubit) and :187-221 (resetQubit) with an
 * injected uniform u in [0,1). Returns the outcome. */
int orc_collapse_qubit(double* psi, int n, int q, double u, int reset) {
    cplx* s = amps(psi);
    const uint64_t N = 1ULL << n, bit = 1ULL << q;
    double p1 = 0.0;
    for (uint64_t i = 0; i < N; ++i)
        if (i & bit) p1 += norm2(s[i]);
    const int outcome = (u < p1) ? 1 : 0;
    const double nf = std::sqrt(outcome ? p1 : 1.0 - p1);
    if (!reset) {
        for (uint64_t i = 0; i < N; ++i) {
            const bool one = (i & bit) != 0;
            if (one != (outcome == 1)) s[i] = 0.0;
            else s[i] /= nf;
        }
    } else {
        const uint64_t half = N >> 1, lmask = bit - 1;
        for (uint64_t i = 0; i < half; ++i) {
            const uint64_t i0 = ((i & ~lmask) << 1) | (i & lmask), i1 = i0 | bit;
            if (outcome == 0) s[i0] /= nf;
            else s[i0] = s[i1] / nf;
            s[i1] = 0.0;
        }
    }
    return outcome;
}

/* A gate list with mid-circuit measurements and resets (GateType::MEASURE
 * = 13, RESET_QUBIT = 14, Structurs.h:10; the reference declares them but its
 * Circuit::execute has no case for them, Circuit.cpp:110-111): unitary ops as
 * orc_run_ops, and each MEASURE/RESET_QUBIT op collapses its target with the
 * next uniform of u in list order (measureQubit / resetQubit semantics,
 * StateVector.cpp:158-221, restated in orc_collapse_qubit); a MEASURE's
 * outcome goes to results[op.control]. Returns 0, or -1 for an unsupported
 * kind, a result slot outside [0, nresults) or too few uniforms. */
int orc_run_measured(double* psi, int n, const qse_op* ops, int nops, const double* params, const double* u, int nu,
                     int* results, int nresults) {
    int k = 0;
    for (int i = 0; i < nops; ++i) {
        const qse_op& op = ops[i];
        if (op.kind == 13 || op.kind == 14) {
            if (k >= nu) return -1;
            const int o = orc_collapse_qubit(psi, n, op.target, u[k++], op.kind == 14);
            if (op.kind == 13) {
                if (op.control < 0 || op.control >= nresults) return -1;
                results[op.control] = o;
            }
        } else if (orc_run_ops(psi, n, &op, 1, params) != 0) {
            return -1;
        }
    }
    return 0;
}

/* Sum_i |a_i|^2 (sequential). */
double orc_norm2(const double* psi, int n) {
    const cplx* s = amps(psi);
    double t = 0.0;
    for (uint64_t i = 0; i < (1ULL << n); ++i) t += norm2(s[i]);
    return t;
}

/* Sum_i |a_i|^2 * sum_k w_k (-1)^popcount(i & mask_k) (no reference counterpart). */
double orc_expect_zstrings(const double* psi, int n, const uint64_t* masks, const double* w, int nterms) {
    const cplx* s = amps(psi);
    double acc = 0.0;
    for (uint64_t i = 0; i < (1ULL << n); ++i) {
        double f = 0.0;
        for (int k = 0; k < nterms; ++k) f += (__builtin_popcountll(i & masks[k]) & 1) ? -w[k] : w[k];
        acc += norm2(s[i]) * f;
    }
    return acc;
}

/* The inverse of one op, applied in place (adjoint sweep below): H, X, Y, Z,
 * CNOT and SWAP are their own inverses, angle kinds take the negated angle,
 * a generic 2x2 its conjugate transpose. */
static int orc_unapply(double* psi, int n, const qse_op& op, const double* params) {
    qse_op inv = op;
    inv.param = -1;
    inv.theta = -op_angle(op, params);
    if (op.kind == QSE_GATE_MATRIX) {
        // (G^dagger)_{rc} = conj(G_{cr})
        const double* m = op.m;
        const double t[8] = {m[0], -m[1], m[4], -m[5], m[2], -m[3], m[6], -m[7]};
        std::memcpy(inv.m, t, sizeof t);
    }
    return orc_run_ops(psi, n, &inv, 1, nullptr);
}

/* Adjoint-method gradient (no reference counterpart; restated from the
 * mathematics as a plain sequential check of the device reverse sweep):
 * psi = U_L ... U_1 |0>, E = <psi|O|psi> with O = sum_t w_t Z(mask_t),
 * lambda = O psi. Walking the gates backwards, at the output psi_k of gate k
 * (lambda_k = U_{k+1}^dag ... U_L^dag O psi) the term of a parameterised gate
 * is, with U = exp(-i theta G / 2) for RX/RY/RZ (G = X, Y, Z) and
 * U = exp(i theta P) for PHASE/CPHASE/MCPHASE (P the projector on the pinned
 * bits being 1):
 *   dE/dtheta = 2 Re <lambda_k| dU/dtheta U^dag |psi_k>
 *             = Im <lambda_k|G|psi_k>       (RX, RY, RZ)
 *             = -2 Im <lambda_k|P|psi_k>     (PHASE, CPHASE, MCPHASE),
 * added to grads[param] times the op's scale; then gate k is un-applied on
 * psi and lambda. Returns 0, or -1 for an unsupported kind. */
int orc_adjoint_expect(int n, const qse_op* ops, int nops, const double* params, int nparams, const uint64_t* masks,
                       const double* w, int nterms, double* energy, double* grads) {
    const uint64_t N = 1ULL << n;
    std::vector<cplx> psi(N, cplx(0.0, 0.0)), lam(N);
    psi[0] = 1.0;
    double* ps = reinterpret_cast<double*>(psi.data());
    double* ls = reinterpret_cast<double*>(lam.data());
    if (orc_run_ops(ps, n, ops, nops, params) != 0) return -1;
    double e = 0.0;
    for (uint64_t i = 0; i < N; ++i) {
        double f = 0.0;
        for (int k = 0; k < nterms; ++k) f += (__builtin_popcountll(i & masks[k]) & 1) ? -w[k] : w[k];
        e += norm2(psi[i]) * f;
        lam[i] = f * psi[i];
    }
    *energy = e;
    for (int j = 0; j < nparams; ++j) grads[j] = 0.0;
    const cplx I(0.0, 1.0);
    for (int k = nops - 1; k >= 0; --k) {
        const qse_op& op = ops[k];
        if (op.param >= 0 && op.param < nparams) {
            const uint64_t m = 1ULL << op.target;
            cplx z(0.0, 0.0);
            double d = 0.0;
            switch (op.kind) {
                case QSE_GATE_RX:
                    for (uint64_t i = 0; i < N; ++i) z += std::conj(lam[i]) * psi[i ^ m];
                    d = z.imag();
                    break;
                case QSE_GATE_RY:  // (Y psi)_i = -i psi_{i|m} (bit 0), +i psi_{i&~m} (bit 1)
                    for (uint64_t i = 0; i < N; ++i) z += std::conj(lam[i]) * ((i & m) ? I : -I) * psi[i ^ m];
                    d = z.imag();
                    break;
                case QSE_GATE_RZ:
                    for (uint64_t i = 0; i < N; ++i) z += std::conj(lam[i]) * ((i & m) ? -psi[i] : psi[i]);
                    d = z.imag();
                    break;
                case QSE_GATE_PHASE:
                case QSE_GATE_CPHASE:
                case QSE_GATE_MCPHASE: {
                    uint64_t all = m;
                    if (op.kind == QSE_GATE_CPHASE) all |= 1ULL << op.control;
                    if (op.kind == QSE_GATE_MCPHASE) all |= op.controls;
                    for (uint64_t i = 0; i < N; ++i)
                        if ((i & all) == all) z += std::conj(lam[i]) * psi[i];
                    d = -2.0 * z.imag();
                    break;
                }
                default: return -1;
            }
            grads[op.param] += op.scale * d;
        }
        if (orc_unapply(ps, n, op, params) != 0 || orc_unapply(ls, n, op, params) != 0) return -1;
    }
    return 0;
}

/* Pauli frame: X on every bit of xmask, then Z on every bit of zmask
 * (Noise.cpp:34-44 applies X_q then Z_q per qubit in ascending q). */
void orc_pauli_frame(double* psi, int n, uint64_t xmask, uint64_t zmask) {
    for (int q = 0; q < n; ++q) {
        if ((xmask >> q) & 1) orc_pauli_x(psi, n, q);
        if ((zmask >> q) & 1) orc_pauli_z(psi, n, q);
    }
}

/* ---- Noise (Noise.cpp:3-44) ---- */
struct orc_noise {
    std::vector<int> bit, phase;
    std::map<int, double> rates;
    double mult;
    std::mt19937 gen;
    std::uniform_real_distribution<double> dist{0.0, 1.0};
};

/* Noise.cpp:3-19 — per-gate-type rates; seed 42 and multiplier 1 in Circuit.cpp:5. */
orc_noise* orc_noise_create(int n, uint32_t seed, double mult) {
    orc_noise* h = new orc_noise;
    h->bit.assign(n, 0);
    h->phase.assign(n, 0);
    h->mult = mult;
    h->gen.seed(seed);
    h->rates[QSE_GATE_X] = 0.001;
    h->rates[QSE_GATE_Y] = 0.001;
    h->rates[QSE_GATE_Z] = 0.001;
    h->rates[QSE_GATE_H] = 0.002;
    h->rates[QSE_GATE_CNOT] = 0.005;
    h->rates[QSE_GATE_PHASE] = 0.001;
    h->rates[QSE_GATE_RX] = 0.001;
    h->rates[QSE_GATE_RY] = 0.001;
    h->rates[QSE_GATE_RZ] = 0.001;
    return h;
}

void orc_noise_destroy(orc_noise* h) { delete h; }

/* Replace the rate of one gate kind (rate < 0 removes the kind). */
void orc_noise_set_rate(orc_noise* h, int kind, double rate) {
    if (rate < 0) h->rates.erase(kind);
    else h->rates[kind] = rate;
}

/* Noise.cpp:21-32 — two draws per eligible gate, bit flip then phase flip. */
void orc_noise_accumulate(orc_noise* h, int kind, int target) {
    auto it = h->rates.find(kind);
    if (it == h->rates.end()) return;
    const double p = it->second * h->mult;
    if (h->dist(h->gen) < p) h->bit[target]++;
    if (h->dist(h->gen) < p) h->phase[target]++;
}

/* Parity masks of the accumulated counts (the frame Noise.cpp:34-44 applies). */
void orc_noise_masks(const orc_noise* h, uint64_t* xmask, uint64_t* zmask, int64_t* totals) {
    uint64_t x = 0, z = 0;
    int64_t tb = 0, tp = 0;
    for (size_t q = 0; q < h->bit.size(); ++q) {
        if (h->bit[q] % 2) x |= 1ULL << q;
        if (h->phase[q] % 2) z |= 1ULL << q;
        tb += h->bit[q];
        tp += h->phase[q];
    }
    *xmask = x;
    *zmask = z;
    if (totals) { totals[0] = tb; totals[1] = tp; }
}

/* Noise.cpp:45-49 (resetStats). */
void orc_noise_reset(orc_noise* h) {
    std::fill(h->bit.begin(), h->bit.end(), 0);
    std::fill(h->phase.begin(), h->phase.end(), 0);
}

/* Draws of one Circuit::execute() over a gate list (Circuit.cpp:67-71:
 * accumulate before each gate with cmd.target), for ntraj generators seeded
 * seed0 + t; rate < 0 keeps the reference rates, else every eligible kind uses
 * `rate`. Writes the per-trajectory parity masks. */
void orc_noise_trajectories(int n, const qse_op* ops, int nops, uint32_t seed0, int ntraj, double rate,
                            double mult, uint64_t* xmask, uint64_t* zmask) {
    for (int t = 0; t < ntraj; ++t) {
        orc_noise* h = orc_noise_create(n, seed0 + (uint32_t)t, mult);
        if (rate >= 0)
            for (auto& kv : h->rates) kv.second = rate;
        for (int k = 0; k < nops; ++k) orc_noise_accumulate(h, ops[k].kind, ops[k].target);
        orc_noise_masks(h, &xmask[t], &zmask[t], nullptr);
        orc_noise_destroy(h);
    }
}

/* The 10000th output of a default-seeded std::mt19937 (C++ standard
 * [rand.predef]: 4123659995); pins the engine the noise draws use. */
uint32_t orc_mt19937_kat(void) {
    std::mt19937 g;
    g.discard(9999);
    return (uint32_t)g();
}

/* Canonical uniforms from std::mt19937(seed), as uniform_real_distribution<double>(0,1). */
void orc_uniforms(uint32_t seed, int64_t count, double* out) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> d(0.0, 1.0);
    for (int64_t i = 0; i < count; ++i) out[i] = d(g);
}

}  // extern "C"

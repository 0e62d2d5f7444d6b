// qse_batch.hpp — work format of the batched small-circuit engine
// (qse_batch_run_small): many circuits of different qubit counts, each run
// from |0...0> through its own fused programs, its observable and, with
// gradients, the adjoint reverse sweep, all by one workgroup of one
// persistent launch. Shared by the host builder (qse_batch_plan.cpp), the
// gfx950 kernel (qse_batch.hip) and the CPU emulation used by the tests.
#pragma once
#include <stdint.h>

#include <vector>

#include "qse_fused.hpp"
#include "../../include/qse_hip.h"

namespace qse {

constexpr int BATCH_MAX_TERMS = 64;  // observable terms (one bit each in BCirc::obs_keep)
constexpr int BATCH_KMAX = 12;       // largest tile of a pass
// Largest tile of a joint (psi, lambda) pass: both tiles share the LDS of one
// complex64 tile pair of BATCH_KMAX qubits (64 KiB).
constexpr int batch_joint_kmax(int prec) { return prec == QSE_C128 ? BATCH_KMAX - 1 : BATCH_KMAX; }

enum BStepKind : int32_t {
    BS_BASIS = 0,   // psi = |0...0>
    BS_PROG = 1,    // run passes [pass_begin, pass_end) on psi (target bit 0) and/or lambda (bit 1)
    BS_OBS = 2,     // lambda_i = f(i) psi_i, f = the circuit's observable
    BS_ENERGY = 3,  // vals[out] = sum_i |psi_i|^2 f(i)
    BS_JOINT = 4    // joint passes [pass_begin, pass_end) on (psi, lambda) with probes: vals[out + v], v < nprobes
};

// All fields 32/64-bit: the kernel reads steps with scalar loads.
struct BStep {
    int32_t kind;
    int32_t target;
    int32_t pass_begin, pass_end;
    int32_t nprobes;
    int32_t out;
};

struct BCirc {
    uint64_t psi_off, lam_off;  // amplitude offsets of psi and lambda in the state arena
    int32_t n, k;               // qubits; tile qubits of its passes (k <= n)
    int32_t step_begin, step_end;
    uint64_t obs_keep;          // bit t: observable term t lies on this circuit's qubits
};

struct BatchGradRef {
    int slot, circ, param;
    double coeff;
};

// Host-side result of planning a batch.
struct BatchPlan {
    HostPlan prog;                     // every program's passes, indices rebased into shared arrays
    std::vector<BCirc> circs;
    std::vector<BStep> steps;
    std::vector<int32_t> order;        // circuits by decreasing estimated cost (work-queue order)
    std::vector<int> energy_slot;      // value slot of each circuit's energy
    std::vector<BatchGradRef> grads;   // grad[circ][param] += coeff * vals[slot]
    int nvals = 0;
    uint64_t arena_amps = 0;           // amplitudes of all psi and lambda vectors
    int kmax = 0;                      // largest tile of any pass
    int max_probes = 0;                // most probe values of one joint step
    double est_bytes = 0.0;            // algorithmic HBM bytes of the whole batch (per amplitude byte = 1)
    double est_bytes_rev = 0.0;        // the part of est_bytes in joint (reverse-sweep) steps
};

// Plans a batch (arguments as qse_batch_run_small). kmax: largest tile the
// kernel is built for; default_L: contiguous low qubits per tile.
int build_batch_plan(int prec, int ncirc, const int* nq, const int64_t* op_off, const qse_op* ops,
                     const int64_t* par_off, const double* params, const uint64_t* obs_masks, int nterms,
                     bool with_grads, int kmax, int default_L, BatchPlan& B);

// Energies (one per circuit) and, when grads != NULL, gradients aligned with
// the params array (grads[par_off[c] + j] = dE_c/dparams[par_off[c] + j]) from
// the values the steps produced.
void batch_results(const BatchPlan& B, const double* vals, const int64_t* par_off, double* energies, double* grads);

}  // namespace qse

// qse_adjoint_plan.hpp — host-side schedule of the adjoint reverse sweep
// (shared by the host-driven adjoint, qse_adjoint.hip, and the batched
// small-circuit engine, qse_batch.hip). No HIP types.
//
// For E(theta) = <psi|O|psi>, psi = U_L ... U_1 |0>, the reverse sweep keeps
// psi_l and lambda_l = U_{l+1}^dag ... U_L^dag O psi_L and, at every
// parameterised gate, adds a generator term Im<lambda|G|psi> (coefficients in
// qse_adjoint.hip). Gates are grouped from the end into blocks of pairwise
// commuting gates; a block's terms are evaluated at the block's output. Blocks
// without parameters need no evaluation point, so consecutive un-applications
// up to the next parameterised block are merged into one segment: the sweep
// evaluates segment s's terms at the current (psi, lambda) and then un-applies
// segment s's gates from both.
#pragma once
#include <stdint.h>

#include <utility>
#include <vector>

#include "../../include/qse_hip.h"

namespace qse {

struct GradTerm {
    int param;
    double coeff;  // d(angle)/d(param) times the generator coefficient
};

struct AdjItem {
    int first, count;         // op range [first, first + count)
    uint64_t nd = 0, dg = 0;  // non-diagonal / diagonal qubit masks
    std::vector<std::pair<uint64_t, GradTerm>> diag;  // Z-string mask -> term
    int pair_q = -1, pair_type = 0;                   // X (0) / Y (1) generator on pair_q
    GradTerm pair{-1, 0.0};
};

// An X (type 0) or Y (type 1) generator term on qubit q, evaluated directly
// as Im<lambda|X_q|psi> or Im<lambda|Y_q|psi> (sums over amplitude pairs).
struct PairTerm {
    int q, type;
    GradTerm t;
};

struct AdjSegment {
    std::vector<std::pair<uint64_t, GradTerm>> diag;  // Z-string terms at this point (rotated frame)
    std::vector<PairTerm> pair;   // X/Y terms left unrotated (pair mode)
    std::vector<qse_op> pre;      // V_s: rotation into this point's frame (from the unrotated state)
    std::vector<qse_op> unapply;  // V_s^dag, the segment's daggers, V_{s+1}, in application order
    std::vector<qse_op> forward;  // the segment's gates, forward order
};

int build_items(const qse_op* ops, int nops, std::vector<AdjItem>& items);
// Blocks of pairwise-commuting items, listed from the end of the circuit.
void build_blocks(const std::vector<AdjItem>& items, std::vector<std::vector<int>>& blocks);
// U^dag of op o with its angle resolved from params.
qse_op dagger(const qse_op& o, const double* params);
// Reverse-sweep segments (see the file comment). prefix receives the gates
// before the first parameterised block (forward order): never un-applied.
// Segment 0's rotation (segs[0].pre) must be applied to psi and lambda once
// lambda = O psi is formed; a psi restored from a checkpoint at segment s needs
// segs[s].pre.
// With pair_mode, X/Y generators stay in seg.pair and no rotation is made
// (pre and the rotations in unapply are empty of them).
int adjoint_segments(const qse_op* ops, int nops, const double* params, std::vector<AdjSegment>& segs,
                     std::vector<qse_op>& prefix, bool pair_mode = false);

}  // namespace qse

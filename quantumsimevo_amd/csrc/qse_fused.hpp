// qse_fused.hpp — program format of the LDS-tiled fused pass (shared by the
// host planner, qse_plan.cpp, and the gfx950 kernel, qse_fused.hip).
//
// A pass reads every amplitude of the state once and writes it once. The state
// is cut into tiles of 2^k amplitudes: the k "tile qubits" vary inside a tile,
// the other n-k ("external") qubits are fixed per tile. Tile qubits always
// include the low qubits 0..L-1, so a tile is made of contiguous runs of 2^L
// amplitudes (coalesced HBM access). Inside the tile (held in LDS), each
// workgroup thread owns 16 amplitudes that differ in 4 "register" tile bits;
// a stage applies gates whose non-diagonal qubit is a register bit, then the
// tile is re-distributed through LDS for the next stage's register bits.
// Diagonal gates are never a reason to end a pass: they are collected as
// monomial phases (multiply amplitude i by d when all bits of a mask are set)
// and flushed in the latest stage allowed (before the next non-diagonal gate
// on one of their qubits), merged by register/thread/external structure.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/qse_hip.h"

namespace qse {

constexpr int FUSE_REG = 4;        // register bits per thread (16 amplitudes)
constexpr int FUSE_MAX_K = 13;     // largest tile (qubits)
constexpr int FUSE_MIN_K = 8;      // smallest tile (16 threads)
constexpr int FUSE_MAX_SLOTS = 256;  // per-tile external factor slots per pass
constexpr int FUSE_MAX_PROBES = 128; // probe accumulators per joint pass

enum FOpType : uint32_t {
    FOP_DENSE = 0,  // 2x2 matrix on register position p
    FOP_H = 1,      // uncontrolled Hadamard with the reference's rounding: s(v0+v1), s(v0-v1)
    FOP_XSWAP = 2,  // exchange (X, CNOT target)
    FOP_MONO = 3,   // multiply by the product of the op's terms
    // Uncontrolled specialisations chosen by the planner (same math as FOP_DENSE):
    FOP_REAL = 4,   // all four entries real
    FOP_RXLIKE = 5, // real diagonal, imaginary off-diagonal
    // Joint passes over (psi, lambda) of an adjoint reverse sweep only: adds
    // Im<lambda|G|psi> over the tile to the pass's accumulator term_begin.
    FOP_PROBE = 6
};

// FOp::mode of a FOP_PROBE: the generator G.
enum ProbeKind : uint32_t {
    PROBE_ZSIGN = 0,  // Z-string: (-1)^(number of set bits among its qubits)
    PROBE_PROJ = 1,   // projector: 1 where all its qubits are set, else 0
    PROBE_X = 2,      // X on register position p
    PROBE_Y = 3       // Y on register position p
};
// A probe's qubits: register positions in rctl, other tile local bits in
// tctl, external qubits (global mask) in ectl.

// All fields are 32/64-bit so the kernel reads them with scalar loads (no
// byte-sized scalar loads on gfx950).
struct FOp {
    uint32_t type;
    uint32_t p;       // register position of the target (dense ops)
    uint32_t rctl;    // register-position mask that must be set
    uint32_t mode;    // unused (0)
    uint32_t tctl;    // dense: non-register tile local bits that must be set;
                      // MONO: thread-id bits that carry a factor in alphas
    uint32_t term_begin, term_end;  // MONO: range of FTerm
    uint32_t alpha_begin;           // MONO: FUSE_NO_ALPHA or index of k-4 complex thread-bit factors
    uint64_t ectl;    // dense: external qubits (global bit mask) that must be set
    double m[8];      // dense: row-major 2x2, (re, im) interleaved; MONO: m[0..1] = constant factor
};

constexpr uint32_t FUSE_NO_ALPHA = 0xffffffffu;

// Planner variants (build_plan flags).
enum PlanFlags : int {
    PLAN_RELABEL = 1,  // SWAPs as a qubit relabelling; last pass out of place
    PLAN_DAG = 2,      // passes by list scheduling over the gate dependency graph
    PLAN_LOOKAHEAD = 4, // with PLAN_DAG: a pass's qubit set grows by the most productive ready gate
    PLAN_CONTIG = 8     // each pass stores out of place with the qubit order that makes the next pass's
                        // tile the contiguous low qubits (its reads and writes then contiguous)
};

// Passes of at least this many stages take the first stage's amplitudes
// straight from HBM and store the last stage's straight to HBM in the run-time
// specialised kernels, where the lane mapping allows (measured on 30-qubit
// complex128 passes: 3.5% faster on the 14-pass random layered circuit, whose
// passes have 3-11 stages; 3% slower on the 3-stage QFT passes, whose lane bits
// then fall on other tile qubits).
constexpr int JIT_DIRECT_MIN_STAGES = 4;

// A plan that writes to the second buffer in any pass.
struct HostPlan;
bool plan_out_of_place(const HostPlan& P);

struct FTerm {
    uint32_t tmask;   // non-register tile local bits that must be set
    int32_t slot;     // >= 0: factor from the per-tile external slot; else d / ang
    uint64_t emask;   // slot < 0: external qubits (global mask) that must be set
    double d[2];      // factor (unit modulus: every monomial comes from a phase gate)
    double ang;       // its angle: d = (cos ang, sin ang)
    double pad;
};

struct FStage {
    uint32_t rmap[FUSE_REG];  // register position -> tile local bit
    uint32_t tmap[16];        // thread-id bit -> tile local bit (k-4 entries)
    uint32_t op_begin, op_end;
};

struct FExtTerm {
    uint64_t emask;   // external qubits (global mask) that must be set
    double ang;       // phase angle contributed when they are
};

struct FSlot {
    uint32_t begin, end;  // range of FExtTerm; slot angle = sum of passing terms' angles
};

struct FPass {
    int32_t k;            // tile qubits
    int32_t next;         // external qubits (n - k)
    int32_t nstages;
    int32_t nslots;
    uint32_t tq[16];       // load: tile position b -> source qubit (ascending)
    uint32_t so_q[16];     // store: store position c -> destination qubit (ascending)
    uint32_t so_l[16];     // store: store position c -> tile local bit holding it
    uint32_t epos[64];     // tile index bit e -> external source qubit (the lowest first, then descending)
    uint32_t eout[64];     // destination qubit of external source qubit epos[e]
    uint32_t stage_begin;
    uint32_t slot_begin;
    int32_t out_of_place; // 1: write to the other buffer (qubit permutation)
    int32_t nprobes;      // FOP_PROBE accumulators of the pass (joint passes)
    int32_t probe_base;   // value index of the pass's first accumulator within its program
    int32_t pad;
};

// Byte offsets of the arrays inside one uploaded program buffer.
struct FProgram {
    const FPass* passes;
    const FStage* stages;
    const FOp* ops;
    const FTerm* terms;
    const FSlot* slots;
    const FExtTerm* eterms;
    const double* alphas;
};

// A probe's gradient term: dE/dparams[param] += coeff * (probed value).
struct ProbeRef {
    int param;
    double coeff;
};

// Host-side plan (built by qse_plan.cpp, serialized and launched by qse_fused.hip).
struct HostPlan {
    int k = 0;
    std::vector<FPass> passes;
    std::vector<FStage> stages;
    std::vector<FOp> ops;
    std::vector<FTerm> terms;
    std::vector<FSlot> slots;
    std::vector<FExtTerm> eterms;
    std::vector<double> alphas;
    std::vector<int32_t> probe_ids;  // host only: caller's id of each probe value, in value order
};

// A host plan flattened into one buffer: the seven program arrays, each
// 256-byte aligned, at off[0..6] (passes, stages, ops, terms, slots, eterms,
// alphas).
struct Serialized {
    std::vector<char> bytes;
    size_t off[7];
};
void serialize(const HostPlan& P, Serialized& S);

// Adjoint reverse sweep of ops (applied to |0>) as joint (psi, lambda) plans:
// from the last op back to the first parameterised one, each parameterised
// op's generator probe (value: Im<lambda|G|psi> after the op) and then the
// op's inverse; lambda must start as O psi. refs[id] is the gradient term of
// probe id (HostPlan::probe_ids); the sweep is split into several plans when
// a pass would hold more than FUSE_MAX_PROBES probes.
int build_reverse_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params,
                       std::vector<HostPlan>& chunks, std::vector<ProbeRef>& refs);

// HIP source of every pass of plan P as run-time specialised kernels
// (qse_jit_gen.cpp); complex128 passes p with (*paired)[p] process tiles in
// adjacent pairs.
std::string jit_source(const HostPlan& P, int prec, const size_t* table_off,
                       const std::vector<bool>* paired = nullptr);

// HIP source of a joint reverse-sweep plan (passes over psi and lambda with
// probes, build_reverse_plan) as kernels qse_jit_jpass<i>(psi, lam, slot_table,
// ntiles, pv, part), coefficients read from the program buffer at table_off.
std::string jit_joint_source(const HostPlan& P, int prec, const size_t* table_off);
std::string jit_joint_kernel_name(int pass);

// The serialized plan with every value field zeroed (gate coefficients, phase
// factors and angles, thread-bit factors): plans that differ only in their
// values (the same circuit at other parameters) share it.
std::vector<char> structure_key(const HostPlan& P);

}  // namespace qse

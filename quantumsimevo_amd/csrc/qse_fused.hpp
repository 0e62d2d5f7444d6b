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

#include <vector>

namespace qse {

constexpr int FUSE_REG = 4;        // register bits per thread (16 amplitudes)
constexpr int FUSE_MAX_K = 13;     // largest tile (qubits)
constexpr int FUSE_MIN_K = 8;      // smallest tile (16 threads)
constexpr int FUSE_MAX_SLOTS = 256;  // per-tile external factor slots per pass

enum FOpType : uint8_t {
    FOP_DENSE = 0,  // 2x2 matrix on register position p
    FOP_H = 1,      // Hadamard with the reference's rounding: s(v0+v1), s(v0-v1)
    FOP_XSWAP = 2,  // exchange (X, CNOT target)
    FOP_MONO = 3    // multiply by the product of the op's terms
};

struct FOp {
    uint8_t type;
    uint8_t p;        // register position of the target (dense ops)
    uint8_t rctl;     // register-position mask that must be set
    uint8_t pad;
    uint32_t tctl;    // dense: non-register tile local bits that must be set
    uint32_t term_begin, term_end;  // MONO: range of FTerm
    uint32_t pad2;
    uint64_t ectl;    // dense: external qubits (global bit mask) that must be set
    double m[8];      // dense: row-major 2x2, (re, im) interleaved
};

struct FTerm {
    uint32_t tmask;   // non-register tile local bits that must be set
    int32_t slot;     // >= 0: factor from the per-tile external slot; else d
    uint64_t emask;   // slot < 0: external qubits (global mask) that must be set
    double d[2];
};

struct FStage {
    uint8_t rmap[FUSE_REG];   // register position -> tile local bit
    uint8_t tmap[16];         // thread-id bit -> tile local bit (k-4 entries)
    uint32_t op_begin, op_end;
};

struct FExtTerm {
    uint64_t emask;   // external qubits (global mask) that must be set
    double d[2];
};

struct FSlot {
    uint32_t begin, end;  // range of FExtTerm; factor = product of passing terms
};

struct FPass {
    int32_t k;            // tile qubits
    int32_t next;         // external qubits (n - k)
    int32_t nstages;
    int32_t nslots;
    uint8_t tq[16];       // load: tile position b -> source qubit (ascending)
    uint8_t so_q[16];     // store: store position c -> destination qubit (ascending)
    uint8_t so_l[16];     // store: store position c -> tile local bit holding it
    uint8_t epos[64];     // external source qubits, ascending (tile index bits)
    uint8_t eout[64];     // destination qubit of external source qubit epos[e]
    uint32_t stage_begin;
    uint32_t slot_begin;
    int32_t out_of_place; // 1: write to the other buffer (qubit permutation)
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
};

}  // namespace qse

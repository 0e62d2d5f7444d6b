// qse_internal.hpp — host-side internals of libqse_hip.so: context/state
// structs, kernel launchers and error plumbing. Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qse_hip.h"
#include "qse_fused.hpp"

namespace qse {

// Kernel families for the timing report (qse_ctx_timing).
enum Family {
    FAM_SWEEP = 0,
    FAM_FUSED = 1,
    FAM_READOUT = 2,
    FAM_ADJOINT = 3,
    FAM_BATCH = 4,
    FAM_DENSE = 5,
    FAM_COUNT = 6
};

struct TimedLaunch {
    int family;
    hipEvent_t a, b;
    double bytes;
};

struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// A fused-pass program, keyed by its serialized bytes and the state precision
// (the same plan serves both precisions, its compiled kernels do not): its
// place in the device program ring while it is there (dev, ring_off), and its
// run-time compiled kernels once it has them (one per pass).
struct ProgEntry {
    std::vector<char> bytes;
    int prec = 0;
    void* dev = nullptr;  // null once the ring has reused its place
    size_t ring_off = 0;
    uint64_t stamp = 0;
    int runs = 0;
    bool jit_failed = false;
    hipModule_t mod = nullptr;
    std::vector<hipFunction_t> fns;
};

// Device program ring: programs are copied to consecutive places of one device
// buffer from the same places of a pinned host buffer, in stream order, so an
// upload neither allocates nor waits for the device. The device copy of a
// place is overwritten only by a later upload on the same stream, after every
// kernel launched before it; the host copy of a place is rewritten once the
// copy that read it has completed (its event).
struct ProgRing {
    char* dev = nullptr;
    char* host = nullptr;
    size_t cap = 0, head = 0;
    struct Copy {
        size_t begin, end;
        hipEvent_t done;
    };
    std::vector<Copy> inflight;  // oldest first
    std::vector<hipEvent_t> spare;
};

// Compiled kernels of one plan structure (coefficients read from the program
// buffer at run time), built on a worker thread: state 0 while compiling,
// 1 ready, 2 failed.
struct JitJob {
    std::atomic<int> state{0};
    hipModule_t mod = nullptr;
    std::vector<hipFunction_t> fns;
    double ms = 0.0;
    std::thread worker;
};
// Joint reverse-sweep kernels (jit_joint_source) of one plan structure,
// compiled on first sight.
struct JointEntry {
    std::vector<char> key;  // structure_key of the joint plan
    int prec = 0;
    uint64_t stamp = 0;
    hipModule_t mod = nullptr;
    std::vector<hipFunction_t> fns;
};
struct StructEntry {
    std::vector<char> key;  // structure_key of the plan
    int prec = 0;
    int seen = 0;
    bool counted = false;
    uint64_t stamp = 0;
    std::shared_ptr<JitJob> job;
};

// Tile shape the shape tuner chose for a gate list, keyed by its structure
// (qubit count, precision, gate kinds, qubits and parameter bindings, with
// angles and matrices zeroed): runs of the same circuit at other angles reuse
// it instead of tuning again.
struct ShapeEntry {
    std::vector<char> key;
    int k = 0, L = 0;
    uint64_t stamp = 0;
};

// Key of a gate list's structure for shape_cache (ops with angles and
// matrices zeroed), and the cached entry for it (null if none).
std::vector<char> gate_structure_key(int n, int prec, const qse_op* ops, int nops);
const ShapeEntry* find_shape(const qse_ctx* ctx, const std::vector<char>& key);
// Candidate tile shapes of a gate list ranked by the static cost model
// (passes x measured relative pass cost of the shape).
struct ShapeCandidate {
    int k, L, passes;
    double est;
};
int rank_shapes(int n, int prec, const qse_op* ops, int nops, const double* params, bool allow_oop,
                std::vector<ShapeCandidate>& out);
// Insert or replace the shape of a structure (least recently used out at 256).
void remember_shape(qse_ctx* ctx, std::vector<char> key, int k, int L);

// A tuned plan for one exact input (qubit count, precision, gate list,
// parameter values, shape settings), so repeated runs skip planning.
struct PlanEntry {
    std::vector<char> key;
    HostPlan plan;
    int k = 0, L = 0;
    uint64_t stamp = 0;
};

}  // namespace qse

struct qse_ctx {
    int device = 0;
    int num_cus = 256;
    size_t lds_per_block = 65536;
    hipStream_t stream = nullptr;
    // Device scratch, grown on demand (partials, flags, cumulative arrays).
    qse::Scratch scratch[12];
    // Recently run fused programs, their device copies in the program ring.
    std::vector<qse::ProgEntry> prog_cache;
    qse::ProgRing ring;
    uint64_t prog_clock = 0;
    // Kernels by plan structure (the adjoint's and batched expectations' runs).
    std::vector<qse::StructEntry> struct_cache;
    // Joint reverse-sweep kernels by plan structure (qse_adjoint_expect's joint schedule).
    std::vector<qse::JointEntry> joint_cache;
    // Tuned tile shapes by gate-list structure (internal runs of the adjoint
    // and batched expectations).
    std::vector<qse::ShapeEntry> shape_cache;
    // Tuned plans of recent qse_run inputs.
    std::vector<qse::PlanEntry> plan_cache;
    // Pinned host staging for small readbacks.
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // How the adjoint evaluates X/Y generator terms (QSE_ADJ_TERMS_*).
    int adjoint_terms = 0;
    // Tile shape of qse_batch_run_small per precision (0: 12-qubit tiles with
    // 3 (complex128) or 4 (complex64) contiguous low qubits); set by
    // qse_batch_autotune.
    int batch_k[2] = {0, 0};
    int batch_L[2] = {0, 0};
    int batch_team[2] = {0, 0};  // workgroups per team for circuits of 16+ qubits (0: 16)
    // Renormalise states after this many passes of qse_run calls (0: never).
    int64_t renorm_every = 0;
    // qse_run sends runs of at least this many ops per qubit on qubits 0..m-1
    // (m = 4..6) to the matrix cores as one dense block (0: never).
    int dense_per_qubit = 20;
    // Fusion shape per precision (tile qubits k, contiguous low qubits L).
    int tile_bits[2] = {0, 0};
    int low_bits[2] = {0, 0};
    // Run-time specialisation (qse_jit_mode) and its counters.
    int jit_mode = QSE_JIT_AUTO;
    int64_t jit_compiled = 0;
    int64_t jit_launches = 0;
    double jit_compile_ms = 0.0;
    // Timing.
    bool timing = false;
    std::vector<qse::TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t launches[qse::FAM_COUNT] = {};
    double total_ms[qse::FAM_COUNT] = {};
    double total_bytes[qse::FAM_COUNT] = {};
};

struct qse_state {
    int n = 0;
    int prec = QSE_C128;
    void* d = nullptr;
    // Second buffer of the same size, allocated on first need: a fused pass
    // that permutes qubits writes here, then the two pointers are exchanged.
    void* shadow = nullptr;
    size_t bytes = 0;
    bool owned = true;
    // Passes of qse_run calls since the last periodic renormalisation.
    int64_t passes_since_renorm = 0;
};

namespace qse {

// Error plumbing: set_error records a message for qse_last_error().
int set_error(int status, const std::string& msg);
int hip_error(hipError_t e, const char* what);
#define QSE_HIP(call)                                            \
    do {                                                         \
        hipError_t _e = (call);                                  \
        if (_e != hipSuccess) return ::qse::hip_error(_e, #call); \
    } while (0)

// Scratch buffer `slot` of at least `bytes` bytes (grows; stream-ordered free).
int scratch(qse_ctx* ctx, int slot, size_t bytes, void** out);
// Pinned host buffer of at least `bytes` bytes.
int pinned(qse_ctx* ctx, size_t bytes, void** out);

// Timing hooks around a launch; `bytes` = algorithmic bytes of that launch.
void launch_begin(qse_ctx* ctx, int family, TimedLaunch* t);
void launch_end(qse_ctx* ctx, TimedLaunch* t, double bytes);
int fold_timing(qse_ctx* ctx);

// Stream-ordered upload of a small host block (programs, matrices) into the
// device program ring; the device copy stays valid until later uploads wrap
// the ring around to its place, which happens only after the kernels launched
// before them. Returns the byte offset in ctx->ring.dev.
int ring_upload(qse_ctx* ctx, const std::vector<char>& bytes, size_t* off);

// Grid for a grid-stride sweep over `work` items with `block` threads.
unsigned grid_for(const qse_ctx* ctx, uint64_t work, int block, int waves_per_cu = 32);

inline size_t amp_bytes(int prec) { return prec == QSE_C128 ? 16 : 8; }

// Run-time specialised kernels for run_plan: none, keyed by the exact program
// (qse_run inputs), or keyed by the plan structure (internal runs whose
// coefficients change from call to call).
enum JitKind { JIT_NONE = 0, JIT_BY_PROGRAM = 1, JIT_BY_STRUCTURE = 2 };

// Smallest state that goes through the fused path (below: one sweep per gate).
constexpr int FUSE_MIN_QUBITS = 10;

// ---- per-gate sweeps (qse_gates.hip) ----
// Pair kernel: for k < nk, base = deposit(k); (i0, i1) = (base|off0, base|off1).
enum PairMode {
    PAIR_DENSE = 0,  // (v0, v1) <- M (v0, v1)
    PAIR_SWAP = 1,   // exchange v0, v1
    PAIR_DIAG = 2,   // v0 *= d0, v1 *= d1
    PAIR_DIAG1 = 3,  // v1 *= d1 only (off0 unused)
    PAIR_HADAMARD = 4 // s*(v0+v1), s*(v0-v1) — the reference's rounding
};
int launch_pairs(qse_ctx* ctx, qse_state* st, int mode, const uint64_t* fixed0, int nfixed0, uint64_t ones,
                 uint64_t off0, uint64_t off1, const double m[8], double algo_bytes);
int launch_frame(qse_ctx* ctx, qse_state* st, uint64_t xmask, uint64_t zmask);
int launch_set_basis(qse_ctx* ctx, qse_state* st, uint64_t index);
int launch_convert_upload(qse_ctx* ctx, qse_state* st, const double* dev_c128, uint64_t count, uint64_t first);
int launch_convert_download(qse_ctx* ctx, const qse_state* st, double* dev_c128, uint64_t count, uint64_t first);
int launch_scale(qse_ctx* ctx, qse_state* st, double s);
int launch_random(qse_ctx* ctx, qse_state* st, uint64_t seed);

// ---- readout helpers (qse_readout.hip) ----
// Stream-ordered FP64 renormalisation; *d_nrm2 (if given) receives the device
// address of the norm^2 measured before scaling.
int renormalize_async(qse_ctx* ctx, qse_state* st, double** d_nrm2);
// Count `passes` against the context's renormalisation period.
int renorm_tick(qse_ctx* ctx, qse_state* st, int64_t passes);

}  // namespace qse

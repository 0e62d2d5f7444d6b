/*
 * qse_hip.h — C ABI of the MI355X (gfx950) state-vector library `libqse_hip.so`.
 *
 * This is the thin, exception-free boundary that the C++ class
 * `CircuitUnitaryOperationHIP` (include/quantumsimevo/CircuitUnitaryOperationHIP.h)
 * calls. Signatures use only plain C types: no torch, no HIP types.
 *
 * Conventions (identical to the reference simulator):
 *  - qubit q is bit q of the amplitude index (qubit 0 = least significant bit),
 *    QSETests/QSETests.cpp:131, StateVector.cpp:262-265;
 *  - a state of n qubits holds 2^n amplitudes, stored contiguously and
 *    interleaved (re, im): FP64 pairs (QSE_C128, 16 B/amplitude) or FP32 pairs
 *    (QSE_C64, 8 B/amplitude). Host buffers crossing this ABI are always FP64
 *    (re, im) pairs, whatever the device precision;
 *  - gate matrices follow QubitUnitaryOperation.cpp:10-71 and
 *    CircuitUnitaryOperation.cpp:10-338.
 *
 * Every function returns a status code (QSE_OK == 0). qse_status_string() gives a
 * message for it and qse_last_error() a detailed message for the last failure on
 * the calling thread. The reference aborts on device errors
 * (CircuitUnitaryOperationGPU.cu:13-21); this ABI reports them instead so callers
 * (C++, Python) can raise.
 *
 * Stream semantics: a context owns one HIP stream. Gate and readout calls are
 * stream-ordered; calls that return host data (download, sample, measure,
 * expectation, gradients) synchronise the context stream before returning.
 * A context is not thread-safe: one host thread drives it at a time, as the
 * reference backend (CircuitUnitaryOperationGPU.h:39-47).
 */
#ifndef QSE_HIP_H
#define QSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSE_ABI_VERSION 1

/* ---- status codes ---- */
enum qse_status {
    QSE_OK = 0,
    QSE_ERR_INVALID_ARG = 1,   /* bad qubit index, null pointer, bad size */
    QSE_ERR_NO_DEVICE = 2,     /* no HIP device visible */
    QSE_ERR_OUT_OF_MEMORY = 3, /* device allocation failed */
    QSE_ERR_HIP = 4,           /* a HIP runtime call failed */
    QSE_ERR_UNSUPPORTED = 5,   /* gate kind / precision / size not supported */
    QSE_ERR_LIMIT = 6          /* request exceeds a documented limit */
};

/* ---- precisions ---- */
enum qse_precision {
    QSE_C128 = 0, /* complex<double> storage and arithmetic */
    QSE_C64 = 1   /* complex<float> storage and arithmetic; FP64 reductions */
};

/* ---- gate kinds: the numbering of GateType, Structurs.h:10 ---- */
enum qse_gate_kind {
    QSE_GATE_IDENTITY = 0,
    QSE_GATE_X = 1,
    QSE_GATE_Y = 2,
    QSE_GATE_Z = 3,
    QSE_GATE_H = 4,
    QSE_GATE_PHASE = 5,
    QSE_GATE_RX = 6,
    QSE_GATE_RY = 7,
    QSE_GATE_RZ = 8,
    QSE_GATE_CNOT = 9,
    QSE_GATE_CPHASE = 10,
    QSE_GATE_MCPHASE = 11,
    QSE_GATE_SWAP = 12,
    /* GateType::MEASURE (13) and RESET_QUBIT (14): declared by the reference
     * but never dispatched (Circuit.cpp:110-111). Only qse_run_measured takes
     * them; the unitary entry points reject them. */
    QSE_GATE_MEASURE = 13,
    QSE_GATE_RESET_QUBIT = 14,
    QSE_GATE_MATRIX = 32 /* generic 2x2, ICircuitUnitaryOperation::applyGate */
};

/*
 * One gate of a gate list; the C counterpart of GateCommand (Structurs.h:12-20).
 *  target   : target qubit (SWAP: first qubit, Circuit.cpp:59-62)
 *  control  : CNOT/CPHASE control, SWAP second qubit; -1 when unused
 *  controls : MCPHASE control mask, bit q = qubit q (GateCommand::controls_list)
 *  angle    : theta, or theta + scale * params[param] when param >= 0
 *             (differentiable RX, RY, RZ, PHASE, CPHASE)
 *  m        : MATRIX kind only: row-major 2x2, (re, im) interleaved (Gate2x2)
 */
typedef struct qse_op {
    int32_t kind;
    int32_t target;
    int32_t control;
    int32_t param;
    uint64_t controls;
    double theta;
    double scale;
    double m[8];
} qse_op; /* 104 bytes */

typedef struct qse_ctx qse_ctx;
typedef struct qse_state qse_state;

/* ---- library / device ---- */
int qse_abi_version(void);
const char* qse_status_string(int status);
const char* qse_last_error(void);
/* Number of HIP devices (0 when none). Replaces the cudaGetDeviceCount probe of
 * CircuitUnitaryOperationFactory.cpp:17-24. */
int qse_device_count(int* count);
/* Free / total device memory of `device` (CircuitUnitaryOperationFactory.cpp:41-63). */
int qse_device_mem_info(int device, size_t* free_bytes, size_t* total_bytes);
/* Device name (the architecture name, e.g. "gfx950:sramecc+:xnack-", when the
 * driver reports no marketing name), compute units and HBM size
 * (HardwareCheck.cpp:35-54). */
int qse_device_info(int device, char* name, size_t name_len, int* compute_units, size_t* total_bytes);

/* ---- context: one device, one stream (CircuitUnitaryOperationGPU.cu:140-156) ---- */
int qse_ctx_create(int device, qse_ctx** out);
int qse_ctx_destroy(qse_ctx* ctx);
int qse_ctx_sync(qse_ctx* ctx);
/* Fusion shape used by qse_run(): tile_bits = qubits per LDS tile (k),
 * low_bits = contiguous low qubits always in the tile (coalescing run of
 * 2^low_bits amplitudes). tile_bits = 0 (the default) lets the shape tuner
 * pick the shape per gate list: every candidate shape is planned and the one
 * of least estimated time (passes x measured relative pass cost) runs;
 * low_bits = 0 with a set tile_bits selects the default run length. No
 * counterpart in the reference (one kernel launch per gate). qse_ctx_get_fusion
 * returns the settings (0 = tuned). */
int qse_ctx_set_fusion(qse_ctx* ctx, int precision, int tile_bits, int low_bits);
/* Measured tile shape (no counterpart in the reference): the shape tuner's
 * best `max_candidates` shapes for this gate list are each compiled and timed
 * on st (two runs, then three timed; st's amplitudes are overwritten); the
 * fastest is kept for the gate list's structure (kinds, qubits, parameter
 * bindings) and used by later qse_run calls of it while the context's tile
 * size is unset. Returns the shape, its time per run in ms, and per candidate
 * tried_k/L/ms (arrays of max_candidates entries, may be NULL) with the count
 * in *ntried. Below 10 qubits nothing is tuned (k = L = 0). */
int qse_autotune(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int max_candidates,
                 int* k, int* L, double* ms, int* tried_k, int* tried_L, double* tried_ms, int* ntried);
int qse_ctx_get_fusion(qse_ctx* ctx, int precision, int* tile_bits, int* low_bits);
/* Run-time specialisation of fused passes (no counterpart in the reference,
 * whose GPU backend launches one fixed kernel per gate,
 * CircuitUnitaryOperationGPU.cu:194-330). A fused program is compiled with
 * hiprtc into straight-line kernels with its coefficients as literals:
 * QSE_JIT_AUTO (default) when the same program (gates and angles) runs a
 * second time, QSE_JIT_ALWAYS on first use, QSE_JIT_OFF never (the generic
 * program-interpreting kernel). Results agree with the generic kernel to
 * rounding. The internal runs of qse_adjoint_expect and qse_batch_expect,
 * whose coefficients change with the parameters, use kernels keyed by the
 * plan's structure instead, which read the coefficients from the program
 * buffer: QSE_JIT_AUTO compiles a structure on a worker thread from its
 * second run on (the generic kernel runs until the module is ready),
 * QSE_JIT_ALWAYS compiles it on first use and waits. */
enum qse_jit_mode { QSE_JIT_OFF = 0, QSE_JIT_AUTO = 1, QSE_JIT_ALWAYS = 2 };
int qse_ctx_set_jit(qse_ctx* ctx, int mode);
/* Programs and structures compiled, passes launched through compiled kernels,
 * and total compile time in milliseconds since the context was created. */
int qse_ctx_jit_stats(qse_ctx* ctx, int64_t* compiled, int64_t* jit_launches, double* compile_ms);
/* Wait until every compilation started on worker threads has finished (no
 * counterpart in the reference). */
int qse_ctx_jit_wait(qse_ctx* ctx);
/* Kernel timing: when enabled, every kernel launched by the context records HIP
 * events; qse_ctx_timing() returns the count and summed milliseconds per kernel
 * family. Families: 0 pair sweep, 1 fused tile pass, 2 readout, 3 adjoint, 4 batched circuits,
 * 5 dense matrix-core block. */
int qse_ctx_enable_timing(qse_ctx* ctx, int enable);
int qse_ctx_timing(qse_ctx* ctx, int family, int64_t* launches, double* total_ms, double* bytes);
int qse_ctx_reset_timing(qse_ctx* ctx);

/* ---- state storage (device mirror of StateVector, StateVector.h:17-47) ---- */
int qse_state_create(qse_ctx* ctx, int nqubits, int precision, qse_state** out);
int qse_state_destroy(qse_ctx* ctx, qse_state* st);
int qse_state_info(const qse_state* st, int* nqubits, int* precision, void** device_ptr, size_t* bytes);
/* Host <-> device copies of the whole state; host layout is FP64 (re, im) pairs.
 * Upload replaces ensureOnDevice's H2D copy (CircuitUnitaryOperationGPU.cu:158-180),
 * download replaces flush's D2H copy (:182-188). Download synchronises. */
int qse_state_upload(qse_ctx* ctx, qse_state* st, const double* host_amps);
int qse_state_download(qse_ctx* ctx, qse_state* st, double* host_amps);
/* Partial download of amplitudes [first, first+count). */
int qse_state_download_range(qse_ctx* ctx, qse_state* st, uint64_t first, uint64_t count, double* host_amps);
/* |index> (StateVector::reset with index 0, StateVector.cpp:34-46). */
int qse_state_set_basis(qse_ctx* ctx, qse_state* st, uint64_t index);
int qse_state_copy(qse_ctx* ctx, qse_state* dst, const qse_state* src);
/* Seeded random normalised state, generated on the device (synthetic input of
 * the large-state benchmarks; no reference counterpart). Amplitude i is
 * r*(cos 2 pi u2, sin 2 pi u2), r = sqrt(-2 ln u1) (Box-Muller), with
 * h1 = splitmix64(seed ^ (2i+1) * 0x9E3779B97F4A7C15), h2 = splitmix64(h1),
 * u1 = ((h1 >> 11) + 1) * 2^-53, u2 = (h2 >> 11) * 2^-53; then the state is
 * divided by its FP64 norm. */
int qse_state_set_random(qse_ctx* ctx, qse_state* st, uint64_t seed);

/* ---- single gates: one coalesced strided sweep per call ----
 * CircuitUnitaryOperation.cpp line ranges each replaces are cited in the
 * implementation; the kernels are independent of the reference's CUDA file. */
int qse_apply_matrix(qse_ctx* ctx, qse_state* st, int q, const double m[8]); /* applyGate */
int qse_apply_hadamard(qse_ctx* ctx, qse_state* st, int q);
int qse_apply_x(qse_ctx* ctx, qse_state* st, int q);
int qse_apply_y(qse_ctx* ctx, qse_state* st, int q);
int qse_apply_z(qse_ctx* ctx, qse_state* st, int q);
int qse_apply_phase(qse_ctx* ctx, qse_state* st, int q, double theta);
int qse_apply_rx(qse_ctx* ctx, qse_state* st, int q, double theta);
int qse_apply_ry(qse_ctx* ctx, qse_state* st, int q, double theta);
int qse_apply_rz(qse_ctx* ctx, qse_state* st, int q, double theta);
int qse_apply_cnot(qse_ctx* ctx, qse_state* st, int control, int target);
int qse_apply_cphase(qse_ctx* ctx, qse_state* st, int control, int target, double theta);
int qse_apply_mcphase(qse_ctx* ctx, qse_state* st, const int* controls, int ncontrols, int target, double theta);
int qse_apply_swap(qse_ctx* ctx, qse_state* st, int q0, int q1);
/* One Pauli frame: psi'[i] = (-1)^popcount(i & zmask) * psi[i ^ xmask]; equal to
 * Noise::applyPauliNoise (Noise.cpp:34-44) with xmask/zmask the odd flip counts. */
int qse_pauli_frame(qse_ctx* ctx, qse_state* st, uint64_t xmask, uint64_t zmask);

/* ---- fused execution of a gate list (Circuit::execute, Circuit.cpp:64-120) ----
 * The list is planned into LDS-tiled passes (gates on tile qubits and every
 * diagonal gate fused into one HBM read+write of the state). params may be NULL
 * when no op has param >= 0. */
int qse_run(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params);
/* The gate list applied `reps` times in a row (repeated identical sweeps, e.g.
 * Trotter steps; no counterpart in the reference, whose Circuit::execute
 * replays its list once per call, Circuit.cpp:64-120). After two eager runs
 * the launches of one run (two when a run exchanges the state's buffers) are
 * captured in a HIP graph and replayed; results equal `reps` calls of qse_run.
 * The capture records kernels only (no memset nodes). */
int qse_run_repeat(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, int reps);
/* A gate list with mid-circuit measurements and resets (QSE_GATE_MEASURE,
 * QSE_GATE_RESET_QUBIT; the reference's GateType::MEASURE / RESET_QUBIT,
 * Structurs.h:10, with the result slot of GateCommand::resultIdx, :18). The
 * unitary stretches between them run as qse_run; each MEASURE or RESET_QUBIT
 * op collapses its target qubit with the next uniform of u, in list order:
 * outcome = (u < P(qubit = 1)), then the state is renormalised
 * (StateVector::measureQubit, StateVector.cpp:158-185); RESET_QUBIT then
 * returns the qubit to |0> (StateVector::resetQubit, :187-221). A MEASURE op
 * writes its outcome to results[op.control], its result slot
 * (0 <= control < nresults). nu must be the number of MEASURE and RESET_QUBIT
 * ops. A collapse onto an outcome of probability 0 cannot happen for
 * u in [0, 1); synchronises at every measurement. */
int qse_run_measured(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params, const double* u,
                     int nu, int32_t* results, int nresults);
/* Same list, one single-gate sweep per op (no fusion): the drop-in per-call
 * path of CircuitUnitaryOperationHIP and the unfused baseline. */
int qse_run_sweeps(qse_ctx* ctx, qse_state* st, const qse_op* ops, int nops, const double* params);
/* Number of HBM passes qse_run() would make for this list (planner report). */
int qse_plan_passes(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, int* passes);
/* The plan qse_run() would execute: passes, tile shape (0, 0 below 10 qubits,
 * where every gate is its own sweep) and total LDS stages. Any out pointer may
 * be NULL. */
int qse_plan_info(qse_ctx* ctx, int nqubits, int precision, const qse_op* ops, int nops, int* passes,
                  int* tile_bits, int* low_bits, int* stages);

/* FP64 closed-form check of a full-size QFT run (no counterpart in the
 * reference): out[2c], out[2c+1] = <indices[c]| QFT |psi> for the normalised
 * state that qse_state_set_random(seed) produces on an nqubits state, with the
 * QFT of QFT.cpp:9-22 (= bit-reversed unitary DFT, exponent +2 pi i x k / N).
 * Evaluated from the seed by one FP64 sweep per 16 indices, without a state
 * vector, so it can check runs whose input no longer fits beside the output.
 * count <= 1024; synchronises. */
int qse_qft_probe(qse_ctx* ctx, int nqubits, uint64_t seed, const uint64_t* indices, int count, double* out);
/* Practical copy ceiling of the device (no counterpart in the reference): GB/s
 * (read + write) of an out-of-place copy of 2^log2_elems 16-byte elements in
 * 64 KiB tiles made of contiguous runs of 2^run_log2 elements, the access
 * pattern of a fused pass with run_log2 contiguous low qubits (complex128).
 * Allocates two buffers of 16 * 2^log2_elems bytes for the call; synchronises. */
int qse_copy_bandwidth(qse_ctx* ctx, int log2_elems, int run_log2, double* gbs);
/* Measurement of a kernel choice (no counterpart in the reference): `reps`
 * layers of RY(0.3 + 0.1 q) on qubits q = 0..5 (in that order) applied in
 * place to a complex128 state of 12 or more qubits in one pass, with the
 * amplitude pairs of those qubits brought together through LDS between
 * register layouts (method 0, the fused pass's exchange) or across the lanes
 * of a wavefront by DPP / ds_swizzle / ds_bpermute (method 1). Both give the
 * same state to rounding; ms = the launch's HIP-event time. Synchronises. */
int qse_exchange_probe(qse_ctx* ctx, qse_state* st, int method, int reps, double* ms);

/* ---- dense blocks on the matrix cores (no counterpart in the reference) ----
 * qse_compose_dense: the 2^m x 2^m unitary of a gate list on qubits 0..m-1
 * (gate semantics of CircuitUnitaryOperation.cpp:10-338), computed on the host;
 * u is row-major, (re, im) interleaved, 2 * 4^m doubles; row = output index.
 * qse_apply_dense: psi <- (U (x) on qubits 0..m-1) psi for m = 4, 5, 6 on a
 * complex128 state of at least m + 4 qubits, one pass over the state with FP64
 * MFMA (v_mfma_f64_16x16x4f64); QSE_ERR_UNSUPPORTED for complex64 states or
 * other m. Stream-ordered. */
int qse_compose_dense(int m, const qse_op* ops, int nops, const double* params, double* u);
int qse_apply_dense(qse_ctx* ctx, qse_state* st, int m, const double* u);
/* Routing of qse_run to the matrix cores (no counterpart in the reference): a
 * maximal run of consecutive ops confined to qubits 0..m-1 (m = 4, 5 or 6,
 * the highest touched qubit + 1, at least 4) with at least ops_per_qubit * m
 * ops, on a complex128 state of m + 4 qubits or more, runs as one
 * qse_apply_dense pass of its composed unitary instead of fused gate passes.
 * Default 20 (about 16 layers of one-qubit gates plus a CNOT brickwork, the
 * measured crossover on MI355X); 0 turns the routing off. qse_run_repeat and
 * qse_run_sweeps never route. */
int qse_ctx_set_dense_blocks(qse_ctx* ctx, int ops_per_qubit);

/* ---- readout (StateVector.cpp:77-156), driven by host-supplied uniforms ----
 * u are canonical uniforms in [0,1) (std::generate_canonical<double,53>); the
 * target of shot s is u[s] * total, total = sum |a_i|^2, and the outcome is the
 * first index whose running probability sum is >= target (std::lower_bound over
 * the sequential cumulative array). Outcomes equal the sequential reference on
 * identical amplitudes: draws whose target lies inside the proven rounding band
 * of the device cumulative sum are re-resolved by an exact sequential sum. */
int qse_sample(qse_ctx* ctx, qse_state* st, const double* u, int64_t nshots, uint64_t* out);
/* Collapsing measurement: outcome as StateVector::measure (first running sum >=
 * u*total, else 0); the state becomes |outcome>. */
int qse_measure(qse_ctx* ctx, qse_state* st, double u, uint64_t* outcome);
/* sum_i |a_i|^2 in FP64. */
int qse_norm2(qse_ctx* ctx, qse_state* st, double* out);
/* FP64 renormalisation on the device (no counterpart in the reference):
 * psi <- psi / sqrt(sum_i |a_i|^2), the norm reduced and the scale factor
 * computed in FP64 whatever the state precision, without a host round trip;
 * stream-ordered unless norm2_before is non-null, which then receives the
 * norm^2 before scaling (synchronises). A state of zero or non-finite norm is
 * left unchanged. */
int qse_state_renormalize(qse_ctx* ctx, qse_state* st, double* norm2_before);
/* Periodic renormalisation (no counterpart in the reference): with
 * every_passes > 0, qse_run, qse_run_repeat and qse_run_sweeps count the passes
 * they make over a state (fused passes, or one per gate on the sweep path) and
 * renormalise it at the end of the call that reaches every_passes passes since
 * the last renormalisation. 0 (the default) turns it off. Bounds the norm drift
 * of long complex64 runs; internal runs (adjoint, batched expectations) are not
 * renormalised. */
int qse_ctx_set_renormalize(qse_ctx* ctx, int every_passes);
/* Probability of qubit q being 1 (StateVector::measureQubit, StateVector.cpp:158-185). */
int qse_prob_one(qse_ctx* ctx, qse_state* st, int q, double* out);
/* Mid-circuit single-qubit collapse with a host uniform u: outcome = (u < P(1));
 * renormalises (StateVector.cpp:158-185). With reset != 0 the qubit is then
 * returned to |0> (StateVector::resetQubit, StateVector.cpp:187-221). */
int qse_collapse_qubit(qse_ctx* ctx, qse_state* st, int q, double u, int reset, int* outcome);

/* ---- expectation values of diagonal observables (FP64 accumulation) ----
 * sum_i |a_i|^2 f(i) with f(i) = sum_k w_k * (-1)^popcount(i & masks[k])
 * (a weighted sum of Z-strings). */
int qse_expect_zstrings(qse_ctx* ctx, qse_state* st, const uint64_t* masks, const double* weights, int nterms, double* out);
/* The same value, computed without allocating any memory (no counterpart in
 * the reference): for callers that cannot allocate, such as tight serving
 * loops, code under HIP graph capture and harnesses that forbid hipMalloc. The
 * kernels and summation order are those of qse_expect_zstrings (bit-identical
 * results); masks, weights, partial sums and the result go through a fixed
 * 32 KiB device array set up per device at context creation. nterms <= 64
 * (QSE_ERR_LIMIT above). The array is shared by every call on one GPU and keeps
 * the last call's contents: calls that overlap on one GPU (several host
 * threads, or several contexts of one device driven concurrently) race.
 * Synchronises the context stream. */
int qse_expect_zstrings_noalloc(qse_ctx* ctx, qse_state* st, const uint64_t* masks, const double* weights, int nterms,
                                double* out);

/* ---- adjoint-method gradients (no counterpart in the reference) ----
 * For `batch` parameter sets params[b*nparams .. +nparams), evolve |0...0> by
 * ops, evaluate E_b = <psi|O|psi> with O the Z-string observable above, and
 * dE_b/dparams by one reverse sweep that un-applies each gate on psi and on
 * lambda = O psi and accumulates 2 Re<lambda|dU/dtheta|psi>. ckpt_every > 0
 * re-loads a forward checkpoint every ckpt_every differentiable gates to bound
 * rounding drift in deep circuits (0 = no checkpoints). energies: batch values;
 * grads: batch*nparams values. */
int qse_adjoint_expect(qse_ctx* ctx, int nqubits, int precision,
                       const qse_op* ops, int nops,
                       const double* params, int nparams, int batch,
                       const uint64_t* obs_masks, const double* obs_weights, int nterms,
                       int ckpt_every,
                       double* energies, double* grads);
/* How qse_adjoint_expect evaluates the terms of X and Y generators (RX, RY)
 * on states of 13 qubits or more (no counterpart in the reference):
 * QSE_ADJ_TERMS_ROTATE turns them into Z-strings by rotating psi and lambda
 * (H, or H S^dag, on their qubits, folded into the un-applications), which
 * costs passes where no un-application absorbs the rotation;
 * QSE_ADJ_TERMS_PAIRS sums amplitude pairs, Im<lambda|X_q|psi>, in read-only
 * passes over both states (one per 12 tile qubits); QSE_ADJ_TERMS_JOINT runs
 * the reverse sweep as joint passes over psi and lambda (run-time compiled,
 * 12-qubit tiles) that un-apply each gate
 * from both and sum every generator term in the tile at its point (needs
 * run-time compilation, not QSE_JIT_OFF, and no checkpoints);
 * QSE_ADJ_TERMS_AUTO (the default) takes the schedule that moves the fewest
 * state passes by the planner's count. The results agree to rounding. */
enum { QSE_ADJ_TERMS_AUTO = 0, QSE_ADJ_TERMS_ROTATE = 1, QSE_ADJ_TERMS_PAIRS = 2, QSE_ADJ_TERMS_JOINT = 3 };
int qse_ctx_set_adjoint_terms(qse_ctx* ctx, int mode);
/* Plain forward expectation for a batch of parameter sets (used by parameter
 * shift and finite differences). */
int qse_batch_expect(qse_ctx* ctx, int nqubits, int precision,
                     const qse_op* ops, int nops,
                     const double* params, int nparams, int batch,
                     const uint64_t* obs_masks, const double* obs_weights, int nterms,
                     double* energies);

/* ---- batched circuits of uneven sizes in persistent launches ----
 * (no counterpart in the reference). ncirc circuits; circuit c has nq[c]
 * qubits (8..30; its states live in HBM) and the ops ops[op_off[c] ..
 * op_off[c+1]), whose param indices refer to params[par_off[c] ..
 * par_off[c+1]). Each circuit starts from |0...0>, runs its ops, and returns
 * energies[c] = <psi|O|psi> with O = sum_t w_t Z-string(m_t) over the terms
 * whose qubits all lie below nq[c] (at most 64 terms, the same list for every
 * circuit); with grads != NULL also grads[par_off[c] + j] = dE_c/dparams[..+j]
 * by the adjoint method (parameterised kinds as qse_adjoint_expect). One
 * workgroup, or a team of workgroups for circuits of 16 or more qubits, runs
 * each circuit's forward passes, energy and observable (first launch) and its
 * reverse sweep (second launch, with gradients only): joint passes over psi
 * and lambda that un-apply the gates and sum each parameter's generator term
 * inside the tile. Circuits are taken largest first from device work queues.
 * Complex64 states accumulate energies and gradients in FP64; sums are formed
 * in a fixed order (bit-reproducible). Synchronises. */
int qse_batch_run_small(qse_ctx* ctx, int precision, int ncirc,
                        const int* nq, const int64_t* op_off, const qse_op* ops,
                        const int64_t* par_off, const double* params,
                        const uint64_t* obs_masks, const double* obs_weights, int nterms,
                        double* energies, double* grads);
/* Measured tile shape for qse_batch_run_small (no counterpart in the
 * reference): runs the given circuits once per candidate shape (tiles of 12,
 * 11 or 10 qubits, 2 to 5 contiguous low qubits, with teams of 16
 * workgroups), then with the best shape once per team size (4, 8, 32
 * workgroups per circuit of 16 or more qubits), compares the engine's kernel
 * times (HIP events; the context's timing totals are left as they were), and
 * keeps the fastest in the context for later batches of this precision (pass
 * a sample of a large batch). with_grads: time energies plus gradients
 * (par_off needed) or energies only. Outputs the chosen shape, team size and
 * kernel milliseconds (may be null). */
int qse_batch_autotune(qse_ctx* ctx, int precision, int ncirc, const int* nq, const int64_t* op_off,
                       const qse_op* ops, const int64_t* par_off, const double* params,
                       const uint64_t* obs_masks, const double* obs_weights, int nterms, int with_grads,
                       int* k_out, int* L_out, int* team_out, double* ms_out);

/* ---- batched Pauli-noise trajectories (Noise.cpp semantics) ----
 * For ntraj frames (xmask[t], zmask[t]) applied to the state, draw
 * nshots_per_traj outcomes per trajectory from u[t*nshots + s], with the same
 * lower_bound rule as qse_sample over the framed amplitudes
 * psi_t[i] = psi[i ^ xmask[t]]; out[t*nshots + s]. The state is not modified. */
int qse_sample_frames(qse_ctx* ctx, qse_state* st, const uint64_t* xmask, const uint64_t* zmask,
                      int ntraj, const double* u, int64_t nshots_per_traj, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* QSE_HIP_H */

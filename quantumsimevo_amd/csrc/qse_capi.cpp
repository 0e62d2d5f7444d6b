// qse_capi.cpp — the extern "C" boundary of libqse_hip.so (include/qse_hip.h):
// device probe, context (one device, one stream), state storage, single-gate
// sweeps, Pauli frames and readout. Fused execution lives in qse_fused.hip,
// gradients and batching in qse_adjoint.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qse_internal.hpp"

namespace qse {

int sample_frames(qse_ctx* ctx, qse_state* st, const uint64_t* xmask, int ntraj, const double* u, int64_t nshots,
                  uint64_t* out);
int reduce(qse_ctx* ctx, const qse_state* st, int mode, int q, const uint64_t* masks, const double* w, int nterms,
           double* out);
int launch_collapse(qse_ctx* ctx, qse_state* st, int q, int outcome, int reset, double nf);
int noalloc_init(int device);
int reduce_zstr_noalloc(qse_ctx* ctx, const qse_state* st, const uint64_t* masks, const double* w, int nterms,
                        double* out);
void jit_release_all(qse_ctx* ctx);

static thread_local std::string g_last_error;

int set_error(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int hip_error(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? QSE_ERR_OUT_OF_MEMORY : QSE_ERR_HIP;
}

int scratch(qse_ctx* ctx, int slot, size_t bytes, void** out) {
    Scratch& s = ctx->scratch[slot];
    if (s.bytes < bytes) {
        if (s.ptr) {
            QSE_HIP(hipStreamSynchronize(ctx->stream));
            QSE_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        const size_t want = std::max(bytes, (size_t)1 << 20);
        QSE_HIP(hipMalloc(&s.ptr, want));
        s.bytes = want;
    }
    *out = s.ptr;
    return QSE_OK;
}

int pinned(qse_ctx* ctx, size_t bytes, void** out) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) {
            QSE_HIP(hipStreamSynchronize(ctx->stream));
            QSE_HIP(hipHostFree(ctx->pinned));
            ctx->pinned = nullptr;
        }
        const size_t want = std::max(bytes, (size_t)4096);
        QSE_HIP(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_bytes = want;
    }
    *out = ctx->pinned;
    return QSE_OK;
}

static hipEvent_t get_event(qse_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void launch_begin(qse_ctx* ctx, int family, TimedLaunch* t) {
    t->family = family;
    t->a = t->b = nullptr;
    if (!ctx->timing) return;
    t->a = get_event(ctx);
    if (t->a) hipEventRecord(t->a, ctx->stream);
}

void launch_end(qse_ctx* ctx, TimedLaunch* t, double bytes) {
    if (!ctx->timing || !t->a) return;
    t->b = get_event(ctx);
    if (!t->b) return;
    hipEventRecord(t->b, ctx->stream);
    t->bytes = bytes;
    ctx->pending.push_back(*t);
    if (ctx->pending.size() > 4096) fold_timing(ctx);
}

int fold_timing(qse_ctx* ctx) {
    if (ctx->pending.empty()) return QSE_OK;
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& t : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            ctx->launches[t.family] += 1;
            ctx->total_ms[t.family] += ms;
            ctx->total_bytes[t.family] += t.bytes;
        }
        ctx->event_pool.push_back(t.a);
        ctx->event_pool.push_back(t.b);
    }
    ctx->pending.clear();
    return QSE_OK;
}

unsigned grid_for(const qse_ctx* ctx, uint64_t work, int block, int waves_per_cu) {
    const uint64_t need = (work + block - 1) / block;
    const uint64_t cap = (uint64_t)ctx->num_cus * waves_per_cu * 64 / block;
    return (unsigned)std::max<uint64_t>(1, std::min(need, cap));
}

}  // namespace qse

using namespace qse;

namespace {

bool valid_q(const qse_state* st, int q) { return st && q >= 0 && q < st->n; }

int check_state(qse_ctx* ctx, qse_state* st) {
    if (!ctx || !st || !st->d) return set_error(QSE_ERR_INVALID_ARG, "null context or state");
    return QSE_OK;
}

}  // namespace

extern "C" {

int qse_abi_version(void) { return QSE_ABI_VERSION; }

const char* qse_status_string(int status) {
    switch (status) {
        case QSE_OK: return "ok";
        case QSE_ERR_INVALID_ARG: return "invalid argument";
        case QSE_ERR_NO_DEVICE: return "no HIP device";
        case QSE_ERR_OUT_OF_MEMORY: return "device out of memory";
        case QSE_ERR_HIP: return "HIP runtime error";
        case QSE_ERR_UNSUPPORTED: return "unsupported";
        case QSE_ERR_LIMIT: return "limit exceeded";
        default: return "unknown status";
    }
}

const char* qse_last_error(void) { return g_last_error.c_str(); }

int qse_device_count(int* count) {
    if (!count) return set_error(QSE_ERR_INVALID_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return QSE_OK;
}

int qse_device_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
    int c = 0;
    qse_device_count(&c);
    if (device < 0 || device >= c) return set_error(QSE_ERR_NO_DEVICE, "device index out of range");
    QSE_HIP(hipSetDevice(device));
    size_t f = 0, t = 0;
    QSE_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return QSE_OK;
}

int qse_device_info(int device, char* name, size_t name_len, int* compute_units, size_t* total_bytes) {
    int c = 0;
    qse_device_count(&c);
    if (device < 0 || device >= c) return set_error(QSE_ERR_NO_DEVICE, "device index out of range");
    hipDeviceProp_t p;
    QSE_HIP(hipGetDeviceProperties(&p, device));
    if (name && name_len) {
        // the marketing name comes from the driver's ID table, which a container may
        // lack; the architecture name always identifies the part then
        const char* src = p.name[0] ? p.name : p.gcnArchName;
        std::strncpy(name, src, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (compute_units) *compute_units = p.multiProcessorCount;
    if (total_bytes) *total_bytes = p.totalGlobalMem;
    return QSE_OK;
}

int qse_ctx_create(int device, qse_ctx** out) {
    if (!out) return set_error(QSE_ERR_INVALID_ARG, "null out");
    int c = 0;
    qse_device_count(&c);
    if (c == 0) return set_error(QSE_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= c) return set_error(QSE_ERR_NO_DEVICE, "device index out of range");
    QSE_HIP(hipSetDevice(device));
    qse_ctx* ctx = new qse_ctx;
    ctx->device = device;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) == hipSuccess) {
        ctx->num_cus = p.multiProcessorCount;
        ctx->lds_per_block = p.sharedMemPerBlock;
    }
    // Scratch address of the allocation-free entry point on this device.
    int rc = noalloc_init(device);
    if (rc) {
        delete ctx;
        return rc;
    }
    // Default-priority stream (no priority variants).
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return hip_error(e, "hipStreamCreateWithFlags");
    }
    *out = ctx;
    return QSE_OK;
}

int qse_ctx_destroy(qse_ctx* ctx) {
    if (!ctx) return QSE_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    qse::jit_release_all(ctx);
    for (auto& t : ctx->pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    for (auto& s : ctx->scratch)
        if (s.ptr) hipFree(s.ptr);
    for (auto& e : ctx->prog_cache)
        if (e.mod) hipModuleUnload(e.mod);
    for (auto& c : ctx->ring.inflight) hipEventDestroy(c.done);
    for (auto e : ctx->ring.spare) hipEventDestroy(e);
    if (ctx->ring.dev) hipFree(ctx->ring.dev);
    if (ctx->ring.host) hipHostFree(ctx->ring.host);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return QSE_OK;
}

int qse_ctx_sync(qse_ctx* ctx) {
    if (!ctx) return set_error(QSE_ERR_INVALID_ARG, "null context");
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    return QSE_OK;
}

int qse_ctx_enable_timing(qse_ctx* ctx, int enable) {
    if (!ctx) return set_error(QSE_ERR_INVALID_ARG, "null context");
    if (!enable) fold_timing(ctx);
    ctx->timing = enable != 0;
    return QSE_OK;
}

int qse_ctx_timing(qse_ctx* ctx, int family, int64_t* launches, double* total_ms, double* bytes) {
    if (!ctx || family < 0 || family >= FAM_COUNT) return set_error(QSE_ERR_INVALID_ARG, "bad family");
    int rc = fold_timing(ctx);
    if (rc) return rc;
    if (launches) *launches = ctx->launches[family];
    if (total_ms) *total_ms = ctx->total_ms[family];
    if (bytes) *bytes = ctx->total_bytes[family];
    return QSE_OK;
}

int qse_ctx_reset_timing(qse_ctx* ctx) {
    if (!ctx) return set_error(QSE_ERR_INVALID_ARG, "null context");
    int rc = fold_timing(ctx);
    for (int f = 0; f < FAM_COUNT; ++f) {
        ctx->launches[f] = 0;
        ctx->total_ms[f] = 0;
        ctx->total_bytes[f] = 0;
    }
    return rc;
}

int qse_state_create(qse_ctx* ctx, int nqubits, int precision, qse_state** out) {
    if (!ctx || !out) return set_error(QSE_ERR_INVALID_ARG, "null argument");
    if (nqubits < 1 || nqubits > 40) return set_error(QSE_ERR_INVALID_ARG, "qubit count must be in [1, 40]");
    if (precision != QSE_C128 && precision != QSE_C64) return set_error(QSE_ERR_INVALID_ARG, "bad precision");
    QSE_HIP(hipSetDevice(ctx->device));
    qse_state* st = new qse_state;
    st->n = nqubits;
    st->prec = precision;
    st->bytes = (size_t)amp_bytes(precision) << nqubits;
    hipError_t e = hipMalloc(&st->d, st->bytes);
    if (e != hipSuccess) {
        delete st;
        (void)hipGetLastError();
        return set_error(QSE_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(1ULL << nqubits) +
                                                    " amplitudes failed: " + hipGetErrorString(e));
    }
    int rc = launch_set_basis(ctx, st, 0);
    if (rc) {
        hipFree(st->d);
        delete st;
        return rc;
    }
    *out = st;
    return QSE_OK;
}

int qse_state_destroy(qse_ctx* ctx, qse_state* st) {
    if (!st) return QSE_OK;
    if (ctx) hipStreamSynchronize(ctx->stream);
    if (st->owned && st->d) hipFree(st->d);
    if (st->shadow) hipFree(st->shadow);
    delete st;
    return QSE_OK;
}

int qse_state_info(const qse_state* st, int* nqubits, int* precision, void** device_ptr, size_t* bytes) {
    if (!st) return set_error(QSE_ERR_INVALID_ARG, "null state");
    if (nqubits) *nqubits = st->n;
    if (precision) *precision = st->prec;
    if (device_ptr) *device_ptr = st->d;
    if (bytes) *bytes = st->bytes;
    return QSE_OK;
}

int qse_state_download_range(qse_ctx* ctx, qse_state* st, uint64_t first, uint64_t count, double* host) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!host || first + count > (1ULL << st->n) || first + count < first)
        return set_error(QSE_ERR_INVALID_ARG, "range outside the state");
    if (st->prec == QSE_C128) {
        QSE_HIP(hipMemcpyAsync(host, (char*)st->d + 16 * first, 16 * count, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        // Convert in device chunks of at most 2^24 amplitudes.
        const uint64_t CH = 1ULL << 24;
        void* buf;
        rc = scratch(ctx, 3, 16 * std::min(CH, count), &buf);
        if (rc) return rc;
        for (uint64_t o = 0; o < count; o += CH) {
            const uint64_t c = std::min(CH, count - o);
            rc = launch_convert_download(ctx, st, (double*)buf, c, first + o);
            if (rc) return rc;
            QSE_HIP(hipMemcpyAsync(host + 2 * o, buf, 16 * c, hipMemcpyDeviceToHost, ctx->stream));
            QSE_HIP(hipStreamSynchronize(ctx->stream));
        }
    }
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    return QSE_OK;
}

int qse_state_download(qse_ctx* ctx, qse_state* st, double* host) {
    if (!st) return set_error(QSE_ERR_INVALID_ARG, "null state");
    return qse_state_download_range(ctx, st, 0, 1ULL << st->n, host);
}

int qse_state_upload(qse_ctx* ctx, qse_state* st, const double* host) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!host) return set_error(QSE_ERR_INVALID_ARG, "null host buffer");
    const uint64_t N = 1ULL << st->n;
    if (st->prec == QSE_C128) {
        QSE_HIP(hipMemcpyAsync(st->d, host, 16 * N, hipMemcpyHostToDevice, ctx->stream));
        QSE_HIP(hipStreamSynchronize(ctx->stream));
        return QSE_OK;
    }
    const uint64_t CH = 1ULL << 24;
    void* buf;
    rc = scratch(ctx, 3, 16 * std::min(CH, N), &buf);
    if (rc) return rc;
    for (uint64_t o = 0; o < N; o += CH) {
        const uint64_t c = std::min(CH, N - o);
        QSE_HIP(hipMemcpyAsync(buf, host + 2 * o, 16 * c, hipMemcpyHostToDevice, ctx->stream));
        rc = launch_convert_upload(ctx, st, (const double*)buf, c, o);
        if (rc) return rc;
        QSE_HIP(hipStreamSynchronize(ctx->stream));
    }
    return QSE_OK;
}

int qse_state_set_basis(qse_ctx* ctx, qse_state* st, uint64_t index) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (index >= (1ULL << st->n)) return set_error(QSE_ERR_INVALID_ARG, "basis index outside the state");
    return launch_set_basis(ctx, st, index);
}

int qse_state_set_random(qse_ctx* ctx, qse_state* st, uint64_t seed) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    rc = launch_random(ctx, st, seed);
    if (rc) return rc;
    double nrm = 0.0;
    rc = reduce(ctx, st, 0, 0, nullptr, nullptr, 0, &nrm);
    if (rc) return rc;
    return launch_scale(ctx, st, 1.0 / std::sqrt(nrm));
}

int qse_state_copy(qse_ctx* ctx, qse_state* dst, const qse_state* src) {
    if (!ctx || !dst || !src) return set_error(QSE_ERR_INVALID_ARG, "null argument");
    if (dst->n != src->n || dst->prec != src->prec) return set_error(QSE_ERR_INVALID_ARG, "shape mismatch");
    QSE_HIP(hipMemcpyAsync(dst->d, src->d, src->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return QSE_OK;
}

// ---- single gates (CircuitUnitaryOperation.cpp:10-338 semantics) ----
// Algorithmic bytes: 2*b*T(g) with T(g) the amplitudes the gate can change.

int qse_apply_matrix(qse_ctx* ctx, qse_state* st, int q, const double m[8]) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q) || !m) return set_error(QSE_ERR_INVALID_ARG, "bad qubit or matrix");
    const uint64_t f0 = 1ULL << q;
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    return launch_pairs(ctx, st, PAIR_DENSE, &f0, 1, 0, 0, 1ULL << q, m, 2 * b * N);
}

int qse_apply_hadamard(qse_ctx* ctx, qse_state* st, int q) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    const double s = 0.7071067811865475244;  // CircuitUnitaryOperation.cpp:86
    const double m[8] = {s, 0, s, 0, s, 0, -s, 0};
    const uint64_t f0 = 1ULL << q;
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    return launch_pairs(ctx, st, PAIR_HADAMARD, &f0, 1, 0, 0, 1ULL << q, m, 2 * b * N);
}

int qse_apply_x(qse_ctx* ctx, qse_state* st, int q) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    const uint64_t f0 = 1ULL << q;
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    const double z[8] = {0};
    return launch_pairs(ctx, st, PAIR_SWAP, &f0, 1, 0, 0, 1ULL << q, z, 2 * b * N);
}

int qse_apply_y(qse_ctx* ctx, qse_state* st, int q) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    // (v0, v1) -> (-i v1, i v0) (CircuitUnitaryOperation.cpp:151-152)
    const double m[8] = {0, 0, 0, -1, 0, 1, 0, 0};
    const uint64_t f0 = 1ULL << q;
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    return launch_pairs(ctx, st, PAIR_DENSE, &f0, 1, 0, 0, 1ULL << q, m, 2 * b * N);
}

static int apply_diag1(qse_ctx* ctx, qse_state* st, uint64_t ones, double re, double im) {
    const double m[8] = {1, 0, 0, 0, 0, 0, re, im};
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    const int m_bits = __builtin_popcountll(ones);
    return launch_pairs(ctx, st, PAIR_DIAG1, nullptr, 0, ones, 0, 0, m, 2 * b * std::ldexp(N, -m_bits));
}

int qse_apply_z(qse_ctx* ctx, qse_state* st, int q) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    return apply_diag1(ctx, st, 1ULL << q, -1.0, 0.0);
}

int qse_apply_phase(qse_ctx* ctx, qse_state* st, int q, double theta) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    return apply_diag1(ctx, st, 1ULL << q, std::cos(theta), std::sin(theta));
}

int qse_apply_rx(qse_ctx* ctx, qse_state* st, int q, double theta) {
    const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
    const double m[8] = {c, 0, 0, -s, 0, -s, c, 0};
    return qse_apply_matrix(ctx, st, q, m);
}

int qse_apply_ry(qse_ctx* ctx, qse_state* st, int q, double theta) {
    const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
    const double m[8] = {c, 0, -s, 0, s, 0, c, 0};
    return qse_apply_matrix(ctx, st, q, m);
}

int qse_apply_rz(qse_ctx* ctx, qse_state* st, int q, double theta) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
    const double m[8] = {c, -s, 0, 0, 0, 0, c, s};
    const uint64_t f0 = 1ULL << q;
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    return launch_pairs(ctx, st, PAIR_DIAG, &f0, 1, 0, 0, 1ULL << q, m, 2 * b * N);
}

int qse_apply_cnot(qse_ctx* ctx, qse_state* st, int control, int target) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, control) || !valid_q(st, target)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    if (control == target) {
        // Reference loop (CircuitUnitaryOperation.cpp:46-53): with c == t the
        // pair guard i < i^t fails whenever the control bit is set: no-op.
        return QSE_OK;
    }
    const uint64_t f0 = 1ULL << target;
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    const double z[8] = {0};
    return launch_pairs(ctx, st, PAIR_SWAP, &f0, 1, 1ULL << control, 0, 1ULL << target, z, b * N);
}

int qse_apply_cphase(qse_ctx* ctx, qse_state* st, int control, int target, double theta) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, control) || !valid_q(st, target)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    return apply_diag1(ctx, st, (1ULL << control) | (1ULL << target), std::cos(theta), std::sin(theta));
}

int qse_apply_mcphase(qse_ctx* ctx, qse_state* st, const int* controls, int ncontrols, int target, double theta) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, target) || ncontrols < 0 || (ncontrols > 0 && !controls))
        return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    uint64_t mask = 1ULL << target;
    for (int i = 0; i < ncontrols; ++i) {
        if (!valid_q(st, controls[i])) return set_error(QSE_ERR_INVALID_ARG, "bad control qubit");
        mask |= 1ULL << controls[i];
    }
    // std::exp(i theta) (CircuitUnitaryOperation.cpp:298)
    return apply_diag1(ctx, st, mask, std::cos(theta), std::sin(theta));
}

int qse_apply_swap(qse_ctx* ctx, qse_state* st, int q0, int q1) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q0) || !valid_q(st, q1)) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    if (q0 == q1) return QSE_OK;  // self-swap is the identity
    const uint64_t f0[2] = {1ULL << q0, 1ULL << q1};
    const double N = std::ldexp(1.0, st->n), b = amp_bytes(st->prec);
    const double z[8] = {0};
    return launch_pairs(ctx, st, PAIR_SWAP, f0, 2, 0, 1ULL << q0, 1ULL << q1, z, b * N);
}

int qse_pauli_frame(qse_ctx* ctx, qse_state* st, uint64_t xmask, uint64_t zmask) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    const uint64_t lim = (st->n >= 64) ? ~0ULL : ((1ULL << st->n) - 1);
    if ((xmask | zmask) & ~lim) return set_error(QSE_ERR_INVALID_ARG, "frame mask outside the register");
    if (!xmask && !zmask) return QSE_OK;
    return launch_frame(ctx, st, xmask, zmask);
}

// ---- readout ----

int qse_sample(qse_ctx* ctx, qse_state* st, const double* u, int64_t nshots, uint64_t* out) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (nshots < 0 || (nshots > 0 && (!u || !out))) return set_error(QSE_ERR_INVALID_ARG, "bad shot buffers");
    if (nshots == 0) return QSE_OK;
    return sample_frames(ctx, st, nullptr, 1, u, nshots, out);
}

int qse_sample_frames(qse_ctx* ctx, qse_state* st, const uint64_t* xmask, const uint64_t* zmask, int ntraj,
                      const double* u, int64_t nshots_per_traj, uint64_t* out) {
    (void)zmask;  // phase flips do not change outcome probabilities
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (ntraj < 1 || ntraj > 65535 || nshots_per_traj < 0 || !xmask || (nshots_per_traj > 0 && (!u || !out)))
        return set_error(QSE_ERR_INVALID_ARG, "bad trajectory arguments");
    const uint64_t lim = (1ULL << st->n) - 1;
    for (int t = 0; t < ntraj; ++t)
        if (xmask[t] & ~lim) return set_error(QSE_ERR_INVALID_ARG, "frame mask outside the register");
    if (nshots_per_traj == 0) return QSE_OK;
    return sample_frames(ctx, st, xmask, ntraj, u, nshots_per_traj, out);
}

int qse_measure(qse_ctx* ctx, qse_state* st, double u, uint64_t* outcome) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!outcome) return set_error(QSE_ERR_INVALID_ARG, "null outcome");
    uint64_t idx = 0;
    rc = sample_frames(ctx, st, nullptr, 1, &u, 1, &idx);
    if (rc) return rc;
    // StateVector.cpp:142-149: zero everything, then [idx] = 1.
    rc = launch_set_basis(ctx, st, idx);
    if (rc) return rc;
    *outcome = idx;
    return QSE_OK;
}

int qse_norm2(qse_ctx* ctx, qse_state* st, double* out) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!out) return set_error(QSE_ERR_INVALID_ARG, "null out");
    return reduce(ctx, st, 0, 0, nullptr, nullptr, 0, out);
}

int qse_state_renormalize(qse_ctx* ctx, qse_state* st, double* norm2_before) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    double* d = nullptr;
    rc = renormalize_async(ctx, st, &d);
    if (rc || !norm2_before) return rc;
    void* pin;
    rc = pinned(ctx, sizeof(double), &pin);
    if (rc) return rc;
    QSE_HIP(hipMemcpyAsync(pin, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QSE_HIP(hipStreamSynchronize(ctx->stream));
    *norm2_before = *(double*)pin;
    return QSE_OK;
}

int qse_ctx_set_renormalize(qse_ctx* ctx, int every_passes) {
    if (!ctx || every_passes < 0) return set_error(QSE_ERR_INVALID_ARG, "bad renormalisation period");
    ctx->renorm_every = every_passes;
    return QSE_OK;
}

int qse_prob_one(qse_ctx* ctx, qse_state* st, int q, double* out) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (!valid_q(st, q) || !out) return set_error(QSE_ERR_INVALID_ARG, "bad qubit");
    return reduce(ctx, st, 1, q, nullptr, nullptr, 0, out);
}

int qse_collapse_qubit(qse_ctx* ctx, qse_state* st, int q, double u, int reset, int* outcome) {
    double p1 = 0.0;
    int rc = qse_prob_one(ctx, st, q, &p1);
    if (rc) return rc;
    const int o = (u < p1) ? 1 : 0;
    const double nf = std::sqrt(o ? p1 : 1.0 - p1);
    if (!(nf > 0.0)) return set_error(QSE_ERR_INVALID_ARG, "collapse onto a zero-probability outcome");
    rc = launch_collapse(ctx, st, q, o, reset, nf);
    if (rc) return rc;
    if (outcome) *outcome = o;
    return QSE_OK;
}

int qse_expect_zstrings(qse_ctx* ctx, qse_state* st, const uint64_t* masks, const double* weights, int nterms,
                        double* out) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (nterms < 0 || (nterms > 0 && (!masks || !weights)) || !out)
        return set_error(QSE_ERR_INVALID_ARG, "bad observable");
    return reduce(ctx, st, 2, 0, masks, weights, nterms, out);
}

int qse_expect_zstrings_noalloc(qse_ctx* ctx, qse_state* st, const uint64_t* masks, const double* weights, int nterms,
                                double* out) {
    int rc = check_state(ctx, st);
    if (rc) return rc;
    if (nterms < 0 || (nterms > 0 && (!masks || !weights)) || !out)
        return set_error(QSE_ERR_INVALID_ARG, "bad observable");
    return reduce_zstr_noalloc(ctx, st, masks, weights, nterms, out);
}

}  // extern "C"

// qse_jit.cpp — run-time specialisation of fused passes with hiprtc.
//
// The generic fused kernel (qse_fused.hip) interprets a pass program: every op
// costs a loop iteration, scalar loads of its fields, a switch on its register
// position and selects on its controls, and every MONO op walks its thread-bit
// factors and terms. When a program runs again (the same circuit on every step
// of a simulation loop), its passes are emitted as straight-line HIP source
// (qse_jit_gen.cpp), compiled as one module for gfx950 with hiprtc here, and
// the module is kept next to the program's device copy (qse_fused.hip drops
// both together).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "qse_fused.hpp"
#include "qse_internal.hpp"

namespace qse {

std::string jit_kernel_name(int pass, bool table);

// Compile `src` for gfx950 and load it; one function per pass, named by `name`.
int jit_compile_named(const std::string& src, int npasses, std::string (*name)(int, bool), bool table,
                      hipModule_t* mod, std::vector<hipFunction_t>& fns) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "qse_jit_program.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return set_error(QSE_ERR_HIP, "hiprtcCreateProgram failed");
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, 6, opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return set_error(QSE_ERR_HIP, std::string("hiprtc compile failed: ") + hiprtcGetErrorString(r) + ": " +
                                          log.substr(0, 2000));
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    QSE_HIP(hipModuleLoadData(mod, code.data()));
    fns.assign(npasses, nullptr);
    for (int p = 0; p < npasses; ++p) {
        const hipError_t e = hipModuleGetFunction(&fns[p], *mod, name(p, table).c_str());
        if (e != hipSuccess) {
            hipModuleUnload(*mod);
            *mod = nullptr;
            fns.clear();
            return hip_error(e, "hipModuleGetFunction");
        }
    }
    return QSE_OK;
}

// One function per pass, named per jit_kernel_name.
int jit_compile(const std::string& src, int npasses, bool table, hipModule_t* mod, std::vector<hipFunction_t>& fns) {
    return jit_compile_named(src, npasses, jit_kernel_name, table, mod, fns);
}

// Tile-index bits from which complex128 passes are first generated with
// paired tiles: 4096 pairs or more (at 12 bits, the QAOA-24 adjoint's passes
// ran 7% slower paired than unpaired).
constexpr int JIT_PAIR_MIN_NEXT = 13;

// Compile plan P's passes (coefficients from the program at toff when given).
// Complex128 passes with enough tiles are generated with paired tiles; a pass
// whose paired kernel spills more than a few registers to scratch memory (its
// two tiles' amplitudes do not fit the registers of two waves per SIMD) is
// generated again unpaired.
constexpr int JIT_PAIR_MAX_SCRATCH = 64;  // bytes per lane
int jit_build(const HostPlan& P, int prec, const size_t* toff, hipModule_t* mod, std::vector<hipFunction_t>& fns) {
    const int np = (int)P.passes.size();
    std::vector<bool> paired(np, false);
    bool any = false;
    for (int p = 0; p < np && prec == QSE_C128; ++p)
        if (P.passes[p].next >= JIT_PAIR_MIN_NEXT) paired[p] = any = true;
    int rc = jit_compile(jit_source(P, prec, toff, &paired), np, toff != nullptr, mod, fns);
    if (rc || !any) return rc;
    bool redo = false;
    for (int p = 0; p < np; ++p) {
        if (!paired[p]) continue;
        int local = 0;
        QSE_HIP(hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fns[p]));
        if (local > JIT_PAIR_MAX_SCRATCH) {
            paired[p] = false;
            redo = true;
        }
    }
    if (!redo) return QSE_OK;
    QSE_HIP(hipModuleUnload(*mod));
    *mod = nullptr;
    fns.clear();
    return jit_compile(jit_source(P, prec, toff, &paired), np, toff != nullptr, mod, fns);
}

int jit_launch(qse_ctx* ctx, hipFunction_t fn, const void* src, void* dst, const void* table, uint64_t ntiles, int K,
               const void* prog) {
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 16);
    unsigned long long nt = ntiles;
    void* args[] = {(void*)&src, (void*)&dst, (void*)&table, (void*)&nt, (void*)&prog};
    QSE_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, 1u << (K - 4), 1, 1, 0, ctx->stream, args, nullptr));
    return QSE_OK;
}

namespace {

constexpr size_t STRUCT_CACHE_MAX = 64;
constexpr int MAX_COMPILE_WORKERS = 4;

void start_compile(qse_ctx* ctx, StructEntry& e, const HostPlan& P, const Serialized& S, int prec) {
    auto job = std::make_shared<JitJob>();
    const int dev = ctx->device;
    JitJob* j = job.get();
    std::vector<size_t> off(S.off, S.off + 7);
    j->worker = std::thread([j, P, off = std::move(off), prec, dev] {
        hipSetDevice(dev);
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = jit_build(P, prec, off.data(), &j->mod, j->fns);
        j->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        j->state.store(rc ? 2 : 1, std::memory_order_release);
    });
    e.job = std::move(job);
}

int running_compiles(const qse_ctx* ctx) {
    int r = 0;
    for (const auto& e : ctx->struct_cache)
        if (e.job && e.job->state.load(std::memory_order_acquire) == 0) ++r;
    return r;
}

// Wait for the entry's worker and release its module (kernels from it may be queued).
int drop_entry(qse_ctx* ctx, StructEntry& e) {
    if (!e.job) return QSE_OK;
    if (e.job->worker.joinable()) e.job->worker.join();
    if (e.job->mod) {
        QSE_HIP(hipStreamSynchronize(ctx->stream));
        QSE_HIP(hipModuleUnload(e.job->mod));
        e.job->mod = nullptr;
    }
    return QSE_OK;
}

}  // namespace

int jit_for_structure(qse_ctx* ctx, const HostPlan& P, const Serialized& S, int prec,
                      const std::vector<hipFunction_t>** fns) {
    *fns = nullptr;
    if (ctx->jit_mode == QSE_JIT_OFF || P.passes.empty()) return QSE_OK;
    std::vector<char> key = structure_key(P);
    StructEntry* hit = nullptr;
    for (auto& e : ctx->struct_cache)
        if (e.prec == prec && e.key.size() == key.size() && std::memcmp(e.key.data(), key.data(), key.size()) == 0) {
            hit = &e;
            break;
        }
    if (!hit) {
        if (ctx->struct_cache.size() >= STRUCT_CACHE_MAX) {
            // least recently used entry whose compilation is not running (waiting for a
            // running one would stall the caller for the compile time); when every
            // entry is compiling, this structure is not cached
            size_t oldest = ctx->struct_cache.size();
            for (size_t i = 0; i < ctx->struct_cache.size(); ++i) {
                const StructEntry& c = ctx->struct_cache[i];
                if (c.job && c.job->state.load(std::memory_order_acquire) == 0) continue;
                if (oldest == ctx->struct_cache.size() || c.stamp < ctx->struct_cache[oldest].stamp) oldest = i;
            }
            if (oldest == ctx->struct_cache.size()) return QSE_OK;
            const int rc = drop_entry(ctx, ctx->struct_cache[oldest]);
            if (rc) return rc;
            ctx->struct_cache.erase(ctx->struct_cache.begin() + oldest);
        }
        StructEntry e;
        e.key = std::move(key);
        e.prec = prec;
        ctx->struct_cache.push_back(std::move(e));
        hit = &ctx->struct_cache.back();
    }
    hit->stamp = ++ctx->prog_clock;
    ++hit->seen;
    if (!hit->job) {
        if (ctx->jit_mode == QSE_JIT_ALWAYS) {
            start_compile(ctx, *hit, P, S, prec);
            hit->job->worker.join();
            if (hit->job->state.load() != 1)
                return set_error(QSE_ERR_HIP, "run-time compilation of a pass structure failed");
        } else if (hit->seen >= 2 && running_compiles(ctx) < MAX_COMPILE_WORKERS) {
            start_compile(ctx, *hit, P, S, prec);
        }
    }
    if (hit->job && hit->job->state.load(std::memory_order_acquire) == 1) {
        if (!hit->counted) {
            hit->counted = true;
            ++ctx->jit_compiled;
            ctx->jit_compile_ms += hit->job->ms;
        }
        *fns = &hit->job->fns;
    }
    return QSE_OK;
}

// Joint reverse-sweep kernels of plan P (the device program S), compiled on
// the first sight of its structure and kept (at most JOINT_CACHE_MAX
// structures, least recently used out).
constexpr size_t JOINT_CACHE_MAX = 16;
int jit_joint_for(qse_ctx* ctx, const HostPlan& P, const Serialized& S, int prec,
                  const std::vector<hipFunction_t>** fns) {
    *fns = nullptr;
    std::vector<char> key = structure_key(P);
    for (auto& e : ctx->joint_cache)
        if (e.prec == prec && e.key == key) {
            e.stamp = ++ctx->prog_clock;
            *fns = &e.fns;
            return QSE_OK;
        }
    if (ctx->joint_cache.size() >= JOINT_CACHE_MAX) {
        size_t oldest = 0;
        for (size_t i = 1; i < ctx->joint_cache.size(); ++i)
            if (ctx->joint_cache[i].stamp < ctx->joint_cache[oldest].stamp) oldest = i;
        QSE_HIP(hipStreamSynchronize(ctx->stream));  // its kernels may still be queued
        if (ctx->joint_cache[oldest].mod) QSE_HIP(hipModuleUnload(ctx->joint_cache[oldest].mod));
        ctx->joint_cache.erase(ctx->joint_cache.begin() + oldest);
    }
    JointEntry e;
    e.key = std::move(key);
    e.prec = prec;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = jit_compile_named(jit_joint_source(P, prec, S.off), (int)P.passes.size(),
                                     [](int p, bool) { return jit_joint_kernel_name(p); }, true, &e.mod, e.fns);
    if (rc) return rc;
    ++ctx->jit_compiled;
    ctx->jit_compile_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e.stamp = ++ctx->prog_clock;
    ctx->joint_cache.push_back(std::move(e));
    *fns = &ctx->joint_cache.back().fns;
    return QSE_OK;
}

// Wait for every background compilation of the context.
int jit_wait_all(qse_ctx* ctx) {
    for (auto& e : ctx->struct_cache)
        if (e.job && e.job->worker.joinable()) e.job->worker.join();
    return QSE_OK;
}

// Release every structure-keyed module (context teardown).
void jit_release_all(qse_ctx* ctx) {
    for (auto& e : ctx->struct_cache) {
        if (!e.job) continue;
        if (e.job->worker.joinable()) e.job->worker.join();
        if (e.job->mod) hipModuleUnload(e.job->mod);
        e.job->mod = nullptr;
    }
    ctx->struct_cache.clear();
    for (auto& e : ctx->joint_cache)
        if (e.mod) hipModuleUnload(e.mod);
    ctx->joint_cache.clear();
}

}  // namespace qse

extern "C" int qse_ctx_jit_wait(qse_ctx* ctx) {
    if (!ctx) return qse::set_error(QSE_ERR_INVALID_ARG, "null context");
    return qse::jit_wait_all(ctx);
}

extern "C" int qse_ctx_set_jit(qse_ctx* ctx, int mode) {
    if (!ctx || mode < QSE_JIT_OFF || mode > QSE_JIT_ALWAYS)
        return qse::set_error(QSE_ERR_INVALID_ARG, "jit mode must be QSE_JIT_OFF, QSE_JIT_AUTO or QSE_JIT_ALWAYS");
    ctx->jit_mode = mode;
    return QSE_OK;
}

extern "C" int qse_ctx_jit_stats(qse_ctx* ctx, int64_t* compiled, int64_t* jit_launches, double* compile_ms) {
    if (!ctx) return qse::set_error(QSE_ERR_INVALID_ARG, "null context");
    if (compiled) *compiled = ctx->jit_compiled;
    if (jit_launches) *jit_launches = ctx->jit_launches;
    if (compile_ms) *compile_ms = ctx->jit_compile_ms;
    return QSE_OK;
}

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
#include <string>
#include <vector>

#include "qse_fused.hpp"
#include "qse_internal.hpp"

namespace qse {

std::string jit_kernel_name(int pass);

// Compile `src` for gfx950 and load it; one function per pass.
int jit_compile(const std::string& src, int npasses, hipModule_t* mod, std::vector<hipFunction_t>& fns) {
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
        const hipError_t e = hipModuleGetFunction(&fns[p], *mod, jit_kernel_name(p).c_str());
        if (e != hipSuccess) {
            hipModuleUnload(*mod);
            *mod = nullptr;
            fns.clear();
            return hip_error(e, "hipModuleGetFunction");
        }
    }
    return QSE_OK;
}

int jit_launch(qse_ctx* ctx, hipFunction_t fn, const void* src, void* dst, const void* table, uint64_t ntiles, int K) {
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 16);
    unsigned long long nt = ntiles;
    void* args[] = {(void*)&src, (void*)&dst, (void*)&table, (void*)&nt};
    QSE_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, 1u << (K - 4), 1, 1, 0, ctx->stream, args, nullptr));
    return QSE_OK;
}

}  // namespace qse

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

// Launch of one generated joint pass (psi and lambda, probe partials) on the
// host: one workgroup of `nthreads` std::threads, as shim_launch.
thread_local shim_dim3 threadIdx;
shim_dim3 blockIdx, gridDim;
static pthread_barrier_t shim_bar;
void shim_syncthreads() { pthread_barrier_wait(&shim_bar); }
static double shim_xor_buf[4096];
double __shfl_xor(double v, int mask, int) {
    const unsigned t = threadIdx.x;
    shim_xor_buf[t] = v;
    shim_syncthreads();
    const double r = shim_xor_buf[t ^ (unsigned)mask];
    shim_syncthreads();
    return r;
}

double qse_wsum64(double v) {
    const unsigned t = threadIdx.x;
    shim_xor_buf[t] = v;
    shim_syncthreads();
    double s = 0.0;
    for (unsigned i = t & ~63u; i < (t & ~63u) + 64u; ++i) s += shim_xor_buf[i];
    shim_syncthreads();
    return s;
}

typedef void (*shim_joint_kernel)(cx*, cx*, const SlotVal*, unsigned long long, const double*, double*);
extern shim_joint_kernel shim_joint_kernels[];

extern "C" int shim_joint_launch(int pass, int nthreads, void* psi, void* lam, const void* table,
                                 unsigned long long ntiles, const void* prog, void* part) {
    pthread_barrier_init(&shim_bar, nullptr, (unsigned)nthreads);
    blockIdx = {0, 0, 0};
    gridDim = {1, 1, 1};
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([=] {
            threadIdx = {(unsigned)t, 0, 0};
            shim_joint_kernels[pass]((cx*)psi, (cx*)lam, (const SlotVal*)table, ntiles, (const double*)prog,
                                     (double*)part);
        });
    for (auto& x : th) x.join();
    pthread_barrier_destroy(&shim_bar);
    return 0;
}

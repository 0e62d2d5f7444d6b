// Launch of one generated pass on the host: one workgroup (gridDim.x = 1, so
// the kernel's tile loop visits every tile) of `nthreads` std::threads.
thread_local shim_dim3 threadIdx;
shim_dim3 blockIdx, gridDim;
static pthread_barrier_t shim_bar;
void shim_syncthreads() { pthread_barrier_wait(&shim_bar); }

typedef void (*shim_kernel)(const cx*, cx*, const SlotVal*, unsigned long long, const double*);
extern shim_kernel shim_kernels[];

extern "C" int shim_launch(int pass, int nthreads, const void* src, void* dst, const void* table,
                           unsigned long long ntiles, const void* prog) {
    pthread_barrier_init(&shim_bar, nullptr, (unsigned)nthreads);
    blockIdx = {0, 0, 0};
    gridDim = {1, 1, 1};
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([=] {
            threadIdx = {(unsigned)t, 0, 0};
            shim_kernels[pass]((const cx*)src, (cx*)dst, (const SlotVal*)table, ntiles, (const double*)prog);
        });
    for (auto& x : th) x.join();
    pthread_barrier_destroy(&shim_bar);
    return 0;
}

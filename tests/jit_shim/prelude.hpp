// Host shim for run-time generated pass kernels (test infrastructure): the
// generated HIP source is compiled with g++ after this prelude, and each
// workgroup thread runs as a std::thread; __syncthreads is a barrier over them
// and __shared__ arrays are function statics shared by those threads.
#include <pthread.h>

#include <thread>
#include <vector>

struct shim_dim3 {
    unsigned x, y, z;
};
extern thread_local shim_dim3 threadIdx;
extern shim_dim3 blockIdx, gridDim;
void shim_syncthreads();

#define __global__
#define __device__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(x)
#define __syncthreads() shim_syncthreads()
// Non-temporal state accesses become plain ones; the two-component vector type
// of the generated code becomes a GCC vector (same [i] subscripts).
#define ext_vector_type(n) vector_size((n) * sizeof(R))
// The generated code's register barrier (keeps the thread id from being
// treated as loop-invariant) in its host form.
#define QSE_OPAQUE(v) __asm__ volatile("" : "+r"(v))
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
#define __popc(x) __builtin_popcount(x)
#define __popcll(x) __builtin_popcountll(x)
// Cross-lane exchange of the joint passes' probe sums (defined with the joint
// launch): lanes are the workgroup's threads, exchanged through a buffer
// between two barriers.
double __shfl_xor(double v, int mask, int width);
// The joint passes' wave sum (DPP on the device): here the sum over the
// thread's group of 64 threads, in every thread.
#define QSE_HOST_SHIM 1
double qse_wsum64(double v);

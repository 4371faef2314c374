// Device helpers shared by the kernel files: value loaders, coherent
// vector accesses and the one-pass deterministic block reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bcm {
namespace {

// Upper bound on the polls of one dependency wait of the sync-free sweep.
constexpr int kSpinLimit = 1 << 22;

// ---- value loaders ----
__device__ __forceinline__ double ldv(const double* p) { return *p; }
__device__ __forceinline__ float ldv(const float* p) { return *p; }
__device__ __forceinline__ float ldv(const uint16_t* p) {
  return __uint_as_float(((uint32_t)*p) << 16);
}
// Streamed (read-once) matrix values: non-temporal loads keep the vectors in cache.
__device__ __forceinline__ double ldnt(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ldnt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ldnt(const uint16_t* p) {
  return __uint_as_float(((uint32_t)__builtin_nontemporal_load(p)) << 16);
}

template <class T> struct AccOf { using type = double; };
template <> struct AccOf<float> { using type = float; };

// Coherent (device-scope) accesses of vectors written inside the same sweep.
template <class T> __device__ __forceinline__ T ld_coherent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void st_coherent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-level reduction of K values followed by the last-block finalisation:
// block sums go to partials[block*K + k]; the block that completes last adds
// them in a fixed order (deterministic) into out[k] and re-arms the counter.
// Orders this wave's LDS writes before its following LDS reads by other
// lanes (and reads before later writes). LDS operations of one wave complete
// in program order, so a wavefront-scope fence, which only stops the compiler
// from moving memory operations across it, is enough.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int K>
__device__ void reduce_finalize(double (&v)[K], double* partials, unsigned* counter, double* out) {
  __shared__ double sm[K][16];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane == 0) sm[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sm[k][w];
      partials[(size_t)blockIdx.x * K + k] = s;
    }
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    is_last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  const int nb = gridDim.x;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
      s += __hip_atomic_load(&partials[(size_t)i * K + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    __syncthreads();
    if (lane == 0) sm[k][wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += sm[k][w];
      out[k] = t;
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
}

}  // namespace
}  // namespace bcm

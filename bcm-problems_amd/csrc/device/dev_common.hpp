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

// Orders this wave's LDS writes before its following LDS reads by other
// lanes (and reads before later writes). LDS operations of one wave complete
// in program order, so a wavefront-scope fence, which only stops the compiler
// from moving memory operations across it, is enough.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Block sums without any cross-block synchronisation: block b writes
// partials[b * K + k]. A later kernel on the same stream adds them with
// sum_partials (the kernel boundary makes the stores visible on every XCD),
// so no block pays for a device-scope fence.
template <int K>
__device__ void block_partials(const double (&v)[K], double* partials) {
  __shared__ double bp_sm[K][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane == 0) bp_sm[k][wave] = x;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += bp_sm[k][w];
    partials[(size_t)blockIdx.x * K + k] = s;
  }
}

// Sum over nb blocks of value k of per-block partials with stride K, in a
// fixed order; called by every thread of one block, returns the sum to all.
__device__ __forceinline__ double sum_partials(const double* partials, int nb, int K, int k) {
  __shared__ double sp_sm[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partials[(size_t)i * K + k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  __syncthreads();
  if (lane == 0) sp_sm[wave] = s;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += sp_sm[w];
  return t;
}


}  // namespace
}  // namespace bcm

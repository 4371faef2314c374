// Preconditioner application that allocates nothing: z = M^-1 r (or M^-T r)
// with the vector between the two sweeps held in a fixed device array,
// for callers that cannot allocate device memory (code under stream capture,
// tight loops that forbid hipMalloc).
//
// The array `scratch` below is one per device, sized once for the process.
// Host code never uses its name as an address: the device address comes from
// hipGetSymbolAddress, looked up once per device with that device current and
// kept in g_scratch[device]. Every call on one device shares that array, so
// calls that overlap on one device (from any stream or context) race.
#include <mutex>
#include <string>

#include "solver.hpp"

namespace bcm {

namespace {

// 256 bytes of sync-free sweep control words, then 32 MiB: an fp64 vector of
// n * nrhs <= 4,194,304 values, or the two fp32 vectors of a low-precision
// application of the same size.
constexpr size_t kCtlBytes = 256;
constexpr size_t kVecBytes = size_t(32) << 20;
constexpr size_t kScratchBytes = kCtlBytes + kVecBytes;
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::once_flag g_once[kMaxDevices];
void* g_scratch[kMaxDevices];
hipError_t g_status[kMaxDevices];

char* scratch_ptr() {
  int dev = 0;
  BCM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(kErrArg, "precond_noalloc: device index out of range");
  std::call_once(g_once[dev], [dev] {
    void* p = nullptr;
    g_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch[dev] = p;
  });
  if (g_status[dev] != hipSuccess || !g_scratch[dev])
    throw Error(kErrHip, std::string("precond_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                             hipGetErrorString(g_status[dev]));
  return static_cast<char*>(g_scratch[dev]);
}

__global__ void k_zero_words(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

size_t precond_noalloc_max_elems() { return kVecBytes / 8; }

// The sweeps follow the context's schedule (so the result is bit for bit
// bcm_precond's on the same context): level by level in mode 0; sync-free in
// modes 1 and 2, with the control words in the scratch array, cleared by a
// kernel (no memset node if the caller captures the stream). The factor's
// device image must already exist: building it would allocate.
void precond_noalloc(Factor& F, Context& ctx, const double* r, double* z, int prec, bool transpose, int nrhs) {
  if (prec < kF64 || prec > kBF16) throw Error(kErrArg, "precond_noalloc: unknown precision");
  const Factor::Image& im = F.img[prec];
  if (!im.has || (transpose && !im.has_t))
    throw Error(kErrArg,
                "precond_noalloc: the factor has no device image in this precision and direction yet (one "
                "bcm_precond call with the same precision and transpose flag uploads it)");
  const DevTri& t1 = transpose ? im.fwdt : im.fwd;
  const DevTri& t2 = transpose ? im.bwdt : im.bwd;
  const size_t elems = (size_t)t1.n * (size_t)nrhs;
  if (elems > precond_noalloc_max_elems())
    throw Error(kErrArg, "precond_noalloc: n * nrhs = " + std::to_string(elems) + " exceeds the scratch limit of " +
                             std::to_string(precond_noalloc_max_elems()) + " values");
  char* ctl = scratch_ptr();
  char* base = ctl + kCtlBytes;
  hipStream_t s = ctx.stream;
  const int64_t ne = (int64_t)elems;
  SweepSync local;
  SweepSync* sync = nullptr;
  if (const SweepSync* cs = ctx.sweep_sync()) {
    local.waves_per_cu = cs->waves_per_cu;
    local.cooperative = cs->cooperative;
    local.level_order = cs->level_order;
    local.ext_ctl = reinterpret_cast<int*>(ctl);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, local.ext_ctl, (int)(kCtlBytes / 4));
    BCM_HIP(hipGetLastError());
    sync = &local;
  }
  if (prec == kF64) {
    double* tmp = reinterpret_cast<double*>(base);
    launch_sweep(t1, r, tmp, kF64, nrhs, s, nullptr, sync);
    launch_sweep(t2, tmp, z, kF64, nrhs, s, nullptr, sync);
  } else {
    float* r32 = reinterpret_cast<float*>(base);
    float* t32 = reinterpret_cast<float*>(base + ((elems * 4 + 255) & ~size_t(255)));  // 256-byte aligned
    launch_convert(ne, r, kF64, r32, kF32, s);
    launch_sweep(t1, r32, t32, kF32, nrhs, s, nullptr, sync);
    launch_sweep(t2, t32, r32, kF32, nrhs, s, nullptr, sync);
    launch_convert(ne, r32, kF32, z, kF64, s);
  }
}

}  // namespace bcm

// Device-side objects of libbcmhip and the launch wrappers of its gfx950 kernels.
//
// Layout in HBM (one GPU, everything resident):
//  * DevBsr: block-CSR with the reference's block order (row-major over block
//    rows, ascending columns, column-major b x b blocks); uniform block size
//    b <= 64 for the fast kernels. Values in fp64, fp32 or bf16.
//  * DevTri: one triangular sweep. Row k lists the rows it depends on and
//    the matching b x b blocks, already transposed where the reference applies
//    the transpose (SolveL on a by-columns factor), plus the inverted diagonal
//    block of every row (again pre-transposed for forward sweeps). Forward
//    sweeps depend on smaller rows, backward sweeps on larger ones.
//  * Scalars of the Krylov recurrences live in device memory, so a solver
//    iteration is a fixed launch sequence with no host round trip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../host/bcm_host.hpp"

#define BCM_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::bcm::Error(::bcm::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace bcm {

enum Prec : int { kF64 = 0, kF32 = 1, kBF16 = 2 };

inline size_t prec_bytes(int p) { return p == kF64 ? 8 : (p == kF32 ? 4 : 2); }

// Stream capture and device set-up on other host threads exclude each other.
// A capture (begin to instantiation) holds the lock exclusively; the
// library's set-up HIP calls (allocation, frees, pinned staging, copies,
// helper streams) hold it shared, one call group at a time, so a build on a
// worker thread never interleaves with a capture on the solving thread.
// Waiting captures go first; set-up calls made by the capturing thread
// itself do not take the lock.
struct SetupGuard {
  bool held;
  SetupGuard();
  ~SetupGuard();
  SetupGuard(const SetupGuard&) = delete;
  SetupGuard& operator=(const SetupGuard&) = delete;
};
struct CaptureGuard {
  CaptureGuard();
  ~CaptureGuard();
  CaptureGuard(const CaptureGuard&) = delete;
  CaptureGuard& operator=(const CaptureGuard&) = delete;
};

// Frees the pinned staging buffers kept for uploads (they are allocated again
// on demand); called when no build is running.
void release_staging();

// Owning device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t nbytes);
  void release();
  void upload(const void* src, size_t nbytes, hipStream_t s = nullptr);
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// b == 0: any block partition (the sizes per block row in sp, the value
// offset of every block in bo, nval values); otherwise every block is b x b.
struct DevBsr {
  int nsup = 0, n = 0, b = 0, nzja = 0;
  int prec = kF64;
  int cplx = 0;  // 1: the real equivalent of a complex matrix, a holding its complex (b/2)^2 blocks
  DevBuf ia, ja, a;
  DevBuf sp, bo;
  int64_t nval = 0;
  int64_t value_bytes() const {
    if (cplx) return (int64_t)nzja * b * b / 2 * 8;
    return (b ? (int64_t)nzja * b * b : nval) * (int64_t)prec_bytes(prec);
  }
  int64_t index_bytes() const { return 4LL * (nzja + nsup + 1) + (b ? 0 : 8LL * nzja + 4LL * (nsup + 1)); }
};

// Builds a device block-CSR from host storage (uniform block size required).
void dev_bsr_upload(const BlockCsr& h, int prec, DevBsr& d);
// The real equivalent h of a complex matrix (every block [[Re C, -Im C],
// [Im C, Re C]]) stored on the device as its complex blocks C, (re, im)
// pairs column-major -- half the bytes -- for the complex-block SpMV.
// Returns false (d untouched) for block sizes without that kernel or blocks
// without that structure.
bool dev_bsr_upload_complex(const BlockCsr& h, DevBsr& d);
// Block-diagonal union of several matrices (same block size), part after part.
void dev_bsr_upload_parts(const std::vector<const BlockCsr*>& parts, int prec, DevBsr& d);
// The same for upper-stored symmetric parts, expanded to full storage on the
// device: only the upper blocks cross the bus. Returns false (and leaves d
// alone) when a part holds a block below its diagonal; full_blocks[q] is the
// full block count of part q.
bool dev_bsr_upload_symm_parts(const std::vector<const BlockCsr*>& uppers, DevBsr& d,
                               std::vector<int64_t>& full_blocks);

struct DevTri {
  int nsup = 0, n = 0, b = 0, ndep = 0;  // b == 0: any block partition (sp, doff, dgo, nval)
  bool forward = true;
  int prec = kF64;
  int levels = 0;
  int cplx = 0;  // 1: real equivalent of a complex factor, off/diag holding its complex (b/2)^2 blocks
  DevBuf lctl;   // complex factors: control words of the per-level launches of the complex sweep kernel
  std::vector<int> level_ptr;  // host: rows of level l are level_rows[level_ptr[l] .. level_ptr[l+1])
  DevBuf level_rows;
  DevBuf ia, ja;      // dependency lists per row (row index space of the sweep)
  DevBuf off;         // ndep blocks, row after row
  DevBuf diag;        // nsup inverted diagonal blocks
  DevBuf sp, doff, dgo;  // variable blocks: partition, value offsets of dependency and diagonal blocks
  int64_t nval = 0;
  int64_t value_bytes() const {
    if (cplx) return ((int64_t)ndep + nsup) * b * b / 2 * 8;
    return (b ? ((int64_t)ndep + nsup) * b * b : nval) * (int64_t)prec_bytes(prec);
  }
  int64_t index_bytes() const { return 4LL * (ndep + nsup + 1) + (b ? 0 : 8LL * (ndep + nsup) + 4LL * (nsup + 1)); }
};

// Block sizes with tuned kernels; matrices and factors with other sizes, or
// with more than one size, take the variable-block kernels.
bool fast_block_size(int b);

// Forward sweep through a factor stored by columns (reference SolveL):
// y_k = D_k^T (x_k - sum_{i<k} F_ik^T y_i).
void dev_tri_forward(const BlockCsr& f, int prec, DevTri& t, bool cplx = false);
// Backward sweep through a factor stored by rows (reference SolveU):
// z_i = D_i (x_i - sum_{j>i} F_ij z_j).
void dev_tri_backward(const BlockCsr& f, int prec, DevTri& t, bool cplx = false);
// cplx: f is the real equivalent of a complex factor (every block
// [[P, -Q], [Q, P]]); FP64 images of block sizes 6, 8, 12, 16, 24, 32 then
// store the complex blocks (P, Q) and sweep them with the complex kernel.

// The same sweeps through the block-diagonal union of several factors.
void dev_tri_forward_parts(const std::vector<const BlockCsr*>& f, int prec, DevTri& t);
void dev_tri_backward_parts(const std::vector<const BlockCsr*>& f, int prec, DevTri& t);

// State of the sync-free sweeps on one stream: ctl[2] is the error flag (a
// dependency wait reached its bound). Launch shape: at most waves_per_cu resident waves
// per CU (never more than the occupancy limit), launched cooperatively
// (residency guaranteed by the runtime) or, for stream capture into a graph,
// as an ordinary launch of a grid that fits the device at once.
struct SweepSync {
  DevBuf ctl;
  int* ext_ctl = nullptr;  // caller-provided control words (at least 16 ints, zeroed) used instead of ctl
  int nsup = 0;
  int waves_per_cu = 8;
  bool cooperative = true;
  bool level_order = true;  // process rows sorted by dependency level (false: natural sweep order)
  void ensure(int nsup, hipStream_t s);
  int error(hipStream_t s);
  void clear(hipStream_t s);  // resets the wait-bound error flag
};

// Per-block partial sums of the reductions. A reduction kernel writes one
// partial per block and value; a later kernel on the stream adds them in a
// fixed order (deterministic, and no block needs a device-scope fence).
// pending_blocks / pending_k describe the partials not yet consumed.
struct RedWork {
  DevBuf partials;  // kRedBlocks * kMaxDots doubles
  int pending_blocks = 0, pending_k = 0;
  void ensure();
};
constexpr int kRedBlocks = 2048;
constexpr int kMaxDots = 8;
// Passed as the output of a reduction: leave the partials in the RedWork for
// the next scalar kernel to add (no finalize launch).
inline double* const kDeferred = reinterpret_cast<double*>(static_cast<uintptr_t>(1));

// out[k] = sum of the pending partials of value k; clears the pending state.
void launch_finalize(RedWork& rw, double* out, hipStream_t s, const int* skip = nullptr);

// ---- launches (all asynchronous on `s`) ----

// y = A x (mode 0) or y = bsub - A x (mode 1); nrhs columns with leading
// dimension n. Optionally computes dot(wdot, y) over all rows (first column
// only) when dot_out != nullptr: into *dot_out, or left as pending partials
// in rw when dot_out == kDeferred.
// Every launch takes an optional device flag `skip`: when *skip != 0 the
// kernel returns at once (a finished solver keeps its queued iterations inert).
// Rows whose part (cell) is inactive can be left out of a product or sweep:
// row_cell[r] is the part of block row r, active[c] != 0 marks the parts
// whose rows are computed (rows of other parts keep their old values).
struct RowGate {
  const int* row_cell = nullptr;
  const int* active = nullptr;
};

// Runs body(i) for i in [0, n) on up to `threads` host threads (each with
// the caller's HIP device current); the first exception is rethrown after
// every thread has joined.
void parallel_for(int n, int threads, const std::function<void(int)>& body);
int host_threads();

void launch_spmv(const DevBsr& A, const double* x, double* y, const double* bsub, int nrhs,
                 double* dot_out, const double* wdot, RedWork& rw, hipStream_t s, const int* skip = nullptr,
                 RowGate gate = RowGate());

// Triangular sweep y = T^-1 x for the given precision of vectors (TX). With
// sync == nullptr: one launch per dependency level; otherwise one sync-free
// launch using `sync`. x and y must not alias.
void launch_sweep(const DevTri& T, const void* x, void* y, int vec_prec, int nrhs, hipStream_t s,
                  const int* skip = nullptr, SweepSync* sync = nullptr, RowGate gate = RowGate());

// out[k] = sum_i x_k[i] * y_k[i] for K pairs (K <= kMaxDots), one pass plus a
// finalize launch; with out == kDeferred the partials stay pending in rw.
void launch_multidot(int n, int k, const double* const* xs, const double* const* ys, double* out,
                     RedWork& rw, hipStream_t s, const int* skip = nullptr);

// Converts between fp64 and fp32 vectors.
void launch_convert(int64_t n, const void* src, int src_prec, void* dst, int dst_prec, hipStream_t s,
                    const int* skip = nullptr);

// Generic element-wise kernels used by the drivers (defined with the kernels).
void launch_axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s);
void launch_copy(int64_t n, const double* x, double* y, hipStream_t s);
void launch_fill(int64_t n, double v, double* y, hipStream_t s);

// Number of compute units of the current device.
int device_cu_count();

}  // namespace bcm

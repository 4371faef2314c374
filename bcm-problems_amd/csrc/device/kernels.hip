// gfx950 kernels of the solve phase: block-CSR SpMV, level-scheduled
// triangular sweeps through inverse-diagonal-block factors, and one-pass
// reductions.
//
// Reference operators: CSMatrixRS::MvmA (mvm.cpp:2530-2569), MvmAt
// (mvm.cpp:2593-2632), MvmASymm (mvm.cpp:3008-3037; run here on the
// symmetric matrix expanded to full storage), SolveL (mvm.cpp:3140-3217),
// SolveU (mvm.cpp:3410-3484), ScProdSVect/DaxpySVect (svector.cpp:321-371).
//
// Mapping: one 64-lane wavefront per block row. For block size B the lanes
// form G = 64/B column groups, lane = r + B*g: group g walks the block
// columns c = g, g+G, ... of the row's blocks, so one load instruction reads
// G*B consecutive doubles of a column-major block (coalesced), and the G
// partial sums of each row element are folded with cross-lane shuffles.
//
// Triangular sweeps are level-scheduled: the host groups the block rows by
// dependency level once, and each level is one launch whose rows depend only
// on rows finished by earlier launches of the same stream. No wave ever waits
// on another wave.
#include "device.hpp"
#include "dev_common.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

namespace bcm {

namespace {

// Maps block index to a row tile so that consecutive tiles stay on one XCD
// (the hardware deals workgroups round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  if (nblk & 7) return bid;
  return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

// ---------------------------------------------------------------------------
// SpMV: y = A x or y = bsub - A x, NR right-hand sides, optional dot(w, y).
//
// One wavefront per block row. A block row's blocks are contiguous, column
// after column, so the row is a stream of ncol = nblk*B columns of B values.
// Each lane loads V consecutive values of a column with one vector load;
// L = B/V lanes cover a column and G = 64/L lanes groups walk the columns
// g, g+G, ... of the stream, so one wave instruction moves G*B values of
// consecutive memory. The G partial sums of every row element are folded
// through LDS.
template <class TA, int V> struct VecOf;
template <> struct VecOf<double, 1> { using T = double; };
template <> struct VecOf<double, 2> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct VecOf<double, 4> { typedef double T __attribute__((ext_vector_type(4))); };
template <> struct VecOf<float, 1> { using T = float; };
template <> struct VecOf<float, 2> { typedef float T __attribute__((ext_vector_type(2))); };
template <> struct VecOf<float, 4> { typedef float T __attribute__((ext_vector_type(4))); };
template <> struct VecOf<uint16_t, 1> { using T = uint16_t; };
template <> struct VecOf<uint16_t, 2> { typedef uint16_t T __attribute__((ext_vector_type(2))); };
template <> struct VecOf<uint16_t, 4> { typedef uint16_t T __attribute__((ext_vector_type(4))); };

template <int V, class TA, class TO>
__device__ __forceinline__ void load_vec(const TA* p, TO (&out)[V]) {
  using VT = typename VecOf<TA, V>::T;
  const VT v = __builtin_nontemporal_load(reinterpret_cast<const VT*>(p));
  if constexpr (V == 1) {
    out[0] = (TO)ldv(&v);
  } else {
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const TA e = v[k];
      out[k] = (TO)ldv(&e);
    }
  }
}

// The same through the cache hierarchy: the multi-RHS SpMV's matrix stream
// measured faster with ordinary loads (the sweeps with non-temporal ones).
template <int V, class TA, class TO>
__device__ __forceinline__ void load_vec_cached(const TA* p, TO (&out)[V]) {
  using VT = typename VecOf<TA, V>::T;
  const VT v = *reinterpret_cast<const VT*>(p);
  if constexpr (V == 1) {
    out[0] = (TO)ldv(&v);
  } else {
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const TA e = v[k];
      out[k] = (TO)ldv(&e);
    }
  }
}

template <int B, int V, int NR, class TA>
__global__ __launch_bounds__(256) void k_spmv(int nsup, int n, const int* __restrict__ ia,
                                              const int* __restrict__ ja, const TA* __restrict__ a,
                                              const double* __restrict__ x, double* __restrict__ y,
                                              const double* __restrict__ bsub, const double* __restrict__ wdot,
                                              double* partials, int rows_per_wave, const int* skip) {
  if (skip && *skip) return;
  constexpr int L = B / V;
  constexpr int G = 64 / L;
  __shared__ double red[4][G][NR][B];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane % L, g = lane / L;
  const bool act = g < G;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  // wave-uniform by construction; readfirstlane makes the row loop a scalar loop
  const int wave_global = __builtin_amdgcn_readfirstlane(tile * (blockDim.x >> 6) + wave);
  double dotv = 0.0;
  for (int q0 = 0; q0 < rows_per_wave; ++q0) {
    const int row = wave_global * rows_per_wave + q0;
    if (row >= nsup) break;
    double acc[NR][V];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) acc[q][k] = 0.0;
    const int jb = ia[row], je = ia[row + 1];
    if (act) {
      const TA* ap = a + (size_t)jb * B * B + lr * V;
      const int ncol = (je - jb) * B;
#pragma unroll 4
      for (int t = g; t < ncol; t += G) {
        const int bk = t / B, c = t - bk * B;
        const int col = ja[jb + bk];
        double av[V];
        load_vec<V>(ap + (size_t)t * B, av);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const double xv = x[(size_t)q * n + (size_t)col * B + c];
#pragma unroll
          for (int k = 0; k < V; ++k) acc[q][k] += av[k] * xv;
        }
      }
#pragma unroll
      for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int k = 0; k < V; ++k) red[wave][g][q][lr * V + k] = acc[q][k];
    }
    wave_lds_sync();
    if (lane < B) {
      const size_t e = (size_t)row * B + lane;
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        double s = 0.0;
#pragma unroll
        for (int h = 0; h < G; ++h) s += red[wave][h][q][lane];
        const double v = bsub ? bsub[(size_t)q * n + e] - s : s;
        y[(size_t)q * n + e] = v;
        if (q == 0 && wdot) dotv += wdot[e] * v;
      }
    }
    wave_lds_sync();
  }
  if (partials) {
    const double vv[1] = {dotv};
    block_partials<1>(vv, partials);
  }
}

// Stages the vector values of nb block columns (cols[0..nb)) for NR
// right-hand sides into this wave's LDS area: xs[(kb*B + c)*NRP + q]. For
// chunks of up to 16 block columns the indices are wave-uniform and read
// first (scalar loads), so no value load waits on an index load of its own:
// each lane issues U value loads back to back before their LDS stores (a
// chunk of up to 64*U values costs one memory round trip). Brackets itself
// with wave-level LDS ordering.
template <int B, int NR, int NRP, int KB, int U, class TV, class TS>
__device__ __forceinline__ void stage_chunk(int lane, int nb, const int* __restrict__ cols, const TV* __restrict__ v,
                                            int n, TS* xs) {
  constexpr bool kScalarIdx = KB <= 16;
  int cidx[kScalarIdx ? KB : 1];
  if constexpr (kScalarIdx) {
#pragma unroll
    for (int k = 0; k < KB; ++k) cidx[k] = cols[k < nb ? k : 0];
  }
  const int cnt = nb * NR * B;
  wave_lds_sync();  // earlier reads of the area are done
  for (int base = 0; base < cnt; base += 64 * U) {
    TV val[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const int e = base + lane + 64 * i;
      if (e < cnt) {
        const int c = e % B, rest = e / B;
        const int q = rest % NR, kb = rest / NR;
        int cj;
        if constexpr (kScalarIdx) {
          cj = cidx[0];
#pragma unroll
          for (int k = 1; k < KB; ++k) cj = kb == k ? cidx[k] : cj;
        } else {
          cj = cols[kb];
        }
        val[i] = v[(size_t)q * n + (size_t)cj * B + c];
      }
    }
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const int e = base + lane + 64 * i;
      if (e < cnt) {
        const int c = e % B, rest = e / B;
        const int q = rest % NR, kb = rest / NR;
        xs[(kb * B + c) * NRP + q] = (TS)val[i];
      }
    }
  }
  wave_lds_sync();
}

// The same staging with lanes laid over the NR*B values of one block column
// (lane slot m = q*B + c): the per-lane offsets are computed once, a block
// column's base is a scalar, and the values of KG block columns are in
// flight together. Used where a block column holds at least a wave's worth
// of values.
template <int B, int NR, int NRP, int KG, class TV, class TS>
__device__ __forceinline__ void stage_cols(int lane, int nb, const int* __restrict__ cols, const TV* __restrict__ v,
                                           int n, TS* xs) {
  constexpr int M = NR * B;
  constexpr int S = (M + 63) / 64;
  size_t goff[S];
  int loff[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int m = lane + 64 * s, q = m / B, c = m - q * B;
    goff[s] = (size_t)q * n + c;
    loff[s] = c * NRP + q;
  }
  wave_lds_sync();  // earlier reads of the area are done
  for (int kb0 = 0; kb0 < nb; kb0 += KG) {
    TV val[KG][S];
#pragma unroll
    for (int k = 0; k < KG; ++k) {
      if (kb0 + k < nb) {
        const TV* vb = v + (size_t)cols[kb0 + k] * B;
#pragma unroll
        for (int s = 0; s < S; ++s)
          if (s + 1 < S || lane + 64 * s < M) val[k][s] = vb[goff[s]];
      }
    }
#pragma unroll
    for (int k = 0; k < KG; ++k) {
      if (kb0 + k < nb) {
#pragma unroll
        for (int s = 0; s < S; ++s)
          if (s + 1 < S || lane + 64 * s < M) xs[(kb0 + k) * B * NRP + loff[s]] = (TS)val[k][s];
      }
    }
  }
  wave_lds_sync();
}

// SpMV for several right-hand sides. The x values of a block row's columns
// are staged through LDS first (one coalesced read of B values per block
// column and right-hand side), so the inner loop issues a single vector load
// of V matrix values per column and reads the NR x values of that column
// from LDS. Lanes, column groups and the per-lane column sequence are those
// of k_spmv; the same LDS area then holds the G partial sums of the row,
// folded QF right-hand sides at a time. The staged chunk is kept to about
// 8 KB per wave (seven 24x24 block columns at six right-hand sides: one
// 7-point stencil row in one chunk), so that LDS does not cap the number of
// resident waves below what the registers allow.
template <int B, int NRP>
constexpr int spmv_mr_chunk() {
  return (8192 / (B * NRP * 8)) > 0 ? 8192 / (B * NRP * 8) : 1;
}

template <int B, int V, int NR, class TA>
__global__ __launch_bounds__(256) void k_spmv_mr(int nsup, int n, const int* __restrict__ ia,
                                                 const int* __restrict__ ja, const TA* __restrict__ a,
                                                 const double* __restrict__ x, double* __restrict__ y,
                                                 const double* __restrict__ bsub, const double* __restrict__ wdot,
                                                 double* partials, int rows_per_wave, const int* skip,
                                                 const int* __restrict__ gcell, const int* __restrict__ gact) {
  if (skip && *skip) return;
  constexpr int L = B / V;
  constexpr int G = 64 / L;
  constexpr int NRP = NR % 2 ? NR + 1 : NR;                          // LDS stride per column (16-byte rows)
  constexpr int KB = spmv_mr_chunk<B, NRP>();                        // block columns per staged chunk
  constexpr int WL0 = KB * B * NRP;
  constexpr int WL = WL0 > G * B ? WL0 : G * B;                      // doubles per wave
  constexpr int QF = WL / (G * B) < NR ? WL / (G * B) : NR;          // right-hand sides per fold pass
  static_assert(QF >= 1, "the fold needs room for one right-hand side");
  constexpr int D = 2;                                               // matrix columns in flight per lane
  __shared__ __align__(16) double lds[4][WL];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane % L, g = lane / L;
  const bool act = g < G;
  double* xs = lds[wave];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int wave_global = __builtin_amdgcn_readfirstlane(tile * (blockDim.x >> 6) + wave);
  double dotv = 0.0;
  for (int q0 = 0; q0 < rows_per_wave; ++q0) {
    const int row = wave_global * rows_per_wave + q0;
    if (row >= nsup) break;
    if (gcell && !gact[gcell[row]]) continue;  // wave-uniform: the row's part is inactive
    double acc[NR][V];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) acc[q][k] = 0.0;
    const int jb = ia[row], nblk = ia[row + 1] - jb;
    const TA* ap = a + (size_t)jb * B * B + lr * V;
    int t = g;  // this lane group's next column of the row's stream
    for (int kb0 = 0; kb0 < nblk; kb0 += KB) {
      const int nb = min(KB, nblk - kb0);
      const int t0 = kb0 * B, t1 = (kb0 + nb) * B;
      // The lane's first D matrix columns of the chunk depend only on the row
      // pointer, so they are in flight while the x values are staged; each
      // column consumed then issues the load of the column D steps ahead.
      double av[D][V];
      if (act) {
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (t + d * G < t1) load_vec_cached<V>(ap + (size_t)(t + d * G) * B, av[d]);
      }
      if constexpr (NR * B >= 64)
        stage_cols<B, NR, NRP, KB>(lane, nb, ja + jb + kb0, x, n, xs);
      else
        stage_chunk<B, NR, NRP, KB, (KB * NR * B + 63) / 64>(lane, nb, ja + jb + kb0, x, n, xs);
      if (act) {
        for (int tc = t; tc < t1; tc += D * G) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const int tt = tc + d * G;
            if (tt < t1) {
              double cur[V];
#pragma unroll
              for (int k = 0; k < V; ++k) cur[k] = av[d][k];
              if (tt + D * G < t1) load_vec_cached<V>(ap + (size_t)(tt + D * G) * B, av[d]);
              const double* xp = xs + (tt - t0) * NRP;
              double xv[NRP];
#pragma unroll
              for (int q = 0; q < NRP; ++q) xv[q] = xp[q];
#pragma unroll
              for (int q = 0; q < NR; ++q)
#pragma unroll
                for (int k = 0; k < V; ++k) acc[q][k] += cur[k] * xv[q];
            }
          }
        }
        if (t < t1) t += (t1 - t + G - 1) / G * G;
      }
    }
#pragma unroll
    for (int q0 = 0; q0 < NR; q0 += QF) {
      wave_lds_sync();
      if (act) {
#pragma unroll
        for (int q = q0; q < q0 + QF && q < NR; ++q)
#pragma unroll
          for (int k = 0; k < V; ++k) xs[(g * QF + (q - q0)) * B + lr * V + k] = acc[q][k];
      }
      wave_lds_sync();
      if (lane < B) {
        const size_t e = (size_t)row * B + lane;
#pragma unroll
        for (int q = q0; q < q0 + QF && q < NR; ++q) {
          double s = 0.0;
#pragma unroll
          for (int h = 0; h < G; ++h) s += xs[(h * QF + (q - q0)) * B + lane];
          const double v = bsub ? bsub[(size_t)q * n + e] - s : s;
          y[(size_t)q * n + e] = v;
          if (q == 0 && wdot) dotv += wdot[e] * v;
        }
      }
    }
  }
  if (partials) {
    const double vv[1] = {dotv};
    block_partials<1>(vv, partials);
  }
}

// ---------------------------------------------------------------------------
// Complex-block SpMV on real-equivalent vectors (DevBsr::cplx): the matrix
// holds complex BC x BC blocks as (re, im) pairs, column-major -- the layout
// of CSMatrixCS (mvmc.cpp:160) -- and each vector column holds, per block
// row, its BC real parts and then its BC imaginary parts, so
//   y_re = sum (Re C x_re - Im C x_im),  y_im = sum (Im C x_re + Re C x_im)
// at half the matrix bytes of the real-equivalent blocks. y = A x or
// bsub - A x for NR columns, optional dot(wdot, y) of the first. Lanes,
// staging and fold are those of k_spmv_mr with complex entries: L = BC/VC
// lanes cover a complex column with VC entries (2*VC doubles, one vector
// load) each, G = 64/L lane groups walk the row's complex columns, the x
// values of a chunk of block columns (2*BC real values each) sit in LDS.
template <int BC>
constexpr int spmv_cx_width() {
  return BC % 2 == 0 ? 2 : 1;
}

template <int BC, int NR>
__global__ __launch_bounds__(256) void k_spmv_cx(int nsup, int n, const int* __restrict__ ia,
                                                 const int* __restrict__ ja, const double* __restrict__ a,
                                                 const double* __restrict__ x, double* __restrict__ y,
                                                 const double* __restrict__ bsub, const double* __restrict__ wdot,
                                                 double* partials, int rows_per_wave, const int* skip) {
  if (skip && *skip) return;
  constexpr int B = 2 * BC;  // real rows of a block row
  constexpr int VC = spmv_cx_width<BC>();
  constexpr int L = BC / VC;
  constexpr int G = 64 / L;
  constexpr int NRP = NR % 2 ? NR + 1 : NR;
  constexpr int KB = spmv_mr_chunk<B, NRP>();
  constexpr int WL0 = KB * B * NRP;
  constexpr int WL = WL0 > G * B ? WL0 : G * B;
  constexpr int QF = WL / (G * B) < NR ? WL / (G * B) : NR;
  static_assert(QF >= 1, "the fold needs room for one right-hand side");
  constexpr int D = 2;
  __shared__ __align__(16) double lds[4][WL];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane % L, g = lane / L;
  const bool act = g < G;
  double* xs = lds[wave];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int wave_global = __builtin_amdgcn_readfirstlane(tile * (blockDim.x >> 6) + wave);
  double dotv = 0.0;
  for (int q0 = 0; q0 < rows_per_wave; ++q0) {
    const int row = wave_global * rows_per_wave + q0;
    if (row >= nsup) break;
    double are[NR][VC], aim[NR][VC];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < VC; ++k) {
        are[q][k] = 0.0;
        aim[q][k] = 0.0;
      }
    const int jb = ia[row], nblk = ia[row + 1] - jb;
    const double* ap = a + (size_t)jb * BC * BC * 2 + lr * VC * 2;
    int t = g;  // this lane group's next complex column of the row's stream
    for (int kb0 = 0; kb0 < nblk; kb0 += KB) {
      const int nb = min(KB, nblk - kb0);
      const int t0 = kb0 * BC, t1 = (kb0 + nb) * BC;
      double av[D][2 * VC];
      if (act) {
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (t + d * G < t1) load_vec_cached<2 * VC>(ap + (size_t)(t + d * G) * BC * 2, av[d]);
      }
      if constexpr (NR * B >= 64)
        stage_cols<B, NR, NRP, KB>(lane, nb, ja + jb + kb0, x, n, xs);
      else
        stage_chunk<B, NR, NRP, KB, (KB * NR * B + 63) / 64>(lane, nb, ja + jb + kb0, x, n, xs);
      if (act) {
        for (int tc = t; tc < t1; tc += D * G) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const int tt = tc + d * G;
            if (tt < t1) {
              double cur[2 * VC];
#pragma unroll
              for (int k = 0; k < 2 * VC; ++k) cur[k] = av[d][k];
              if (tt + D * G < t1) load_vec_cached<2 * VC>(ap + (size_t)(tt + D * G) * BC * 2, av[d]);
              const int kc = tt - t0, kb = kc / BC, c = kc - kb * BC;
              const double* xr = xs + (kb * B + c) * NRP;  // real part of x_c, then BC rows on the imaginary
              const double* xi = xr + BC * NRP;
              double vr[NRP], vi[NRP];
#pragma unroll
              for (int q = 0; q < NRP; ++q) {
                vr[q] = xr[q];
                vi[q] = xi[q];
              }
#pragma unroll
              for (int q = 0; q < NR; ++q)
#pragma unroll
                for (int k = 0; k < VC; ++k) {
                  are[q][k] += cur[2 * k] * vr[q] - cur[2 * k + 1] * vi[q];
                  aim[q][k] += cur[2 * k + 1] * vr[q] + cur[2 * k] * vi[q];
                }
            }
          }
        }
        if (t < t1) t += (t1 - t + G - 1) / G * G;
      }
    }
#pragma unroll
    for (int q0 = 0; q0 < NR; q0 += QF) {
      wave_lds_sync();
      if (act) {
#pragma unroll
        for (int q = q0; q < q0 + QF && q < NR; ++q)
#pragma unroll
          for (int k = 0; k < VC; ++k) {
            xs[(g * QF + (q - q0)) * B + lr * VC + k] = are[q][k];
            xs[(g * QF + (q - q0)) * B + BC + lr * VC + k] = aim[q][k];
          }
      }
      wave_lds_sync();
      if (lane < B) {
        const size_t e = (size_t)row * B + lane;
#pragma unroll
        for (int q = q0; q < q0 + QF && q < NR; ++q) {
          double sm = 0.0;
#pragma unroll
          for (int h = 0; h < G; ++h) sm += xs[(h * QF + (q - q0)) * B + lane];
          const double v = bsub ? bsub[(size_t)q * n + e] - sm : sm;
          y[(size_t)q * n + e] = v;
          if (q == 0 && wdot) dotv += wdot[e] * v;
        }
      }
    }
  }
  if (partials) {
    const double vv[1] = {dotv};
    block_partials<1>(vv, partials);
  }
}

// ---------------------------------------------------------------------------
// Level-scheduled triangular sweep: one launch per dependency level; every
// row of a level depends only on rows of earlier levels, which earlier
// launches on the same stream have completed, so no wave ever waits on
// another. One wavefront per block row. A level's waves start together and
// each does little work, so a level costs about one wave's chain of
// dependent memory round trips: the row's right-hand side values and
// inverted diagonal block are loaded as soon as the row is known, the first
// factor columns while the dependencies are staged, and the staging itself
// takes one round trip for up to SKG block columns.
template <int B, int V, int NR, class TA, class TX>
__global__ __launch_bounds__(256) void k_sweep_level(int nrows, const int* __restrict__ rows, int n,
                                                     const int* __restrict__ ia, const int* __restrict__ ja,
                                                     const TA* __restrict__ off, const TA* __restrict__ diag,
                                                     const TX* __restrict__ x, TX* __restrict__ y,
                                                     const int* skip, const int* __restrict__ gcell,
                                                     const int* __restrict__ gact) {
  if (skip && *skip) return;
  using Acc = typename AccOf<TX>::type;
  static_assert(sizeof(Acc) == sizeof(TX), "staged values and partial sums share the LDS area");
  constexpr int L = B / V;
  constexpr int G = 64 / L;
  constexpr int NRP = NR % 2 ? NR + 1 : NR;
  constexpr int KB = (G * NR * B + B * NRP - 1) / (B * NRP);
  constexpr int WL = KB * B * NRP;
  constexpr int SD = 2, SKG = 4;  // factor columns in flight per lane; block columns staged per round trip
  __shared__ __align__(16) Acc lds[4][WL + NR * B];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane % L, g = lane / L;
  const bool act = g < G;
  Acc* xs = lds[wave];
  Acc* vs = lds[wave] + WL;
  const int idx = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  if (idx >= nrows) return;  // wave-uniform; no block-level synchronisation below
  const int row = rows[idx];
  // the row's right-hand side values and its inverted diagonal block are loaded up front
  TX xv[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) xv[q] = lane < B ? x[(size_t)q * n + (size_t)row * B + lane] : TX(0);
  constexpr int DC = (B + G - 1) / G;
  double dvv[DC][V];
  {
    const TA* dp0 = diag + (size_t)row * B * B + lr * V;
#pragma unroll
    for (int i = 0; i < DC; ++i) {
      const int c = g + i * G;
      if (act && c < B) load_vec<V>(dp0 + (size_t)c * B, dvv[i]);
    }
  }
  // v = x_row - sum_k F_rk y_k, with the y values of each chunk of block columns staged in LDS
  Acc acc[NR][V];
#pragma unroll
  for (int q = 0; q < NR; ++q)
#pragma unroll
    for (int k = 0; k < V; ++k) acc[q][k] = Acc(0);
  const int jb = ia[row], nblk = ia[row + 1] - jb;
  if (gcell && !gact[gcell[row]]) return;  // wave-uniform: the row's part is inactive
  const TA* ap = off + (size_t)jb * B * B + lr * V;
  int t = g;
  for (int kb0 = 0; kb0 < nblk; kb0 += KB) {
    const int nb = min(KB, nblk - kb0);
    const int t0 = kb0 * B, t1 = (kb0 + nb) * B;
    // The lane's first SD factor columns of the chunk are in flight while the
    // y values are staged; each column consumed issues the load SD ahead.
    Acc av[SD][V];
    if (act) {
#pragma unroll
      for (int d = 0; d < SD; ++d)
        if (t + d * G < t1) load_vec<V>(ap + (size_t)(t + d * G) * B, av[d]);
    }
    if constexpr (NR * B >= 64)
      stage_cols<B, NR, NRP, SKG>(lane, nb, ja + jb + kb0, y, n, xs);
    else
      stage_chunk<B, NR, NRP, KB, (KB * NR * B + 63) / 64>(lane, nb, ja + jb + kb0, y, n, xs);
    if (act) {
      for (int tc = t; tc < t1; tc += SD * G) {
#pragma unroll
        for (int d = 0; d < SD; ++d) {
          const int tt = tc + d * G;
          if (tt < t1) {
            Acc cur[V];
#pragma unroll
            for (int k = 0; k < V; ++k) cur[k] = av[d][k];
            if (tt + SD * G < t1) load_vec<V>(ap + (size_t)(tt + SD * G) * B, av[d]);
            const Acc* xp = xs + (tt - t0) * NRP;
#pragma unroll
            for (int q = 0; q < NR; ++q)
#pragma unroll
              for (int k = 0; k < V; ++k) acc[q][k] += cur[k] * xp[q];
          }
        }
      }
      if (t < t1) t += (t1 - t + G - 1) / G * G;
    }
  }
  wave_lds_sync();
  if (act) {
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) xs[(g * NR + q) * B + lr * V + k] = acc[q][k];
  }
  wave_lds_sync();
  if (lane < B) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      Acc sm = Acc(0);
#pragma unroll
      for (int h = 0; h < G; ++h) sm += xs[(h * NR + q) * B + lane];
      vs[q * B + lane] = (Acc)xv[q] - sm;
    }
  }
  wave_lds_sync();
  // y_row = D v (inverted diagonal block, column-major as stored for this sweep)
#pragma unroll
  for (int q = 0; q < NR; ++q)
#pragma unroll
    for (int k = 0; k < V; ++k) acc[q][k] = Acc(0);
  if (act) {
#pragma unroll
    for (int i = 0; i < DC; ++i) {
      const int c = g + i * G;
      if (c < B) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
#pragma unroll
          for (int k = 0; k < V; ++k) acc[q][k] += (Acc)dvv[i][k] * vs[q * B + c];
      }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) xs[(g * NR + q) * B + lr * V + k] = acc[q][k];
  }
  wave_lds_sync();
  if (lane < B) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      Acc sm = Acc(0);
#pragma unroll
      for (int h = 0; h < G; ++h) sm += xs[(h * NR + q) * B + lane];
      y[(size_t)q * n + (size_t)row * B + lane] = (TX)sm;
    }
  }
}

// ---------------------------------------------------------------------------
// Sync-free triangular sweep (one cooperative launch per sweep). Wave w owns
// positions w, w+W, w+2W, ... of a topological order of the rows -- the
// natural sweep order, or the rows sorted by dependency level, which
// interleaves independent chains (e.g. the cells of a block-diagonal union)
// -- and processes them in that order; every wave of the grid is resident
// (cooperative launch), so the smallest unfinished position always belongs
// to a wave whose dependencies are complete and the sweep cannot deadlock.
//
// The output vector is its own completion record: before the sweep it is
// filled with a sentinel NaN pattern that no arithmetic produces (a signaling
// NaN with a fixed payload); a row's values are final once they differ from
// it. A wave puts its row's factor columns in flight, stages the values of
// the rows it depends on through LDS (each value loaded once; the ones still
// holding the sentinel re-polled, bounded, until a wave vote finds none),
// computes, and stores its row. Values cross XCDs through
// device-scope relaxed atomic loads/stores (coherent across the 8 L2s, the
// protocol rocPRIM's look-back scan uses on gfx942/gfx950); there is no
// separate flag and no fence on the critical path. A wait that exceeds its
// bound raises ctl[2], which stops every other wait; the host reports it.
//
// The row loop is wave-uniform (blockIdx/gridDim scalars) and every wait
// ends on a wave vote; lane-divergent code inside holds no cross-lane
// operation.
template <class T> struct Sentinel;
template <> struct Sentinel<double> {
  using Bits = unsigned long long;
  static constexpr Bits bits = 0x7FF4DEADC0DE0001ull;
};
template <> struct Sentinel<float> {
  using Bits = unsigned int;
  static constexpr Bits bits = 0x7FA0DEADu;
};

template <class TX>
__device__ __forceinline__ bool is_sentinel(TX v) {
  using Bits = typename Sentinel<TX>::Bits;
  return __builtin_bit_cast(Bits, v) == Sentinel<TX>::bits;
}

template <class TX>
__global__ void k_sweep_fill(int64_t n, TX* __restrict__ y, const int* skip) {
  if (skip && *skip) return;
  using Bits = typename Sentinel<TX>::Bits;
  Bits* yb = reinterpret_cast<Bits*>(y);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) yb[i] = Sentinel<TX>::bits;
}

// Stages the values of dependency blocks kb0 .. kb0+nb-1 of a row (block
// columns ja[jb + k], every right-hand side) into ys[((k - kb0) * NR + q) * B
// + c]. Each value is loaded once, by one of the 64 lanes (the columns a lane
// multiplies are shared by the B lanes of its column group, so loading them
// per lane would repeat every load B times), and the lanes re-poll the values
// still holding the sentinel until a wave vote finds none: one coherent round
// trip once the dependency rows are done, and a waiting wave polls only the
// values it still lacks. The loop exit is a wave vote (uniform); a wait past
// its bound raises ctl[2], which ends every other wait.
template <int B, int NR, int KD, class TX>
__device__ __forceinline__ void stage_deps(int kb0, int nb, int jb, const int* __restrict__ ja, TX* y, int n,
                                           TX* ys, int* ctl) {
  constexpr int SV = KD * B * NR, SJ = (SV + 63) / 64;
  const int lane = threadIdx.x;
  const int nv = nb * B * NR;
  int64_t src[SJ];  // element offsets in y (-1: beyond the chunk)
  TX val[SJ];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < SJ; ++j) {
    const int e = lane + 64 * j;
    src[j] = -1;
    val[j] = TX(0);
    if (e < nv) {
      const int k = e / (B * NR), rem = e - k * (B * NR), q = rem / B, c = rem - q * B;
      src[j] = (int64_t)q * n + (int64_t)ja[jb + kb0 + k] * B + c;
      val[j] = ld_coherent(y + src[j]);
      ok = ok && !is_sentinel(val[j]);
    }
  }
  int spins = 0;
  while (!__all(ok)) {
    __builtin_amdgcn_s_sleep(1);
    ok = true;
#pragma unroll
    for (int j = 0; j < SJ; ++j)
      if (src[j] >= 0 && is_sentinel(val[j])) {
        val[j] = ld_coherent(y + src[j]);
        ok = ok && !is_sentinel(val[j]);
      }
    ++spins;
    if (__any(spins > kSpinLimit || ((spins & 1023) == 0 && ld_coherent(&ctl[2]) != 0))) {
      if (lane == 0) st_coherent(&ctl[2], 1);
      break;
    }
  }
#pragma unroll
  for (int j = 0; j < SJ; ++j)
    if (src[j] >= 0) ys[lane + 64 * j] = val[j];
}

// Rows per lane of the sync-free sweep: two (one 16-byte load per factor
// column for FP64) for even blocks of 12 and more, so that more lanes work
// and fewer registers hold the columns in flight; one for small blocks, whose
// lane groups would otherwise grow the cross-group sums.
template <int B>
constexpr int sweep_free_width() {
  return (B % 2 == 0 && B >= 12) ? 2 : 1;
}

template <int B, int NR, class TA, class TX>
__global__ __launch_bounds__(64) void k_sweep_free(int nsup, int n, int forward, const int* __restrict__ order,
                                                   const int* __restrict__ ia,
                                                   const int* __restrict__ ja, const TA* __restrict__ off,
                                                   const TA* __restrict__ diag, const TX* __restrict__ x, TX* y,
                                                   int* ctl, const int* skip) {
  if (skip && *skip) return;
  using Acc = typename AccOf<TX>::type;
  constexpr int V = sweep_free_width<B>();  // consecutive rows per lane (one vector load per column)
  constexpr int L = B / V;                  // lanes per column
  constexpr int G = 64 / L;                 // column groups
  constexpr int KD = 3;                     // dependency blocks per staged chunk (one round trip for most rows)
  constexpr int CK = (KD * B + G - 1) / G;  // columns of a chunk per lane
  constexpr int PD = (B + G - 1) / G;       // diagonal-block columns per lane
  __shared__ Acc v[NR][B];
  __shared__ TX ys[KD * B * NR];
  const int lane = threadIdx.x;
  const int lr = lane % L, g = lane / L;
  const bool act = g < G;
  for (int t = blockIdx.x; t < nsup; t += gridDim.x) {
    const int row = order ? order[t] : (forward ? t : nsup - 1 - t);
    const int jb = ia[row], ndep = ia[row + 1] - jb;
    const TA* blk0 = off + (size_t)jb * B * B + lr * V;
    const TA* d = diag + (size_t)row * B * B + lr * V;
    // the diagonal block and the right-hand side do not depend on the wait
    Acc pd[PD][V];
#pragma unroll
    for (int i = 0; i < PD; ++i) {
      if (act && g + i * G < B) {
        load_vec<V>(d + (g + i * G) * B, pd[i]);
      } else {
#pragma unroll
        for (int k = 0; k < V; ++k) pd[i][k] = Acc(0);
      }
    }
    Acc xv[NR][V];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k)
        xv[q][k] = lane < L ? (Acc)x[(size_t)q * n + (size_t)row * B + lr * V + k] : Acc(0);
    Acc acc[NR][V];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) acc[q][k] = Acc(0);
    // Chunks of KD dependency blocks. Column tt of the row's block stream
    // (block tt / B, column tt % B) starts at blk0 + tt * B; the lane group g
    // multiplies the columns tt = g (mod G), in ascending order.
    for (int kb0 = 0; kb0 < ndep; kb0 += KD) {
      const int nb = min(KD, ndep - kb0);
      const int c0 = kb0 * B, c1 = (kb0 + nb) * B;
      const int tf = c0 + ((g - c0 % G) % G + G) % G;
      // the chunk's factor values do not depend on the wait: in flight first
      Acc av[CK][V];
#pragma unroll
      for (int i = 0; i < CK; ++i) {
        const int tt = tf + i * G;
        if (act && tt < c1) {
          load_vec<V>(blk0 + (size_t)tt * B, av[i]);
        } else {
#pragma unroll
          for (int k = 0; k < V; ++k) av[i][k] = Acc(0);
        }
      }
      stage_deps<B, NR, KD>(kb0, nb, jb, ja, y, n, ys, ctl);
      __syncthreads();
      if (act) {
#pragma unroll
        for (int i = 0; i < CK; ++i) {
          const int tt = tf + i * G;
          if (tt < c1) {
            const int k = tt / B - kb0, c = tt % B;
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const Acc yv = (Acc)ys[(k * NR + q) * B + c];
#pragma unroll
              for (int e = 0; e < V; ++e) acc[q][e] += av[i][e] * yv;
            }
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) {
        Acc s = acc[q][k];
#pragma unroll
        for (int h = 1; h < G; ++h) s += __shfl(acc[q][k], lane + h * L);
        if (lane < L) v[q][lr * V + k] = xv[q][k] - s;
      }
    __syncthreads();
    Acc z[NR][V];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) z[q][k] = Acc(0);
    if (act) {
#pragma unroll
      for (int i = 0; i < PD; ++i) {
        const int c = g + i * G;
        if (c < B) {
#pragma unroll
          for (int q = 0; q < NR; ++q) {
            const Acc vc = v[q][c];
#pragma unroll
            for (int k = 0; k < V; ++k) z[q][k] += pd[i][k] * vc;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
      for (int k = 0; k < V; ++k) {
        Acc s = z[q][k];
#pragma unroll
        for (int h = 1; h < G; ++h) s += __shfl(z[q][k], lane + h * L);
        if (lane < L) st_coherent(&y[(size_t)q * n + (size_t)row * B + lr * V + k], (TX)s);
      }
    __syncthreads();
  }
}

// The sync-free sweep of a complex factor (DevTri::cplx): factor blocks are
// complex BC x BC, (re, im) pairs column-major, vectors real-equivalent
// ([Re z; Im z] per block row). The protocol -- row order, sentinel
// completion, bounded waits, dependency staging -- is k_sweep_free's; lane
// (r, g) multiplies complex row r of the columns g, g + G, ... of the
// dependency stream (one 16-byte load each), then of the diagonal block:
//   v = x_row - sum_k F_k y_k,  y_row = D v,  (P + iQ)(a + ib) = (Pa - Qb) + i(Qa + Pb).
// With order given and every dependency complete before the launch (one
// launch per level) no wait ever spins.
__device__ __forceinline__ VecOf<double, 2>::T ldnt_c(const double* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const VecOf<double, 2>::T*>(p));
}

template <int BC, int NR>
__global__ __launch_bounds__(64) void k_sweep_free_cx(int nsup, int n, int forward, const int* __restrict__ order,
                                                      const int* __restrict__ ia, const int* __restrict__ ja,
                                                      const double* __restrict__ off,
                                                      const double* __restrict__ diag,
                                                      const double* __restrict__ x, double* y, int* ctl,
                                                      const int* skip) {
  if (skip && *skip) return;
  using C2 = VecOf<double, 2>::T;
  constexpr int B = 2 * BC;                   // real rows of a block row
  constexpr int G = 64 / BC;
  constexpr int KD = 3;                       // dependency blocks per staged chunk
  constexpr int CK = (KD * BC + G - 1) / G;   // complex columns of a chunk per lane
  constexpr int PD = (BC + G - 1) / G;        // diagonal-block columns per lane
  __shared__ double v[NR][B];
  __shared__ double ys[KD * B * NR];
  const int lane = threadIdx.x;
  const int r = lane % BC, g = lane / BC;
  const bool act = g < G;
  const C2 zero2 = {0.0, 0.0};
  for (int t = blockIdx.x; t < nsup; t += gridDim.x) {
    const int row = order ? order[t] : (forward ? t : nsup - 1 - t);
    const int jb = ia[row], ndep = ia[row + 1] - jb;
    const double* blk0 = off + (size_t)jb * BC * BC * 2;
    const double* d = diag + (size_t)row * BC * BC * 2;
    C2 pd[PD];
#pragma unroll
    for (int i = 0; i < PD; ++i) pd[i] = (act && g + i * G < BC) ? ldnt_c(d + ((g + i * G) * BC + r) * 2) : zero2;
    double xr[NR], xi[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      xr[q] = lane < BC ? x[(size_t)q * n + (size_t)row * B + lane] : 0.0;
      xi[q] = lane < BC ? x[(size_t)q * n + (size_t)row * B + BC + lane] : 0.0;
    }
    double ar[NR], ai[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      ar[q] = 0.0;
      ai[q] = 0.0;
    }
    for (int kb0 = 0; kb0 < ndep; kb0 += KD) {
      const int nb = min(KD, ndep - kb0);
      const int c0 = kb0 * BC, c1 = (kb0 + nb) * BC;
      const int tf = c0 + ((g - c0 % G) % G + G) % G;
      C2 av[CK];
#pragma unroll
      for (int i = 0; i < CK; ++i) {
        const int tt = tf + i * G;
        av[i] = (act && tt < c1) ? ldnt_c(blk0 + ((size_t)tt * BC + r) * 2) : zero2;
      }
      stage_deps<B, NR, KD>(kb0, nb, jb, ja, y, n, ys, ctl);
      __syncthreads();
      if (act) {
#pragma unroll
        for (int i = 0; i < CK; ++i) {
          const int tt = tf + i * G;
          if (tt < c1) {
            const int k = tt / BC - kb0, c = tt % BC;
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const double yr = ys[(k * NR + q) * B + c], yi = ys[(k * NR + q) * B + BC + c];
              ar[q] += av[i].x * yr - av[i].y * yi;
              ai[q] += av[i].y * yr + av[i].x * yi;
            }
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      double sr = ar[q], si = ai[q];
#pragma unroll
      for (int h = 1; h < G; ++h) {
        sr += __shfl(ar[q], lane + h * BC);
        si += __shfl(ai[q], lane + h * BC);
      }
      if (lane < BC) {
        v[q][lane] = xr[q] - sr;
        v[q][BC + lane] = xi[q] - si;
      }
    }
    __syncthreads();
    double zr[NR], zi[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      zr[q] = 0.0;
      zi[q] = 0.0;
    }
    if (act) {
#pragma unroll
      for (int i = 0; i < PD; ++i) {
        const int c = g + i * G;
        if (c < BC) {
#pragma unroll
          for (int q = 0; q < NR; ++q) {
            const double vr = v[q][c], vi = v[q][BC + c];
            zr[q] += pd[i].x * vr - pd[i].y * vi;
            zi[q] += pd[i].y * vr + pd[i].x * vi;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      double sr = zr[q], si = zi[q];
#pragma unroll
      for (int h = 1; h < G; ++h) {
        sr += __shfl(zr[q], lane + h * BC);
        si += __shfl(zi[q], lane + h * BC);
      }
      if (lane < BC) {
        st_coherent(&y[(size_t)q * n + (size_t)row * B + lane], sr);
        st_coherent(&y[(size_t)q * n + (size_t)row * B + BC + lane], si);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// BLAS-1

struct DotArgs {
  const double* x[kMaxDots];
  const double* y[kMaxDots];
};

template <int K>
__global__ __launch_bounds__(256) void k_multidot(int64_t n, DotArgs args, double* partials, const int* skip) {
  if (skip && *skip) return;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += args.x[k][i] * args.y[k][i];
  }
  block_partials<K>(v, partials);
}

// out[k] = sum of the per-block partials of value k (one block per value).
__global__ __launch_bounds__(256) void k_finalize(const double* partials, int nb, int K, double* out,
                                                  const int* skip) {
  if (skip && *skip) return;
  const double v = sum_partials(partials, nb, K, blockIdx.x);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

__global__ void k_axpby(int64_t n, double alpha, const double* __restrict__ x, double beta, double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = alpha * x[i] + beta * y[i];
}

__global__ void k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = x[i];
}

__global__ void k_fill(int64_t n, double v, double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = v;
}

template <class S, class D>
__global__ void k_convert(int64_t n, const S* __restrict__ x, D* __restrict__ y, const int* skip) {
  if (skip && *skip) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = (D)x[i];
}

int grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

}  // namespace

// ---------------------------------------------------------------------------
// Buffers and host-side construction

namespace {
// writer-preferring shared lock (a waiting capture blocks new set-up calls)
struct CaptureLock {
  std::mutex mu;
  std::condition_variable cv;
  int readers = 0, writers_waiting = 0;
  bool writer = false;
};
CaptureLock g_capture;
thread_local int t_capturing = 0;
thread_local int t_setup = 0;  // nesting depth: only the outermost guard takes the lock
}  // namespace

SetupGuard::SetupGuard() : held(t_capturing == 0 && t_setup == 0) {
  ++t_setup;
  if (!held) return;
  std::unique_lock<std::mutex> l(g_capture.mu);
  g_capture.cv.wait(l, [] { return !g_capture.writer && g_capture.writers_waiting == 0; });
  ++g_capture.readers;
}

SetupGuard::~SetupGuard() {
  --t_setup;
  if (!held) return;
  std::lock_guard<std::mutex> l(g_capture.mu);
  if (--g_capture.readers == 0) g_capture.cv.notify_all();
}

CaptureGuard::CaptureGuard() {
  if (t_setup > 0) throw Error(kErrInternal, "stream capture inside device set-up");
  std::unique_lock<std::mutex> l(g_capture.mu);
  ++g_capture.writers_waiting;
  g_capture.cv.wait(l, [] { return !g_capture.writer && g_capture.readers == 0; });
  --g_capture.writers_waiting;
  g_capture.writer = true;
  ++t_capturing;
}

CaptureGuard::~CaptureGuard() {
  --t_capturing;
  std::lock_guard<std::mutex> l(g_capture.mu);
  g_capture.writer = false;
  g_capture.cv.notify_all();
}

void DevBuf::alloc(size_t nbytes) {
  release();
  if (nbytes == 0) return;
  SetupGuard g;
  BCM_HIP(hipMalloc(&p, nbytes));
  bytes = nbytes;
}

void DevBuf::release() {
  if (p) {
    SetupGuard g;
    (void)hipFree(p);
  }
  p = nullptr;
  bytes = 0;
}

void DevBuf::upload(const void* src, size_t nbytes, hipStream_t s) {
  if (nbytes > bytes) alloc(nbytes);
  if (!nbytes) return;
  SetupGuard g;
  BCM_HIP(hipMemcpyAsync(p, src, nbytes, hipMemcpyHostToDevice, s));
  BCM_HIP(hipStreamSynchronize(s));
}

int device_cu_count() {
  // one GPU per process: the first query holds for every later launch
  static const int cus = [] {
    int dev = 0, c = 0;
    BCM_HIP(hipGetDevice(&dev));
    BCM_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
    return c;
  }();
  return cus;
}

int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, h ? h : 1u));
}

void parallel_for(int n, int threads, const std::function<void(int)>& body) {
  const int nt = std::max(1, std::min(n, threads));
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) body(i);
    return;
  }
  int dev = 0;
  BCM_HIP(hipGetDevice(&dev));
  std::atomic<int> next{0};
  std::exception_ptr err;
  std::mutex mu;
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      try {
        BCM_HIP(hipSetDevice(dev));
        for (int i = next++; i < n; i = next++) body(i);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu);
        if (!err) err = std::current_exception();
        next = n;  // the other threads stop at their next item
      }
    });
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

static uint16_t to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  const uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)((u + rounding) >> 16);
}

// Pinned host staging buffers for the image uploads, shared by the host
// threads of a build: a thread takes one for a part and returns it after.
// They are allocated on first use and kept until the process ends.
namespace {
constexpr size_t kStageBytes = 32u << 20;
std::mutex g_stage_mu;
std::vector<void*> g_stage_free;

void* stage_take() {
  {
    std::lock_guard<std::mutex> g(g_stage_mu);
    if (!g_stage_free.empty()) {
      void* p = g_stage_free.back();
      g_stage_free.pop_back();
      return p;
    }
  }
  void* p = nullptr;
  SetupGuard g;
  BCM_HIP(hipHostMalloc(&p, kStageBytes, hipHostMallocDefault));
  return p;
}

void stage_give(void* p) {
  std::lock_guard<std::mutex> g(g_stage_mu);
  g_stage_free.push_back(p);
}
}  // namespace

void release_staging() {
  std::vector<void*> bufs;
  {
    std::lock_guard<std::mutex> g(g_stage_mu);
    bufs.swap(g_stage_free);
  }
  SetupGuard g;
  for (void* p : bufs) (void)hipHostFree(p);
}

// Converts `count` fp64 values to `prec` and copies them to dst (device),
// through a pinned staging buffer one chunk at a time.
static void upload_values_at(const double* v, size_t count, int prec, void* dst) {
  if (count == 0) return;
  const size_t es = prec_bytes(prec), chunk = kStageBytes / es;
  void* st = stage_take();
  try {
    for (size_t i0 = 0; i0 < count; i0 += chunk) {
      const size_t m = std::min(chunk, count - i0);
      if (prec == kF64) {
        std::memcpy(st, v + i0, m * 8);
      } else if (prec == kF32) {
        float* f = static_cast<float*>(st);
        for (size_t i = 0; i < m; ++i) f[i] = (float)v[i0 + i];
      } else {
        uint16_t* h = static_cast<uint16_t*>(st);
        for (size_t i = 0; i < m; ++i) h[i] = to_bf16((float)v[i0 + i]);
      }
      SetupGuard g;
      BCM_HIP(hipMemcpy(static_cast<char*>(dst) + i0 * es, st, m * es, hipMemcpyHostToDevice));
    }
  } catch (...) {
    stage_give(st);
    throw;
  }
  stage_give(st);
}

// ---------------------------------------------------------------------------
// Any block partition (CSPRNDS gives one size per block row; blocks up to 64
// rows): products and sweeps with the sizes read per row, the path for block
// sizes the tuned kernels above are not instantiated for and for variable
// partitions. One 64-lane wave per block row; lane = g * b_i + r with
// G = 64 / b_i lane groups, group g taking the block columns c = g, g + G, ...
// of every block of the row in ja order; the G partial sums of a row element
// are folded in group order, so results do not depend on the launch shape.

// y = A x or y = bsub - A x over nrhs columns (leading dimension n).
__global__ __launch_bounds__(256) void k_spmv_var(int nsup, int64_t n, const int* __restrict__ sp,
                                                  const int* __restrict__ ia, const int* __restrict__ ja,
                                                  const int64_t* __restrict__ bo, const double* __restrict__ a,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  const double* __restrict__ bsub, int nrhs, const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = blockIdx.x * 4 + wave;
  if (row >= nsup) return;
  const int r0 = sp[row], bi = sp[row + 1] - r0;
  const int G = 64 / bi;
  const int g = lane / bi, r = lane - g * bi;
  const bool act = g < G;
  for (int q = 0; q < nrhs; ++q) {
    const double* xq = x + (size_t)q * n;
    double acc = 0.0;
    if (act) {
      for (int j = ia[row]; j < ia[row + 1]; ++j) {
        const int col = ja[j];
        const int c0 = sp[col], bj = sp[col + 1] - c0;
        const double* blk = a + bo[j];
        for (int c = g; c < bj; c += G) acc += blk[(size_t)c * bi + r] * xq[c0 + c];
      }
    }
    red[wave][lane] = acc;
    wave_lds_sync();
    if (lane < bi) {
      double s = 0.0;
      for (int h = 0; h < G; ++h) s += red[wave][h * bi + lane];
      const size_t o = (size_t)q * n + r0 + lane;
      y[o] = bsub ? bsub[o] - s : s;
    }
    wave_lds_sync();
  }
}

// One dependency level of a sweep: y_k = D_k (x_k - sum_i G_ki y_i) for the
// rows k of the level, every block stored with b_k rows (column-major).
template <class TA, class TX>
__global__ __launch_bounds__(256) void k_sweep_var(int nrows, const int* __restrict__ rows, int64_t n,
                                                   const int* __restrict__ sp, const int* __restrict__ ia,
                                                   const int* __restrict__ ja, const int64_t* __restrict__ doff,
                                                   const TA* __restrict__ off, const int64_t* __restrict__ dgo,
                                                   const TA* __restrict__ diag, const TX* __restrict__ x,
                                                   TX* __restrict__ y, int nrhs, const int* skip) {
  if (skip && *skip) return;
  using Acc = typename AccOf<TX>::type;
  __shared__ Acc red[4][64];
  __shared__ Acc vs[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = blockIdx.x * 4 + wave;
  if (t >= nrows) return;
  const int k = rows[t];
  const int r0 = sp[k], bk = sp[k + 1] - r0;
  const int G = 64 / bk;
  const int g = lane / bk, r = lane - g * bk;
  const bool act = g < G;
  const TA* d = diag + dgo[k];
  for (int q = 0; q < nrhs; ++q) {
    const TX* xq = x + (size_t)q * n;
    TX* yq = y + (size_t)q * n;
    Acc acc = Acc(0);
    if (act) {
      for (int j = ia[k]; j < ia[k + 1]; ++j) {
        const int i = ja[j];
        const int c0 = sp[i], bi = sp[i + 1] - c0;
        const TA* blk = off + doff[j];
        for (int c = g; c < bi; c += G) acc += (Acc)ldv(blk + (size_t)c * bk + r) * (Acc)yq[c0 + c];
      }
    }
    red[wave][lane] = acc;
    wave_lds_sync();
    if (lane < bk) {
      Acc s = Acc(0);
      for (int h = 0; h < G; ++h) s += red[wave][h * bk + lane];
      vs[wave][lane] = (Acc)xq[r0 + lane] - s;
    }
    wave_lds_sync();
    acc = Acc(0);
    if (act)
      for (int c = g; c < bk; c += G) acc += (Acc)ldv(d + (size_t)c * bk + r) * vs[wave][c];
    red[wave][lane] = acc;
    wave_lds_sync();
    if (lane < bk) {
      Acc s = Acc(0);
      for (int h = 0; h < G; ++h) s += red[wave][h * bk + lane];
      yq[r0 + lane] = (TX)s;
    }
    wave_lds_sync();
  }
}

// Device image of a matrix with any block partition (values fp64, as stored).
static void dev_bsr_upload_var(const BlockCsr& h, DevBsr& d) {
  for (int i = 0; i < h.nsup; ++i)
    if (h.bla(i) < 1 || h.bla(i) > 64) throw Error(kErrArg, "device matrices: blocks of 1..64 rows");
  d.nsup = h.nsup;
  d.n = h.n();
  d.b = 0;
  d.nzja = h.nzja();
  d.prec = kF64;
  d.nval = h.nza();
  d.cplx = 0;
  d.ia.upload(h.ia.data(), h.ia.size() * 4);
  d.ja.upload(h.ja.data(), std::max<size_t>(h.ja.size(), 1) * 4);
  d.sp.upload(h.sprnd.data(), h.sprnd.size() * 4);
  d.bo.upload(h.bsa.data(), std::max<size_t>(h.bsa.size(), 1) * 8);
  d.a.alloc(std::max<size_t>(h.a.size(), 1) * 8);
  upload_values_at(h.a.data(), h.a.size(), kF64, d.a.p);
}

// Sweep image for any block partition: the same dependency lists and levels
// as the uniform builder, blocks of row k stored with b_k rows.
static void build_tri_var(const BlockCsr& f, bool forward, int prec, DevTri& t) {
  const int ns = f.nsup;
  for (int i = 0; i < ns; ++i)
    if (f.bla(i) < 1 || f.bla(i) > 64) throw Error(kErrArg, "device factors: blocks of 1..64 rows");
  t.nsup = ns;
  t.n = f.n();
  t.b = 0;
  t.cplx = 0;
  t.forward = forward;
  t.prec = prec;
  t.ndep = f.nzja() - ns;
  std::vector<int> ia(ns + 1, 0), ja;
  std::vector<int64_t> doff, dgo(ns);
  std::vector<double> off, diag;
  int64_t dpos = 0;
  for (int i = 0; i < ns; ++i) {
    if (f.ia[i] == f.ia[i + 1] || f.ja[f.ia[i]] != i) throw Error(kErrArg, "factor rows must start with the diagonal block");
    const int b = f.bla(i);
    const double* d = &f.a[f.bsa[f.ia[i]]];
    dgo[i] = dpos;
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) diag.push_back(forward ? d[r * b + c] : d[c * b + r]);
    dpos += (int64_t)b * b;
  }
  if (!forward) {
    for (int i = 0; i < ns; ++i) {
      for (int j = f.ia[i] + 1; j < f.ia[i + 1]; ++j) {
        ja.push_back(f.ja[j]);
        doff.push_back((int64_t)off.size());
        const size_t sz = (size_t)f.bla(i) * f.bla(f.ja[j]);
        off.insert(off.end(), f.a.begin() + f.bsa[j], f.a.begin() + f.bsa[j] + sz);
      }
      ia[i + 1] = (int)ja.size();
    }
  } else {
    // row k gathers G_ki = F_ik^T (b_k x b_i) from the factor stored by columns
    std::vector<int> cnt(ns + 1, 0);
    for (int i = 0; i < ns; ++i)
      for (int j = f.ia[i] + 1; j < f.ia[i + 1]; ++j) cnt[f.ja[j] + 1]++;
    for (int k = 0; k < ns; ++k) ia[k + 1] = ia[k] + cnt[k + 1];
    ja.resize(ia[ns]);
    std::vector<std::vector<int>> src(ns);
    for (int i = 0; i < ns; ++i)
      for (int j = f.ia[i] + 1; j < f.ia[i + 1]; ++j) src[f.ja[j]].push_back(j);
    doff.resize(ia[ns]);
    for (int k = 0; k < ns; ++k) {
      const int bk = f.bla(k);
      int p = ia[k];
      for (int j : src[k]) {
        const int i = std::upper_bound(f.ia.begin(), f.ia.end(), j) - f.ia.begin() - 1;
        const int bi = f.bla(i);
        ja[p] = i;
        doff[p] = (int64_t)off.size();
        const double* s = &f.a[f.bsa[j]];  // F_ik: b_i x b_k
        for (int c = 0; c < bi; ++c)
          for (int r = 0; r < bk; ++r) off.push_back(s[r * bi + c]);
        ++p;
      }
    }
  }
  std::vector<int> lev(ns, 0);
  int mx = 0;
  for (int s = 0; s < ns; ++s) {
    const int row = forward ? s : ns - 1 - s;
    int l = 0;
    for (int j = ia[row]; j < ia[row + 1]; ++j) l = std::max(l, lev[ja[j]] + 1);
    lev[row] = l;
    mx = std::max(mx, l);
  }
  t.levels = ns ? mx + 1 : 0;
  t.level_ptr.assign(t.levels + 1, 0);
  for (int i = 0; i < ns; ++i) t.level_ptr[lev[i] + 1]++;
  for (int l = 0; l < t.levels; ++l) t.level_ptr[l + 1] += t.level_ptr[l];
  std::vector<int> rows(ns), fill(t.level_ptr.begin(), t.level_ptr.end() - 1);
  for (int s = 0; s < ns; ++s) {
    const int row = forward ? s : ns - 1 - s;
    rows[fill[lev[row]]++] = row;
  }
  const size_t es = prec_bytes(prec);
  t.nval = (int64_t)off.size() + (int64_t)diag.size();
  t.off.alloc(std::max<size_t>(off.size() * es, es));
  t.diag.alloc(std::max<size_t>(diag.size() * es, es));
  upload_values_at(off.data(), off.size(), prec, t.off.p);
  upload_values_at(diag.data(), diag.size(), prec, t.diag.p);
  t.level_rows.upload(rows.data(), std::max<size_t>(rows.size(), 1) * 4);
  t.ia.upload(ia.data(), ia.size() * 4);
  t.ja.upload(ja.data(), std::max<size_t>(ja.size(), 1) * 4);
  t.sp.upload(f.sprnd.data(), f.sprnd.size() * 4);
  t.doff.upload(doff.data(), std::max<size_t>(doff.size(), 1) * 8);
  t.dgo.upload(dgo.data(), dgo.size() * 8);
}

template <class TA, class TX>
static void sweep_var_levels(const DevTri& T, const void* x, void* y, int nrhs, hipStream_t s, const int* skip) {
  const int* rows = T.level_rows.as<int>();
  for (int l = 0; l < T.levels; ++l) {
    const int r0 = T.level_ptr[l], nr = T.level_ptr[l + 1] - r0;
    hipLaunchKernelGGL((k_sweep_var<TA, TX>), dim3((nr + 3) / 4), dim3(256), 0, s, nr, rows + r0, (int64_t)T.n,
                       T.sp.as<int>(), T.ia.as<int>(), T.ja.as<int>(), T.doff.as<int64_t>(), T.off.as<TA>(),
                       T.dgo.as<int64_t>(), T.diag.as<TA>(), static_cast<const TX*>(x), static_cast<TX*>(y), nrhs,
                       skip);
  }
  BCM_HIP(hipGetLastError());
}

static void launch_sweep_var(const DevTri& T, const void* x, void* y, int vec_prec, int nrhs, hipStream_t s,
                             const int* skip) {
  if (T.prec == kF64 && vec_prec == kF64) sweep_var_levels<double, double>(T, x, y, nrhs, s, skip);
  else if (T.prec == kF32 && vec_prec == kF32) sweep_var_levels<float, float>(T, x, y, nrhs, s, skip);
  else if (T.prec == kBF16 && vec_prec == kF32) sweep_var_levels<uint16_t, float>(T, x, y, nrhs, s, skip);
  else throw Error(kErrArg, "unsupported factor/vector precision pair (fp64/fp64, fp32/fp32, bf16/fp32)");
}

static void launch_spmv_var(const DevBsr& A, const double* x, double* y, const double* bsub, int nrhs,
                            hipStream_t s, const int* skip) {
  hipLaunchKernelGGL(k_spmv_var, dim3((A.nsup + 3) / 4), dim3(256), 0, s, A.nsup, (int64_t)A.n, A.sp.as<int>(),
                     A.ia.as<int>(), A.ja.as<int>(), A.bo.as<int64_t>(), A.a.as<double>(), x, y, bsub, nrhs, skip);
  BCM_HIP(hipGetLastError());
}

bool fast_block_size(int b) {
  return b == 3 || b == 6 || b == 8 || b == 9 || b == 12 || b == 16 || b == 24 || b == 32;
}

static int uniform_block(const BlockCsr& h) {
  const int b = h.nsup ? h.bla(0) : 0;
  for (int i = 0; i < h.nsup; ++i)
    if (h.bla(i) != b) return 0;
  return b;
}

static int parts_block(const std::vector<const BlockCsr*>& parts, const char* what) {
  int b = -1;
  for (const BlockCsr* p : parts) {
    const int bp = uniform_block(*p);
    if (bp == 0 || bp > 64 || (b >= 0 && bp != b))
      throw Error(kErrArg, std::string(what) + " need one uniform block size of at most 64");
    b = bp;
  }
  if (b <= 0) throw Error(kErrArg, std::string(what) + ": empty");
  return b;
}

void dev_bsr_upload_parts(const std::vector<const BlockCsr*>& parts, int prec, DevBsr& d) {
  const int b = parts_block(parts, "device matrices");
  int64_t nsup = 0, nzja = 0;
  for (const BlockCsr* p : parts) { nsup += p->nsup; nzja += p->nzja(); }
  if (nsup * b > INT32_MAX || nzja > INT32_MAX) throw Error(kErrArg, "device matrix exceeds 32-bit indexing");
  d.nsup = (int)nsup;
  d.n = (int)(nsup * b);
  d.b = b;
  d.nzja = (int)nzja;
  d.prec = prec;
  d.cplx = 0;
  const size_t bb = (size_t)b * b, es = prec_bytes(prec);
  std::vector<int> ia(nsup + 1, 0), ja(std::max<int64_t>(nzja, 1), 0);
  d.a.alloc(std::max<size_t>(nzja * bb * es, es));
  const int np = (int)parts.size();
  std::vector<int64_t> r0(np + 1, 0), j0(np + 1, 0);  // first row and block of each part
  for (int q = 0; q < np; ++q) {
    r0[q + 1] = r0[q] + parts[q]->nsup;
    j0[q + 1] = j0[q] + parts[q]->nzja();
  }
  // parts are independent: index shifts, value conversion and the copy to the
  // device run on host threads
  parallel_for(np, host_threads(), [&](int q) {
    const BlockCsr* p = parts[q];
    for (int i = 0; i < p->nsup; ++i) ia[r0[q] + i + 1] = (int)(j0[q] + p->ia[i + 1]);
    for (int j = 0; j < p->nzja(); ++j) ja[j0[q] + j] = (int)(r0[q] + p->ja[j]);
    upload_values_at(p->a.data(), p->a.size(), prec, static_cast<char*>(d.a.p) + (size_t)j0[q] * bb * es);
  });
  d.ia.upload(ia.data(), ia.size() * 4);
  d.ja.upload(ja.data(), ja.size() * 4);
}

void dev_bsr_upload(const BlockCsr& h, int prec, DevBsr& d) {
  if (!fast_block_size(uniform_block(h))) {
    if (prec != kF64) throw Error(kErrArg, "device matrices of variable or untuned block sizes are kept in fp64");
    dev_bsr_upload_var(h, d);
    return;
  }
  dev_bsr_upload_parts({&h}, prec, d);
}

bool dev_bsr_upload_complex(const BlockCsr& h, DevBsr& d) {
  const int b = uniform_block(h);
  const int bc = b / 2;
  if (b % 2 || !(bc == 3 || bc == 4 || bc == 6 || bc == 8 || bc == 12 || bc == 16)) return false;
  const int64_t nzja = h.nzja();
  if ((int64_t)h.nsup * b > INT32_MAX || nzja > INT32_MAX) return false;
  const size_t bb = (size_t)b * b, cb = (size_t)bc * bc;
  if (h.a.size() != (size_t)nzja * bb) return false;
  // every block [[P, -Q], [Q, P]]: complex entry (r, c) = (P(r, c), Q(r, c))
  std::vector<double> c((size_t)std::max<int64_t>(nzja, 1) * cb * 2, 0.0);
  std::atomic<bool> ok{true};
  const int nt = host_threads();
  parallel_for(nt, nt, [&](int w) {
    const int64_t j0 = nzja * w / nt, j1 = nzja * (w + 1) / nt;
    for (int64_t j = j0; j < j1 && ok.load(std::memory_order_relaxed); ++j) {
      const double* blk = h.a.data() + (size_t)j * bb;
      double* dst = c.data() + (size_t)j * cb * 2;
      for (int cc = 0; cc < bc; ++cc)
        for (int r = 0; r < bc; ++r) {
          const double p = blk[(size_t)cc * b + r], q = blk[(size_t)cc * b + r + bc];
          if (blk[(size_t)(cc + bc) * b + r + bc] != p || blk[(size_t)(cc + bc) * b + r] != -q) {
            ok.store(false, std::memory_order_relaxed);
            return;
          }
          dst[((size_t)cc * bc + r) * 2] = p;
          dst[((size_t)cc * bc + r) * 2 + 1] = q;
        }
    }
  });
  if (!ok.load()) return false;
  std::vector<int> ia(h.ia.begin(), h.ia.end()), ja(h.ja.begin(), h.ja.end());
  if (ja.empty()) ja.push_back(0);
  d.nsup = h.nsup;
  d.n = h.nsup * b;
  d.b = b;
  d.nzja = (int)nzja;
  d.prec = kF64;
  d.cplx = 1;
  d.a.upload(c.data(), c.size() * 8);
  d.ia.upload(ia.data(), ia.size() * 4);
  d.ja.upload(ja.data(), ja.size() * 4);
  return true;
}

// Full storage of a symmetric matrix from its upper blocks: full block k is
// upper block src[k] (src[k] >= 0) or the transpose of upper block ~src[k]
// (a block below the diagonal). Blocks are b x b, column-major.
__global__ __launch_bounds__(256) void k_expand_symm(int64_t nblk, int b, const int64_t* __restrict__ src,
                                                     const double* __restrict__ up, double* __restrict__ full) {
  const int bb = b * b;
  for (int64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
    const int64_t sk = src[k];
    const bool tr = sk < 0;
    const double* sp = up + (tr ? ~sk : sk) * bb;
    double* dp = full + k * bb;
    for (int e = threadIdx.x; e < bb; e += blockDim.x) {
      const int c = e / b, r = e - c * b;  // element (r, c) of the full block
      dp[e] = tr ? sp[r * b + c] : sp[e];
    }
  }
}

bool dev_bsr_upload_symm_parts(const std::vector<const BlockCsr*>& uppers, DevBsr& d,
                               std::vector<int64_t>& full_blocks) {
  const int b = parts_block(uppers, "device matrices");
  const int np = (int)uppers.size();
  for (const BlockCsr* u : uppers)
    for (int i = 0; i < u->nsup; ++i)
      for (int p = u->ia[i]; p < u->ia[i + 1]; ++p)
        if (u->ja[p] < i) return false;
  // per part: full row i = the transposes of upper blocks (j, i), j < i, in
  // ascending j, then upper row i (the same order as the host expansion)
  std::vector<std::vector<int>> pia(np), pja(np);
  std::vector<std::vector<int64_t>> psrc(np);
  parallel_for(np, host_threads(), [&](int q) {
    const BlockCsr& u = *uppers[q];
    const int ns = u.nsup;
    std::vector<int> cnt(ns + 1, 0);
    for (int i = 0; i < ns; ++i)
      for (int p = u.ia[i]; p < u.ia[i + 1]; ++p)
        if (u.ja[p] > i) cnt[u.ja[p]]++;
    std::vector<int>& ia = pia[q];
    ia.assign(ns + 1, 0);
    for (int i = 0; i < ns; ++i) ia[i + 1] = ia[i] + cnt[i] + (u.ia[i + 1] - u.ia[i]);
    pja[q].assign(ia[ns], 0);
    psrc[q].assign(ia[ns], 0);
    std::vector<int> fill(ia.begin(), ia.end() - 1);
    for (int i = 0; i < ns; ++i)  // lower parts, rows j receive (j, i) in ascending i
      for (int p = u.ia[i]; p < u.ia[i + 1]; ++p)
        if (u.ja[p] > i) {
          const int j = u.ja[p], at = fill[j]++;
          pja[q][at] = i;
          psrc[q][at] = ~(int64_t)p;
        }
    for (int i = 0; i < ns; ++i)
      for (int p = u.ia[i]; p < u.ia[i + 1]; ++p) {
        const int at = fill[i]++;
        pja[q][at] = u.ja[p];
        psrc[q][at] = p;
      }
  });
  std::vector<int64_t> r0(np + 1, 0), f0(np + 1, 0), u0(np + 1, 0);
  for (int q = 0; q < np; ++q) {
    r0[q + 1] = r0[q] + uppers[q]->nsup;
    f0[q + 1] = f0[q] + pia[q][uppers[q]->nsup];
    u0[q + 1] = u0[q] + uppers[q]->nzja();
  }
  const int64_t nsup = r0[np], nfull = f0[np], nup = u0[np];
  if (nsup * b > INT32_MAX || nfull > INT32_MAX) throw Error(kErrArg, "device matrix exceeds 32-bit indexing");
  const size_t bb = (size_t)b * b;
  std::vector<int> ia(nsup + 1, 0), ja(std::max<int64_t>(nfull, 1), 0);
  std::vector<int64_t> src(std::max<int64_t>(nfull, 1), 0);
  DevBuf up;
  up.alloc(std::max<size_t>(nup * bb * 8, 8));
  full_blocks.assign(np, 0);
  parallel_for(np, host_threads(), [&](int q) {
    const BlockCsr& u = *uppers[q];
    for (int i = 0; i < u.nsup; ++i) ia[r0[q] + i + 1] = (int)(f0[q] + pia[q][i + 1]);
    for (size_t k = 0; k < pja[q].size(); ++k) {
      ja[f0[q] + k] = (int)(r0[q] + pja[q][k]);
      const int64_t sk = psrc[q][k];
      src[f0[q] + k] = sk < 0 ? ~(u0[q] + ~sk) : u0[q] + sk;
    }
    full_blocks[q] = (int64_t)pja[q].size();
    upload_values_at(u.a.data(), u.a.size(), kF64, static_cast<char*>(up.p) + (size_t)u0[q] * bb * 8);
  });
  d.nsup = (int)nsup;
  d.n = (int)(nsup * b);
  d.b = b;
  d.nzja = (int)nfull;
  d.prec = kF64;
  d.a.alloc(std::max<size_t>(nfull * bb * 8, 8));
  DevBuf dsrc;
  dsrc.upload(src.data(), src.size() * 8);
  if (nfull > 0) {
    // a stream of its own, so that the build does not wait for work other
    // streams run meanwhile (e.g. the solve of the previous batch)
    SetupGuard g;
    hipStream_t xs = nullptr;
    BCM_HIP(hipStreamCreateWithFlags(&xs, hipStreamNonBlocking));
    hipLaunchKernelGGL(k_expand_symm, dim3((unsigned)std::min<int64_t>(nfull, 1 << 16)), dim3(256), 0, xs, nfull,
                       b, dsrc.as<int64_t>(), up.as<double>(), d.a.as<double>());
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(xs);
    (void)hipStreamDestroy(xs);
    BCM_HIP(le);
    BCM_HIP(se);
  }
  d.ia.upload(ia.data(), ia.size() * 4);
  d.ja.upload(ja.data(), ja.size() * 4);
  return true;
}

// Sweep image of the block-diagonal union of factors (one per part, rows of
// part p after those of parts < p). Forward: rows gather the transposed
// blocks of the column-stored factor; backward: rows of the row-stored factor.
// Diagonal blocks are stored transposed for the forward sweep. Dependency
// levels (a row's level is one more than the deepest row it depends on) are
// merged over the parts for the level-scheduled launches.
// Real-equivalent blocks [[P, -Q], [Q, P]] (b = 2 bc, column-major) to complex
// bc x bc blocks of (P, Q) pairs, column-major; false if a block lacks the
// structure.
static bool complex_blocks(const std::vector<double>& src, int b, std::vector<double>& dst) {
  const int bc = b / 2;
  const size_t bb = (size_t)b * b, nblk = src.size() / bb;
  dst.assign(nblk * bc * bc * 2, 0.0);
  for (size_t k = 0; k < nblk; ++k) {
    const double* blk = src.data() + k * bb;
    double* o = dst.data() + k * bc * bc * 2;
    for (int c = 0; c < bc; ++c)
      for (int r = 0; r < bc; ++r) {
        const double p = blk[(size_t)c * b + r], q = blk[(size_t)c * b + r + bc];
        if (blk[(size_t)(c + bc) * b + r + bc] != p || blk[(size_t)(c + bc) * b + r] != -q) return false;
        o[((size_t)c * bc + r) * 2] = p;
        o[((size_t)c * bc + r) * 2 + 1] = q;
      }
  }
  return true;
}

static bool complex_sweep_block(int b) {
  return b == 6 || b == 8 || b == 12 || b == 16 || b == 24 || b == 32;
}

// cplx: the parts are real equivalents of complex factors; their blocks are
// stored complex (half the bytes) for k_sweep_free_cx when every block has
// the structure, else the image is built real.
static void build_tri_parts(const std::vector<const BlockCsr*>& parts, bool forward, int prec, DevTri& t,
                            bool cplx = false) {
  const int b = parts_block(parts, "device factors");
  if (cplx && (prec != kF64 || !complex_sweep_block(b))) cplx = false;
  const size_t bb = (size_t)b * b, es = prec_bytes(prec);
  const size_t vb = cplx ? bb / 2 : bb;  // stored doubles per block
  int64_t nsup = 0, ndep = 0;
  for (const BlockCsr* f : parts) { nsup += f->nsup; ndep += f->nzja() - f->nsup; }
  if (nsup * b > INT32_MAX || ndep > INT32_MAX) throw Error(kErrArg, "device factor exceeds 32-bit indexing");
  t.nsup = (int)nsup;
  t.n = (int)(nsup * b);
  t.b = b;
  t.forward = forward;
  t.prec = prec;
  t.ndep = (int)ndep;
  t.cplx = cplx ? 1 : 0;
  std::vector<int> ia_all(nsup + 1, 0), ja_all(std::max<int64_t>(ndep, 1), 0), lev_all(nsup, 0);
  t.off.alloc(std::max<size_t>(ndep * vb * es, es));
  t.diag.alloc(nsup * vb * es);
  std::atomic<bool> cx_ok{true};
  const int np = (int)parts.size();
  std::vector<int64_t> rs(np + 1, 0), ds(np + 1, 0);  // first row and dependency block of each part
  for (int q = 0; q < np; ++q) {
    rs[q + 1] = rs[q] + parts[q]->nsup;
    ds[q + 1] = ds[q] + parts[q]->nzja() - parts[q]->nsup;
  }
  std::vector<int> part_mx(np, 0);
  // parts are independent (a row depends on rows of its own part only): the
  // transposition, levels, conversion and copy of each run on host threads
  parallel_for(np, host_threads(), [&](int q) {
    const BlockCsr& f = *parts[q];
    const int64_t r0 = rs[q], d0 = ds[q];
    int mx = 0;
    const int ns = f.nsup;
    std::vector<int> ia(ns + 1, 0), ja;
    std::vector<double> off, diag((size_t)ns * bb);
    for (int i = 0; i < ns; ++i) {
      if (f.ia[i] == f.ia[i + 1] || f.ja[f.ia[i]] != i)
        throw Error(kErrArg, "factor rows must start with the diagonal block");
      const double* d = &f.a[f.bsa[f.ia[i]]];
      double* o = &diag[(size_t)i * bb];
      for (int r = 0; r < b; ++r)
        for (int c = 0; c < b; ++c) o[c * b + r] = forward ? d[r * b + c] : d[c * b + r];
    }
    if (!forward) {
      for (int i = 0; i < ns; ++i) {
        for (int j = f.ia[i] + 1; j < f.ia[i + 1]; ++j) {
          ja.push_back(f.ja[j]);
          off.insert(off.end(), f.a.begin() + f.bsa[j], f.a.begin() + f.bsa[j] + bb);
        }
        ia[i + 1] = (int)ja.size();
      }
    } else {
      std::vector<int> cnt(ns + 1, 0);
      for (int i = 0; i < ns; ++i)
        for (int j = f.ia[i] + 1; j < f.ia[i + 1]; ++j) cnt[f.ja[j] + 1]++;
      for (int k = 0; k < ns; ++k) ia[k + 1] = ia[k] + cnt[k + 1];
      ja.resize(ia[ns]);
      off.resize((size_t)ia[ns] * bb);
      std::vector<int> pos(ia.begin(), ia.end() - 1);
      for (int i = 0; i < ns; ++i)
        for (int j = f.ia[i] + 1; j < f.ia[i + 1]; ++j) {
          const int k = f.ja[j];
          const int p = pos[k]++;
          ja[p] = i;
          const double* src = &f.a[f.bsa[j]];
          double* o = &off[(size_t)p * bb];
          for (int r = 0; r < b; ++r)
            for (int c = 0; c < b; ++c) o[c * b + r] = src[r * b + c];
        }
    }
    for (int s = 0; s < ns; ++s) {
      const int row = forward ? s : ns - 1 - s;
      int l = 0;
      for (int j = ia[row]; j < ia[row + 1]; ++j) l = std::max(l, lev_all[r0 + ja[j]] + 1);
      lev_all[r0 + row] = l;
      mx = std::max(mx, l);
    }
    if ((int64_t)ja.size() != ds[q + 1] - d0) throw Error(kErrInternal, "factor image: dependency count");
    for (int i = 0; i < ns; ++i) ia_all[r0 + i + 1] = (int)(d0 + ia[i + 1]);
    for (size_t j = 0; j < ja.size(); ++j) ja_all[d0 + j] = (int)(r0 + ja[j]);
    if (cplx) {
      std::vector<double> co, cd;
      if (!complex_blocks(off, b, co) || !complex_blocks(diag, b, cd)) {
        cx_ok.store(false);
      } else {
        off.swap(co);
        diag.swap(cd);
      }
    }
    if (!cplx || cx_ok.load()) {
      upload_values_at(off.data(), off.size(), prec, static_cast<char*>(t.off.p) + (size_t)d0 * vb * es);
      upload_values_at(diag.data(), diag.size(), prec, static_cast<char*>(t.diag.p) + (size_t)r0 * vb * es);
    }
    part_mx[q] = mx;
  });
  if (cplx && !cx_ok.load()) {  // a block without the complex structure: the real image
    build_tri_parts(parts, forward, prec, t, false);
    return;
  }
  if (cplx) {
    const int zeros[16] = {};
    t.lctl.upload(zeros, sizeof(zeros));
  }
  const int mx = np ? *std::max_element(part_mx.begin(), part_mx.end()) : 0;
  t.levels = nsup ? mx + 1 : 0;
  t.level_ptr.assign(t.levels + 1, 0);
  for (int64_t i = 0; i < nsup; ++i) t.level_ptr[lev_all[i] + 1]++;
  for (int l = 0; l < t.levels; ++l) t.level_ptr[l + 1] += t.level_ptr[l];
  std::vector<int> rows(nsup), fill(t.level_ptr.begin(), t.level_ptr.end() - 1);
  for (int q = 0; q < np; ++q) {
    const int ns = parts[q]->nsup;
    for (int s = 0; s < ns; ++s) {
      const int row = (int)rs[q] + (forward ? s : ns - 1 - s);
      rows[fill[lev_all[row]]++] = row;
    }
  }
  t.level_rows.upload(rows.data(), std::max<size_t>(rows.size(), 1) * 4);
  t.ia.upload(ia_all.data(), ia_all.size() * 4);
  t.ja.upload(ja_all.data(), ja_all.size() * 4);
}

void dev_tri_forward(const BlockCsr& f, int prec, DevTri& t, bool cplx) {
  if (!fast_block_size(uniform_block(f))) build_tri_var(f, true, prec, t);
  else build_tri_parts({&f}, true, prec, t, cplx);
}
void dev_tri_backward(const BlockCsr& f, int prec, DevTri& t, bool cplx) {
  if (!fast_block_size(uniform_block(f))) build_tri_var(f, false, prec, t);
  else build_tri_parts({&f}, false, prec, t, cplx);
}
void dev_tri_forward_parts(const std::vector<const BlockCsr*>& f, int prec, DevTri& t) {
  build_tri_parts(f, true, prec, t);
}
void dev_tri_backward_parts(const std::vector<const BlockCsr*>& f, int prec, DevTri& t) {
  build_tri_parts(f, false, prec, t);
}

void SweepSync::ensure(int n, hipStream_t s) {
  if (ext_ctl || ctl.p) return;
  ctl.alloc(64);
  BCM_HIP(hipMemsetAsync(ctl.p, 0, 64, s));
  BCM_HIP(hipStreamSynchronize(s));
  nsup = n;
}

void SweepSync::clear(hipStream_t s) {
  if (!ctl.p) return;
  BCM_HIP(hipMemsetAsync(ctl.p, 0, 64, s));
  BCM_HIP(hipStreamSynchronize(s));
}

int SweepSync::error(hipStream_t s) {
  if (!ctl.p) return 0;
  int h[4] = {0, 0, 0, 0};
  BCM_HIP(hipMemcpyAsync(h, ctl.p, sizeof(h), hipMemcpyDeviceToHost, s));
  BCM_HIP(hipStreamSynchronize(s));
  return h[2];
}

void RedWork::ensure() {
  if (partials.p) return;
  partials.alloc((size_t)kRedBlocks * kMaxDots * 8);
}

void launch_finalize(RedWork& rw, double* out, hipStream_t s, const int* skip) {
  if (rw.pending_k < 1) throw Error(kErrInternal, "finalize: no pending reduction");
  hipLaunchKernelGGL(k_finalize, dim3(rw.pending_k), dim3(256), 0, s, rw.partials.as<double>(), rw.pending_blocks,
                     rw.pending_k, out, skip);
  BCM_HIP(hipGetLastError());
  rw.pending_k = 0;
}

// ---------------------------------------------------------------------------
// Launch wrappers

// Vector width of the SpMV loads: four values where the block size allows.
template <int B>
constexpr int spmv_width() {
  return B % 4 == 0 ? 4 : (B % 2 == 0 ? 2 : 1);
}

template <int B, int NR, class TA>
static void spmv_b(const DevBsr& A, const double* x, double* y, const double* bsub, double* dot_out,
                   const double* wdot, RedWork& rw, hipStream_t s, const int* skip, RowGate gate) {
  constexpr int kWavesPerBlock = 4;
  // several right-hand sides: one row per wave, so the rows an XCD has in
  // flight stay close together and their x columns are reused from its L2
  const int nwaves_max = (NR == 1 || dot_out) ? 2048 * kWavesPerBlock : (1 << 30);
  const int rows_per_wave = (A.nsup + nwaves_max - 1) / nwaves_max;
  const int nwaves = (A.nsup + rows_per_wave - 1) / rows_per_wave;
  int grid = (nwaves + kWavesPerBlock - 1) / kWavesPerBlock;
  if (grid > 8) grid = (grid + 7) / 8 * 8;  // whole XCD rounds, so the tile map is a permutation
  double* partials = nullptr;
  if (dot_out) {
    rw.ensure();
    if (grid > kRedBlocks * kMaxDots) throw Error(kErrInternal, "spmv: too many blocks for the fused dot");
    partials = rw.partials.as<double>();
  }
  if constexpr (NR == 1)
    hipLaunchKernelGGL((k_spmv<B, spmv_width<B>(), 1, TA>), dim3(grid), dim3(64 * kWavesPerBlock), 0, s, A.nsup,
                       A.n, A.ia.as<int>(), A.ja.as<int>(), A.a.as<TA>(), x, y, bsub, wdot, partials, rows_per_wave,
                       skip);
  else
    hipLaunchKernelGGL((k_spmv_mr<B, spmv_width<B>(), NR, TA>), dim3(grid), dim3(64 * kWavesPerBlock), 0, s, A.nsup,
                       A.n, A.ia.as<int>(), A.ja.as<int>(), A.a.as<TA>(), x, y, bsub, wdot, partials, rows_per_wave,
                       skip, gate.row_cell, gate.active);
  BCM_HIP(hipGetLastError());
  if (dot_out) {
    rw.pending_blocks = grid;
    rw.pending_k = 1;
    if (dot_out != kDeferred) launch_finalize(rw, dot_out, s, skip);
  }
}

template <int B, int NR>
static void spmv_prec(const DevBsr& A, const double* x, double* y, const double* bsub, double* dot_out,
                      const double* wdot, RedWork& rw, hipStream_t s, const int* skip, RowGate gate) {
  if (A.prec == kF64) spmv_b<B, NR, double>(A, x, y, bsub, dot_out, wdot, rw, s, skip, gate);
  else if (A.prec == kF32) spmv_b<B, NR, float>(A, x, y, bsub, dot_out, wdot, rw, s, skip, gate);
  else spmv_b<B, NR, uint16_t>(A, x, y, bsub, dot_out, wdot, rw, s, skip, gate);
}

template <int B>
static void spmv_nr(const DevBsr& A, const double* x, double* y, const double* bsub, int nrhs, double* dot_out,
                    const double* wdot, RedWork& rw, hipStream_t s, const int* skip, RowGate gate) {
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const size_t o = (size_t)done * A.n;
    const double* bs = bsub ? bsub + o : nullptr;
    double* dd = done == 0 ? dot_out : nullptr;
    if (left >= 6) { spmv_prec<B, 6>(A, x + o, y + o, bs, dd, wdot, rw, s, skip, gate); done += 6; }
    else if (left >= 3) { spmv_prec<B, 3>(A, x + o, y + o, bs, dd, wdot, rw, s, skip, gate); done += 3; }
    else if (left == 2) { spmv_prec<B, 2>(A, x + o, y + o, bs, dd, wdot, rw, s, skip, gate); done += 2; }
    else { spmv_prec<B, 1>(A, x + o, y + o, bs, dd, wdot, rw, s, skip, RowGate()); done += 1; }
  }
}

template <int BC, int NR>
static void spmv_cx_b(const DevBsr& A, const double* x, double* y, const double* bsub, double* dot_out,
                      const double* wdot, RedWork& rw, hipStream_t s, const int* skip) {
  constexpr int kWavesPerBlock = 4;
  const int nwaves_max = (NR == 1 || dot_out) ? 2048 * kWavesPerBlock : (1 << 30);
  const int rows_per_wave = (A.nsup + nwaves_max - 1) / nwaves_max;
  const int nwaves = (A.nsup + rows_per_wave - 1) / rows_per_wave;
  int grid = (nwaves + kWavesPerBlock - 1) / kWavesPerBlock;
  if (grid > 8) grid = (grid + 7) / 8 * 8;
  double* partials = nullptr;
  if (dot_out) {
    rw.ensure();
    if (grid > kRedBlocks * kMaxDots) throw Error(kErrInternal, "spmv: too many blocks for the fused dot");
    partials = rw.partials.as<double>();
  }
  hipLaunchKernelGGL((k_spmv_cx<BC, NR>), dim3(grid), dim3(64 * kWavesPerBlock), 0, s, A.nsup, A.n,
                     A.ia.as<int>(), A.ja.as<int>(), A.a.as<double>(), x, y, bsub, wdot, partials, rows_per_wave,
                     skip);
  BCM_HIP(hipGetLastError());
  if (dot_out) {
    rw.pending_blocks = grid;
    rw.pending_k = 1;
    if (dot_out != kDeferred) launch_finalize(rw, dot_out, s, skip);
  }
}

template <int BC>
static void spmv_cx_nr(const DevBsr& A, const double* x, double* y, const double* bsub, int nrhs, double* dot_out,
                       const double* wdot, RedWork& rw, hipStream_t s, const int* skip) {
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const size_t o = (size_t)done * A.n;
    const double* bs = bsub ? bsub + o : nullptr;
    double* dd = done == 0 ? dot_out : nullptr;
    if (left >= 6) { spmv_cx_b<BC, 6>(A, x + o, y + o, bs, dd, wdot, rw, s, skip); done += 6; }
    else if (left >= 3) { spmv_cx_b<BC, 3>(A, x + o, y + o, bs, dd, wdot, rw, s, skip); done += 3; }
    else if (left == 2) { spmv_cx_b<BC, 2>(A, x + o, y + o, bs, dd, wdot, rw, s, skip); done += 2; }
    else { spmv_cx_b<BC, 1>(A, x + o, y + o, bs, dd, wdot, rw, s, skip); done += 1; }
  }
}

void launch_spmv(const DevBsr& A, const double* x, double* y, const double* bsub, int nrhs, double* dot_out,
                 const double* wdot, RedWork& rw, hipStream_t s, const int* skip, RowGate gate) {
  if (A.cplx) {
    if (gate.active) throw Error(kErrArg, "row-gated products are not defined for complex matrices");
    switch (A.b / 2) {
      case 3: spmv_cx_nr<3>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip); break;
      case 4: spmv_cx_nr<4>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip); break;
      case 6: spmv_cx_nr<6>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip); break;
      case 8: spmv_cx_nr<8>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip); break;
      case 12: spmv_cx_nr<12>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip); break;
      case 16: spmv_cx_nr<16>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip); break;
      default: throw Error(kErrInternal, "complex SpMV: unsupported block size");
    }
    return;
  }
  if (A.b == 0) {
    if (gate.active) throw Error(kErrArg, "row-gated products need a uniform tuned block size");
    launch_spmv_var(A, x, y, bsub, nrhs, s, skip);
    if (dot_out) {  // the fused dot of the first column, as the tuned kernels leave it
      const double* xs[1] = {wdot};
      const double* ys[1] = {y};
      launch_multidot(A.n, 1, xs, ys, dot_out, rw, s, skip);
    }
    return;
  }
  switch (A.b) {
    case 3: spmv_nr<3>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 6: spmv_nr<6>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 8: spmv_nr<8>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 9: spmv_nr<9>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 12: spmv_nr<12>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 16: spmv_nr<16>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 24: spmv_nr<24>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    case 32: spmv_nr<32>(A, x, y, bsub, nrhs, dot_out, wdot, rw, s, skip, gate); break;
    default: throw Error(kErrArg, "unsupported block size for the device SpMV (use 3,6,8,9,12,16,24,32)");
  }
}

template <int B, int NR, class TA, class TX>
static void sweep_b(const DevTri& T, const void* x, void* y, hipStream_t s, const int* skip, SweepSync* sync,
                    RowGate gate) {
  if (sync) {
    // every wave must be resident: cooperative launch of at most the occupancy limit
    auto kern = k_sweep_free<B, NR, TA, TX>;
    // resident one-wave blocks per CU of this instantiation (queried once)
    static const int per_cu = [&] {
      int v = 0;
      BCM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kern, 64, 0));
      return v;
    }();
    const int waves = std::max(1, std::min(T.nsup, std::min(per_cu, sync->waves_per_cu) * device_cu_count()));
    const int64_t ny = (int64_t)T.n * NR;
    TX* yy = static_cast<TX*>(y);
    hipLaunchKernelGGL((k_sweep_fill<TX>), dim3(std::max<int64_t>(1, std::min<int64_t>((ny + 255) / 256, 2048))),
                       dim3(256), 0, s, ny, yy, skip);
    int nsup = T.nsup, n = T.n, fwd = T.forward ? 1 : 0;
    const int* ia = T.ia.as<int>();
    const int* ja = T.ja.as<int>();
    const TA* off = T.off.as<TA>();
    const TA* dg = T.diag.as<TA>();
    const TX* xx = static_cast<const TX*>(x);
    int* ct = sync->ext_ctl ? sync->ext_ctl : sync->ctl.as<int>();
    const int* ord = sync->level_order ? T.level_rows.as<int>() : nullptr;
    void* args[] = {&nsup, &n, &fwd, &ord, &ia, &ja, &off, &dg, &xx, &yy, &ct, (void*)&skip};
    if (sync->cooperative)
      BCM_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(waves), dim3(64), args, 0, s));
    else
      BCM_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kern), dim3(waves), dim3(64), args, 0, s));
    return;
  }
  const int* rows = T.level_rows.as<int>();
  for (int l = 0; l < T.levels; ++l) {
    const int r0 = T.level_ptr[l], nr = T.level_ptr[l + 1] - r0;
    hipLaunchKernelGGL((k_sweep_level<B, spmv_width<B>(), NR, TA, TX>), dim3((nr + 3) / 4), dim3(256), 0, s, nr,
                       rows + r0, T.n,
                       T.ia.as<int>(), T.ja.as<int>(), T.off.as<TA>(), T.diag.as<TA>(), static_cast<const TX*>(x),
                       static_cast<TX*>(y), skip, gate.row_cell, gate.active);
  }
  BCM_HIP(hipGetLastError());
}

template <int B, int NR>
static void sweep_prec(const DevTri& T, const void* x, void* y, int vec_prec, hipStream_t s, const int* skip,
                       SweepSync* sync, RowGate gate) {
  if (T.prec == kF64 && vec_prec == kF64) sweep_b<B, NR, double, double>(T, x, y, s, skip, sync, gate);
  else if (T.prec == kF32 && vec_prec == kF32) sweep_b<B, NR, float, float>(T, x, y, s, skip, sync, gate);
  else if (T.prec == kBF16 && vec_prec == kF32) sweep_b<B, NR, uint16_t, float>(T, x, y, s, skip, sync, gate);
  else throw Error(kErrArg, "unsupported factor/vector precision pair (fp64/fp64, fp32/fp32, bf16/fp32)");
}

template <int B>
static void sweep_nr(const DevTri& T, const void* x, void* y, int vec_prec, int nrhs, hipStream_t s,
                     const int* skip, SweepSync* sync, RowGate gate) {
  const size_t es = prec_bytes(vec_prec);
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const char* xs = static_cast<const char*>(x) + (size_t)done * T.n * es;
    char* ys = static_cast<char*>(y) + (size_t)done * T.n * es;
    if (left >= 6) { sweep_prec<B, 6>(T, xs, ys, vec_prec, s, skip, sync, gate); done += 6; }
    else if (left >= 3) { sweep_prec<B, 3>(T, xs, ys, vec_prec, s, skip, sync, gate); done += 3; }
    else if (left == 2) { sweep_prec<B, 2>(T, xs, ys, vec_prec, s, skip, sync, gate); done += 2; }
    else { sweep_prec<B, 1>(T, xs, ys, vec_prec, s, skip, sync, gate); done += 1; }
  }
}

template <int BC, int NR>
static void sweep_cx_b(const DevTri& T, const double* x, double* y, hipStream_t s, const int* skip,
                       SweepSync* sync) {
  auto kern = k_sweep_free_cx<BC, NR>;
  int n = T.n, fwd = T.forward ? 1 : 0;
  const int* ia = T.ia.as<int>();
  const int* ja = T.ja.as<int>();
  const double* off = T.off.as<double>();
  const double* dg = T.diag.as<double>();
  if (sync) {  // as the real sync-free sweep: every wave resident, y filled with the sentinel first
    static const int per_cu = [&] {
      int v = 0;
      BCM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kern, 64, 0));
      return v;
    }();
    const int waves = std::max(1, std::min(T.nsup, std::min(per_cu, sync->waves_per_cu) * device_cu_count()));
    const int64_t ny = (int64_t)T.n * NR;
    hipLaunchKernelGGL((k_sweep_fill<double>), dim3(std::max<int64_t>(1, std::min<int64_t>((ny + 255) / 256, 2048))),
                       dim3(256), 0, s, ny, y, skip);
    int nsup = T.nsup;
    int* ct = sync->ext_ctl ? sync->ext_ctl : sync->ctl.as<int>();
    const int* ord = sync->level_order ? T.level_rows.as<int>() : nullptr;
    void* args[] = {&nsup, &n, &fwd, &ord, &ia, &ja, &off, &dg, &x, &y, &ct, (void*)&skip};
    if (sync->cooperative)
      BCM_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(waves), dim3(64), args, 0, s));
    else
      BCM_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kern), dim3(waves), dim3(64), args, 0, s));
    return;
  }
  // level by level: a launch's rows depend only on rows of earlier launches, so its waits never spin
  const int* rows = T.level_rows.as<int>();
  int* ct = T.lctl.as<int>();
  for (int l = 0; l < T.levels; ++l) {
    const int r0 = T.level_ptr[l], nr = T.level_ptr[l + 1] - r0;
    hipLaunchKernelGGL(kern, dim3(nr), dim3(64), 0, s, nr, n, fwd, rows + r0, ia, ja, off, dg, x, y, ct, skip);
  }
  BCM_HIP(hipGetLastError());
}

template <int BC>
static void sweep_cx_nr(const DevTri& T, const double* x, double* y, int nrhs, hipStream_t s, const int* skip,
                        SweepSync* sync) {
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const size_t o = (size_t)done * T.n;
    if (left >= 6) { sweep_cx_b<BC, 6>(T, x + o, y + o, s, skip, sync); done += 6; }
    else if (left >= 3) { sweep_cx_b<BC, 3>(T, x + o, y + o, s, skip, sync); done += 3; }
    else if (left == 2) { sweep_cx_b<BC, 2>(T, x + o, y + o, s, skip, sync); done += 2; }
    else { sweep_cx_b<BC, 1>(T, x + o, y + o, s, skip, sync); done += 1; }
  }
}

void launch_sweep(const DevTri& T, const void* x, void* y, int vec_prec, int nrhs, hipStream_t s,
                  const int* skip, SweepSync* sync, RowGate gate) {
  if (x == y) throw Error(kErrArg, "triangular sweep: input and output must not alias");
  if (T.cplx) {
    if (gate.active) throw Error(kErrArg, "row-gated sweeps are not defined for complex factors");
    if (vec_prec != kF64) throw Error(kErrArg, "complex factors sweep FP64 vectors");
    if (sync) sync->ensure(T.nsup, s);
    const double* xd = static_cast<const double*>(x);
    double* yd = static_cast<double*>(y);
    switch (T.b / 2) {
      case 3: sweep_cx_nr<3>(T, xd, yd, nrhs, s, skip, sync); break;
      case 4: sweep_cx_nr<4>(T, xd, yd, nrhs, s, skip, sync); break;
      case 6: sweep_cx_nr<6>(T, xd, yd, nrhs, s, skip, sync); break;
      case 8: sweep_cx_nr<8>(T, xd, yd, nrhs, s, skip, sync); break;
      case 12: sweep_cx_nr<12>(T, xd, yd, nrhs, s, skip, sync); break;
      case 16: sweep_cx_nr<16>(T, xd, yd, nrhs, s, skip, sync); break;
      default: throw Error(kErrInternal, "complex sweep: unsupported block size");
    }
    return;
  }
  if (T.b == 0) {  // variable blocks: level by level whatever the schedule
    if (gate.active) throw Error(kErrArg, "row-gated sweeps need a uniform tuned block size");
    launch_sweep_var(T, x, y, vec_prec, nrhs, s, skip);
    return;
  }
  if (sync) sync->ensure(T.nsup, s);
  switch (T.b) {
    case 3: sweep_nr<3>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 6: sweep_nr<6>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 8: sweep_nr<8>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 9: sweep_nr<9>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 12: sweep_nr<12>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 16: sweep_nr<16>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 24: sweep_nr<24>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    case 32: sweep_nr<32>(T, x, y, vec_prec, nrhs, s, skip, sync, gate); break;
    default: throw Error(kErrArg, "unsupported block size for the device sweep (use 3,6,8,9,12,16,24,32)");
  }
}

void launch_multidot(int n, int k, const double* const* xs, const double* const* ys, double* out, RedWork& rw,
                     hipStream_t s, const int* skip) {
  if (k < 1 || k > kMaxDots) throw Error(kErrArg, "multidot: 1..8 products");
  rw.ensure();
  DotArgs args{};
  for (int i = 0; i < k; ++i) { args.x[i] = xs[i]; args.y[i] = ys[i]; }
  const int grid = std::min(grid_for(n, 256), kRedBlocks);
  double* p = rw.partials.as<double>();
  switch (k) {
    case 1: hipLaunchKernelGGL(k_multidot<1>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    case 2: hipLaunchKernelGGL(k_multidot<2>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    case 3: hipLaunchKernelGGL(k_multidot<3>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    case 4: hipLaunchKernelGGL(k_multidot<4>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    case 5: hipLaunchKernelGGL(k_multidot<5>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    case 6: hipLaunchKernelGGL(k_multidot<6>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    case 7: hipLaunchKernelGGL(k_multidot<7>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
    default: hipLaunchKernelGGL(k_multidot<8>, dim3(grid), dim3(256), 0, s, (int64_t)n, args, p, skip); break;
  }
  BCM_HIP(hipGetLastError());
  rw.pending_blocks = grid;
  rw.pending_k = k;
  if (out != kDeferred) launch_finalize(rw, out, s, skip);
}

void launch_convert(int64_t n, const void* src, int sp, void* dst, int dp, hipStream_t s, const int* skip) {
  const int grid = grid_for(n, 256);
  if (sp == kF64 && dp == kF32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid), dim3(256), 0, s, n, (const double*)src, (float*)dst,
                       skip);
  else if (sp == kF32 && dp == kF64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid), dim3(256), 0, s, n, (const float*)src, (double*)dst,
                       skip);
  else
    throw Error(kErrArg, "convert: unsupported precision pair");
  BCM_HIP(hipGetLastError());
}

void launch_axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256)), dim3(256), 0, s, n, alpha, x, beta, y);
  BCM_HIP(hipGetLastError());
}

void launch_copy(int64_t n, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x, y);
  BCM_HIP(hipGetLastError());
}

void launch_fill(int64_t n, double v, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, n, v, y);
  BCM_HIP(hipGetLastError());
}

}  // namespace bcm

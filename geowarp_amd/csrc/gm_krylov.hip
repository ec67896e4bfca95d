// Block-CSR Krylov solver for the Newton correction and the consistent-mass
// projection (gfx950).  Replaces the reference's Warp `bsr_set_from_triplets` +
// `preconditioner('diag')` + `bicgstab` (qs2d:201-204, qc3d:335-338, 410-412)
// and its host SciPy/pyamg/pypardiso options.
//
//  * SpMV: G lanes per block row, each lane walks every G-th block of the row;
//    lane-group reduction with cross-lane shuffles; optional fused dot products
//    written as per-workgroup partials (deterministic two-stage reduction).
//  * Preconditioner: exact inverse of every nb x nb diagonal block (block Jacobi).
//  * BiCGStab (van der Vorst 1992), right preconditioned, with all scalars kept
//    on the device; the host reads ||r|| only every `check_every` iterations.
//  * Matrix and vectors in FP64 or FP32; reductions always in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "gm_common.h"
#include "gm_types.h"

#define SPMV_T 256
#define RED_T 256

// ------------------------------------------------------------------ SpMV
template <int NB, int G, typename TM, typename TX, typename TY, int T = SPMV_T>
__global__ void __launch_bounds__(T) k_spmv(int n_rows, const int* row_ptr, const int* col_idx, const TM* vals,
                                            const TX* x, TY* y, const TY* w1, const TY* w2, double* part,
                                            const double* gate) {
  if (gate && *gate != 0.0) return;  // a gated launch after its solve has converged
  const int rows_per_blk = T / G;
  int lr = threadIdx.x / G;
  int lane = threadIdx.x % G;
  int row = blockIdx.x * rows_per_blk + lr;
  double acc[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) acc[r] = 0.0;
  if (NB == 1 && row < n_rows) {
    // scalar rows: four entries per lane and step, G apart (each load instruction reads a
    // contiguous stretch), their loads issued before they are accumulated in the
    // one-entry-at-a-time order
    const int s = row_ptr[row], e = row_ptr[row + 1];
    int k = s + lane;
    for (; k + 3 * G < e; k += 4 * G) {
      int c[4];
      TM v[4];
      double xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        c[u] = col_idx[k + u * G];
        v[u] = vals[k + u * G];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = (double)x[c[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[0] += (double)v[u] * xv[u];
    }
    for (; k < e; k += G) acc[0] += (double)vals[k] * (double)x[col_idx[k]];
  } else if (row < n_rows) {
    int s = row_ptr[row], e = row_ptr[row + 1];
    for (int k = s + lane; k < e; k += G) {
      int c = col_idx[k];
      const TM* blk = vals + (long)k * NB * NB;
      double xv[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) xv[j] = (double)x[c * NB + j];
#pragma unroll
      for (int r = 0; r < NB; ++r)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[r] += (double)blk[r * NB + j] * xv[j];
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < NB; ++r) acc[r] += __shfl_xor(acc[r], off, G);
  double d1 = 0.0, d2 = 0.0;
  if (row < n_rows && lane == 0) {
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      TY v = (TY)acc[r];
      y[row * NB + r] = v;
      if (w1) d1 += (double)v * (double)w1[row * NB + r];
      if (w2) d2 += (double)v * (double)w2[row * NB + r];
    }
  }
  if (part) {
    __shared__ double sh1[T], sh2[T];
    sh1[threadIdx.x] = d1;
    sh2[threadIdx.x] = d2;
    __syncthreads();
    for (int o = T / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        sh1[threadIdx.x] += sh1[threadIdx.x + o];
        sh2[threadIdx.x] += sh2[threadIdx.x + o];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = sh1[0];
      part[2 * blockIdx.x + 1] = sh2[0];
    }
  }
}

// Workgroup size of the SpMV per (block size, matrix precision), chosen by
// gm_spmv_autotune (256 until tuned).  Launches with fused dot products keep
// 256 threads so their per-workgroup partials, and hence their round-off, do
// not depend on the tuning; without them the workgroup size does not change
// any result bit (each row's lanes and reduction order stay the same).
static int g_spmv_threads[5][2] = {{256, 256}, {256, 256}, {256, 256}, {256, 256}, {256, 256}};

// Lanes per block row: 3D u-p rows hold 27-125 blocks (16 lanes measured 4-6 % faster than 8
// on the 64x64x32-cell indentation Jacobian in both precisions); 2D rows hold at most 25
// blocks; scalar rows (nb = 1) are the consistent-mass matrices, 125 entries in 3D.
static constexpr int spmv_lanes(int nb) { return nb >= 4 ? 16 : (nb == 3 ? 8 : (nb == 2 ? 4 : 16)); }

template <int NB, typename TM, typename TX, typename TY>
static void launch_spmv(int n_rows, const int* rp, const int* ci, const TM* v, const TX* x, TY* y, const TY* w1,
                        const TY* w2, double* part, hipStream_t s, int threads = 0,
                        const double* gate = nullptr) {
  const int G = spmv_lanes(NB);
  int t = part ? SPMV_T : (threads ? threads : g_spmv_threads[NB][sizeof(TM) == 4 ? 1 : 0]);
  switch (t) {
    case 64:
      k_spmv<NB, G, TM, TX, TY, 64><<<gm_blocks(n_rows, 64 / G), 64, 0, s>>>(n_rows, rp, ci, v, x, y, w1, w2, part,
                                                                               gate);
      break;
    case 128:
      k_spmv<NB, G, TM, TX, TY, 128><<<gm_blocks(n_rows, 128 / G), 128, 0, s>>>(n_rows, rp, ci, v, x, y, w1, w2,
                                                                                 part, gate);
      break;
    case 512:
      k_spmv<NB, G, TM, TX, TY, 512><<<gm_blocks(n_rows, 512 / G), 512, 0, s>>>(n_rows, rp, ci, v, x, y, w1, w2,
                                                                                 part, gate);
      break;
    default:
      k_spmv<NB, G, TM, TX, TY, 256><<<gm_blocks(n_rows, 256 / G), 256, 0, s>>>(n_rows, rp, ci, v, x, y, w1, w2,
                                                                                 part, gate);
  }
}

// ------------------------------------------------------------------ FP16 smoother operator
// An FP32 level matrix stored again in FP16, each (scalar row, column dof)
// sub-row of a block row scaled by its largest magnitude, so that every stored
// value lies in [-1, 1] and the small pressure-pressure entries of u-p systems
// keep their own range:
//     A[k][r][c] ~ lo[k][r][c] * scale[(row * NB + r) * NB + c].
// The V-cycle's smoothing and residual products read it (half the matrix
// bytes of FP32); the Krylov operator itself stays FP32.
typedef _Float16 gm_half;

template <int NB, int G, int T, typename TM>
__global__ void __launch_bounds__(T) k_lo_convert(int n_rows, const int* row_ptr, const TM* vals, gm_half* lo,
                                                  float* scale) {
  const int rows_per_blk = T / G;
  int row = blockIdx.x * rows_per_blk + threadIdx.x / G;
  int lane = threadIdx.x % G;
  float mx[NB * NB];
#pragma unroll
  for (int q = 0; q < NB * NB; ++q) mx[q] = 0.0f;
  int s = 0, e = 0;
  if (row < n_rows) {
    s = row_ptr[row];
    e = row_ptr[row + 1];
    for (int k = s + lane; k < e; k += G)
#pragma unroll
      for (int q = 0; q < NB * NB; ++q) mx[q] = fmaxf(mx[q], fabsf((float)vals[(long)k * NB * NB + q]));
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) mx[q] = fmaxf(mx[q], __shfl_xor(mx[q], off, G));
  if (row >= n_rows) return;
  float inv[NB * NB];
#pragma unroll
  for (int q = 0; q < NB * NB; ++q) {
    float sc = (mx[q] > 0.0f && mx[q] < 3.0e38f) ? mx[q] : 1.0f;
    inv[q] = 1.0f / sc;
    if (lane == 0) scale[(long)row * NB * NB + q] = sc;
  }
  for (int k = s + lane; k < e; k += G)
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) lo[(long)k * NB * NB + q] = (gm_half)((float)vals[(long)k * NB * NB + q] * inv[q]);
}

// The same conversion with G lanes per row and up to C blocks per lane kept in registers
// between the row maximum and the conversion, so that a row of at most G C blocks is read
// from memory once (the FP16 values and scales are those of k_lo_convert).
template <int NB, int G, int T, typename TM, int C>
__global__ void __launch_bounds__(T) k_lo_convert_cached(int n_rows, const int* row_ptr, const TM* vals, gm_half* lo,
                                                         float* scale) {
  const int row = blockIdx.x * (T / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  float mx[NB * NB], cv[C][NB * NB];
#pragma unroll
  for (int q = 0; q < NB * NB; ++q) mx[q] = 0.0f;
  int s = 0, e = 0;
  if (row < n_rows) {
    s = row_ptr[row];
    e = row_ptr[row + 1];
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const int k = s + lane + u * G;
      if (k < e) {
#pragma unroll
        for (int q = 0; q < NB * NB; ++q) {
          cv[u][q] = (float)vals[(long)k * NB * NB + q];
          mx[q] = fmaxf(mx[q], fabsf(cv[u][q]));
        }
      }
    }
    for (int k = s + lane + C * G; k < e; k += G)
#pragma unroll
      for (int q = 0; q < NB * NB; ++q) mx[q] = fmaxf(mx[q], fabsf((float)vals[(long)k * NB * NB + q]));
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) mx[q] = fmaxf(mx[q], __shfl_xor(mx[q], off, G));
  if (row >= n_rows) return;
  float inv[NB * NB];
#pragma unroll
  for (int q = 0; q < NB * NB; ++q) {
    float sc = (mx[q] > 0.0f && mx[q] < 3.0e38f) ? mx[q] : 1.0f;
    inv[q] = 1.0f / sc;
    if (lane == 0) scale[(long)row * NB * NB + q] = sc;
  }
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int k = s + lane + u * G;
    if (k < e)
#pragma unroll
      for (int q = 0; q < NB * NB; ++q) lo[(long)k * NB * NB + q] = (gm_half)(cv[u][q] * inv[q]);
  }
  for (int k = s + lane + C * G; k < e; k += G)
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) lo[(long)k * NB * NB + q] = (gm_half)((float)vals[(long)k * NB * NB + q] * inv[q]);
}

// Row product of a level matrix, 16 (8, 4) lanes per block row; every lane of the row
// ends with the row's full product (butterfly reduction in a fixed order).
template <int NB, int G, typename TM, typename TX>
__device__ __forceinline__ void row_acc(int row, int lane, int n_rows, const int* row_ptr, const int* col_idx,
                                        const TM* vals, const TX* x, double acc[NB]) {
#pragma unroll
  for (int r = 0; r < NB; ++r) acc[r] = 0.0;
  if (row < n_rows) {
    int s = row_ptr[row], e = row_ptr[row + 1];
    for (int k = s + lane; k < e; k += G) {
      int c = col_idx[k];
      const TM* blk = vals + (long)k * NB * NB;
      double xv[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) xv[j] = (double)x[(long)c * NB + j];
#pragma unroll
      for (int r = 0; r < NB; ++r)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[r] += (double)blk[r * NB + j] * xv[j];
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int r = 0; r < NB; ++r) acc[r] += __shfl_xor(acc[r], off, G);
}

// The same from the FP16 copy: FP32 partial sums per (r, c), scaled after the reduction.
template <int NB, int G>
__device__ __forceinline__ void row_acc_lo(int row, int lane, int n_rows, const int* row_ptr, const int* col_idx,
                                           const gm_half* lo, const float* scale, const float* x, double acc[NB]) {
  float a[NB * NB];
#pragma unroll
  for (int q = 0; q < NB * NB; ++q) a[q] = 0.0f;
  if (row < n_rows) {
    int s = row_ptr[row], e = row_ptr[row + 1];
    int k = s + lane;
    // two blocks per step: both blocks' loads are issued before either is accumulated (the
    // accumulation order is the one-block-at-a-time order)
    for (; k + G < e; k += 2 * G) {
      int c0 = col_idx[k], c1 = col_idx[k + G];
      const gm_half* b0 = lo + (long)k * NB * NB;
      const gm_half* b1 = lo + (long)(k + G) * NB * NB;
      gm_half h0[NB * NB], h1[NB * NB];
      float x0[NB], x1[NB];
#pragma unroll
      for (int q = 0; q < NB * NB; ++q) {
        h0[q] = b0[q];
        h1[q] = b1[q];
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        x0[j] = x[(long)c0 * NB + j];
        x1[j] = x[(long)c1 * NB + j];
      }
#pragma unroll
      for (int r = 0; r < NB; ++r)
#pragma unroll
        for (int j = 0; j < NB; ++j) a[r * NB + j] += (float)h0[r * NB + j] * x0[j];
#pragma unroll
      for (int r = 0; r < NB; ++r)
#pragma unroll
        for (int j = 0; j < NB; ++j) a[r * NB + j] += (float)h1[r * NB + j] * x1[j];
    }
    for (; k < e; k += G) {
      int c = col_idx[k];
      const gm_half* blk = lo + (long)k * NB * NB;
      float xv[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) xv[j] = x[(long)c * NB + j];
#pragma unroll
      for (int r = 0; r < NB; ++r)
#pragma unroll
        for (int j = 0; j < NB; ++j) a[r * NB + j] += (float)blk[r * NB + j] * xv[j];
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) a[q] += __shfl_xor(a[q], off, G);
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    double v = 0.0;
    if (row < n_rows)
#pragma unroll
      for (int j = 0; j < NB; ++j) v += (double)scale[((long)row * NB + r) * NB + j] * (double)a[r * NB + j];
    acc[r] = v;
  }
}

// One smoother-side product of a multigrid level, fused with what consumes it:
//   MODE 1 (residual):      out = b - A x
//   MODE 2 (Jacobi sweep):  out = x + omega Dinv (b - A x)      (out != x)
// A is the level's matrix (TM) or, with LO, its FP16 copy (FP32 hierarchies).
template <int NB, int G, int MODE, typename TM, typename T, bool LO>
__global__ void __launch_bounds__(256) k_level(int n_rows, const int* row_ptr, const int* col_idx, const TM* vals,
                                               const gm_half* lo, const float* scale, const T* x, const T* b,
                                               const T* dinv, double omega, T* out) {
  int row = blockIdx.x * (256 / G) + threadIdx.x / G;
  int lane = threadIdx.x % G;
  double acc[NB];
  if constexpr (LO)
    row_acc_lo<NB, G>(row, lane, n_rows, row_ptr, col_idx, lo, scale, (const float*)x, acc);
  else
    row_acc<NB, G, TM, T>(row, lane, n_rows, row_ptr, col_idx, vals, x, acc);
  if (row >= n_rows || lane != 0) return;
  double res[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) res[r] = (double)b[(long)row * NB + r] - acc[r];
  if (MODE == 1) {
#pragma unroll
    for (int r = 0; r < NB; ++r) out[(long)row * NB + r] = (T)res[r];
  } else {
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) s += (double)dinv[(long)row * NB * NB + r * NB + j] * res[j];
      out[(long)row * NB + r] = (T)((double)x[(long)row * NB + r] + omega * s);
    }
  }
}

template <int NB, typename TM>
static void launch_lo_convert(int n_rows, const int* rp, const TM* vals, gm_half* lo, float* scale,
                              hipStream_t s) {
  if constexpr (NB == 4 && sizeof(TM) == 4) {
    // 3D u-p rows hold at most 125 blocks: 32 lanes with 4 blocks each read every row once
    k_lo_convert_cached<NB, 32, 256, TM, 4><<<gm_blocks(n_rows, 256 / 32), 256, 0, s>>>(n_rows, rp, vals, lo,
                                                                                        scale);
  } else {
    constexpr int G = spmv_lanes(NB);
    k_lo_convert<NB, G, 256, TM><<<gm_blocks(n_rows, 256 / G), 256, 0, s>>>(n_rows, rp, vals, lo, scale);
  }
}

// workgroups of a launch with fused dot products (SPMV_T threads): the partial-sum count
static int spmv_blocks(int n_rows, int nb) { return gm_blocks(n_rows, SPMV_T / spmv_lanes(nb)); }

// ------------------------------------------------------------------ reductions
// Sum `n` partial pairs (stride 2) deterministically into out[0], out[1].
__global__ void k_sum_parts(const double* part, int n, double* out0, double* out1) {
  __shared__ double a[RED_T], b[RED_T];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < n; i += RED_T) {
    s0 += part[2 * i];
    s1 += part[2 * i + 1];
  }
  a[threadIdx.x] = s0;
  b[threadIdx.x] = s1;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (out0) *out0 = a[0];
    if (out1) *out1 = b[0];
  }
}

// per-block partial dots of (u.v) and (u2.v2) over n entries
template <typename T>
__global__ void k_dot2(long n, const T* u, const T* v, const T* u2, const T* v2, double* part) {
  __shared__ double a[RED_T], b[RED_T];
  double s0 = 0.0, s1 = 0.0;
  for (long i = blockIdx.x * (long)RED_T + threadIdx.x; i < n; i += (long)gridDim.x * RED_T) {
    s0 += (double)u[i] * (double)v[i];
    if (u2) s1 += (double)u2[i] * (double)v2[i];
  }
  a[threadIdx.x] = s0;
  b[threadIdx.x] = s1;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = b[0];
  }
}

static int dot_blocks(long n) {
  long b = (n + RED_T - 1) / RED_T;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

// ------------------------------------------------------------------ block Jacobi
template <int NB, typename TM, typename TD = TM>
__global__ void k_diag_inv(int n_rows, const int* row_ptr, const int* col_idx, const TM* vals, TD* dinv) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  int s = row_ptr[row], e = row_ptr[row + 1];
  int k = -1;
  // binary search (the patterns list a row's columns in ascending order), then a linear
  // scan should a row not be sorted
  for (int lo = s, hi = e - 1; lo <= hi;) {
    const int mid = (lo + hi) >> 1, c = col_idx[mid];
    if (c == row) {
      k = mid;
      break;
    }
    if (c < row)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  if (k < 0)
    for (int q = s; q < e; ++q)
      if (col_idx[q] == row) {
        k = q;
        break;
      }
  double M[NB][2 * NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      M[i][j] = (k >= 0) ? (double)vals[(long)k * NB * NB + i * NB + j] : (i == j ? 1.0 : 0.0);
      M[i][NB + j] = (i == j) ? 1.0 : 0.0;
    }
  // Gauss-Jordan with partial pivoting; a singular block falls back to identity rows
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    int piv = c;
    double best = fabs(M[c][c]);
#pragma unroll
    for (int i = c + 1; i < NB; ++i)
      if (fabs(M[i][c]) > best) {
        best = fabs(M[i][c]);
        piv = i;
      }
    if (best == 0.0) {
      M[c][c] = 1.0;
      piv = c;
    }
    if (piv != c)
#pragma unroll
      for (int j = 0; j < 2 * NB; ++j) {
        double t = M[c][j];
        M[c][j] = M[piv][j];
        M[piv][j] = t;
      }
    double inv = 1.0 / M[c][c];
#pragma unroll
    for (int j = 0; j < 2 * NB; ++j) M[c][j] *= inv;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i == c) continue;
      double f = M[i][c];
#pragma unroll
      for (int j = 0; j < 2 * NB; ++j) M[i][j] -= f * M[c][j];
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) dinv[(long)row * NB * NB + i * NB + j] = (TD)M[i][NB + j];
}

template <int NB, typename TM, typename TV>
__device__ __forceinline__ void apply_dinv(const TM* dinv, int row, const TV* in, TV* out) {
  double v[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) v[j] = (double)in[row * NB + j];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) s += (double)dinv[(long)row * NB * NB + i * NB + j] * v[j];
    out[row * NB + i] = (TV)s;
  }
}

// ------------------------------------------------------------------ BiCGStab kernels
// scalar slots
enum { S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_RR, S_BB, S_RV, S_TS, S_TT, S_FLAG, S_RHONEW, S_NSC };

template <int NB, typename TM, typename TV>
__global__ void k_bicg_p(int n_rows, const double* sc, const TV* r, TV* p, const TV* v, const TM* dinv, TV* y) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  double beta = sc[S_BETA], omega = sc[S_OMEGA];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    int i = row * NB + j;
    p[i] = (TV)((double)r[i] + beta * ((double)p[i] - omega * (double)v[i]));
  }
  apply_dinv<NB, TM, TV>(dinv, row, p, y);
}

template <int NB, typename TM, typename TV>
__global__ void k_bicg_s(int n_rows, const double* sc, const TV* r, const TV* v, TV* s, const TM* dinv, TV* z) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  double alpha = sc[S_ALPHA];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    int i = row * NB + j;
    s[i] = (TV)((double)r[i] - alpha * (double)v[i]);
  }
  apply_dinv<NB, TM, TV>(dinv, row, s, z);
}

template <typename TV>
__global__ void k_bicg_xr(long n, const double* sc, TV* x, const TV* y, const TV* z, TV* r, const TV* s, const TV* t,
                          const TV* rhat, double* part) {
  __shared__ double a[RED_T], b[RED_T];
  double alpha = sc[S_ALPHA], omega = sc[S_OMEGA];
  double s0 = 0.0, s1 = 0.0;
  for (long i = blockIdx.x * (long)RED_T + threadIdx.x; i < n; i += (long)gridDim.x * RED_T) {
    x[i] = (TV)((double)x[i] + alpha * (double)y[i] + omega * (double)z[i]);
    TV rv = (TV)((double)s[i] - omega * (double)t[i]);
    r[i] = rv;
    s0 += (double)rhat[i] * (double)rv;
    s1 += (double)rv * (double)rv;
  }
  a[threadIdx.x] = s0;
  b[threadIdx.x] = s1;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = b[0];
  }
}

// scalar updates (single thread)
__global__ void k_fin_alpha(double* sc) {
  double rv = sc[S_RV];
  if (rv == 0.0 || !(rv == rv)) {
    sc[S_ALPHA] = 0.0;
    sc[S_FLAG] = 1.0;
  } else {
    sc[S_ALPHA] = sc[S_RHO] / rv;
  }
}
__global__ void k_fin_omega(double* sc) {
  double tt = sc[S_TT];
  sc[S_OMEGA] = (tt > 0.0) ? sc[S_TS] / tt : 0.0;
}
__global__ void k_fin_rho(double* sc) {
  double rn = sc[S_RHONEW];
  double om = sc[S_OMEGA];
  double rho = sc[S_RHO];
  if (om == 0.0 || rho == 0.0 || !(rn == rn)) {
    sc[S_BETA] = 0.0;
    sc[S_FLAG] = 1.0;
  } else {
    sc[S_BETA] = (rn / rho) * (sc[S_ALPHA] / om);
  }
  sc[S_RHO] = rn;
}
__global__ void k_bicg_init_sc(double* sc) {
  sc[S_ALPHA] = 1.0;
  sc[S_OMEGA] = 1.0;
  sc[S_BETA] = 0.0;
  sc[S_FLAG] = 0.0;
  sc[S_RHO] = sc[S_BB];
  sc[S_RR] = sc[S_BB];
}

// restart from the current x: r = rhat = b - Ax (ax holds Ax), p = v = 0
template <typename TV>
__global__ void k_bicg_restart_vecs(long n, const TV* b, const TV* ax, TV* r, TV* rhat, TV* p, TV* v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    TV rv = (TV)((double)b[i] - (double)ax[i]);
    r[i] = rv;
    rhat[i] = rv;
    p[i] = (TV)0;
    v[i] = (TV)0;
  }
}
__global__ void k_bicg_restart_sc(double* sc) {
  sc[S_ALPHA] = 1.0;
  sc[S_OMEGA] = 1.0;
  sc[S_BETA] = 0.0;
  sc[S_FLAG] = 0.0;
  sc[S_RHO] = sc[S_RR];
}

template <typename TV>
__global__ void k_copy_zero(long n, const TV* b, TV* r, TV* rhat, TV* p, TV* v, TV* x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    TV bv = b[i];
    r[i] = bv;
    rhat[i] = bv;
    p[i] = (TV)0;
    v[i] = (TV)0;
    x[i] = (TV)0;
  }
}

// ------------------------------------------------------------------ host drivers

template <int NB, typename TM, typename TV>
static int bicgstab_impl(int n_rows, const int* rp, const int* ci, const TM* vals, const TV* b, TV* x,
                         gm_krylov_ws* ws, double tol, int maxiter, int check_every, int* iters_out,
                         double* relres_out, hipStream_t st) {
  long n = (long)n_rows * NB;
  TV *r = (TV*)ws->r, *rhat = (TV*)ws->rhat, *p = (TV*)ws->p, *v = (TV*)ws->v, *s = (TV*)ws->s, *t = (TV*)ws->t,
     *y = (TV*)ws->y, *z = (TV*)ws->z;
  TM* dinv = (TM*)ws->dinv;
  double* sc = ws->sc;
  double* part = ws->part;
  int rb = gm_blocks(n_rows, 256);
  int db = dot_blocks(n);
  int sb = spmv_blocks(n_rows, NB);
  k_diag_inv<NB, TM><<<rb, 256, 0, st>>>(n_rows, rp, ci, vals, dinv);
  k_copy_zero<TV><<<db, RED_T, 0, st>>>(n, b, r, rhat, p, v, x);
  k_dot2<TV><<<db, RED_T, 0, st>>>(n, b, b, nullptr, nullptr, part);
  k_sum_parts<<<1, RED_T, 0, st>>>(part, db, sc + S_BB, nullptr);
  k_bicg_init_sc<<<1, 1, 0, st>>>(sc);
  GM_LAUNCH_CHECK();
  double hs[S_NSC];
  GM_HIP(hipMemcpyAsync(hs, sc, sizeof(double) * S_NSC, hipMemcpyDeviceToHost, st));
  GM_HIP(hipStreamSynchronize(st));
  double bnorm = sqrt(hs[S_BB]);
  *iters_out = 0;
  *relres_out = 0.0;
  if (bnorm == 0.0 || !(bnorm == bnorm)) {
    if (!(bnorm == bnorm)) {
      gm::set_error("bicgstab: right-hand side is not finite");
      return -5;
    }
    return 0;
  }
  int it = 0, restarts = 0;
  double rel = 1.0;
  while (it < maxiter) {
    int nrun = check_every;
    if (it + nrun > maxiter) nrun = maxiter - it;
    for (int k = 0; k < nrun; ++k) {
      k_bicg_p<NB, TM, TV><<<rb, 256, 0, st>>>(n_rows, sc, r, p, v, dinv, y);
      launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, y, v, rhat, nullptr, part, st);
      k_sum_parts<<<1, RED_T, 0, st>>>(part, sb, sc + S_RV, nullptr);
      k_fin_alpha<<<1, 1, 0, st>>>(sc);
      k_bicg_s<NB, TM, TV><<<rb, 256, 0, st>>>(n_rows, sc, r, v, s, dinv, z);
      launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, z, t, s, t, part, st);
      k_sum_parts<<<1, RED_T, 0, st>>>(part, sb, sc + S_TS, sc + S_TT);
      k_fin_omega<<<1, 1, 0, st>>>(sc);
      k_bicg_xr<TV><<<db, RED_T, 0, st>>>(n, sc, x, y, z, r, s, t, rhat, part);
      k_sum_parts<<<1, RED_T, 0, st>>>(part, db, sc + S_RHONEW, sc + S_RR);
      k_fin_rho<<<1, 1, 0, st>>>(sc);
    }
    GM_LAUNCH_CHECK();
    it += nrun;
    GM_HIP(hipMemcpyAsync(hs, sc, sizeof(double) * S_NSC, hipMemcpyDeviceToHost, st));
    GM_HIP(hipStreamSynchronize(st));
    rel = sqrt(hs[S_RR]) / bnorm;
    if (!(rel == rel)) {
      gm::set_error("bicgstab: residual became NaN after %d iterations", it);
      *iters_out = it;
      *relres_out = rel;
      return -5;
    }
    if (rel <= tol) break;
    if (hs[S_FLAG] != 0.0) {
      // breakdown (rho, (rhat, v) or omega vanished): restart from the current iterate
      if (++restarts > 64) break;
      launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, x, t, nullptr, nullptr, nullptr, st);
      k_bicg_restart_vecs<TV><<<db, RED_T, 0, st>>>(n, b, t, r, rhat, p, v);
      k_dot2<TV><<<db, RED_T, 0, st>>>(n, r, r, nullptr, nullptr, part);
      k_sum_parts<<<1, RED_T, 0, st>>>(part, db, sc + S_RR, nullptr);
      k_bicg_restart_sc<<<1, 1, 0, st>>>(sc);
      GM_LAUNCH_CHECK();
    }
  }
  *iters_out = it;
  *relres_out = rel;
  return 0;
}

extern "C" int gm_bicgstab(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                           const void* b, void* x, gm_krylov_ws* ws, double tol, int maxiter, int check_every,
                           int* iters, double* relres, hipStream_t s) {
  GM_REQUIRE(ws && iters && relres, "gm_bicgstab: null argument");
  if (check_every < 1) check_every = 1;
#define GM_BICG(NB)                                                                                              \
  if (nb == NB)                                                                                                   \
    return fp32 ? bicgstab_impl<NB, float, float>(n_rows, row_ptr, col_idx, (const float*)vals, (const float*)b,   \
                                                  (float*)x, ws, tol, maxiter, check_every, iters, relres, s)      \
                : bicgstab_impl<NB, double, double>(n_rows, row_ptr, col_idx, (const double*)vals, (const double*)b, \
                                                    (double*)x, ws, tol, maxiter, check_every, iters, relres, s);
  GM_BICG(1)
  GM_BICG(2)
  GM_BICG(3)
  GM_BICG(4)
#undef GM_BICG
  gm::set_error("gm_bicgstab: unsupported block size %d", nb);
  return -1;
}

// ------------------------------------------------------------------ PCG
// Conjugate gradients (Hestenes & Stiefel 1952) with a diagonal-block preconditioner for
// symmetric positive definite systems (the consistent-mass projection): per iteration
// the SpMV q = A p with fused p.q partials, one workgroup summing them into alpha, the
// update x += alpha p, r -= alpha q, z = M^-1 r with fused r.z and r.r partials, and
// p = z + beta p, every workgroup of which sums those partials itself (in one fixed
// order) into beta; every scalar stays on the device and the host reads ||r|| every
// `check_every` iterations.  C_RZN holds the newest r.z, C_RZ the one alpha last used.
// The iteration stops on the device: the p update's workgroup 0 counts the iteration
// (C_ITS) and sets C_DONE once ||r|| meets the tolerance; the SpMV, alpha and the x, r
// update of later iterations then return at once (they read C_DONE, written by an earlier
// launch), and alpha copies C_DONE into C_GATE, which gates the p update, so that no
// launch reads a flag it writes itself.
enum { C_RZ = 0, C_RZN, C_PQ, C_RR, C_BB, C_ALPHA, C_BETA, C_FLAG, C_DONE, C_GATE, C_ITS, C_NSC };
#define PCG_WIDE 1024

// x = 0 (cold start) or kept (warm start, ax = A x), r = b - ax, z = Dinv r, p = z;
// per-workgroup partials of r.z and r.r
template <int NB, typename TM, typename TV>
__global__ void k_pcg_start(int n_rows, const TV* b, const TV* ax, const TM* dinv, TV* x, TV* r, TV* z, TV* p,
                            double* part) {
  __shared__ double a[RED_T], c[RED_T];
  double s0 = 0.0, s1 = 0.0;
  for (int row = blockIdx.x * RED_T + threadIdx.x; row < n_rows; row += gridDim.x * RED_T) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int i = row * NB + j;
      if (ax) {
        r[i] = (TV)((double)b[i] - (double)ax[i]);
      } else {
        x[i] = (TV)0;
        r[i] = b[i];
      }
    }
    apply_dinv<NB, TM, TV>(dinv, row, r, z);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double rv = (double)r[row * NB + j], zv = (double)z[row * NB + j];
      p[row * NB + j] = z[row * NB + j];
      s0 += rv * zv;
      s1 += rv * rv;
    }
  }
  a[threadIdx.x] = s0;
  c[threadIdx.x] = s1;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      c[threadIdx.x] += c[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = c[0];
  }
}

__global__ void k_pcg_init_sc(double* sc) {
  sc[C_FLAG] = 0.0;
  sc[C_DONE] = 0.0;
  sc[C_GATE] = 0.0;
  sc[C_ITS] = 0.0;
}

// one workgroup: the sums of n partial pairs in the order of k_sum_parts
__device__ void sum_pairs(const double* part, int n, double* s0, double* s1) {
  __shared__ double a[RED_T], b[RED_T];
  double u = 0.0, v = 0.0;
  for (int i = threadIdx.x; i < n; i += RED_T) {
    u += part[2 * i];
    v += part[2 * i + 1];
  }
  a[threadIdx.x] = u;
  b[threadIdx.x] = v;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  *s0 = a[0];
  *s1 = b[0];
}

// the first of every partial pair, summed by a workgroup of PCG_WIDE threads: four
// independent accumulators per thread (loads in flight), then a tree in LDS; one fixed
// order for a given n
__device__ double sum_firsts_wide(const double* part, int n) {
  __shared__ double a[PCG_WIDE];
  double u0 = 0.0, u1 = 0.0, u2 = 0.0, u3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * PCG_WIDE < n; i += 4 * PCG_WIDE) {
    u0 += part[2 * i];
    u1 += part[2 * (i + PCG_WIDE)];
    u2 += part[2 * (i + 2 * PCG_WIDE)];
    u3 += part[2 * (i + 3 * PCG_WIDE)];
  }
  for (; i < n; i += PCG_WIDE) u0 += part[2 * i];
  a[threadIdx.x] = (u0 + u1) + (u2 + u3);
  __syncthreads();
  for (int o = PCG_WIDE / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) a[threadIdx.x] += a[threadIdx.x + o];
    __syncthreads();
  }
  return a[0];
}

// p.q from the SpMV's partials; alpha = (r.z) / (p.q) with r.z the newest (C_RZN, kept
// in C_RZ for the next beta); a non-positive or non-finite p.q stops the iteration
// (flag, alpha = 0)
__global__ void __launch_bounds__(PCG_WIDE) k_pcg_alpha(const double* part, int n, double* sc) {
  if (sc[C_DONE] != 0.0) {
    if (threadIdx.x == 0) sc[C_GATE] = 1.0;
    return;
  }
  const double pq = sum_firsts_wide(part, n);
  if (threadIdx.x == 0) {
    const double rz = sc[C_RZN];
    sc[C_PQ] = pq;
    sc[C_RZ] = rz;
    const bool ok = pq > 0.0 && pq < 1e308 && sc[C_FLAG] == 0.0;
    sc[C_ALPHA] = ok ? rz / pq : 0.0;
    if (!ok) sc[C_FLAG] = 1.0;
  }
}

// x += alpha p, r -= alpha q, z = Dinv r; per-workgroup partials of r.z and r.r
template <int NB, typename TM, typename TV>
__global__ void k_pcg_xr(int n_rows, const double* sc, TV* x, TV* r, const TV* p, const TV* q, const TM* dinv,
                         TV* z, double* part) {
  if (sc[C_DONE] != 0.0) return;
  __shared__ double a[RED_T], c[RED_T];
  const double alpha = sc[C_ALPHA];
  double s0 = 0.0, s1 = 0.0;
  for (int row = blockIdx.x * RED_T + threadIdx.x; row < n_rows; row += gridDim.x * RED_T) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int i = row * NB + j;
      x[i] = (TV)((double)x[i] + alpha * (double)p[i]);
      r[i] = (TV)((double)r[i] - alpha * (double)q[i]);
    }
    apply_dinv<NB, TM, TV>(dinv, row, r, z);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double rv = (double)r[row * NB + j];
      s0 += rv * (double)z[row * NB + j];
      s1 += rv * rv;
    }
  }
  a[threadIdx.x] = s0;
  c[threadIdx.x] = s1;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      c[threadIdx.x] += c[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = c[0];
  }
}

// r.z and r.r from the update's nparts partials (every workgroup sums them in the same
// order), beta = (r.z)_new / (r.z), p = z + beta p; workgroup 0 stores r.z, r.r and beta,
// counts the iteration and sets C_DONE once r.r <= thr2 (the squared absolute tolerance)
template <typename TV>
__global__ void __launch_bounds__(RED_T) k_pcg_p(long n, const double* part, int nparts, double* sc, const TV* z,
                                                 TV* p, double thr2) {
  if (sc[C_GATE] != 0.0) return;
  double rzn, rr;
  sum_pairs(part, nparts, &rzn, &rr);
  const double rz = sc[C_RZ];
  const double beta = (rz != 0.0) ? rzn / rz : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[C_RZN] = rzn;
    sc[C_RR] = rr;
    sc[C_BETA] = beta;
    sc[C_ITS] += 1.0;
    if (rr <= thr2) sc[C_DONE] = 1.0;
  }
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = (TV)((double)z[i] + beta * (double)p[i]);
}

// lumped (row-sum) diagonal preconditioner of a scalar matrix: dinv = 1 / sum_k A_ik
// (a consistent-mass matrix has non-negative entries and row sums equal to the lumped
// nodal masses)
template <typename TM>
__global__ void k_rowsum_inv(int n_rows, const int* rp, const TM* vals, TM* dinv) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  double s = 0.0;
  for (int k = rp[row]; k < rp[row + 1]; ++k) s += (double)vals[k];
  dinv[row] = (TM)((s != 0.0) ? 1.0 / s : 1.0);
}

template <int NB, typename TM, typename TV>
static int pcg_impl(int n_rows, const int* rp, const int* ci, const TM* vals, const TV* b, TV* x, gm_krylov_ws* ws,
                    double tol, int maxiter, int check_every, int options, int* iters_out, double* relres_out,
                    hipStream_t st) {
  const bool lumped = options & 1, warm = options & 2;
  long n = (long)n_rows * NB;
  TV *r = (TV*)ws->r, *p = (TV*)ws->p, *q = (TV*)ws->v, *z = (TV*)ws->z;
  TM* dinv = (TM*)ws->dinv;
  double* sc = ws->sc;
  double* part = ws->part;
  const int rb = gm_blocks(n_rows, 256);
  const int db = std::min(1024, gm_blocks(n_rows, RED_T));
  const int sb = spmv_blocks(n_rows, NB);
  if (NB == 1 && lumped)
    k_rowsum_inv<TM><<<rb, 256, 0, st>>>(n_rows, rp, vals, dinv);
  else
    k_diag_inv<NB, TM><<<rb, 256, 0, st>>>(n_rows, rp, ci, vals, dinv);
  k_pcg_init_sc<<<1, 1, 0, st>>>(sc);
  k_dot2<TV><<<dot_blocks(n), RED_T, 0, st>>>(n, b, b, nullptr, nullptr, part);
  k_sum_parts<<<1, RED_T, 0, st>>>(part, dot_blocks(n), sc + C_BB, nullptr);
  if (warm) launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, x, q, nullptr, nullptr, nullptr, st);
  k_pcg_start<NB, TM, TV><<<db, RED_T, 0, st>>>(n_rows, b, warm ? q : nullptr, dinv, x, r, z, p, part);
  k_sum_parts<<<1, RED_T, 0, st>>>(part, db, sc + C_RZN, sc + C_RR);
  GM_LAUNCH_CHECK();
  double hs[C_NSC];
  GM_HIP(hipMemcpyAsync(hs, sc, sizeof(double) * C_NSC, hipMemcpyDeviceToHost, st));
  GM_HIP(hipStreamSynchronize(st));
  const double bnorm = sqrt(hs[C_BB]);
  *iters_out = 0;
  *relres_out = 0.0;
  if (!(bnorm == bnorm) || bnorm > 1e300) {
    gm::set_error("pcg: right-hand side is not finite");
    return -5;
  }
  if (bnorm == 0.0) {
    if (warm) k_copy_zero<TV><<<dot_blocks(n), RED_T, 0, st>>>(n, b, r, r, r, r, x);  // x = 0
    GM_LAUNCH_CHECK();
    return 0;
  }
  double rel = sqrt(fmax(hs[C_RR], 0.0)) / bnorm;
  const double thr2 = (tol * bnorm) * (tol * bnorm);
  int it = 0;
  while (it < maxiter && rel > tol) {
    const int nrun = std::min(check_every, maxiter - it);
    for (int k = 0; k < nrun; ++k) {
      launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, p, q, p, nullptr, part, st, 0, sc + C_DONE);
      k_pcg_alpha<<<1, PCG_WIDE, 0, st>>>(part, sb, sc);
      k_pcg_xr<NB, TM, TV><<<db, RED_T, 0, st>>>(n_rows, sc, x, r, p, q, dinv, z, part);
      k_pcg_p<TV><<<gm_blocks(n, RED_T), RED_T, 0, st>>>(n, part, db, sc, z, p, thr2);
    }
    GM_LAUNCH_CHECK();
    GM_HIP(hipMemcpyAsync(hs, sc, sizeof(double) * C_NSC, hipMemcpyDeviceToHost, st));
    GM_HIP(hipStreamSynchronize(st));
    it = (int)hs[C_ITS];
    rel = sqrt(fmax(hs[C_RR], 0.0)) / bnorm;
    if (!(hs[C_RR] == hs[C_RR])) {
      gm::set_error("pcg: residual became NaN after %d iterations", it);
      *iters_out = it;
      *relres_out = hs[C_RR];
      return -5;
    }
    if (hs[C_FLAG] != 0.0) break;  // p.Ap <= 0: the caller sees rel
    if (hs[C_DONE] != 0.0) break;  // converged on the device (rel may round to just above tol)
  }
  *iters_out = it;
  *relres_out = rel;
  return 0;
}

extern "C" int gm_pcg(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                      const void* b, void* x, gm_krylov_ws* ws, double tol, int maxiter, int check_every, int options,
                      int* iters, double* relres, hipStream_t s) {
  GM_REQUIRE(ws && iters && relres, "gm_pcg: null argument");
  GM_REQUIRE(!(options & 1) || nb == 1, "gm_pcg: the lumped preconditioner needs a scalar matrix (nb = 1)");
  const int lumped = options;
  if (check_every < 1) check_every = 1;
#define GM_PCG(NB)                                                                                              \
  if (nb == NB)                                                                                                 \
    return fp32 ? pcg_impl<NB, float, float>(n_rows, row_ptr, col_idx, (const float*)vals, (const float*)b,     \
                                             (float*)x, ws, tol, maxiter, check_every, lumped, iters, relres, s) \
                : pcg_impl<NB, double, double>(n_rows, row_ptr, col_idx, (const double*)vals, (const double*)b, \
                                               (double*)x, ws, tol, maxiter, check_every, lumped, iters, relres, s);
  GM_PCG(1)
  GM_PCG(2)
  GM_PCG(3)
  GM_PCG(4)
#undef GM_PCG
  gm::set_error("gm_pcg: unsupported block size %d", nb);
  return -1;
}

// y = A x  (A in FP64 or FP32, x and y in FP64 or FP32; FP64 accumulation)
extern "C" int gm_bsr_spmv(int nb, int mat_fp32, int vec_fp32, int n_rows, const int* row_ptr, const int* col_idx,
                           const void* vals, const void* x, void* y, hipStream_t s) {
#define GM_SPMV(NB)                                                                                              \
  if (nb == NB) {                                                                                                \
    if (mat_fp32 && vec_fp32)                                                                                    \
      launch_spmv<NB, float, float, float>(n_rows, row_ptr, col_idx, (const float*)vals, (const float*)x,         \
                                           (float*)y, nullptr, nullptr, nullptr, s);                             \
    else if (mat_fp32)                                                                                           \
      launch_spmv<NB, float, double, double>(n_rows, row_ptr, col_idx, (const float*)vals, (const double*)x,      \
                                             (double*)y, nullptr, nullptr, nullptr, s);                          \
    else                                                                                                         \
      launch_spmv<NB, double, double, double>(n_rows, row_ptr, col_idx, (const double*)vals, (const double*)x,    \
                                              (double*)y, nullptr, nullptr, nullptr, s);                         \
    GM_LAUNCH_CHECK();                                                                                           \
    return 0;                                                                                                    \
  }
  GM_SPMV(1)
  GM_SPMV(2)
  GM_SPMV(3)
  GM_SPMV(4)
#undef GM_SPMV
  gm::set_error("gm_bsr_spmv: unsupported block size %d", nb);
  return -1;
}

// squared 2-norm of an FP64 vector into out (device scalar); part >= 2*1024 doubles
extern "C" int gm_norm2(long n, const double* v, double* part, double* out, hipStream_t s) {
  int db = dot_blocks(n);
  k_dot2<double><<<db, RED_T, 0, s>>>(n, v, v, nullptr, nullptr, part);
  k_sum_parts<<<1, RED_T, 0, s>>>(part, db, out, nullptr);
  GM_LAUNCH_CHECK();
  return 0;
}

// sum of an FP64 vector (deterministic) into out
__global__ void k_sum1(long n, const double* v, double* part) {
  __shared__ double a[RED_T];
  double s0 = 0.0;
  for (long i = blockIdx.x * (long)RED_T + threadIdx.x; i < n; i += (long)gridDim.x * RED_T) s0 += v[i];
  a[threadIdx.x] = s0;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) a[threadIdx.x] += a[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

extern "C" int gm_sum(long n, const double* v, double* part, double* out, hipStream_t s) {
  int db = dot_blocks(n);
  k_sum1<<<db, RED_T, 0, s>>>(n, v, part);
  k_sum_parts<<<1, RED_T, 0, s>>>(part, db, out, nullptr);
  GM_LAUNCH_CHECK();
  return 0;
}

// y = a*x + y over FP64 (Newton update u += delta) or mixed (x FP32)
__global__ void k_axpy_d(long n, double a, const double* x, double* y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] += a * x[i];
}
__global__ void k_axpy_f(long n, double a, const float* x, double* y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] += a * (double)x[i];
}
extern "C" int gm_axpy(long n, double a, const void* x, int x_fp32, double* y, hipStream_t s) {
  int b = dot_blocks(n);
  if (x_fp32) k_axpy_f<<<b, RED_T, 0, s>>>(n, a, (const float*)x, y);
  else k_axpy_d<<<b, RED_T, 0, s>>>(n, a, (const double*)x, y);
  GM_LAUNCH_CHECK();
  return 0;
}

// out = b - A x  in FP64 (refinement residual; A may be FP32), and its squared norm
__global__ void k_sub(long n, const double* b, const double* ax, double* out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = b[i] - ax[i];
}
extern "C" int gm_residual_refine(int nb, int mat_fp32, int n_rows, const int* row_ptr, const int* col_idx,
                                  const void* vals, const double* b, const double* x, double* ax, double* out,
                                  hipStream_t s) {
  int rc = gm_bsr_spmv(nb, mat_fp32, 0, n_rows, row_ptr, col_idx, vals, x, ax, s);
  if (rc) return rc;
  long n = (long)n_rows * nb;
  k_sub<<<dot_blocks(n), RED_T, 0, s>>>(n, b, ax, out);
  GM_LAUNCH_CHECK();
  return 0;
}

// FP64 <-> FP32 conversions
__global__ void k_d2f(long n, const double* a, float* b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (float)a[i];
}
extern "C" int gm_d2f(long n, const double* a, float* b, hipStream_t s) {
  k_d2f<<<dot_blocks(n), RED_T, 0, s>>>(n, a, b);
  GM_LAUNCH_CHECK();
  return 0;
}

// identity rows for empty diagonals of a scalar (nb=1) matrix; zero rhs there
__global__ void k_fix_empty_diag(int n_rows, const int* row_ptr, const int* col_idx, double* vals, double* rhs) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  const int s = row_ptr[row], e = row_ptr[row + 1];
  int k = -1;
  // binary search (ascending columns), then a linear scan should a row not be sorted
  for (int lo = s, hi = e - 1; lo <= hi;) {
    const int mid = (lo + hi) >> 1, c = col_idx[mid];
    if (c == row) {
      k = mid;
      break;
    }
    if (c < row)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  if (k < 0)
    for (int q = s; q < e; ++q)
      if (col_idx[q] == row) {
        k = q;
        break;
      }
  if (k >= 0 && vals[k] == 0.0) {
    vals[k] = 1.0;
    rhs[row] = 0.0;
  }
}
extern "C" int gm_fix_empty_diag(int n_rows, const int* row_ptr, const int* col_idx, double* vals, double* rhs,
                                 hipStream_t s) {
  k_fix_empty_diag<<<gm_blocks(n_rows, 256), 256, 0, s>>>(n_rows, row_ptr, col_idx, vals, rhs);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== restarted GMRES(m)
// Right-preconditioned GMRES(m) (Saad & Schultz 1986) with block-Jacobi M:
// Arnoldi by classical Gram-Schmidt with selective reorthogonalisation on
// device-resident basis vectors; the small Hessenberg least-squares problem (Givens rotations)
// is solved on the host, which reads one column of H per iteration.
// x <- x + M^{-1} V y at the end of every cycle; restart from r = b - A x.

// partial dots of w with basis vectors V[0..k): part[i*nchunk + c]
// gate: when given, the launch does nothing unless *gate != 0 (a decision taken on the
// device inside an enqueued Arnoldi step)
template <typename TV>
__global__ void k_multidot(const double* gate, long n, int k, const TV* V, long ldv, const TV* w, int nchunk,
                           double* part) {
  if (gate && *gate == 0.0) return;
  __shared__ double sh[RED_T];
  int i = blockIdx.y;
  double s = 0.0;
  const TV* vi = V + (long)i * ldv;
  for (long r = blockIdx.x * (long)RED_T + threadIdx.x; r < n; r += (long)nchunk * RED_T)
    s += (double)vi[r] * (double)w[r];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)i * nchunk + blockIdx.x] = sh[0];
}

// h[i] = sum_c part[i*nchunk + c]  (one block per basis vector, fixed order)
__global__ void k_multidot_fin(const double* gate, int nchunk, const double* part, double* h) {
  if (gate && *gate == 0.0) return;
  __shared__ double sh[RED_T];
  int i = blockIdx.x;
  double s = 0.0;
  for (int c = threadIdx.x; c < nchunk; c += RED_T) s += part[(long)i * nchunk + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) h[i] = sh[0];
}

// w -= sum_i V_i h_i   (h on the device); one row per thread, four independent partial
// sums over the basis vectors so that four loads are in flight per thread
template <typename TV>
__global__ void __launch_bounds__(256) k_multiaxpy(const double* gate, long n, int k, const TV* V, long ldv,
                                                   const double* h, TV* w) {
  if (gate && *gate == 0.0) return;
  long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = 0;
  for (; i + 3 < k; i += 4) {
    a0 += h[i] * (double)V[(long)i * ldv + r];
    a1 += h[i + 1] * (double)V[(long)(i + 1) * ldv + r];
    a2 += h[i + 2] * (double)V[(long)(i + 2) * ldv + r];
    a3 += h[i + 3] * (double)V[(long)(i + 3) * ldv + r];
  }
  for (; i < k; ++i) a0 += h[i] * (double)V[(long)i * ldv + r];
  w[r] = (TV)((double)w[r] - ((a0 + a1) + (a2 + a3)));
}

// out = sum_i V_i y_i   (y on the device); one row per thread, four partial sums
template <typename TV>
__global__ void __launch_bounds__(256) k_combine(long n, int k, const TV* V, long ldv, const double* y, TV* out) {
  long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = 0;
  for (; i + 3 < k; i += 4) {
    a0 += y[i] * (double)V[(long)i * ldv + r];
    a1 += y[i + 1] * (double)V[(long)(i + 1) * ldv + r];
    a2 += y[i + 2] * (double)V[(long)(i + 2) * ldv + r];
    a3 += y[i + 3] * (double)V[(long)(i + 3) * ldv + r];
  }
  for (; i < k; ++i) a0 += y[i] * (double)V[(long)i * ldv + r];
  out[r] = (TV)((a0 + a1) + (a2 + a3));
}

// selective reorthogonalisation test (Daniel, Gragg, Kaufman & Stewart 1976) on the
// device: gate = 1 when the first pass removed more than half of ||w||^2 (or a norm is NaN)
__global__ void k_reorth_gate(const double* nrm_after, const double* nrm_before, double* gate) {
  *gate = (*nrm_after >= 0.5 * *nrm_before) ? 0.0 : 1.0;
}

// b = a / sqrt(*nrm2) (zero when the norm is not positive: a breakdown the host sees)
template <typename TV>
__global__ void k_scale_norm(long n, const TV* a, const double* nrm2, TV* b) {
  const double q = *nrm2;
  const double s = (q > 0.0) ? 1.0 / sqrt(q) : 0.0;
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
    b[r] = (TV)((double)a[r] * s);
}

template <typename TV>
__global__ void k_scale_copy(long n, const TV* a, double s, TV* b) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
    b[r] = (TV)((double)a[r] * s);
}

template <typename TV>
__global__ void k_sub_into(long n, const TV* b, const TV* ax, TV* r) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    r[i] = (TV)((double)b[i] - (double)ax[i]);
}

template <int NB, typename TM, typename TV>
__global__ void k_apply_dinv_add(int n_rows, const TM* dinv, const TV* in, TV* x) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  double v[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) v[j] = (double)in[row * NB + j];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) s += (double)dinv[(long)row * NB * NB + i * NB + j] * v[j];
    x[row * NB + i] = (TV)((double)x[row * NB + i] + s);
  }
}

template <int NB, typename TM, typename TV>
__global__ void k_apply_dinv(int n_rows, const TM* dinv, const TV* in, TV* out) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  apply_dinv<NB, TM, TV>(dinv, row, in, out);
}

template <typename TV>
__global__ void k_add_into(long n, const TV* a, TV* x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = (TV)((double)x[i] + (double)a[i]);
}

// ------------------------------------------------------------------ Arnoldi step graphs
// One Arnoldi step (preconditioner, operator, both Gram-Schmidt passes, normalisation,
// copy of the Hessenberg column to pinned host memory) is a fixed sequence of launches
// for a given step index j.  It is captured once per j as a HIP graph and replayed: one
// launch per step instead of one per kernel, which matters because a multigrid cycle is
// dozens of short launches.  A graph keeps the pointers and sizes it was captured with,
// so the cache is keyed on all of them (the preconditioner's included); a solve with
// another key takes another cache slot (four, least recently used dropped).  The eager
// path enqueues the same sequence, so both give the same bits.
template <typename T>
static void put_key(std::vector<char>& k, const T& v) {
  const char* p = reinterpret_cast<const char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

namespace {
int g_graphs_on = -1;  // -1: taken from GM_KRYLOV_GRAPHS ("0" turns the graphs off), default on
bool graphs_on() {
  if (g_graphs_on < 0) {
    const char* e = getenv("GM_KRYLOV_GRAPHS");
    g_graphs_on = (e && e[0] == '0') ? 0 : 1;
  }
  return g_graphs_on == 1;
}
struct StepGraphs {
  std::vector<char> key;  // empty: free slot
  std::vector<hipGraphExec_t> exec;
  unsigned long last_use = 0;
};
constexpr int N_GRAPH_SLOTS = 4;
StepGraphs g_sg[N_GRAPH_SLOTS];
unsigned long g_sg_clock = 0;
hipStream_t g_cap_stream = nullptr;
double* g_pinned = nullptr;
size_t g_pinned_n = 0;

void drop_graphs(StepGraphs& e) {
  for (auto x : e.exec)
    if (x) (void)hipGraphExecDestroy(x);
  e.exec.clear();
  e.key.clear();
}

// pinned host buffer of at least n doubles for the Hessenberg columns
int pinned_buffer(size_t n, double** out) {
  if (g_pinned_n < n) {
    for (auto& e : g_sg) drop_graphs(e);  // they copy into the old buffer
    if (g_pinned) (void)hipHostFree(g_pinned);
    g_pinned = nullptr;
    g_pinned_n = 0;
    const size_t want = std::max<size_t>(n, 1024);
    GM_HIP(hipHostMalloc((void**)&g_pinned, want * sizeof(double), hipHostMallocDefault));
    g_pinned_n = want;
  }
  *out = g_pinned;
  return 0;
}

StepGraphs* graphs_for(const std::vector<char>& key) {
  for (auto& e : g_sg)
    if (!e.key.empty() && e.key == key) {
      e.last_use = ++g_sg_clock;
      return &e;
    }
  StepGraphs* victim = nullptr;
  for (auto& e : g_sg)
    if (e.key.empty()) {
      victim = &e;
      break;
    }
  if (!victim) {
    victim = &g_sg[0];
    for (auto& e : g_sg)
      if (e.last_use < victim->last_use) victim = &e;
  }
  drop_graphs(*victim);
  victim->key = key;
  victim->last_use = ++g_sg_clock;
  return victim;
}

// run the enqueued sequence of step j on st: directly, or from its captured graph
template <typename F>
int run_step(StepGraphs* sg, int j, hipStream_t st, F&& enqueue) {
  if (!sg) return enqueue(st);
  if ((int)sg->exec.size() <= j) sg->exec.resize(j + 1, nullptr);
  if (!sg->exec[j]) {
    if (!g_cap_stream) GM_HIP(hipStreamCreateWithFlags(&g_cap_stream, hipStreamNonBlocking));
    GM_HIP(hipStreamBeginCapture(g_cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(g_cap_stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(g_cap_stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess || !graph) {
      gm::set_error("Arnoldi step capture: %s", hipGetErrorString(e));
      return -2;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      gm::set_error("Arnoldi step graph: %s", hipGetErrorString(ei));
      return -2;
    }
    sg->exec[j] = ex;
  }
  GM_HIP(hipGraphLaunch(sg->exec[j], st));
  return 0;
}
}  // namespace

// Turns the Arnoldi step graphs on (1) or off (0) for later solves; -1 only reads the
// setting.  Returns the previous setting.
extern "C" int gm_krylov_graphs(int on) {
  const int prev = graphs_on() ? 1 : 0;
  if (on == 0 || on == 1) g_graphs_on = on;
  return prev;
}

// Block-Jacobi right preconditioner: out = Dinv in; x += Dinv in.
template <int NB, typename TM, typename TV>
struct JacobiPrec {
  int n_rows;
  const TM* dinv;
  void key(std::vector<char>& k) const {
    put_key(k, 1);
    put_key(k, n_rows);
    put_key(k, dinv);
  }
  int apply(const TV* in, TV* out, hipStream_t st) const {
    k_apply_dinv<NB, TM, TV><<<gm_blocks(n_rows, 256), 256, 0, st>>>(n_rows, dinv, in, out);
    GM_LAUNCH_CHECK();
    return 0;
  }
  int apply_add(const TV* in, TV* x, TV* /*tmp*/, hipStream_t st) const {
    k_apply_dinv_add<NB, TM, TV><<<gm_blocks(n_rows, 256), 256, 0, st>>>(n_rows, dinv, in, x);
    GM_LAUNCH_CHECK();
    return 0;
  }
};

// Restarted GMRES(m), right preconditioned (prec.apply / prec.apply_add), classical
// Gram-Schmidt Arnoldi (second pass when needed) with device dot products, Givens
// rotations on the host.
// With Zs (m vectors) the method is flexible GMRES (Saad 1993): the
// preconditioned directions z_j = M^-1 v_j are kept and x += Z y, so the
// residual estimate stays exact when M^-1 is not applied identically each time
// (round-off growth inside a multigrid cycle).
template <int NB, typename TM, typename TV, typename Prec>
static int gmres_impl(int n_rows, const int* rp, const int* ci, const TM* vals, const TV* b, TV* x, TV* V, TV* z,
                      TV* w, const Prec& prec, double* part, double* dh, int m, double tol, int maxiter,
                      int* iters_out, double* relres_out, hipStream_t st, TV* Zs = nullptr) {
  long n = (long)n_rows * NB;
  long ldv = n;
  int vb = dot_blocks(n);
  // multidot chunks per basis vector: up to 256 while few vectors exist (fills the chip), within
  // the 64 (m + 1) partial sums of the workspace
  auto chunks = [m](int cols) { return std::min(256, 64 * (m + 1) / cols); };
  GM_HIP(hipMemsetAsync(x, 0, sizeof(TV) * n, st));
  // ||b||
  k_dot2<TV><<<vb, RED_T, 0, st>>>(n, b, b, nullptr, nullptr, part);
  k_sum_parts<<<1, RED_T, 0, st>>>(part, vb, dh, nullptr);
  GM_LAUNCH_CHECK();
  double hb = 0.0;
  GM_HIP(hipMemcpyAsync(&hb, dh, sizeof(double), hipMemcpyDeviceToHost, st));
  GM_HIP(hipStreamSynchronize(st));
  double bnorm = sqrt(hb);
  *iters_out = 0;
  *relres_out = 0.0;
  if (!(bnorm == bnorm) || bnorm > 1e300) {
    gm::set_error("gmres: right-hand side is not finite");
    return -5;
  }
  if (bnorm == 0.0) return 0;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hcol(m + 2);
  const int nsl = 2 * (m + 2), gate = 2 * m + 6;
  double* host = nullptr;
  {
    int prc = pinned_buffer(3 * (m + 2), &host);
    if (prc) return prc;
  }
  StepGraphs* sg = nullptr;
  if (graphs_on()) {
    std::vector<char> key;
    put_key(key, NB);
    put_key(key, (int)sizeof(TM));
    put_key(key, (int)sizeof(TV));
    put_key(key, n_rows);
    put_key(key, rp);
    put_key(key, ci);
    put_key(key, vals);
    put_key(key, V);
    put_key(key, Zs);
    put_key(key, z);
    put_key(key, w);
    put_key(key, part);
    put_key(key, dh);
    put_key(key, m);
    put_key(key, host);
    prec.key(key);
    sg = graphs_for(key);
  }
  int total = 0;
  double rel = 1.0;
  bool first = true;
  while (total < maxiter) {
    // r = b - A x  -> V[0]; on the first cycle x = 0, so r = b and ||r||^2 is the ||b||^2
    // already in hb (the same reduction of the same values)
    if (first) {
      GM_HIP(hipMemcpyAsync(V, b, sizeof(TV) * n, hipMemcpyDeviceToDevice, st));
    } else {
      launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, x, w, nullptr, nullptr, nullptr, st);
      k_sub_into<TV><<<vb, RED_T, 0, st>>>(n, b, w, V);
      k_dot2<TV><<<vb, RED_T, 0, st>>>(n, V, V, nullptr, nullptr, part);
      k_sum_parts<<<1, RED_T, 0, st>>>(part, vb, dh, nullptr);
      GM_LAUNCH_CHECK();
      GM_HIP(hipMemcpyAsync(&hb, dh, sizeof(double), hipMemcpyDeviceToHost, st));
      GM_HIP(hipStreamSynchronize(st));
    }
    double beta = sqrt(hb);
    rel = beta / bnorm;
    if (!(rel == rel)) {
      gm::set_error("gmres: residual became NaN after %d iterations", total);
      *iters_out = total;
      *relres_out = rel;
      return -5;
    }
    if (rel <= tol) break;
    first = false;
    k_scale_copy<TV><<<vb, RED_T, 0, st>>>(n, V, 1.0 / beta, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && total < maxiter; ++j, ++total) {
      // one Arnoldi step on the device: z_j = M^-1 v_j, w = A z_j, classical Gram-Schmidt
      // with selective reorthogonalisation (Daniel, Gragg, Kaufman & Stewart 1976: a second
      // pass only when the first removed more than half of ||w||^2, decided on the device),
      // v_{j+1} = w / ||w||, and the column of H copied to pinned host memory.  dh: [0, m+2)
      // pass-1 coefficients, [m+2, 2m+4) pass-2 coefficients, [2m+4] ||w||^2 after the
      // passes, [2m+5] before them, [2m+6] the second-pass flag.
      auto enqueue = [&, j](hipStream_t q) -> int {
        TV* vj = V + (long)j * ldv;
        TV* vj1 = V + (long)(j + 1) * ldv;
        TV* zj = Zs ? Zs + (long)j * ldv : z;
        int prc = prec.apply(vj, zj, q);
        if (prc) return prc;
        launch_spmv<NB, TM, TV, TV>(n_rows, rp, ci, vals, zj, w, nullptr, nullptr, nullptr, q);
        k_dot2<TV><<<vb, RED_T, 0, q>>>(n, w, w, nullptr, nullptr, part);
        k_sum_parts<<<1, RED_T, 0, q>>>(part, vb, dh + nsl + 1, nullptr);
        const int nchunk = chunks(j + 1);
        k_multidot<TV><<<dim3(nchunk, j + 1), RED_T, 0, q>>>(nullptr, n, j + 1, V, ldv, w, nchunk, part);
        k_multidot_fin<<<j + 1, RED_T, 0, q>>>(nullptr, nchunk, part, dh);
        k_multiaxpy<TV><<<gm_blocks(n, 256), 256, 0, q>>>(nullptr, n, j + 1, V, ldv, dh, w);
        k_dot2<TV><<<vb, RED_T, 0, q>>>(n, w, w, nullptr, nullptr, part);
        k_sum_parts<<<1, RED_T, 0, q>>>(part, vb, dh + nsl, nullptr);
        k_reorth_gate<<<1, 1, 0, q>>>(dh + nsl, dh + nsl + 1, dh + gate);
        k_multidot<TV><<<dim3(nchunk, j + 1), RED_T, 0, q>>>(dh + gate, n, j + 1, V, ldv, w, nchunk, part);
        k_multidot_fin<<<j + 1, RED_T, 0, q>>>(dh + gate, nchunk, part, dh + (m + 2));
        k_multiaxpy<TV><<<gm_blocks(n, 256), 256, 0, q>>>(dh + gate, n, j + 1, V, ldv, dh + (m + 2), w);
        const int nc1 = chunks(1);
        k_multidot<TV><<<dim3(nc1, 1), RED_T, 0, q>>>(dh + gate, n, 1, w, ldv, w, nc1, part);
        k_multidot_fin<<<1, RED_T, 0, q>>>(dh + gate, nc1, part, dh + nsl);
        k_scale_norm<TV><<<vb, RED_T, 0, q>>>(n, w, dh + nsl, vj1);
        GM_LAUNCH_CHECK();
        GM_HIP(hipMemcpyAsync(host, dh, sizeof(double) * 3 * (m + 2), hipMemcpyDeviceToHost, q));
        return 0;
      };
      int prc = run_step(sg, j, st, enqueue);
      if (prc) return prc;
      GM_HIP(hipStreamSynchronize(st));
      const double* hh = host;
      const bool second = hh[gate] != 0.0;
      for (int i = 0; i <= j; ++i) hcol[i] = second ? hh[i] + hh[(m + 2) + i] : hh[i];
      double hn = sqrt(fmax(hh[nsl], 0.0));
      if (!(hn == hn) || hn > 1e300 || !(hcol[j] == hcol[j])) {
        gm::set_error("gmres: Arnoldi vector became non-finite after %d iterations (preconditioner overflow)",
                      total);
        *iters_out = total;
        *relres_out = hn;
        return -6;
      }
      hcol[j + 1] = hn;
      // apply previous rotations, then a new one
      for (int i = 0; i < j; ++i) {
        double t = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
        hcol[i] = t;
      }
      double den = sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
      cs[j] = den > 0.0 ? hcol[j] / den : 1.0;
      sn[j] = den > 0.0 ? hcol[j + 1] / den : 0.0;
      hcol[j] = den;
      hcol[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hcol[i];
      rel = fabs(g[j + 1]) / bnorm;
      if (rel <= tol || hn == 0.0) {
        ++j;
        ++total;
        break;
      }
    }
    // y = H^{-1} g (upper triangular j x j), x += M^{-1} V y
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int k2 = i + 1; k2 < j; ++k2) s -= H[(size_t)i * m + k2] * y[k2];
      y[i] = s / H[(size_t)i * m + i];
    }
    GM_HIP(hipMemcpyAsync(dh, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, st));
    if (Zs) {
      k_combine<TV><<<gm_blocks(n, 256), 256, 0, st>>>(n, j, Zs, ldv, dh, w);
      k_add_into<TV><<<vb, RED_T, 0, st>>>(n, w, x);
      GM_LAUNCH_CHECK();
    } else {
      k_combine<TV><<<gm_blocks(n, 256), 256, 0, st>>>(n, j, V, ldv, dh, w);
      GM_LAUNCH_CHECK();
      int prc = prec.apply_add(w, x, z, st);
      if (prc) return prc;
    }
    GM_HIP(hipStreamSynchronize(st));
    if (rel <= tol) break;
  }
  *iters_out = total;
  *relres_out = rel;
  return 0;
}

// V: (m+1)*n_rows*nb entries of the solve precision; z, w: n_rows*nb each;
// dinv: n_rows*nb*nb; part >= 64*(m+1) doubles and >= 2048; dh >= 3*(m+2) doubles.
extern "C" int gm_gmres(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                        const void* b, void* x, void* V, void* z, void* w, void* dinv, double* part, double* dh, int m,
                        double tol, int maxiter, int* iters, double* relres, hipStream_t s) {
  GM_REQUIRE(iters && relres && m >= 1 && m <= 256, "gm_gmres: bad argument");
#define GM_GMRES(NB)                                                                                          \
  if (nb == NB) {                                                                                              \
    if (fp32) {                                                                                                \
      k_diag_inv<NB, float><<<gm_blocks(n_rows, 256), 256, 0, s>>>(n_rows, row_ptr, col_idx, (const float*)vals, \
                                                                   (float*)dinv);                              \
      JacobiPrec<NB, float, float> pc{n_rows, (const float*)dinv};                                             \
      return gmres_impl<NB, float, float>(n_rows, row_ptr, col_idx, (const float*)vals, (const float*)b,        \
                                          (float*)x, (float*)V, (float*)z, (float*)w, pc, part, dh, m, tol,     \
                                          maxiter, iters, relres, s);                                          \
    }                                                                                                          \
    k_diag_inv<NB, double><<<gm_blocks(n_rows, 256), 256, 0, s>>>(n_rows, row_ptr, col_idx, (const double*)vals, \
                                                                  (double*)dinv);                              \
    JacobiPrec<NB, double, double> pc{n_rows, (const double*)dinv};                                            \
    return gmres_impl<NB, double, double>(n_rows, row_ptr, col_idx, (const double*)vals, (const double*)b,      \
                                          (double*)x, (double*)V, (double*)z, (double*)w, pc, part, dh, m, tol, \
                                          maxiter, iters, relres, s);                                          \
  }
  GM_GMRES(1)
  GM_GMRES(2)
  GM_GMRES(3)
  GM_GMRES(4)
#undef GM_GMRES
  gm::set_error("gm_gmres: unsupported block size %d", nb);
  return -1;
}

// ================================================================== geometric multigrid
// Preconditioner for the Newton systems: V-cycles on a hierarchy of background
// grids (every second node per axis), trilinear (bilinear in 2D) prolongation P
// restricted to free dofs, Galerkin coarse operators A_c = P^T A P built on the
// device in two gathers (AP per fine row, then P^T(AP) per coarse row, fixed
// summation order), damped block-Jacobi smoothing with the exact nb x nb
// diagonal-block inverses.  The cycle is a fixed linear operator, so it serves as
// the right preconditioner of GMRES.
namespace mg {

__device__ __forceinline__ int floordiv2(int a) { return (a >= 0) ? (a >> 1) : -((-a + 1) >> 1); }

template <int DIM>
__device__ __forceinline__ void coords(int node, int nx, int ny, int* c) {
  c[0] = node % (nx + 1);
  int t = node / (nx + 1);
  if (DIM == 2) {
    c[1] = t;
    c[2] = 0;
  } else {
    c[1] = t % (ny + 1);
    c[2] = t / (ny + 1);
  }
}

template <int DIM>
__device__ __forceinline__ int node_of(const int* c, int nx, int ny) {
  return (DIM == 2) ? c[0] + c[1] * (nx + 1) : c[0] + c[1] * (nx + 1) + c[2] * (nx + 1) * (ny + 1);
}

// block index of window bit `bit` in a row, or -1 when the block is absent
__device__ __forceinline__ int block_at(const int* rp, const uint64_t* mask, int row, int bit) {
  uint64_t m0 = mask[2 * row], m1 = mask[2 * row + 1];
  if (bit < 64) {
    if (!((m0 >> bit) & 1ull)) return -1;
    return rp[row] + __popcll(m0 & ((1ull << bit) - 1ull));
  }
  int b = bit - 64;
  if (!((m1 >> b) & 1ull)) return -1;
  return rp[row] + __popcll(m0) + __popcll(m1 & ((1ull << b) - 1ull));
}

struct Dims {
  int n[3];   // cells per axis
};

template <int DIM, int NB>
__global__ void k_free(Dims f, Dims c, int n_coarse, const uint8_t* freem_f, uint8_t* freem_c) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_coarse) return;
  int A[3];
  coords<DIM>(I, c.n[0], c.n[1], A);
  uint8_t fr[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) fr[k] = 0;
  const int oz = (DIM == 3) ? 1 : 0;
  for (int dz = -oz; dz <= oz; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int a[3] = {2 * A[0] + dx, 2 * A[1] + dy, 2 * A[2] + dz};
        if (a[0] < 0 || a[0] > f.n[0] || a[1] < 0 || a[1] > f.n[1]) continue;
        if (DIM == 3 && (a[2] < 0 || a[2] > f.n[2])) continue;
        int i = node_of<DIM>(a, f.n[0], f.n[1]);
#pragma unroll
        for (int k = 0; k < NB; ++k) fr[k] |= freem_f[(long)i * NB + k];
      }
#pragma unroll
  for (int k = 0; k < NB; ++k) freem_c[(long)I * NB + k] = fr[k] ? 1 : 0;
}

// AP[i, t] = sum_j A[i, j] P[j, J(t)] (free columns only); t indexes the 4^DIM coarse
// nodes J with J_axis = floor((i_axis - 2) / 2) + t_axis.
template <int DIM, int NB, typename T>
__global__ void k_ap(Dims f, Dims c, int n_fine, const int* rp, const uint64_t* mask, const T* vals,
                     const uint8_t* freem_f, double* AP) {
  const int TT = (DIM == 2) ? 16 : 64;
  long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (gid >= (long)n_fine * TT) return;
  int i = (int)(gid / TT), t = (int)(gid % TT);
  int a[3];
  coords<DIM>(i, f.n[0], f.n[1], a);
  int tt[3] = {t & 3, (t >> 2) & 3, (t >> 4) & 3};
  int J[3] = {0, 0, 0};
  bool ok = true;
#pragma unroll
  for (int ax = 0; ax < DIM; ++ax) {
    J[ax] = floordiv2(a[ax] - 2) + tt[ax];
    if (J[ax] < 0 || J[ax] > c.n[ax]) ok = false;
  }
  double acc[NB * NB];
#pragma unroll
  for (int k = 0; k < NB * NB; ++k) acc[k] = 0.0;
  if (ok) {
    const int oz = (DIM == 3) ? 1 : 0;
    for (int dz = -oz; dz <= oz; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int j[3] = {2 * J[0] + dx, 2 * J[1] + dy, 2 * J[2] + dz};
          int o[3] = {j[0] - a[0], j[1] - a[1], j[2] - a[2]};
          bool in = true;
#pragma unroll
          for (int ax = 0; ax < DIM; ++ax)
            if (j[ax] < 0 || j[ax] > f.n[ax] || o[ax] < -2 || o[ax] > 2) in = false;
          if (!in) continue;
          int bit = (o[0] + 2) + 5 * (o[1] + 2) + ((DIM == 3) ? 25 * (o[2] + 2) : 0);
          int q = block_at(rp, mask, i, bit);
          if (q < 0) continue;
          double w = (dx == 0 ? 1.0 : 0.5) * (dy == 0 ? 1.0 : 0.5) * ((DIM == 3 && dz != 0) ? 0.5 : 1.0);
          int jn = node_of<DIM>(j, f.n[0], f.n[1]);
          const T* blk = vals + (long)q * NB * NB;
#pragma unroll
          for (int cc = 0; cc < NB; ++cc) {
            if (!freem_f[(long)jn * NB + cc]) continue;
#pragma unroll
            for (int r = 0; r < NB; ++r) acc[r * NB + cc] += w * (double)blk[r * NB + cc];
          }
        }
  }
  double* out = AP + gid * NB * NB;
#pragma unroll
  for (int k = 0; k < NB * NB; ++k) out[k] = acc[k];
}

// A_c[I, J] = sum_i P[i, I] AP[i, J] over free fine rows; identity on coarse dofs without free fine dofs.
template <int DIM, int NB>
__global__ void k_rap(Dims f, Dims c, int n_coarse, const int* rp_c, const uint64_t* mask_c, double* vals_c,
                      const uint8_t* freem_f, const uint8_t* freem_c, const double* AP) {
  const int TT = (DIM == 2) ? 16 : 64;
  const int W = (DIM == 2) ? 25 : 125;
  long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (gid >= (long)n_coarse * W) return;
  int I = (int)(gid / W), s = (int)(gid % W);
  int q = block_at(rp_c, mask_c, I, s);
  if (q < 0) return;
  int A[3];
  coords<DIM>(I, c.n[0], c.n[1], A);
  int so[3] = {s % 5 - 2, (s / 5) % 5 - 2, (DIM == 3) ? s / 25 - 2 : 0};
  int J[3] = {A[0] + so[0], A[1] + so[1], A[2] + so[2]};
  double acc[NB * NB];
#pragma unroll
  for (int k = 0; k < NB * NB; ++k) acc[k] = 0.0;
  const int oz = (DIM == 3) ? 1 : 0;
  for (int dz = -oz; dz <= oz; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int a[3] = {2 * A[0] + dx, 2 * A[1] + dy, 2 * A[2] + dz};
        bool in = true;
        int t = 0, mul = 1;
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
          if (a[ax] < 0 || a[ax] > f.n[ax]) in = false;
          int ta = J[ax] - floordiv2(a[ax] - 2);
          if (ta < 0 || ta > 3) in = false;
          t += ta * mul;
          mul *= 4;
        }
        if (!in) continue;
        double w = (dx == 0 ? 1.0 : 0.5) * (dy == 0 ? 1.0 : 0.5) * ((DIM == 3 && dz != 0) ? 0.5 : 1.0);
        int i = node_of<DIM>(a, f.n[0], f.n[1]);
        const double* blk = AP + ((long)i * TT + t) * NB * NB;
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          if (!freem_f[(long)i * NB + r]) continue;
#pragma unroll
          for (int cc = 0; cc < NB; ++cc) acc[r * NB + cc] += w * (double)blk[r * NB + cc];
        }
      }
  if (so[0] == 0 && so[1] == 0 && so[2] == 0)
#pragma unroll
    for (int r = 0; r < NB; ++r)
      if (!freem_c[(long)I * NB + r]) acc[r * NB + r] = 1.0;
  double* out = vals_c + (long)q * NB * NB;
#pragma unroll
  for (int k = 0; k < NB * NB; ++k) out[k] = acc[k];
}

// b_c = P^T r_f (free fine dofs only)
template <int DIM, int NB, typename T>
// one thread per coarse dof (consecutive threads: the dofs of a node, then the next node)
__global__ void k_restrict(Dims f, Dims c, int n_coarse, const uint8_t* freem_f, const T* r_f, T* b_c) {
  long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (gid >= (long)n_coarse * NB) return;
  int I = (int)(gid / NB), k = (int)(gid % NB);
  int A[3];
  coords<DIM>(I, c.n[0], c.n[1], A);
  double acc = 0.0;
  const int oz = (DIM == 3) ? 1 : 0;
  for (int dz = -oz; dz <= oz; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int a[3] = {2 * A[0] + dx, 2 * A[1] + dy, 2 * A[2] + dz};
        if (a[0] < 0 || a[0] > f.n[0] || a[1] < 0 || a[1] > f.n[1]) continue;
        if (DIM == 3 && (a[2] < 0 || a[2] > f.n[2])) continue;
        double w = (dx == 0 ? 1.0 : 0.5) * (dy == 0 ? 1.0 : 0.5) * ((DIM == 3 && dz != 0) ? 0.5 : 1.0);
        int i = node_of<DIM>(a, f.n[0], f.n[1]);
        if (freem_f[(long)i * NB + k]) acc += w * (double)r_f[(long)i * NB + k];
      }
  b_c[gid] = (T)acc;
}

// x_f += P x_c (free fine dofs only); one thread per fine dof
template <int DIM, int NB, typename T>
__global__ void k_prolong_add(Dims f, Dims c, int n_fine, const uint8_t* freem_f, const T* x_c, T* x_f) {
  long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (gid >= (long)n_fine * NB) return;
  int i = (int)(gid / NB), k = (int)(gid % NB);
  if (!freem_f[gid]) return;
  int a[3];
  coords<DIM>(i, f.n[0], f.n[1], a);
  int lo[3], hi[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    lo[ax] = a[ax] >> 1;
    hi[ax] = (a[ax] + 1) >> 1;  // == lo for even a
  }
  double acc = 0.0;
  const int zc = (DIM == 3) ? 2 : 1;
  for (int pz = 0; pz < zc; ++pz)
    for (int py = 0; py < 2; ++py)
      for (int px = 0; px < 2; ++px) {
        int A[3] = {px ? hi[0] : lo[0], py ? hi[1] : lo[1], pz ? hi[2] : lo[2]};
        // even coordinate: one parent (skip the duplicate), odd: two parents of weight 0.5
        if ((px && hi[0] == lo[0]) || (py && hi[1] == lo[1]) || (DIM == 3 && pz && hi[2] == lo[2])) continue;
        double w = ((a[0] & 1) ? 0.5 : 1.0) * ((a[1] & 1) ? 0.5 : 1.0) * ((DIM == 3 && (a[2] & 1)) ? 0.5 : 1.0);
        int J = node_of<DIM>(A, c.n[0], c.n[1]);
        acc += w * (double)x_c[(long)J * NB + k];
      }
  x_f[gid] = (T)((double)x_f[gid] + acc);
}

// x = omega Dinv b; one thread per dof
template <int NB, typename T>
__global__ void k_jac_first(int n_rows, const T* dinv, const T* b, double omega, T* x) {
  long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (gid >= (long)n_rows * NB) return;
  long row = gid / NB;
  int r = (int)(gid % NB);
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < NB; ++j) s += (double)dinv[row * NB * NB + r * NB + j] * (double)b[row * NB + j];
  x[gid] = (T)(omega * s);
}

// x = Minv b for the coarsest level's dense inverse: one wavefront per row, lanes
// striding the row (coalesced), fixed-order shuffle reduction
template <typename T>
__global__ void __launch_bounds__(256) k_dense_mv(int n, const double* Minv, const T* b, T* x) {
  int i = blockIdx.x * 4 + threadIdx.x / 64;
  int lane = threadIdx.x % 64;
  double s = 0.0;
  if (i < n)
    for (int j = lane; j < n; j += 64) s += Minv[(long)i * n + j] * (double)b[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (i < n && lane == 0) x[i] = (T)s;
}

template <typename T>
__global__ void k_axpy_t(long n, const T* a, T* x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = (T)((double)x[i] + (double)a[i]);
}


inline Dims dims_of(const gm_mg_level& l) {
  Dims d;
  d.n[0] = l.nx;
  d.n[1] = l.ny;
  d.n[2] = l.nz;
  return d;
}

struct Cycle {
  int dim, n_levels, nu_pre, nu_post, coarse_sweeps;
  double omega;
  const gm_mg_level* L;
};

// everything a captured cycle depends on: its parameters and every level's pointers and
// sizes (a level's nnzb only sizes host-side work, so it is left out)
inline void cycle_key(const Cycle& cy, std::vector<char>& k) {
  put_key(k, cy.dim);
  put_key(k, cy.n_levels);
  put_key(k, cy.nu_pre);
  put_key(k, cy.nu_post);
  put_key(k, cy.coarse_sweeps);
  put_key(k, cy.omega);
  for (int l = 0; l < cy.n_levels; ++l) {
    const gm_mg_level& v = cy.L[l];
    put_key(k, v.nx);
    put_key(k, v.ny);
    put_key(k, v.nz);
    put_key(k, v.n_nodes);
    put_key(k, v.row_ptr);
    put_key(k, v.col_idx);
    put_key(k, v.row_mask);
    put_key(k, v.vals);
    put_key(k, v.dinv);
    put_key(k, v.freem);
    put_key(k, v.x);
    put_key(k, v.b);
    put_key(k, v.r);
    put_key(k, v.dense_inv);
    put_key(k, v.gamma);
    put_key(k, v.vals_lo);
    put_key(k, v.row_scale);
  }
}

// A fused level operation (k_level MODE) on level l: from the FP16 copy where the level has
// one (FP32 hierarchies); level 0 holds the Jacobian in the solve precision, coarser levels FP64
template <int NB, int MODE, typename T>
static void level_op(const gm_mg_level& lv, int l, const T* x, const T* b, const T* dinv, double omega, T* out,
                     hipStream_t st) {
  constexpr int G = spmv_lanes(NB);
  int blocks = gm_blocks(lv.n_nodes, 256 / G);
  if constexpr (std::is_same<T, float>::value) {
    if (lv.vals_lo) {
      k_level<NB, G, MODE, float, float, true><<<blocks, 256, 0, st>>>(
          lv.n_nodes, lv.row_ptr, lv.col_idx, nullptr, (const gm_half*)lv.vals_lo, lv.row_scale, x, b, dinv, omega,
          out);
      return;
    }
  }
  if (l == 0)
    k_level<NB, G, MODE, T, T, false><<<blocks, 256, 0, st>>>(lv.n_nodes, lv.row_ptr, lv.col_idx, (const T*)lv.vals,
                                                               nullptr, nullptr, x, b, dinv, omega, out);
  else
    k_level<NB, G, MODE, double, T, false><<<blocks, 256, 0, st>>>(lv.n_nodes, lv.row_ptr, lv.col_idx,
                                                                    (const double*)lv.vals, nullptr, nullptr, x, b,
                                                                    dinv, omega, out);
}

// `count` damped block-Jacobi sweeps on x, alternating between x and the level's work
// vector r (a sweep cannot update in place); the result ends in x.
template <int NB, typename T>
static int smooth(const gm_mg_level& lv, int l, const T* b, T* x, int count, double omega, hipStream_t st) {
  T* src = x;
  T* dst = (T*)lv.r;
  const T* dinv = (const T*)lv.dinv;
  for (int k = 0; k < count; ++k) {
    level_op<NB, 2, T>(lv, l, src, b, dinv, omega, dst, st);
    T* t = src;
    src = dst;
    dst = t;
  }
  GM_LAUNCH_CHECK();
  if (src != x) GM_HIP(hipMemcpyAsync(x, src, sizeof(T) * (size_t)lv.n_nodes * NB, hipMemcpyDeviceToDevice, st));
  return 0;
}

template <int DIM, int NB, typename T>
static int vcycle(const Cycle& cy, int l, const T* b, T* x, hipStream_t st) {
  const gm_mg_level& lv = cy.L[l];
  int n = lv.n_nodes;
  const T* dinv = (const T*)lv.dinv;
  T* r = (T*)lv.r;
  bool coarsest = (l == cy.n_levels - 1);
  if (coarsest && lv.dense_inv) {
    int nd = n * NB;
    k_dense_mv<T><<<gm_blocks(nd, 4), 256, 0, st>>>(nd, (const double*)lv.dense_inv, b, x);
    GM_LAUNCH_CHECK();
    return 0;
  }
  int sweeps = coarsest ? cy.coarse_sweeps : cy.nu_pre;
  k_jac_first<NB, T><<<gm_blocks((long)n * NB, 256), 256, 0, st>>>(n, dinv, b, cy.omega, x);
  GM_LAUNCH_CHECK();
  int rc = smooth<NB, T>(lv, l, b, x, sweeps - 1, cy.omega, st);
  if (rc) return rc;
  if (coarsest) return 0;
  const gm_mg_level& cv = cy.L[l + 1];
  const int gamma = lv.gamma > 1 ? lv.gamma : 1;  // 1: V-cycle, 2: W-cycle at this level
  for (int g = 0; g < gamma; ++g) {
    level_op<NB, 1, T>(lv, l, x, b, nullptr, 0.0, r, st);  // r = b - A x
    k_restrict<DIM, NB, T><<<gm_blocks((long)cv.n_nodes * NB, 256), 256, 0, st>>>(
        dims_of(lv), dims_of(cv), cv.n_nodes, lv.freem, r, (T*)cv.b);
    GM_LAUNCH_CHECK();
    rc = vcycle<DIM, NB, T>(cy, l + 1, (const T*)cv.b, (T*)cv.x, st);
    if (rc) return rc;
    k_prolong_add<DIM, NB, T><<<gm_blocks((long)n * NB, 256), 256, 0, st>>>(dims_of(lv), dims_of(cv), n, lv.freem,
                                                                             (const T*)cv.x, x);
  }
  return smooth<NB, T>(lv, l, b, x, cy.nu_post, cy.omega, st);
}

template <int DIM, int NB, typename T>
static int setup(int n_levels, gm_mg_level* L, void* ap_work, hipStream_t st) {
  for (int l = 0; l + 1 < n_levels; ++l) {
    gm_mg_level& f = L[l];
    gm_mg_level& c = L[l + 1];
    Dims df = dims_of(f), dc = dims_of(c);
    k_free<DIM, NB><<<gm_blocks(c.n_nodes, 256), 256, 0, st>>>(df, dc, c.n_nodes, f.freem, c.freem);
    const int TT = (DIM == 2) ? 16 : 64;
    const int W = (DIM == 2) ? 25 : 125;
    long nap = (long)f.n_nodes * TT;
    // Galerkin products in FP64 (the AP workspace and the coarse operators are FP64 in
    // both precisions: FP32 rounding of the cancelling sums corrupts deep coarse levels)
    if (l == 0)
      k_ap<DIM, NB, T><<<gm_blocks(nap, 256), 256, 0, st>>>(df, dc, f.n_nodes, f.row_ptr, f.row_mask,
                                                             (const T*)f.vals, f.freem, (double*)ap_work);
    else
      k_ap<DIM, NB, double><<<gm_blocks(nap, 256), 256, 0, st>>>(df, dc, f.n_nodes, f.row_ptr, f.row_mask,
                                                                  (const double*)f.vals, f.freem, (double*)ap_work);
    long nrap = (long)c.n_nodes * W;
    k_rap<DIM, NB><<<gm_blocks(nrap, 256), 256, 0, st>>>(df, dc, c.n_nodes, c.row_ptr, c.row_mask,
                                                          (double*)c.vals, f.freem, c.freem, (const double*)ap_work);
    GM_LAUNCH_CHECK();
  }
  for (int l = 0; l < n_levels; ++l) {
    int nl = L[l].n_nodes;
    if (l == 0)
      k_diag_inv<NB, T, T><<<gm_blocks(nl, 256), 256, 0, st>>>(nl, L[l].row_ptr, L[l].col_idx, (const T*)L[l].vals,
                                                                (T*)L[l].dinv);
    else
      k_diag_inv<NB, double, T><<<gm_blocks(nl, 256), 256, 0, st>>>(nl, L[l].row_ptr, L[l].col_idx,
                                                                     (const double*)L[l].vals, (T*)L[l].dinv);
  }
  if constexpr (std::is_same<T, float>::value) {
    for (int l = 0; l < n_levels; ++l) {
      if (!L[l].vals_lo) continue;
      if (l == 0)
        launch_lo_convert<NB, float>(L[l].n_nodes, L[l].row_ptr, (const float*)L[l].vals, (gm_half*)L[l].vals_lo,
                                     L[l].row_scale, st);
      else
        launch_lo_convert<NB, double>(L[l].n_nodes, L[l].row_ptr, (const double*)L[l].vals,
                                      (gm_half*)L[l].vals_lo, L[l].row_scale, st);
    }
  }
  GM_LAUNCH_CHECK();
  return 0;
}

}  // namespace mg

static int mg_check(int dim, int nb, int n_levels, const gm_mg_level* levels, int fp32) {
  GM_REQUIRE(levels && n_levels >= 1 && n_levels <= 16, "multigrid: bad level list");
  GM_REQUIRE(dim == 2 || dim == 3, "multigrid: dim must be 2 or 3");
  GM_REQUIRE(nb >= 1 && nb <= 4, "multigrid: nb must be 1..4");
  for (int l = 0; l < n_levels; ++l) {
    const gm_mg_level& v = levels[l];
    GM_REQUIRE(v.row_ptr && v.col_idx && v.row_mask && v.vals && v.dinv && v.freem && v.r,
               "multigrid: level array missing");
    long expect = (long)(v.nx + 1) * (v.ny + 1) * (dim == 3 ? (v.nz + 1) : 1);
    GM_REQUIRE(expect == v.n_nodes, "multigrid: n_nodes does not match the level's grid");
    GM_REQUIRE(!v.vals_lo || (fp32 && v.row_scale), "multigrid: an FP16 level copy needs an FP32 hierarchy "
               "and row scales");
    if (l > 0) {
      const gm_mg_level& f = levels[l - 1];
      GM_REQUIRE(v.x && v.b, "multigrid: coarse level vectors missing");
      GM_REQUIRE(v.nx == (f.nx + 1) / 2 && v.ny == (f.ny + 1) / 2 && (dim == 2 || v.nz == (f.nz + 1) / 2),
                 "multigrid: coarse grid is not ceil(n/2) of the finer one");
    }
  }
  return 0;
}

#define GM_MG_DISPATCH(CALL)                                  \
  do {                                                        \
    if (dim == 2) {                                           \
      if (nb == 1) CALL(2, 1);                                \
      if (nb == 2) CALL(2, 2);                                \
      if (nb == 3) CALL(2, 3);                                \
      if (nb == 4) CALL(2, 4);                                \
    } else {                                                  \
      if (nb == 1) CALL(3, 1);                                \
      if (nb == 2) CALL(3, 2);                                \
      if (nb == 3) CALL(3, 3);                                \
      if (nb == 4) CALL(3, 4);                                \
    }                                                         \
  } while (0)

extern "C" int gm_mg_setup(int dim, int nb, int fp32, int n_levels, gm_mg_level* levels, void* ap_work,
                           hipStream_t s) {
  int rc = mg_check(dim, nb, n_levels, levels, fp32);
  if (rc) return rc;
  GM_REQUIRE(n_levels == 1 || ap_work, "gm_mg_setup: missing AP workspace");
#define GM_SETUP(D, B)                                                                      \
  return fp32 ? mg::setup<D, B, float>(n_levels, levels, ap_work, s)                       \
              : mg::setup<D, B, double>(n_levels, levels, ap_work, s)
  GM_MG_DISPATCH(GM_SETUP);
#undef GM_SETUP
  return -1;
}

// One smoother-side operation on one level (the fused kernel the V-cycle runs):
// mode 1: out = b - A x; mode 2: out = x + omega Dinv (b - A x).
extern "C" int gm_mg_level_op(int nb, int fp32, const gm_mg_level* level, int level_index, int mode, const void* x,
                              const void* b, double omega, void* out, hipStream_t s) {
  GM_REQUIRE(level && x && b && out && level->row_ptr && level->col_idx && level->vals,
             "gm_mg_level_op: null argument");
  GM_REQUIRE(nb >= 1 && nb <= 4, "gm_mg_level_op: nb must be 1..4");
  GM_REQUIRE(mode == 1 || mode == 2, "gm_mg_level_op: mode must be 1 (residual) or 2 (Jacobi sweep)");
  GM_REQUIRE(mode == 1 || level->dinv, "gm_mg_level_op: the Jacobi sweep needs the level's dinv");
  GM_REQUIRE(!level->vals_lo || (fp32 && level->row_scale), "gm_mg_level_op: FP16 copy needs FP32 and scales");
  GM_REQUIRE(x != out, "gm_mg_level_op: out must not alias x");
#define GM_LOP(B)                                                                                              \
  do {                                                                                                         \
    if (fp32) {                                                                                                \
      if (mode == 1)                                                                                           \
        mg::level_op<B, 1, float>(*level, level_index, (const float*)x, (const float*)b, nullptr, omega,       \
                                  (float*)out, s);                                                             \
      else                                                                                                     \
        mg::level_op<B, 2, float>(*level, level_index, (const float*)x, (const float*)b,                       \
                                  (const float*)level->dinv, omega, (float*)out, s);                           \
    } else {                                                                                                   \
      if (mode == 1)                                                                                           \
        mg::level_op<B, 1, double>(*level, level_index, (const double*)x, (const double*)b, nullptr, omega,    \
                                   (double*)out, s);                                                           \
      else                                                                                                     \
        mg::level_op<B, 2, double>(*level, level_index, (const double*)x, (const double*)b,                    \
                                   (const double*)level->dinv, omega, (double*)out, s);                        \
    }                                                                                                          \
  } while (0)
  if (nb == 1) GM_LOP(1);
  if (nb == 2) GM_LOP(2);
  if (nb == 3) GM_LOP(3);
  if (nb == 4) GM_LOP(4);
#undef GM_LOP
  GM_LAUNCH_CHECK();
  return 0;
}

extern "C" int gm_mg_vcycle(int dim, int nb, int fp32, int n_levels, const gm_mg_level* levels, int nu_pre,
                            int nu_post, int coarse_sweeps, double omega, const void* b, void* x, hipStream_t s) {
  int rc = mg_check(dim, nb, n_levels, levels, fp32);
  if (rc) return rc;
  GM_REQUIRE(b && x && nu_pre >= 1 && nu_post >= 0 && coarse_sweeps >= 1, "gm_mg_vcycle: bad argument");
  mg::Cycle cy{dim, n_levels, nu_pre, nu_post, coarse_sweeps, omega, levels};
#define GM_CYC(D, B)                                                                          \
  return fp32 ? mg::vcycle<D, B, float>(cy, 0, (const float*)b, (float*)x, s)                \
              : mg::vcycle<D, B, double>(cy, 0, (const double*)b, (double*)x, s)
  GM_MG_DISPATCH(GM_CYC);
#undef GM_CYC
  return -1;
}

// ---- GMRES with the multigrid V-cycle as right preconditioner
namespace mg {
template <int DIM, int NB, typename T>
struct MGPrec {
  Cycle cy;
  long n;
  void key(std::vector<char>& k) const {
    put_key(k, 2);
    cycle_key(cy, k);
    put_key(k, n);
  }
  int apply(const T* in, T* out, hipStream_t st) const { return vcycle<DIM, NB, T>(cy, 0, in, out, st); }
  int apply_add(const T* in, T* x, T* tmp, hipStream_t st) const {
    int rc = vcycle<DIM, NB, T>(cy, 0, in, tmp, st);
    if (rc) return rc;
    k_axpy_t<T><<<dot_blocks(n), RED_T, 0, st>>>(n, tmp, x);
    GM_LAUNCH_CHECK();
    return 0;
  }
};
}  // namespace mg

extern "C" int gm_gmres_mg(int dim, int nb, int fp32, int n_levels, const gm_mg_level* levels, const void* op_vals,
                           int nu_pre, int nu_post, int coarse_sweeps, double omega, const void* b, void* x, void* V,
                           void* Z, void* z, void* w, double* part, double* dh, int m, double tol, int maxiter,
                           int* iters, double* relres, hipStream_t s) {
  int rc = mg_check(dim, nb, n_levels, levels, fp32);
  if (rc) return rc;
  GM_REQUIRE(iters && relres && m >= 1 && m <= 256 && b && x && V && Z && z && w, "gm_gmres_mg: bad argument");
  GM_REQUIRE(nu_pre >= 1 && nu_post >= 0 && coarse_sweeps >= 1, "gm_gmres_mg: bad smoother counts");
  mg::Cycle cy{dim, n_levels, nu_pre, nu_post, coarse_sweeps, omega, levels};
  const gm_mg_level& l0 = levels[0];
  long n = (long)l0.n_nodes * nb;
  // the Krylov operator: op_vals (same pattern as level 0) when the hierarchy is built from
  // a modified copy of the system matrix, else level 0's own matrix
  const void* opv = op_vals ? op_vals : l0.vals;
#define GM_GMG(D, B)                                                                                            \
  do {                                                                                                          \
    if (fp32) {                                                                                                 \
      mg::MGPrec<D, B, float> pc{cy, n};                                                                        \
      return gmres_impl<B, float, float>(l0.n_nodes, l0.row_ptr, l0.col_idx, (const float*)opv,                 \
                                         (const float*)b, (float*)x, (float*)V, (float*)z, (float*)w, pc, part, \
                                         dh, m, tol, maxiter, iters, relres, s, (float*)Z);                     \
    }                                                                                                           \
    mg::MGPrec<D, B, double> pc{cy, n};                                                                         \
    return gmres_impl<B, double, double>(l0.n_nodes, l0.row_ptr, l0.col_idx, (const double*)opv,                \
                                         (const double*)b, (double*)x, (double*)V, (double*)z, (double*)w, pc,  \
                                         part, dh, m, tol, maxiter, iters, relres, s, (double*)Z);              \
  } while (0)
  GM_MG_DISPATCH(GM_GMG);
#undef GM_GMG
  return -1;
}

// ================================================================== u-p block preconditioner
// For the coupled (u-p) Newton systems K = [[A, G], [D, C]] (A: nbu x nbu
// displacement blocks, one pressure dof per node), a block lower-triangular
// right preconditioner:
//     z_u = MG(A)^-1 v_u,     z_p = S^-1 (v_p - D z_u),
// with the displacement block solved by a multigrid V-cycle (elasticity) and
// the Schur complement S = C - D A^-1 G approximated by its diagonal with A
// replaced by its block diagonal (per node: C_II - sum_J D_IJ A_JJ^-1 G_JI).
namespace up {

// u-block of every block (rows/cols 0..NBU-1) and the u-dof free mask
template <int NB, typename T>
__global__ void k_extract(int n_rows, int nnzb, const T* vals, const uint8_t* freem, T* vals_u, uint8_t* freem_u) {
  constexpr int NBU = NB - 1;
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < nnzb) {
#pragma unroll
    for (int r = 0; r < NBU; ++r)
#pragma unroll
      for (int c = 0; c < NBU; ++c) vals_u[q * NBU * NBU + r * NBU + c] = vals[q * NB * NB + r * NB + c];
  }
  if (q < n_rows)
#pragma unroll
    for (int c = 0; c < NBU; ++c) freem_u[q * NBU + c] = freem[q * NB + c];
}

// sinv[I] = 1 / (C_II - sum_J D_IJ Ainv_JJ G_JI); fixed / inactive pressure dofs (identity rows) give 1.
template <int NB, typename T>
__global__ void k_schur_diag(int n_rows, const int* rp, const int* ci, const T* vals, const T* dinv_u,
                             const uint8_t* freem, T* sinv) {
  constexpr int NBU = NB - 1;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_rows) return;
  double s = 0.0, cii = 1.0;
  for (int q = rp[I]; q < rp[I + 1]; ++q) {
    int J = ci[q];
    const T* b = vals + (long)q * NB * NB;
    if (J == I) cii = (double)b[(NB - 1) * NB + (NB - 1)];
    double d[NBU];
    bool any = false;
#pragma unroll
    for (int c = 0; c < NBU; ++c) {
      d[c] = (double)b[(NB - 1) * NB + c];
      any |= (d[c] != 0.0);
    }
    if (!any) continue;
    // G_JI: block (J, I), column NB-1, rows 0..NBU-1 (pattern is structurally symmetric)
    int lo = rp[J], hi = rp[J + 1] - 1, qt = -1;
    while (lo <= hi) {
      int mid = (lo + hi) >> 1;
      int cc = ci[mid];
      if (cc == I) {
        qt = mid;
        break;
      }
      if (cc < I) lo = mid + 1;
      else hi = mid - 1;
    }
    if (qt < 0) continue;
    const T* bt = vals + (long)qt * NB * NB;
    const T* ai = dinv_u + (long)J * NBU * NBU;
#pragma unroll
    for (int r = 0; r < NBU; ++r) {
      double t = 0.0;
#pragma unroll
      for (int c = 0; c < NBU; ++c) t += (double)ai[r * NBU + c] * (double)bt[c * NB + (NB - 1)];
      s += d[r] * t;
    }
  }
  double S = freem[(long)I * NB + NB - 1] ? cii - s : 1.0;
  sinv[I] = (T)((S != 0.0) ? 1.0 / S : 1.0);
}

template <int NB, typename T>
__global__ void k_split(int n_rows, const T* v, T* vu) {
  constexpr int NBU = NB - 1;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_rows) return;
#pragma unroll
  for (int c = 0; c < NBU; ++c) vu[(long)I * NBU + c] = v[(long)I * NB + c];
}

// z_p = sinv (v_p - D z_u); z = (z_u, z_p) interleaved (z_u may alias nothing in z)
template <int NB, typename T>
__global__ void k_pressure(int n_rows, const int* rp, const int* ci, const T* vals, const T* sinv, const T* v,
                           const T* zu, T* z) {
  constexpr int NBU = NB - 1;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_rows) return;
  double s = 0.0;
  for (int q = rp[I]; q < rp[I + 1]; ++q) {
    int J = ci[q];
    const T* b = vals + (long)q * NB * NB + (NB - 1) * NB;
#pragma unroll
    for (int c = 0; c < NBU; ++c) s += (double)b[c] * (double)zu[(long)J * NBU + c];
  }
#pragma unroll
  for (int c = 0; c < NBU; ++c) z[(long)I * NB + c] = zu[(long)I * NBU + c];
  z[(long)I * NB + NB - 1] = (T)((double)sinv[I] * ((double)v[(long)I * NB + NB - 1] - s));
}

template <int DIM, int NB, typename T>
struct BlockPrec {
  mg::Cycle cy;  // multigrid on the displacement block (nb = NB - 1)
  int n_rows;
  const int* rp;
  const int* ci;
  const T* vals;
  const T* sinv;
  T* vu;  // n_rows * (NB-1)
  T* zu;  // n_rows * (NB-1)
  long n;
  void key(std::vector<char>& k) const {
    put_key(k, 3);
    mg::cycle_key(cy, k);
    put_key(k, n_rows);
    put_key(k, rp);
    put_key(k, ci);
    put_key(k, vals);
    put_key(k, sinv);
    put_key(k, vu);
    put_key(k, zu);
    put_key(k, n);
  }
  int apply(const T* in, T* out, hipStream_t st) const {
    int rb = gm_blocks(n_rows, 256);
    k_split<NB, T><<<rb, 256, 0, st>>>(n_rows, in, vu);
    GM_LAUNCH_CHECK();
    int rc = mg::vcycle<DIM, NB - 1, T>(cy, 0, vu, zu, st);
    if (rc) return rc;
    k_pressure<NB, T><<<rb, 256, 0, st>>>(n_rows, rp, ci, vals, sinv, in, zu, out);
    GM_LAUNCH_CHECK();
    return 0;
  }
  int apply_add(const T* in, T* x, T* tmp, hipStream_t st) const {
    int rc = apply(in, tmp, st);
    if (rc) return rc;
    mg::k_axpy_t<T><<<dot_blocks(n), RED_T, 0, st>>>(n, tmp, x);
    GM_LAUNCH_CHECK();
    return 0;
  }
};

}  // namespace up

// Extract the displacement block into u-level 0 (levels[0].vals / freem) of a
// multigrid hierarchy with nb - 1 dofs per node, build that hierarchy, and the
// inverse diagonal Schur approximation sinv (n_rows values).
extern "C" int gm_up_setup(int dim, int nb, int fp32, int n_rows, int nnzb, const int* row_ptr, const int* col_idx,
                           const void* vals, const uint8_t* freem, int n_levels, gm_mg_level* ulevels, void* ap_work,
                           void* sinv, hipStream_t s) {
  GM_REQUIRE(nb == 3 || nb == 4, "gm_up_setup: nb must be 3 (2D u-p) or 4 (3D u-p)");
  GM_REQUIRE(dim == nb - 1, "gm_up_setup: nb must be dim + 1");
  GM_REQUIRE(row_ptr && col_idx && vals && freem && sinv && ulevels, "gm_up_setup: null argument");
  int rc = mg_check(dim, nb - 1, n_levels, ulevels, fp32);
  if (rc) return rc;
  GM_REQUIRE(ulevels[0].n_nodes == n_rows && ulevels[0].row_ptr == row_ptr && ulevels[0].col_idx == col_idx,
             "gm_up_setup: u-level 0 must share the system's pattern");
  long nt = nnzb > n_rows ? nnzb : n_rows;
#define GM_UPS(NB_, T)                                                                                              \
  do {                                                                                                              \
    up::k_extract<NB_, T><<<gm_blocks(nt, 256), 256, 0, s>>>(n_rows, nnzb, (const T*)vals, freem,                   \
                                                             (T*)ulevels[0].vals, ulevels[0].freem);               \
    GM_LAUNCH_CHECK();                                                                                              \
    rc = mg::setup<NB_ - 1, NB_ - 1, T>(n_levels, ulevels, ap_work, s);                                             \
    if (rc) return rc;                                                                                              \
    up::k_schur_diag<NB_, T><<<gm_blocks(n_rows, 256), 256, 0, s>>>(n_rows, row_ptr, col_idx, (const T*)vals,       \
                                                                    (const T*)ulevels[0].dinv, freem, (T*)sinv);   \
    GM_LAUNCH_CHECK();                                                                                              \
    return 0;                                                                                                       \
  } while (0)
  if (nb == 3) {
    if (fp32) GM_UPS(3, float);
    GM_UPS(3, double);
  }
  if (fp32) GM_UPS(4, float);
  GM_UPS(4, double);
#undef GM_UPS
}

// Flexible GMRES(m) for a u-p system with the block preconditioner above.
// vu, zu: n_rows * (nb-1) work vectors; other arguments as gm_gmres_mg.
extern "C" int gm_gmres_up(int dim, int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx,
                           const void* vals, const void* sinv, int n_levels, const gm_mg_level* ulevels, int nu_pre,
                           int nu_post, int coarse_sweeps, double omega, void* vu, void* zu, const void* b, void* x,
                           void* V, void* Z, void* z, void* w, double* part, double* dh, int m, double tol,
                           int maxiter, int* iters, double* relres, hipStream_t s) {
  GM_REQUIRE(nb == 3 || nb == 4, "gm_gmres_up: nb must be 3 or 4");
  GM_REQUIRE(dim == nb - 1, "gm_gmres_up: nb must be dim + 1");
  int rc = mg_check(dim, nb - 1, n_levels, ulevels, fp32);
  if (rc) return rc;
  GM_REQUIRE(iters && relres && m >= 1 && m <= 256 && b && x && V && Z && z && w && vu && zu && sinv,
             "gm_gmres_up: bad argument");
  GM_REQUIRE(nu_pre >= 1 && nu_post >= 0 && coarse_sweeps >= 1, "gm_gmres_up: bad smoother counts");
  mg::Cycle cy{dim, n_levels, nu_pre, nu_post, coarse_sweeps, omega, ulevels};
  long n = (long)n_rows * nb;
#define GM_GUP(D, NB_, T)                                                                                      \
  do {                                                                                                         \
    up::BlockPrec<D, NB_, T> pc{cy, n_rows, row_ptr, col_idx, (const T*)vals, (const T*)sinv, (T*)vu, (T*)zu, n}; \
    return gmres_impl<NB_, T, T>(n_rows, row_ptr, col_idx, (const T*)vals, (const T*)b, (T*)x, (T*)V, (T*)z,    \
                                 (T*)w, pc, part, dh, m, tol, maxiter, iters, relres, s, (T*)Z);               \
  } while (0)
  if (nb == 3) {
    if (fp32) GM_GUP(2, 3, float);
    GM_GUP(2, 3, double);
  }
  if (fp32) GM_GUP(3, 4, float);
  GM_GUP(3, 4, double);
#undef GM_GUP
}

// ================================================================== row equilibration
// Scale every row of K (and of the right-hand side) by 1/|K_ii|, so that
// displacement and pressure equations, whose magnitudes differ by many orders,
// weigh alike in the Krylov residual norm; the solution is unchanged.
template <int NB, typename TM>
__global__ void k_row_scale(int n_rows, const int* rp, const int* ci, TM* vals, const double* rhs_in,
                            double* rhs_out) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_rows) return;
  int qd = -1;
  for (int q = rp[I]; q < rp[I + 1]; ++q)
    if (ci[q] == I) {
      qd = q;
      break;
    }
  double sc[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    double d = (qd >= 0) ? fabs((double)vals[(long)qd * NB * NB + r * NB + r]) : 0.0;
    sc[r] = (d > 0.0) ? 1.0 / d : 1.0;
  }
  for (int q = rp[I]; q < rp[I + 1]; ++q)
#pragma unroll
    for (int r = 0; r < NB; ++r)
#pragma unroll
      for (int c = 0; c < NB; ++c) vals[(long)q * NB * NB + r * NB + c] = (TM)((double)vals[(long)q * NB * NB + r * NB + c] * sc[r]);
#pragma unroll
  for (int r = 0; r < NB; ++r) rhs_out[(long)I * NB + r] = rhs_in[(long)I * NB + r] * sc[r];
}

extern "C" int gm_row_scale(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, void* vals,
                            const double* rhs_in, double* rhs_out, hipStream_t s) {
  GM_REQUIRE(row_ptr && col_idx && vals && rhs_in && rhs_out, "gm_row_scale: null argument");
  if (n_rows == 0) return 0;
#define GM_RS(NB)                                                                                               \
  if (nb == NB) {                                                                                               \
    if (fp32)                                                                                                   \
      k_row_scale<NB, float><<<gm_blocks(n_rows, 256), 256, 0, s>>>(n_rows, row_ptr, col_idx, (float*)vals,     \
                                                                     rhs_in, rhs_out);                          \
    else                                                                                                        \
      k_row_scale<NB, double><<<gm_blocks(n_rows, 256), 256, 0, s>>>(n_rows, row_ptr, col_idx, (double*)vals,   \
                                                                      rhs_in, rhs_out);                         \
    GM_LAUNCH_CHECK();                                                                                          \
    return 0;                                                                                                   \
  }
  GM_RS(1)
  GM_RS(2)
  GM_RS(3)
  GM_RS(4)
#undef GM_RS
  gm::set_error("gm_row_scale: unsupported block size %d", nb);
  return -1;
}

// ================================================================== SpMV autotuner
// Times the SpMV on a given matrix for each workgroup size (64, 128, 256, 512
// threads; `reps` launches each, HIP events on the stream), stores the fastest
// for later launches with this block size and precision, and returns it with
// the four average times in milliseconds.
extern "C" int gm_spmv_autotune(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx,
                                const void* vals, const void* x, void* y, int reps, int* best, float* ms,
                                hipStream_t s) {
  GM_REQUIRE(nb >= 1 && nb <= 4 && row_ptr && col_idx && vals && x && y && best && ms && reps >= 1,
             "gm_spmv_autotune: bad argument");
  const int cand[4] = {64, 128, 256, 512};
  hipEvent_t e0, e1;
  GM_HIP(hipEventCreate(&e0));
  GM_HIP(hipEventCreate(&e1));
  int bi = -1;
  for (int c = 0; c < 4; ++c) {
    auto launch = [&]() {
#define GM_TUNE(NB)                                                                                              \
  if (nb == NB) {                                                                                                \
    if (fp32)                                                                                                    \
      launch_spmv<NB, float, float, float>(n_rows, row_ptr, col_idx, (const float*)vals, (const float*)x,       \
                                           (float*)y, nullptr, nullptr, nullptr, s, cand[c]);                  \
    else                                                                                                         \
      launch_spmv<NB, double, double, double>(n_rows, row_ptr, col_idx, (const double*)vals, (const double*)x,  \
                                              (double*)y, nullptr, nullptr, nullptr, s, cand[c]);              \
  }
      GM_TUNE(1) GM_TUNE(2) GM_TUNE(3) GM_TUNE(4)
#undef GM_TUNE
    };
    launch();  // warm-up
    GM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch();
    GM_HIP(hipEventRecord(e1, s));
    GM_HIP(hipEventSynchronize(e1));
    GM_LAUNCH_CHECK();
    float t = 0.0f;
    GM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[c] = t / reps;
    if (bi < 0 || ms[c] < ms[bi]) bi = c;
  }
  GM_HIP(hipEventDestroy(e0));
  GM_HIP(hipEventDestroy(e1));
  *best = cand[bi];
  g_spmv_threads[nb][fp32 ? 1 : 0] = cand[bi];
  return 0;
}

// cpGIMP implicit MPM kernels for gfx950 (FP64 state, FP64 or FP32 Jacobian).
//
// Design (one load step, weights fixed because particles do not move inside
// the Newton loop):
//   1. gm_bin_particles   sort particle ids by their stencil base node (counting
//                         sort + per-bucket id sort => deterministic order)
//   2. gm_bsr_pattern     per node row, OR the support boxes of the particles
//                         touching it into a 5^d window mask -> block-CSR pattern
//   3. per Newton iteration:
//      gm_particle_stage  one particle per lane: gather du, p; kinematics;
//                         constitutive update; particle stress vector Y and
//                         tangent D = dY/dz (z = grad(du), p, pbar, grad p ...)
//      gm_node_residual   one node per lane: r_I = sum_p L(geom_I) Y_p
//                         (deterministic gather instead of the reference's atomics)
//      gm_node_jacobian   one wavefront per node row: K_IJ = -sum_p L_I D_p M_J
//                         into the block-CSR slots, masked + identity rows
//   The reference builds the same matrix with 5^d x nb coloured reverse sweeps
//   (solvers/sparse_differentiation.py:213-401, qc3d:285-321); here it is
//   assembled analytically in one pass.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "gm_common.h"
#include "gm_device.h"
#include "gm_types.h"

namespace gm {
static thread_local char g_err[1024] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace gm

extern "C" const char* gm_last_error(void) { return gm::g_err; }
extern "C" int gm_version(void) { return 1; }

using namespace gm;

// ------------------------------------------------------------------ layouts
template <int DIM, bool COUPLED>
struct Lay {
  static constexpr int NB = DIM + (COUPLED ? 1 : 0);
  static constexpr int NH = DIM * DIM;
  static constexpr int Z_DU = NH;  // 3D u-p only: z-displacement for the indenter penalty
  static constexpr int Z_P = NH + ((DIM == 3 && COUPLED) ? 1 : 0);
  static constexpr int Z_PB = Z_P + 1;
  static constexpr int Z_G = Z_P + 2;
  static constexpr int NZ = COUPLED ? Z_G + DIM : NH;
  static constexpr int Y_T = NH;
  static constexpr int Y_Q = NH + DIM;
  static constexpr int Y_M = Y_Q + 1;
  static constexpr int Y_S = Y_M + DIM;
  static constexpr int NY = COUPLED ? Y_S + 1 : NH + DIM;
};

template <int DIM>
__device__ __forceinline__ void load_axes(const gm_grid& g, const double* x, const double* lp, int p, Axis* ax) {
#pragma unroll
  for (int a = 0; a < DIM; ++a) ax[a] = gimp_axis(x[p * DIM + a], lp[p * DIM + a], g.h[a], g.inv_h[a]);
}

template <int DIM>
__device__ __forceinline__ int node_index(const gm_grid& g, const int* c) {
  if (DIM == 2) return c[0] + c[1] * (g.nx + 1);
  return c[0] + c[1] * (g.nx + 1) + c[2] * (g.nx + 1) * (g.ny + 1);
}

template <int DIM>
__device__ __forceinline__ void node_coords(const gm_grid& g, int n, int* c) {
  c[0] = n % (g.nx + 1);
  if (DIM == 2) {
    c[1] = n / (g.nx + 1);
  } else {
    c[1] = (n / (g.nx + 1)) % (g.ny + 1);
    c[2] = n / ((g.nx + 1) * (g.ny + 1));
  }
}

// ================================================================== binning
__global__ void k_bin_count(gm_grid g, const double* x, const double* lp, int n, int* pbase, uint8_t* ext,
                            int* count, int* err) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int ncell[3] = {g.nx, g.ny, g.nz};
  int lo[3] = {0, 0, 0};
  uint8_t e = 0;
  bool ok = true;
  for (int a = 0; a < g.dim; ++a) {
    Axis ax = gimp_axis(x[p * g.dim + a], lp[p * g.dim + a], g.h[a], g.inv_h[a]);
    lo[a] = ax.lo;
    if (ax.ext == 3) e |= (uint8_t)(1u << a);
    if (ax.lo < 0 || ax.lo + ax.ext - 1 > ncell[a]) ok = false;
    double xv = x[p * g.dim + a];
    if (!(xv == xv)) ok = false;
  }
  if (!ok) {
    atomicAdd(err, 1);
    pbase[p] = -1;
    ext[p] = 0;
    return;
  }
  int b = (g.dim == 2) ? lo[0] + lo[1] * (g.nx + 1) : lo[0] + lo[1] * (g.nx + 1) + lo[2] * (g.nx + 1) * (g.ny + 1);
  pbase[p] = b;
  ext[p] = e;
  atomicAdd(&count[b], 1);
}

__global__ void k_bin_scatter(int n, const int* pbase, const int* start, int* cursor, int* sorted) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int b = pbase[p];
  if (b < 0) return;
  int pos = start[b] + atomicAdd(&cursor[b], 1);
  sorted[pos] = p;
}

// bin-sorted copies of the per-particle geometry (index k = position in `sorted`)
__global__ void k_bin_finalize(int n, int dim, const int* sorted, const double* x, const double* lp,
                               const uint8_t* ext, uint8_t* ext_s, double* x_s, double* lp_s) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int p = sorted[k];
  ext_s[k] = ext[p];
  for (int a = 0; a < dim; ++a) {
    x_s[k * dim + a] = x[p * dim + a];
    lp_s[k * dim + a] = lp[p * dim + a];
  }
}

__global__ void k_bin_sort(int n_nodes, const int* start, int* sorted) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_nodes) return;
  int s = start[c], e = start[c + 1];
  for (int i = s + 1; i < e; ++i) {
    int v = sorted[i];
    int j = i - 1;
    while (j >= s && sorted[j] > v) {
      sorted[j + 1] = sorted[j];
      --j;
    }
    sorted[j + 1] = v;
  }
}

// ---- exclusive scan of n ints (n <= 4M), three phases
#define SCAN_T 1024
#define SCAN_ITEMS 4
__global__ void k_scan_blocks(const int* in, int* out, int n, int* sums) {
  __shared__ int sh[SCAN_T];
  int base = blockIdx.x * SCAN_T * SCAN_ITEMS + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int loc = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    loc += v[i];
  }
  sh[threadIdx.x] = loc;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    int t = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int excl = sh[threadIdx.x] - loc;
  int run = excl;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == SCAN_T - 1) sums[blockIdx.x] = sh[SCAN_T - 1];
}

__global__ void k_scan_sums(int* sums, int nb) {
  __shared__ int sh[SCAN_T];
  int v = (threadIdx.x < nb) ? sums[threadIdx.x] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    int t = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (threadIdx.x < nb) sums[threadIdx.x] = sh[threadIdx.x] - v;
  if (threadIdx.x == SCAN_T - 1) sums[SCAN_T] = sh[SCAN_T - 1];
}

__global__ void k_scan_add(int* out, int n, const int* sums) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] += sums[i / (SCAN_T * SCAN_ITEMS)];
}

// exclusive scan of in[0..n) into out[0..n]; out[n] = total.  tmp >= SCAN_T+1 ints.
static int exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t s) {
  int nblk = gm_blocks(n, SCAN_T * SCAN_ITEMS);
  GM_REQUIRE(nblk <= SCAN_T, "scan: array too large");
  k_scan_blocks<<<nblk, SCAN_T, 0, s>>>(in, out, n, tmp);
  k_scan_sums<<<1, SCAN_T, 0, s>>>(tmp, nblk);
  k_scan_add<<<gm_blocks(n, 256), 256, 0, s>>>(out, n, tmp);
  GM_HIP(hipMemcpyAsync(out + n, tmp + SCAN_T, sizeof(int), hipMemcpyDeviceToDevice, s));
  GM_LAUNCH_CHECK();
  return 0;
}

extern "C" int gm_exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t s) {
  return exclusive_scan(in, out, n, tmp, s);
}

// pbase: n_particles ints of scratch; err: one device int (count of particles off the grid)
extern "C" int gm_bin_particles(const gm_grid* g, const double* x, const double* lp, gm_bins* b, int* pbase,
                                int* err, hipStream_t s) {
  GM_REQUIRE(g && b && x && lp, "gm_bin_particles: null argument");
  int n = b->n_particles;
  GM_HIP(hipMemsetAsync(b->cell_count, 0, sizeof(int) * (g->n_nodes + 1), s));
  GM_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  if (n > 0) {
    k_bin_count<<<gm_blocks(n, 256), 256, 0, s>>>(*g, x, lp, n, pbase, b->ext, b->cell_count, err);
    GM_LAUNCH_CHECK();
  }
  int rc = exclusive_scan(b->cell_count, b->cell_start, g->n_nodes, b->scan_tmp, s);
  if (rc) return rc;
  GM_HIP(hipMemsetAsync(b->cell_count, 0, sizeof(int) * (g->n_nodes + 1), s));
  if (n > 0) {
    k_bin_scatter<<<gm_blocks(n, 256), 256, 0, s>>>(n, pbase, b->cell_start, b->cell_count, b->sorted);
    GM_LAUNCH_CHECK();
  }
  k_bin_sort<<<gm_blocks(g->n_nodes, 256), 256, 0, s>>>(g->n_nodes, b->cell_start, b->sorted);
  if (n > 0)
    k_bin_finalize<<<gm_blocks(n, 256), 256, 0, s>>>(n, g->dim, b->sorted, x, lp, b->ext, b->ext_s, b->x_s,
                                                     b->lp_s);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== candidate walk
// Visit every particle whose 3^d stencil contains node I (base cell I - o,
// o in {0,1,2}^d), in a fixed order.  f(k, o, b, in_support) with k the
// particle's bin-sorted position (ext_s is bin-sorted).
template <int DIM, typename F>
__device__ __forceinline__ void for_candidates(const gm_grid& g, const int* ci, const int* cell_start,
                                               const uint8_t* ext_s, F&& f) {
  int ncell[3] = {g.nx, g.ny, g.nz};
  const int oz_n = (DIM == 3) ? 3 : 1;
  for (int oz = 0; oz < oz_n; ++oz) {
    for (int oy = 0; oy < 3; ++oy) {
      for (int ox = 0; ox < 3; ++ox) {
        int o[3] = {ox, oy, oz};
        int b[3];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          b[a] = ci[a] - o[a];
          if (b[a] < 0 || b[a] > ncell[a]) ok = false;
        }
        if (!ok) continue;
        int bn = node_index<DIM>(g, b);
        int s = cell_start[bn], e = cell_start[bn + 1];
        for (int k = s; k < e; ++k) {
          uint8_t ex = ext_s[k];
          bool in = true;
#pragma unroll
          for (int a = 0; a < DIM; ++a) in = in && (o[a] < (((ex >> a) & 1) ? 3 : 2));
          f(k, o, b, in);
        }
      }
    }
  }
}

// ================================================================== pattern
template <int DIM>
__global__ void k_pattern_count(gm_grid g, const int* cell_start, const uint8_t* ext_s, uint64_t* mask,
                                int* row_nnzb) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= g.n_nodes) return;
  int ci[3] = {0, 0, 0};
  node_coords<DIM>(g, I, ci);
  uint64_t m0 = 0, m1 = 0;
  // centre bit
  int cb = (DIM == 2) ? (2 + 5 * 2) : (2 + 5 * 2 + 25 * 2);
  if (cb < 64) m0 |= (1ull << cb); else m1 |= (1ull << (cb - 64));
  for_candidates<DIM>(g, ci, cell_start, ext_s, [&](int k, const int* o, const int* b, bool in) {
    if (!in) return;
    uint8_t ex = ext_s[k];
    int e[3];
#pragma unroll
    for (int a = 0; a < DIM; ++a) e[a] = ((ex >> a) & 1) ? 3 : 2;
    const int kz_n = (DIM == 3) ? e[2] : 1;
    for (int kz = 0; kz < kz_n; ++kz)
      for (int ky = 0; ky < e[1]; ++ky)
        for (int kx = 0; kx < e[0]; ++kx) {
          int rx = kx - o[0] + 2, ry = ky - o[1] + 2, rz = (DIM == 3) ? (kz - o[2] + 2) : 0;
          int bit = rx + 5 * ry + 25 * rz;
          if (bit < 64) m0 |= (1ull << bit); else m1 |= (1ull << (bit - 64));
        }
  });
  mask[2 * I] = m0;
  mask[2 * I + 1] = m1;
  row_nnzb[I] = __popcll(m0) + __popcll(m1);
}

// The same row masks from each bin's set of support shapes: bit e of extset[b] is set
// when a particle of bin b has ext value e.  A particle's contribution to a row depends
// only on its bin (base node) and its ext value, so the union over the bin's particles
// is the union over the shapes present, at most 2^DIM per bin instead of one per particle.
__global__ void k_bin_extset(int n_bins, const int* cell_start, const uint8_t* ext_s, int* extset) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_bins) return;
  int s = 0;
  for (int k = cell_start[b]; k < cell_start[b + 1]; ++k) s |= 1 << ext_s[k];
  extset[b] = s;
}

template <int DIM>
__global__ void k_pattern_count_sets(gm_grid g, const int* extset, uint64_t* mask, int* row_nnzb) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= g.n_nodes) return;
  int ci[3] = {0, 0, 0};
  node_coords<DIM>(g, I, ci);
  const int ncell[3] = {g.nx, g.ny, g.nz};
  uint64_t m0 = 0, m1 = 0;
  int cb = (DIM == 2) ? (2 + 5 * 2) : (2 + 5 * 2 + 25 * 2);
  if (cb < 64) m0 |= (1ull << cb); else m1 |= (1ull << (cb - 64));
  const int oz_n = (DIM == 3) ? 3 : 1;
  for (int oz = 0; oz < oz_n; ++oz)
    for (int oy = 0; oy < 3; ++oy)
      for (int ox = 0; ox < 3; ++ox) {
        const int o[3] = {ox, oy, oz};
        int bc[3] = {0, 0, 0};
        bool ok = true;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          bc[a] = ci[a] - o[a];
          if (bc[a] < 0 || bc[a] > ncell[a]) ok = false;
        }
        if (!ok) continue;
        const int es = extset[node_index<DIM>(g, bc)];
        for (int ex = 0; ex < (1 << DIM); ++ex) {
          if (!((es >> ex) & 1)) continue;
          int e[3] = {1, 1, 1};
          bool in = true;
#pragma unroll
          for (int a = 0; a < DIM; ++a) {
            e[a] = ((ex >> a) & 1) ? 3 : 2;
            in = in && o[a] < e[a];
          }
          if (!in) continue;
          for (int kz = 0; kz < e[2]; ++kz)
            for (int ky = 0; ky < e[1]; ++ky)
              for (int kx = 0; kx < e[0]; ++kx) {
                int rx = kx - o[0] + 2, ry = ky - o[1] + 2, rz = (DIM == 3) ? (kz - o[2] + 2) : 0;
                int bit = rx + 5 * ry + 25 * rz;
                if (bit < 64) m0 |= (1ull << bit); else m1 |= (1ull << (bit - 64));
              }
        }
      }
  mask[2 * I] = m0;
  mask[2 * I + 1] = m1;
  row_nnzb[I] = __popcll(m0) + __popcll(m1);
}

template <int DIM>
__global__ void k_pattern_fill(gm_grid g, const uint64_t* mask, const int* row_ptr, int* col_idx) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= g.n_nodes) return;
  int ci[3] = {0, 0, 0};
  node_coords<DIM>(g, I, ci);
  int k = row_ptr[I];
  const int nbits = (DIM == 2) ? 25 : 125;
  for (int bit = 0; bit < nbits; ++bit) {
    uint64_t w = (bit < 64) ? mask[2 * I] : mask[2 * I + 1];
    if (!((w >> (bit & 63)) & 1ull)) continue;
    int rx = bit % 5 - 2, ry = (bit / 5) % 5 - 2, rz = (DIM == 3) ? (bit / 25 - 2) : 0;
    int c[3] = {ci[0] + rx, ci[1] + ry, ci[2] + rz};
    col_idx[k++] = node_index<DIM>(g, c);
  }
}

// Row masks from the bins' support-shape sets (1, the default) or particle by particle (0).
static int g_pattern_by_sets = 1;

extern "C" int gm_pattern_kernel(int by_sets) {
  g_pattern_by_sets = by_sets ? 1 : 0;
  return 0;
}

// row_nnzb: n_nodes + 1 ints of scratch.  On return pat->nnzb holds the block count
// (one host synchronisation, once per load step).
extern "C" int gm_bsr_pattern(const gm_grid* g, const gm_bins* b, gm_pattern* pat, int* row_nnzb,
                              int max_nnzb, hipStream_t s) {
  GM_REQUIRE(g && b && pat, "gm_bsr_pattern: null argument");
  int n = g->n_nodes;
  if (g_pattern_by_sets) {
    // the bins' support-shape sets in the binning's scratch counts (free once the particles are sorted)
    k_bin_extset<<<gm_blocks(n, 256), 256, 0, s>>>(n, b->cell_start, b->ext_s, b->cell_count);
    if (g->dim == 2)
      k_pattern_count_sets<2><<<gm_blocks(n, 128), 128, 0, s>>>(*g, b->cell_count, pat->row_mask, row_nnzb);
    else
      k_pattern_count_sets<3><<<gm_blocks(n, 128), 128, 0, s>>>(*g, b->cell_count, pat->row_mask, row_nnzb);
  } else if (g->dim == 2) {
    k_pattern_count<2><<<gm_blocks(n, 128), 128, 0, s>>>(*g, b->cell_start, b->ext_s, pat->row_mask, row_nnzb);
  } else {
    k_pattern_count<3><<<gm_blocks(n, 128), 128, 0, s>>>(*g, b->cell_start, b->ext_s, pat->row_mask, row_nnzb);
  }
  GM_LAUNCH_CHECK();
  int rc = exclusive_scan(row_nnzb, pat->row_ptr, n, b->scan_tmp, s);
  if (rc) return rc;
  int nnzb = 0;
  GM_HIP(hipMemcpyAsync(&nnzb, pat->row_ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
  GM_HIP(hipStreamSynchronize(s));
  pat->nnzb = nnzb;
  if (nnzb > max_nnzb) {
    gm::set_error("gm_bsr_pattern: pattern needs %d blocks, capacity %d", nnzb, max_nnzb);
    return -4;
  }
  if (g->dim == 2)
    k_pattern_fill<2><<<gm_blocks(n, 128), 128, 0, s>>>(*g, pat->row_mask, pat->row_ptr, pat->col_idx);
  else
    k_pattern_fill<3><<<gm_blocks(n, 128), 128, 0, s>>>(*g, pat->row_mask, pat->row_ptr, pat->col_idx);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== particle stage
// NHONLY: an instance for the Neo-Hookean skeleton alone, compiled without the
// Hencky-family return maps (eigen-decomposition, local Newton solves) and so with
// a fraction of the general instance's registers.
template <int DIM, bool COUPLED, typename TD, bool NHONLY = false>
__global__ void __launch_bounds__(128) k_particle_stage(gm_grid g, gm_stage_args a) {
  typedef Lay<DIM, COUPLED> L;
  const bool neo_hookean = NHONLY || a.mat_kind == MAT_NEOHOOKEAN;
  int kq = blockIdx.x * blockDim.x + threadIdx.x;  // bin-sorted position: Y/D row
  if (kq >= a.n_particles) return;
  int p = a.sorted ? a.sorted[kq] : kq;               // particle id: state arrays
  Axis ax[DIM];
  load_axes<DIM>(g, a.x, a.lp, p, ax);
  double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double dup[3] = {0, 0, 0};
  double pp = 0, ppo = 0, ppa = 0, ppoa = 0;
  double graw[3] = {0, 0, 0};
  const int NB = L::NB;
  // gather over the support, reference loop order (i outer ... k inner)
  const int ez = (DIM == 3) ? ax[2].ext : 1;
  for (int i = 0; i < ax[0].ext; ++i)
    for (int j = 0; j < ax[1].ext; ++j)
      for (int k = 0; k < ez; ++k) {
        int o[3] = {i, j, k};
        int c[3] = {ax[0].lo + i, ax[1].lo + j, (DIM == 3) ? ax[2].lo + k : 0};
        int node = node_index<DIM>(g, c);
        double w, G[3], wa;
        node_geom<DIM>(ax, o, w, G, wa);
        double du[3];
#pragma unroll
        for (int cc = 0; cc < DIM; ++cc) {
          du[cc] = a.u_new[node * NB + cc];
          if (COUPLED) du[cc] -= a.u_old[node * NB + cc];
        }
#pragma unroll
        for (int aa = 0; aa < DIM; ++aa)
#pragma unroll
          for (int cc = 0; cc < DIM; ++cc) H[aa][cc] += G[aa] * du[cc];
        if (COUPLED) {
#pragma unroll
          for (int cc = 0; cc < DIM; ++cc) dup[cc] += w * du[cc];
          double pn = a.u_new[node * NB + DIM];
          double po = a.u_old[node * NB + DIM];
          pp += w * pn;
          ppo += w * po;
          ppa += wa * pn;
          ppoa += wa * po;
#pragma unroll
          for (int aa = 0; aa < DIM; ++aa) graw[aa] += G[aa] * pn;
        }
      }
  // kinematics: incr_F = I + delta_u_GRAD (mpm_utils.py:510-517, 1199-1206)
  M3 f = m3_eye();
#pragma unroll
  for (int aa = 0; aa < DIM; ++aa)
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) f.a[aa][cc] += H[aa][cc];
  M3 A;
  if (DIM == 2) {
    double d = f.a[0][0] * f.a[1][1] - f.a[0][1] * f.a[1][0];
    A = m3_eye();
    A.a[0][0] = f.a[1][1] / d;
    A.a[0][1] = -f.a[0][1] / d;
    A.a[1][0] = -f.a[1][0] / d;
    A.a[1][1] = f.a[0][0] / d;
  } else {
    A = m3_inv(f);
  }
  M3 Fo = m3_load(a.F_old + 9 * p);
  M3 Fn = m3_mul(f, Fo);
  double J = m3_det(Fn);
  double Jo = m3_det(Fo);
  M3 bo = m3_load(a.b_old + 9 * p);
  M3 btr = m3_mul(f, m3_mul(bo, m3_T(f)));
  MatParams mp;
  mp.kind = a.mat_kind;
  mp.dp_apex = a.dp_apex;
  mp.lam = a.lam;
  mp.mu = a.mu;
  mp.E = a.E;
  mp.kappa = a.kappa;
  mp.H = a.H;
  mp.phi = a.p_phi ? a.p_phi[p] : a.phi;
  mp.psi = a.psi;
  mp.coh = a.p_coh ? a.p_coh[p] : a.coh;
  mp.shape = a.shape;
  mp.ns = a.ns;
  mp.lnJ = log(J);
  M3 tau, bn;
  SpectralState ss;
  double alpha_o = a.alpha_old ? a.alpha_old[p] : 0.0;
  double alpha_n = alpha_o;
  M3 sig;
  if (neo_hookean) {
    // sigma = mu/J (F F^T - I) + lambda/J ln J I (mpm_utils.py:915)
    M3 FFt = m3_mul(Fn, m3_T(Fn));
    double lnJ = log(J);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double id = (i == j) ? 1.0 : 0.0;
        sig.a[i][j] = a.mu / J * (FFt.a[i][j] - id) + a.lam / J * lnJ * id;
        tau.a[i][j] = sig.a[i][j] * J;
      }
    bn = btr;
  } else if constexpr (!NHONLY) {
    hencky_family(btr, mp, alpha_o, ss, tau, bn, alpha_n);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) sig.a[i][j] = tau.a[i][j] / J;
  }
  m3_store(a.F_new + 9 * p, Fn);
  m3_store(a.b_new + 9 * p, bn);
  m3_store(a.sigma + 9 * p, sig);
  if (a.alpha_new) a.alpha_new[p] = alpha_n;

  // ---- particle stress vector Y
  double V = a.p_vol * J;
  double Y[L::NY];
  // P[b][c] = -V sum_a A[a][b] (sigma - p I)[a][c]    (rhs_value, mpm_utils.py:588, 994, 1271)
#pragma unroll
  for (int bb = 0; bb < DIM; ++bb)
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) {
      double s = 0.0;
#pragma unroll
      for (int aa = 0; aa < DIM; ++aa) s += A.a[aa][bb] * (sig.a[aa][cc] - ((COUPLED && aa == cc) ? pp : 0.0));
      Y[bb * DIM + cc] = -s * V;
    }
  double pen_z = 0.0;
  bool pen_in = false;
  if (!COUPLED) {
    // single phase: rho_new V_new g + f_ext (mpm_utils.py:564-571)
    double rho = a.p_rho / J;
    bool ef = a.ext_flag && a.ext_flag[p];
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) Y[L::Y_T + cc] = rho * a.gravity[cc] * V + (ef ? a.point_load[cc] : 0.0);
  } else {
    double npor = 1.0 - (1.0 - a.phi0) / J;
    double rho_mix = npor * 1.0 + (1.0 - npor) * a.p_rho;
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) Y[L::Y_T + cc] = rho_mix * a.gravity[cc] * V;
    if (DIM == 2) {
      bool ef = a.ext_flag && a.ext_flag[p];
      bool tf = a.trac_flag && a.trac_flag[p];
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc)
        Y[L::Y_T + cc] += (ef ? a.point_load[cc] : 0.0) + (tf ? a.traction[cc] : 0.0);
    } else if (a.penalty) {
      // penalty indenter (mpm_utils.py:1235-1248)
      double xz = a.x[p * 3 + 2], xx = a.x[p * 3], xy = a.x[p * 3 + 1];
      if (xz >= a.footing_height && xx > a.footing[0] && xx < a.footing[1] && xy > a.footing[2] &&
          xy < a.footing[3]) {
        pen_in = true;
        double pen = a.footing_height - (xz + dup[2]);
        pen_z = -pen * a.pen_par;
        Y[L::Y_T + 2] += pen_z;
      }
    }
    if (a.pen_force) a.pen_force[p] = pen_in ? -pen_z : 0.0;
    // mass balance and PPP (mpm_utils.py:1006-1015, 1284-1294)
    double lnJ = log(J), lnJo = log(Jo);
    Y[L::Y_Q] = (lnJ - lnJo) * V;
    double gsp[3];
#pragma unroll
    for (int aa = 0; aa < DIM; ++aa) {
      if (DIM == 2) {
        double s = 0.0;
#pragma unroll
        for (int bb = 0; bb < DIM; ++bb) s += A.a[aa][bb] * graw[bb];
        gsp[aa] = s;  // spatial gradient in 2D (mpm_utils.py:933, 950)
      } else {
        gsp[aa] = graw[aa];  // material gradient in 3D (mpm_utils.py:1169, 1194)
      }
    }
#pragma unroll
    for (int bb = 0; bb < DIM; ++bb) {
      double s = 0.0;
#pragma unroll
      for (int aa = 0; aa < DIM; ++aa) s += A.a[aa][bb] * (a.mobility * gsp[aa]);
      Y[L::Y_M + bb] = s * a.dt * V;
    }
    double tau_s = 0.5 / a.mu;
    double cpp = (pp - ppa) - (ppo - ppoa);
    Y[L::Y_S] = tau_s * cpp * V;
  }
#pragma unroll
  for (int y = 0; y < L::NY; ++y) a.Y[(long)kq * L::NY + y] = Y[y];
  if (!a.want_tangent) return;

  // ---- tangent D = dY/dz, written column by column (every entry exactly once)
  TD* D = (TD*)a.D + (long)kq * L::NY * L::NZ;
  auto put = [&](int y, int z, double v) { D[y * L::NZ + z] = (TD)v; };
  // M = tau - p J I; P = -V0 A^T M
  M3 Mm = tau;
  if (COUPLED)
#pragma unroll
    for (int i = 0; i < 3; ++i) Mm.a[i][i] -= pp * J;
  M3 AtM;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) s += A.a[k][i] * Mm.a[k][j];
      AtM.a[i][j] = s;
    }
  M3 Bnh;  // F_old F^T for Neo-Hookean, b_old f^T for the Hencky family
  if (neo_hookean) Bnh = m3_mul(Fo, m3_T(Fn));
  else Bnh = m3_mul(bo, m3_T(f));
  double lnJ = log(J), lnJo = log(Jo);
  // A^T A, A g, A^T g, A^T A g (mass-balance terms)
  double AtA[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double Ag[3] = {0, 0, 0}, Atg[3] = {0, 0, 0}, AtAg[3] = {0, 0, 0};
  if (COUPLED) {
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += A.a[k][i] * A.a[k][j];
        AtA[i][j] = s;
        Ag[i] += A.a[i][j] * graw[j];
        Atg[i] += A.a[j][i] * graw[j];
      }
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int k = 0; k < DIM; ++k) AtAg[i] += A.a[k][i] * Ag[k];
  }
  const double kdt = a.mobility * a.dt * a.p_vol;
  for (int aa = 0; aa < DIM; ++aa)
    for (int bb = 0; bb < DIM; ++bb) {
      int z = aa * DIM + bb;  // column of dH_ab
      double dJ = J * A.a[bb][aa];
      M3 dtau;
      if (neo_hookean) {
        // d tau = mu (dH B + B^T dH^T) + lambda tr(A dH) I, B = F_old F^T
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            dtau.a[i][j] = a.mu * ((i == aa ? Bnh.a[bb][j] : 0.0) + (j == aa ? Bnh.a[bb][i] : 0.0)) +
                           ((i == j) ? a.lam * A.a[bb][aa] : 0.0);
      } else {
        // d b_tr = dH C + C^T dH^T, C = b_old f^T; spectral tangent
        M3 db;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) db.a[i][j] = (i == aa ? Bnh.a[bb][j] : 0.0) + (j == aa ? Bnh.a[bb][i] : 0.0);
        dtau = spectral_dtau(ss, db);
        if (a.mat_kind == MAT_NORSAND) {
          // the return also depends on ln J (specific volume): d ln J / d H_ab = A_ba
          const double dlnJ = A.a[bb][aa];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              double s = 0.0;
#pragma unroll
              for (int k = 0; k < 3; ++k) s += ss.V.a[i][k] * ss.TJ[k] * ss.V.a[j][k];
              dtau.a[i][j] += dlnJ * s;
            }
        }
      }
      // dP = -V0 [ dA^T M + A^T (dtau - p dJ I) ],  (dA^T X)[i][c] = -A[b][i] (A^T X)[a][c]
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          double t1 = -A.a[bb][i] * AtM.a[aa][c];
          double t2 = 0.0;
#pragma unroll
          for (int k = 0; k < DIM; ++k) t2 += A.a[k][i] * (dtau.a[k][c] - ((COUPLED && k == c) ? pp * dJ : 0.0));
          put(i * DIM + c, z, -a.p_vol * (t1 + t2));
        }
      if (!COUPLED) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) put(L::Y_T + c, z, 0.0);
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) put(L::Y_T + c, z, a.gravity[c] * a.p_vol * dJ);
        put(L::Y_Q, z, a.p_vol * (1.0 + lnJ - lnJo) * dJ);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          double v;
          if (DIM == 3) {
            // m = kdt J A^T g
            v = kdt * (J * (-A.a[bb][i] * Atg[aa]) + dJ * Atg[i]);
          } else {
            // m = kdt J A^T A g ; d(A^T A g) = dA^T (A g) + A^T (dA g), (dA g)[k] = -A[k][a] (A g)[b]
            double dAtAg = -A.a[bb][i] * AtAg[aa];
            double AtdAg = 0.0;
#pragma unroll
            for (int k = 0; k < DIM; ++k) AtdAg += A.a[k][i] * (-A.a[k][aa] * Ag[bb]);
            v = kdt * (dJ * AtAg[i] + J * (dAtAg + AtdAg));
          }
          put(L::Y_M + i, z, v);
        }
        put(L::Y_S, z, Y[L::Y_S] * A.a[bb][aa]);
      }
    }
  if (COUPLED) {
    double ts = 0.5 / a.mu * a.p_vol * J;
    double kdtJ = kdt * J;
    // column du_z (3D indenter penalty)
    if (DIM == 3) {
#pragma unroll
      for (int y = 0; y < L::NY; ++y) put(y, L::Z_DU, (y == L::Y_T + 2 && pen_in) ? a.pen_par : 0.0);
    }
    // column p: dP/dp = V0 J A^T, ds/dp = ts
#pragma unroll
    for (int y = 0; y < L::NY; ++y) {
      double v = 0.0;
      if (y < L::NH) {
        int i = y / DIM, c = y % DIM;
        v = a.p_vol * J * A.a[c][i];
      } else if (y == L::Y_S) {
        v = ts;
      }
      put(y, L::Z_P, v);
    }
    // column pbar: ds/dpbar = -ts
#pragma unroll
    for (int y = 0; y < L::NY; ++y) put(y, L::Z_PB, (y == L::Y_S) ? -ts : 0.0);
    // columns grad p: dm/dg = kdt J A^T (3D) or kdt J A^T A (2D)
#pragma unroll
    for (int k = 0; k < DIM; ++k)
#pragma unroll
      for (int y = 0; y < L::NY; ++y) {
        double v = 0.0;
        if (y >= L::Y_M && y < L::Y_M + DIM) {
          int i = y - L::Y_M;
          v = kdtJ * (DIM == 3 ? A.a[k][i] : AtA[i][k]);
        }
        put(y, L::Z_G + k, v);
      }
  }
}

extern "C" int gm_particle_stage(const gm_grid* g, const gm_stage_args* a, hipStream_t s) {
  GM_REQUIRE(g && a, "gm_particle_stage: null argument");
  int n = a->n_particles;
  if (n == 0) return 0;
  GM_REQUIRE(!a->want_tangent || a->D, "gm_particle_stage: tangent requested without D");
  int blk = gm_blocks(n, 128);
  const bool nh = a->mat_kind == MAT_NEOHOOKEAN;
#define GM_STAGE(D, C)                                                                               \
  do {                                                                                              \
    if (nh) {                                                                                       \
      if (a->tangent_fp32) k_particle_stage<D, C, float, true><<<blk, 128, 0, s>>>(*g, *a);         \
      else k_particle_stage<D, C, double, true><<<blk, 128, 0, s>>>(*g, *a);                        \
    } else {                                                                                        \
      if (a->tangent_fp32) k_particle_stage<D, C, float, false><<<blk, 128, 0, s>>>(*g, *a);        \
      else k_particle_stage<D, C, double, false><<<blk, 128, 0, s>>>(*g, *a);                       \
    }                                                                                               \
  } while (0)
  if (g->dim == 2 && !a->coupled) {
    GM_STAGE(2, false);
  } else if (g->dim == 2 && a->coupled) {
    GM_STAGE(2, true);
  } else if (g->dim == 3 && a->coupled) {
    GM_STAGE(3, true);
  } else {
    gm::set_error("gm_particle_stage: 3D single-phase is not a supported form");
    return -1;
  }
#undef GM_STAGE
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== node residual
template <int DIM, bool COUPLED>
__global__ void __launch_bounds__(128) k_node_residual(gm_grid g, const double* x_s, const double* lp_s,
                                                       const int* cell_start, const uint8_t* ext_s, const double* Y,
                                                       const uint8_t* freem, double* rhs) {
  typedef Lay<DIM, COUPLED> L;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= g.n_nodes) return;
  int ci[3] = {0, 0, 0};
  node_coords<DIM>(g, I, ci);
  double r[L::NB];
#pragma unroll
  for (int k = 0; k < L::NB; ++k) r[k] = 0.0;
  for_candidates<DIM>(g, ci, cell_start, ext_s, [&](int k, const int* o, const int* b, bool in) {
    if (!in) return;
    Axis ax[DIM];
    load_axes<DIM>(g, x_s, lp_s, k, ax);
    double w, G[3], wa;
    node_geom<DIM>(ax, o, w, G, wa);
    const double* yp = Y + (long)k * L::NY;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      double s = 0.0;
#pragma unroll
      for (int bb = 0; bb < DIM; ++bb) s += G[bb] * yp[bb * DIM + c];
      r[c] += s + w * yp[L::Y_T + c];
    }
    if (COUPLED) {
      double s = w * yp[L::Y_Q];
#pragma unroll
      for (int bb = 0; bb < DIM; ++bb) s += G[bb] * yp[L::Y_M + bb];
      s += (w - wa) * yp[L::Y_S];
      r[DIM] += s;
    }
  });
#pragma unroll
  for (int k = 0; k < L::NB; ++k) rhs[I * L::NB + k] = freem[I * L::NB + k] ? r[k] : 0.0;
}

extern "C" int gm_node_residual(const gm_grid* g, const gm_bins* b, const double* Y, const uint8_t* freem,
                                double* rhs, int coupled, hipStream_t s) {
  GM_REQUIRE(g && b, "gm_node_residual: null argument");
  int n = g->n_nodes;
  int blk = gm_blocks(n, 128);
  if (g->dim == 2 && !coupled)
    k_node_residual<2, false><<<blk, 128, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s, Y, freem, rhs);
  else if (g->dim == 2)
    k_node_residual<2, true><<<blk, 128, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s, Y, freem, rhs);
  else if (coupled)
    k_node_residual<3, true><<<blk, 128, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s, Y, freem, rhs);
  else {
    gm::set_error("gm_node_residual: 3D single-phase is not a supported form");
    return -1;
  }
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== node Jacobian
// One 64-lane wavefront per block row.  Lane l owns compressed slots l and l+64.
template <int DIM, bool COUPLED, typename TD, typename TV>
__global__ void __launch_bounds__(64, 3) k_node_jacobian(gm_grid g, const double* x_s, const double* lp_s,
                                                      const int* cell_start, const uint8_t* ext_s, const TD* D,
                                                      const uint8_t* freem, const int* row_ptr,
                                                      const uint64_t* row_mask, TV* vals) {
  typedef Lay<DIM, COUPLED> L;
  const int NB = L::NB, NZ = L::NZ, NY = L::NY;
  // the particles of a bin are taken JB at a time: their axis data, support flags and
  // E = L_I D_p rows are formed in parallel, then every slot lane accumulates them in
  // bin order (the order and arithmetic of one-particle-at-a-time processing)
  constexpr int JB = 8;
  __shared__ double E[JB][NB * NZ];
  __shared__ Axis AX[JB][DIM];
  __shared__ int INS[JB];
  __shared__ uint8_t EXS[JB];
  int I = blockIdx.x;
  if (I >= g.n_nodes) return;
  int lane = threadIdx.x;
  int ci[3] = {0, 0, 0};
  node_coords<DIM>(g, I, ci);
  uint64_t m0 = row_mask[2 * I], m1 = row_mask[2 * I + 1];
  int nnz = __popcll(m0) + __popcll(m1);
  int rp = row_ptr[I];
  // window offsets of my two slots
  int myoff[2][3];
  bool has[2];
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2) {
    int slot = lane + 64 * s2;
    has[s2] = slot < nnz;
    myoff[s2][0] = myoff[s2][1] = myoff[s2][2] = 0;
    if (has[s2]) {
      int cnt = 0, bit = 0;
      const int nbits = (DIM == 2) ? 25 : 125;
      for (bit = 0; bit < nbits; ++bit) {
        uint64_t w = (bit < 64) ? m0 : m1;
        if ((w >> (bit & 63)) & 1ull) {
          if (cnt == slot) break;
          ++cnt;
        }
      }
      myoff[s2][0] = bit % 5 - 2;
      myoff[s2][1] = (bit / 5) % 5 - 2;
      myoff[s2][2] = (DIM == 3) ? bit / 25 - 2 : 0;
    }
  }
  double acc[2][NB * NB];
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
    for (int k = 0; k < NB * NB; ++k) acc[s2][k] = 0.0;

  int ncell[3] = {g.nx, g.ny, g.nz};
  const int oz_n = (DIM == 3) ? 3 : 1;
  for (int oz = 0; oz < oz_n; ++oz)
    for (int oy = 0; oy < 3; ++oy)
      for (int ox = 0; ox < 3; ++ox) {
        // base cell I - o: its particles have node I at stencil offset o (candidate walk order)
        int o[3] = {ox, oy, oz};
        int bc[3];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          bc[a] = ci[a] - o[a];
          if (bc[a] < 0 || bc[a] > ncell[a]) ok = false;
        }
        if (!ok) continue;
        int bn = node_index<DIM>(g, bc);
        int ks = cell_start[bn], ke = cell_start[bn + 1];
        for (int k0 = ks; k0 < ke; k0 += JB) {
          const int nj = min(JB, ke - k0);
          if (lane < nj) {
            int k = k0 + lane;
            uint8_t ex = ext_s[k];
            bool in = true;
#pragma unroll
            for (int a = 0; a < DIM; ++a) in = in && (o[a] < (((ex >> a) & 1) ? 3 : 2));
            INS[lane] = in ? 1 : 0;
            EXS[lane] = ex;
            Axis ax[DIM];
            load_axes<DIM>(g, x_s, lp_s, k, ax);
#pragma unroll
            for (int a = 0; a < DIM; ++a) AX[lane][a] = ax[a];
          }
          __syncthreads();
          // E = L_I D_p  (NB x NZ) for each particle of the batch
          for (int q = lane; q < nj * NB * NZ; q += 64) {
            int pj = q / (NB * NZ), qq = q % (NB * NZ);
            if (!INS[pj]) continue;
            double wI, GI[3], waI;
            node_geom<DIM>(AX[pj], o, wI, GI, waI);
            const TD* Dp = D + (long)(k0 + pj) * NY * NZ;
            int r = qq / NZ, z = qq % NZ;
            double sE = 0.0;
            if (r < DIM) {
#pragma unroll
              for (int bb = 0; bb < DIM; ++bb) sE += GI[bb] * (double)Dp[(bb * DIM + r) * NZ + z];
              sE += wI * (double)Dp[(L::Y_T + r) * NZ + z];
            } else {
              sE = wI * (double)Dp[L::Y_Q * NZ + z];
#pragma unroll
              for (int bb = 0; bb < DIM; ++bb) sE += GI[bb] * (double)Dp[(L::Y_M + bb) * NZ + z];
              sE += (wI - waI) * (double)Dp[L::Y_S * NZ + z];
            }
            E[pj][qq] = sE;
          }
          __syncthreads();
          for (int pj = 0; pj < nj; ++pj) {
            if (!INS[pj]) continue;  // uniform across the workgroup
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
              if (!has[s2]) continue;
              int oJ[3];
              bool inJ = true;
#pragma unroll
              for (int aa = 0; aa < DIM; ++aa) {
                oJ[aa] = o[aa] + myoff[s2][aa];  // offset of J inside this particle's stencil
                int e = ((EXS[pj] >> aa) & 1) ? 3 : 2;
                inJ = inJ && (oJ[aa] >= 0) && (oJ[aa] < e);
              }
              if (!inJ) continue;
              double wJ, GJ[3], waJ;
              node_geom<DIM>(AX[pj], oJ, wJ, GJ, waJ);
#pragma unroll
              for (int r = 0; r < NB; ++r) {
#pragma unroll
                for (int e = 0; e < DIM; ++e) {
                  double sv = 0.0;
#pragma unroll
                  for (int aa = 0; aa < DIM; ++aa) sv += E[pj][r * NZ + aa * DIM + e] * GJ[aa];
                  if (DIM == 3 && COUPLED && e == 2) sv += E[pj][r * NZ + L::Z_DU] * wJ;
                  acc[s2][r * NB + e] += sv;
                }
                if (COUPLED) {
                  double sv = E[pj][r * NZ + L::Z_P] * wJ + E[pj][r * NZ + L::Z_PB] * waJ;
#pragma unroll
                  for (int aa = 0; aa < DIM; ++aa) sv += E[pj][r * NZ + L::Z_G + aa] * GJ[aa];
                  acc[s2][r * NB + DIM] += sv;
                }
              }
            }
          }
          __syncthreads();
        }
      }
  // masks, sign (the system matrix is -dR/du) and identity on fixed/inactive dofs
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2) {
    if (!has[s2]) continue;
    int cj[3] = {ci[0] + myoff[s2][0], ci[1] + myoff[s2][1], ci[2] + myoff[s2][2]};
    int Jn = node_index<DIM>(g, cj);
    bool diag = (Jn == I);
    TV* out = vals + (long)(rp + lane + 64 * s2) * NB * NB;
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      bool fr = freem[I * NB + r];
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        bool fc = freem[Jn * NB + c];
        double v = (fr && fc) ? -acc[s2][r * NB + c] : 0.0;
        if (diag && r == c && !fr) v = 1.0;
        out[r * NB + c] = (TV)v;
      }
    }
  }
}

// 3D u-p node Jacobian: one 128-lane workgroup per block row I.  The candidate
// particles of row I come in 27 base bins b = I - o; every particle of bin o
// reaches the same 3x3x3 box of columns J = b + k, so lane (k, c) (108 of 128
// lanes: 27 columns x 4 column dofs) owns entry column c of block (I, b + k) for
// the whole bin and adds the bin's particles to it in bin order.  The row's 125
// window blocks are accumulated in LDS (window-bit order), loaded into registers
// per bin and stored back, so every entry receives the same sums in the same
// order as in the one-wavefront-per-row form (k_node_jacobian) -- bit for bit the
// same matrix -- while no lane waits on columns outside the particle's support.
template <typename TD, typename TV>
__global__ void __launch_bounds__(128) k_node_jacobian_up3(gm_grid g, const double* x_s, const double* lp_s,
                                                           const int* cell_start, const uint8_t* ext_s,
                                                           const TD* D, const uint8_t* freem, const int* row_ptr,
                                                           const uint64_t* row_mask, TV* vals) {
  typedef Lay<3, true> L;
  constexpr int NB = 4, NZ = L::NZ, NY = L::NY, T = 128, JB = 8;
  __shared__ double ACC[125 * NB * NB];
  __shared__ double E[JB][NB * NZ];
  __shared__ Axis AX[JB][3];
  __shared__ int INS[JB];
  __shared__ uint8_t EXS[JB];
  const int I = blockIdx.x;
  if (I >= g.n_nodes) return;
  const int t = threadIdx.x;
  for (int q = t; q < 125 * NB * NB; q += T) ACC[q] = 0.0;
  int ci[3];
  node_coords<3>(g, I, ci);
  const int kk = t >> 2, c = t & 3;
  const int kx = kk % 3, ky = (kk / 3) % 3, kz = kk / 9;
  const bool lane_on = t < 27 * NB;
  const int ncell[3] = {g.nx, g.ny, g.nz};
  __syncthreads();
  for (int oz = 0; oz < 3; ++oz)
    for (int oy = 0; oy < 3; ++oy)
      for (int ox = 0; ox < 3; ++ox) {
        const int o[3] = {ox, oy, oz};
        int bc[3];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          bc[a] = ci[a] - o[a];
          if (bc[a] < 0 || bc[a] > ncell[a]) ok = false;
        }
        if (!ok) continue;  // uniform across the workgroup
        const int bn = node_index<3>(g, bc);
        const int ks = cell_start[bn], ke = cell_start[bn + 1];
        // window bit of column J = b + k in row I (offset k - o per axis)
        const int bit = (kx - ox + 2) + 5 * (ky - oy + 2) + 25 * (kz - oz + 2);
        for (int k0 = ks; k0 < ke; k0 += JB) {
          const int nj = min(JB, ke - k0);
          if (t < nj) {
            const int k = k0 + t;
            const uint8_t ex = ext_s[k];
            bool in = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) in = in && (o[a] < (((ex >> a) & 1) ? 3 : 2));
            INS[t] = in ? 1 : 0;
            EXS[t] = ex;
            Axis ax[3];
            load_axes<3>(g, x_s, lp_s, k, ax);
#pragma unroll
            for (int a = 0; a < 3; ++a) AX[t][a] = ax[a];
          }
          __syncthreads();
          // E = L_I D_p (NB x NZ) of each particle of the batch (same expressions as k_node_jacobian)
          for (int q = t; q < nj * NB * NZ; q += T) {
            const int pj = q / (NB * NZ), qq = q % (NB * NZ);
            if (!INS[pj]) continue;
            double wI, GI[3], waI;
            node_geom<3>(AX[pj], o, wI, GI, waI);
            const TD* Dp = D + (long)(k0 + pj) * NY * NZ;
            const int r = qq / NZ, z = qq % NZ;
            double sE = 0.0;
            if (r < 3) {
#pragma unroll
              for (int bb = 0; bb < 3; ++bb) sE += GI[bb] * (double)Dp[(bb * 3 + r) * NZ + z];
              sE += wI * (double)Dp[(L::Y_T + r) * NZ + z];
            } else {
              sE = wI * (double)Dp[L::Y_Q * NZ + z];
#pragma unroll
              for (int bb = 0; bb < 3; ++bb) sE += GI[bb] * (double)Dp[(L::Y_M + bb) * NZ + z];
              sE += (wI - waI) * (double)Dp[L::Y_S * NZ + z];
            }
            E[pj][qq] = sE;
          }
          __syncthreads();
          if (lane_on) {
            double acc[NB];
#pragma unroll
            for (int r = 0; r < NB; ++r) acc[r] = ACC[bit * NB * NB + r * NB + c];
            const int kJ[3] = {kx, ky, kz};
            for (int pj = 0; pj < nj; ++pj) {
              if (!INS[pj]) continue;
              bool inJ = true;
#pragma unroll
              for (int a = 0; a < 3; ++a) inJ = inJ && (kJ[a] < (((EXS[pj] >> a) & 1) ? 3 : 2));
              if (!inJ) continue;
              double wJ, GJ[3], waJ;
              node_geom<3>(AX[pj], kJ, wJ, GJ, waJ);
#pragma unroll
              for (int r = 0; r < NB; ++r) {
                double sv;
                if (c < 3) {
                  sv = 0.0;
#pragma unroll
                  for (int aa = 0; aa < 3; ++aa) sv += E[pj][r * NZ + aa * 3 + c] * GJ[aa];
                  if (c == 2) sv += E[pj][r * NZ + L::Z_DU] * wJ;
                } else {
                  sv = E[pj][r * NZ + L::Z_P] * wJ + E[pj][r * NZ + L::Z_PB] * waJ;
#pragma unroll
                  for (int aa = 0; aa < 3; ++aa) sv += E[pj][r * NZ + L::Z_G + aa] * GJ[aa];
                }
                acc[r] += sv;
              }
            }
#pragma unroll
            for (int r = 0; r < NB; ++r) ACC[bit * NB * NB + r * NB + c] = acc[r];
          }
          __syncthreads();
        }
      }
  // flush the window blocks present in the row pattern: masks, sign (-dR/du), identity rows
  const uint64_t m0 = row_mask[2 * I], m1 = row_mask[2 * I + 1];
  const int rp = row_ptr[I];
  for (int b = t; b < 125; b += T) {
    const uint64_t w = (b < 64) ? m0 : m1;
    const int sh = b & 63;
    if (!((w >> sh) & 1ull)) continue;
    const int below = (b < 64) ? __popcll(sh ? (m0 & ((1ull << sh) - 1ull)) : 0ull)
                               : __popcll(m0) + __popcll(sh ? (m1 & ((1ull << sh) - 1ull)) : 0ull);
    const int cj[3] = {ci[0] + b % 5 - 2, ci[1] + (b / 5) % 5 - 2, ci[2] + b / 25 - 2};
    const int Jn = node_index<3>(g, cj);
    const bool diag = (Jn == I);
    TV* out = vals + (long)(rp + below) * NB * NB;
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const bool fr = freem[I * NB + r];
#pragma unroll
      for (int cc = 0; cc < NB; ++cc) {
        const bool fc = freem[Jn * NB + cc];
        double v = (fr && fc) ? -ACC[b * NB * NB + r * NB + cc] : 0.0;
        if (diag && r == cc && !fr) v = 1.0;
        out[r * NB + cc] = (TV)v;
      }
    }
  }
}

extern "C" int gm_node_jacobian(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, const void* D,
                                int tangent_fp32, const uint8_t* freem, void* vals, int vals_fp32, int coupled,
                                hipStream_t s) {
  GM_REQUIRE(g && b && pat, "gm_node_jacobian: null argument");
  int n = g->n_nodes;
#define GM_JAC(DIM, CP, TD, TV)                                                                           \
  k_node_jacobian<DIM, CP, TD, TV><<<n, 64, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s,          \
                                                    (const TD*)D, freem, pat->row_ptr, pat->row_mask, (TV*)vals)
#define GM_JAC_T(DIM, CP)                                  \
  do {                                                     \
    if (tangent_fp32 && vals_fp32) GM_JAC(DIM, CP, float, float);   \
    else if (tangent_fp32) GM_JAC(DIM, CP, float, double);          \
    else if (vals_fp32) GM_JAC(DIM, CP, double, float);             \
    else GM_JAC(DIM, CP, double, double);                           \
  } while (0)
  if (g->dim == 2 && !coupled) GM_JAC_T(2, false);
  else if (g->dim == 2) GM_JAC_T(2, true);
  else if (coupled && !getenv("GM_JACOBIAN_ROW_WAVE")) {
    // lanes per (column, column dof) of a bin (k_node_jacobian_up3); GM_JACOBIAN_ROW_WAVE=1
    // selects the one-wavefront-per-row form, which assembles the same matrix
#define GM_JAC3(TD, TV)                                                                                   \
  k_node_jacobian_up3<TD, TV><<<n, 128, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s, (const TD*)D, \
                                                freem, pat->row_ptr, pat->row_mask, (TV*)vals)
    if (tangent_fp32 && vals_fp32) GM_JAC3(float, float);
    else if (tangent_fp32) GM_JAC3(float, double);
    else if (vals_fp32) GM_JAC3(double, float);
    else GM_JAC3(double, double);
#undef GM_JAC3
  } else if (coupled) GM_JAC_T(3, true);
  else {
    gm::set_error("gm_node_jacobian: 3D single-phase is not a supported form");
    return -1;
  }
#undef GM_JAC_T
#undef GM_JAC
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== indenter penalty, preconditioner copy
// The reference's indenter term r_z += -(h - (z + du_z)) pen_par w_I (mpm_utils.py:1240-1246)
// enters the Newton matrix K = -dR/du as -pen_par w_I w_J on the z-dofs of every
// penetrating particle's support: a negative stiffness that outweighs the soil's
// under the indenter and makes K indefinite there.  The multigrid hierarchy is
// built from a copy of K with that term's sign reversed (coef = 2 adds
// 2 pen_par w_I w_J), which is positive definite under the indenter; the Krylov
// operator keeps the exact K.  One thread per node row of the box
// [lo, lo + cnt) that holds every support node of the penetrating particles;
// the particles and the predicate are those of the particle stage
// (x_s is the bin-sorted copy of x), and only free z-dofs are touched.
template <typename TV>
__global__ void __launch_bounds__(128) k_penalty_pc(gm_grid g, const double* x_s, const double* lp_s,
                                                    const int* cell_start, const uint8_t* ext_s,
                                                    const uint8_t* freem, const int* row_ptr,
                                                    const uint64_t* row_mask, TV* vals, int lo0, int lo1, int lo2,
                                                    int c0, int c1, int c2, double fx0, double fx1, double fy0,
                                                    double fy1, double height, double coef) {
  const int NB = 4;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c0 * c1 * c2) return;
  int ci[3] = {lo0 + t % c0, lo1 + (t / c0) % c1, lo2 + t / (c0 * c1)};
  if (ci[0] > g.nx || ci[1] > g.ny || ci[2] > g.nz) return;
  int I = node_index<3>(g, ci);
  if (!freem[I * NB + 2]) return;
  const int rs = row_ptr[I];
  const uint64_t m0 = row_mask[2 * I], m1 = row_mask[2 * I + 1];
  for_candidates<3>(g, ci, cell_start, ext_s, [&](int k, const int* o, const int* b, bool in) {
    if (!in) return;
    double xx = x_s[3 * k], xy = x_s[3 * k + 1], xz = x_s[3 * k + 2];
    if (!(xz >= height && xx > fx0 && xx < fx1 && xy > fy0 && xy < fy1)) return;
    Axis ax[3];
    load_axes<3>(g, x_s, lp_s, k, ax);
    double wI = sel3(ax[0].w, o[0]) * sel3(ax[1].w, o[1]) * sel3(ax[2].w, o[2]);
    for (int kz = 0; kz < ax[2].ext; ++kz)
      for (int ky = 0; ky < ax[1].ext; ++ky)
        for (int kx = 0; kx < ax[0].ext; ++kx) {
          int cj[3] = {b[0] + kx, b[1] + ky, b[2] + kz};
          int Jn = node_index<3>(g, cj);
          if (!freem[Jn * NB + 2]) continue;
          double wJ = sel3(ax[0].w, kx) * sel3(ax[1].w, ky) * sel3(ax[2].w, kz);
          int bit = (kx - o[0] + 2) + 5 * (ky - o[1] + 2) + 25 * (kz - o[2] + 2);
          // compressed slot: set window bits below `bit`
          int below;
          if (bit < 64) {
            below = __popcll(bit ? (m0 & ((1ull << bit) - 1ull)) : 0ull);
          } else {
            int hb = bit - 64;
            below = __popcll(m0) + __popcll(hb ? (m1 & ((1ull << hb) - 1ull)) : 0ull);
          }
          long slot = (long)(rs + below) * (NB * NB) + 2 * NB + 2;
          vals[slot] = (TV)((double)vals[slot] + coef * wI * wJ);
        }
  });
}

extern "C" int gm_penalty_pc(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, const uint8_t* freem,
                             void* vals, int vals_fp32, const int* box_lo, const int* box_cnt,
                             const double* footing, double height, double coef, hipStream_t s) {
  GM_REQUIRE(g && b && pat && freem && vals && box_lo && box_cnt && footing, "gm_penalty_pc: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_penalty_pc: the indenter term exists in the 3D u-p form only");
  long cnt = (long)box_cnt[0] * box_cnt[1] * box_cnt[2];
  GM_REQUIRE(box_lo[0] >= 0 && box_lo[1] >= 0 && box_lo[2] >= 0 && box_cnt[0] >= 0 && box_cnt[1] >= 0 &&
                 box_cnt[2] >= 0 && cnt < (1L << 30),
             "gm_penalty_pc: bad node box");
  if (cnt == 0) return 0;
  int blk = gm_blocks(cnt, 128);
  if (vals_fp32)
    k_penalty_pc<float><<<blk, 128, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s, freem, pat->row_ptr,
                                            pat->row_mask, (float*)vals, box_lo[0], box_lo[1], box_lo[2],
                                            box_cnt[0], box_cnt[1], box_cnt[2], footing[0], footing[1],
                                            footing[2], footing[3], height, coef);
  else
    k_penalty_pc<double><<<blk, 128, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->ext_s, freem, pat->row_ptr,
                                             pat->row_mask, (double*)vals, box_lo[0], box_lo[1], box_lo[2],
                                             box_cnt[0], box_cnt[1], box_cnt[2], footing[0], footing[1],
                                             footing[2], footing[3], height, coef);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== P2G
// Node-gather form of P2G_2d / P2G_coupled_2d / P2G_coupled_3d (mpm_utils.py:83-383).
//  act:     activation (threshold < 0: every support node, as P2G_2d; else w > threshold)
//  bflag:   pressure-Dirichlet flags are OR-ed into the p dof (coupled)
//  mvals:   consistent mass in the node pattern (b = 1), may be null
//  mrhs:    consistent-mass rhs, may be null
template <int DIM>
__global__ void k_p2g(gm_grid g, int nb, const double* x_s, const double* lp_s, const int* cell_start,
                      const int* sorted, const uint8_t* ext_s, double threshold, const uint8_t* pbc_flag,
                      double node_h_last, double p_mass, const double* p_part, const int* row_ptr,
                      const uint64_t* row_mask, uint8_t* act, uint8_t* bflag, double* mvals, double* mrhs) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= g.n_nodes) return;
  int ci[3] = {0, 0, 0};
  node_coords<DIM>(g, I, ci);
  bool a_on = false, pb = false;
  double rhs = 0.0;
  int rp = row_ptr ? row_ptr[I] : 0;
  uint64_t m0 = row_mask ? row_mask[2 * I] : 0, m1 = row_mask ? row_mask[2 * I + 1] : 0;
  for_candidates<DIM>(g, ci, cell_start, ext_s, [&](int k, const int* o, const int* b, bool in) {
    int p = sorted[k];
    // pressure boundary: 2D loops over the support, 3D over all 27 nodes (mpm_utils.py:225-235, 349-352)
    if (pbc_flag && pbc_flag[p] && (in || DIM == 3)) {
      double nc = (double)ci[DIM - 1] * node_h_last;
      if (nc >= x_s[k * DIM + DIM - 1]) pb = true;
    }
    if (!in) return;
    Axis ax[DIM];
    load_axes<DIM>(g, x_s, lp_s, k, ax);
    double w, G[3], wa;
    node_geom<DIM>(ax, o, w, G, wa);
    if (threshold < 0.0 || w > threshold) a_on = true;
    if (DIM == 3 && !mvals && mrhs) {
      // sum of the particle's weights over its support, one axis at a time (the
      // 3D mass matrix comes from gm_mass_binned)
      double wsum = 1.0;
#pragma unroll
      for (int aa = 0; aa < 3; ++aa) wsum *= ax[aa].w[0] + ax[aa].w[1] + ax[aa].w[2];
      rhs += w * wsum * p_mass * p_part[p];
    } else if (mvals || mrhs) {
      double wsum = 0.0;
      uint8_t ex = ext_s[k];
      int e[3];
#pragma unroll
      for (int aa = 0; aa < DIM; ++aa) e[aa] = ((ex >> aa) & 1) ? 3 : 2;
      const int kz_n = (DIM == 3) ? e[2] : 1;
      for (int kx = 0; kx < e[0]; ++kx)
        for (int ky = 0; ky < e[1]; ++ky)
          for (int kz = 0; kz < kz_n; ++kz) {
            int oJ[3] = {kx, ky, kz};
            double wJ, GJ[3], waJ;
            node_geom<DIM>(ax, oJ, wJ, GJ, waJ);
            wsum += wJ;
            if (mvals) {
              int rx = kx - o[0] + 2, ry = ky - o[1] + 2, rz = (DIM == 3) ? kz - o[2] + 2 : 0;
              int bit = rx + 5 * ry + 25 * rz;
              int before = (bit < 64) ? __popcll(m0 & ((1ull << bit) - 1ull))
                                      : __popcll(m0) + __popcll(m1 & ((1ull << (bit - 64)) - 1ull));
              mvals[rp + before] += w * wJ * p_mass;
            }
          }
      if (mrhs) rhs += w * wsum * p_mass * p_part[p];
    }
  });
  if (a_on)
    for (int k = 0; k < nb; ++k) act[I * nb + k] = 1;
  if (pb && bflag) bflag[I * nb + DIM] = 1;
  if (mrhs) mrhs[I] = rhs;
}

extern "C" int gm_p2g(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, double threshold,
                      const uint8_t* pbc_flag, double node_h_last, double p_mass, const double* p_part, uint8_t* act,
                      uint8_t* bflag, double* mvals, double* mrhs, hipStream_t s) {
  GM_REQUIRE(g && b, "gm_p2g: null argument");
  int n = g->n_nodes;
  if (mvals) {
    GM_REQUIRE(pat, "gm_p2g: mass matrix needs the pattern");
    GM_HIP(hipMemsetAsync(mvals, 0, sizeof(double) * pat->nnzb, s));
  }
  const int* rp = pat ? pat->row_ptr : nullptr;
  const uint64_t* rm = pat ? pat->row_mask : nullptr;
  if (g->dim == 2)
    k_p2g<2><<<gm_blocks(n, 128), 128, 0, s>>>(*g, g->nb, b->x_s, b->lp_s, b->cell_start, b->sorted, b->ext_s, threshold,
                                               pbc_flag, node_h_last, p_mass, p_part, rp, rm, act, bflag, mvals,
                                               mrhs);
  else
    k_p2g<3><<<gm_blocks(n, 128), 128, 0, s>>>(*g, g->nb, b->x_s, b->lp_s, b->cell_start, b->sorted, b->ext_s, threshold,
                                               pbc_flag, node_h_last, p_mass, p_part, rp, rm, act, bflag, mvals,
                                               mrhs);
  GM_LAUNCH_CHECK();
  return 0;
}

// free mask = !boundary && active, node-interleaved
__global__ void k_free_mask(int n, const uint8_t* bflag, const uint8_t* act, uint8_t* freem) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  freem[i] = (!bflag[i] && act[i]) ? 1 : 0;
}

extern "C" int gm_free_mask(int n, const uint8_t* bflag, const uint8_t* act, uint8_t* freem, hipStream_t s) {
  if (n == 0) return 0;
  k_free_mask<<<gm_blocks(n, 256), 256, 0, s>>>(n, bflag, act, freem);
  GM_LAUNCH_CHECK();
  return 0;
}

// P2G_set_nodal_solution(_3d) (mpm_utils.py:387-414)
__global__ void k_set_nodal_p(int n_nodes, int nb, int pdof, const uint8_t* bflag, const uint8_t* act,
                              int require_active, const double* xp, double* old_s, double* new_s) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_nodes) return;
  int d = I * nb + pdof;
  if (!bflag[d] && (!require_active || act[d])) {
    old_s[d] = xp[I];
    new_s[d] = xp[I];
  }
}

extern "C" int gm_set_nodal_pressure(int n_nodes, int nb, const uint8_t* bflag, const uint8_t* act,
                                     int require_active, const double* xp, double* old_s, double* new_s,
                                     hipStream_t s) {
  k_set_nodal_p<<<gm_blocks(n_nodes, 256), 256, 0, s>>>(n_nodes, nb, nb - 1, bflag, act, require_active, xp, old_s,
                                                        new_s);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== G2P + lp update
// G2P_2d / G2P_coupled_2d / G2P_coupled_3d (mpm_utils.py:1351-1590)
template <int DIM, bool COUPLED>
__global__ void k_g2p(gm_grid g, int n, double* x, const double* lp, const double* u_new, const double* u_old,
                      double* F_old, const double* F_new, double* b_old, const double* b_new, double* alpha_old,
                      const double* alpha_new, double* p_part) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int NB = DIM + (COUPLED ? 1 : 0);
  Axis ax[DIM];
  load_axes<DIM>(g, x, lp, p, ax);
  double du[3] = {0, 0, 0};
  double pn = 0.0;
  const int ez = (DIM == 3) ? ax[2].ext : 1;
  for (int i = 0; i < ax[0].ext; ++i)
    for (int j = 0; j < ax[1].ext; ++j)
      for (int k = 0; k < ez; ++k) {
        int o[3] = {i, j, k};
        int c[3] = {ax[0].lo + i, ax[1].lo + j, (DIM == 3) ? ax[2].lo + k : 0};
        int node = node_index<DIM>(g, c);
        double w, G[3], wa;
        node_geom<DIM>(ax, o, w, G, wa);
#pragma unroll
        for (int cc = 0; cc < DIM; ++cc) {
          double d = u_new[node * NB + cc];
          if (COUPLED) d -= u_old[node * NB + cc];
          du[cc] += w * d;
        }
        if (COUPLED) pn += w * u_new[node * NB + DIM];
      }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F_old[9 * p + i] = F_new[9 * p + i];
    b_old[9 * p + i] = b_new[9 * p + i];
  }
  if (alpha_old && alpha_new) alpha_old[p] = alpha_new[p];
#pragma unroll
  for (int cc = 0; cc < DIM; ++cc) x[p * DIM + cc] += du[cc];
  if (COUPLED && p_part) p_part[p] = pn;
}

extern "C" int gm_g2p(const gm_grid* g, int n, double* x, const double* lp, const double* u_new, const double* u_old,
                      double* F_old, const double* F_new, double* b_old, const double* b_new, double* alpha_old,
                      const double* alpha_new, double* p_part, int coupled, hipStream_t s) {
  if (n == 0) return 0;
  int blk = gm_blocks(n, 128);
  if (g->dim == 2 && !coupled)
    k_g2p<2, false><<<blk, 128, 0, s>>>(*g, n, x, lp, u_new, u_old, F_old, F_new, b_old, b_new, alpha_old, alpha_new,
                                        p_part);
  else if (g->dim == 2)
    k_g2p<2, true><<<blk, 128, 0, s>>>(*g, n, x, lp, u_new, u_old, F_old, F_new, b_old, b_new, alpha_old, alpha_new,
                                       p_part);
  else if (coupled)
    k_g2p<3, true><<<blk, 128, 0, s>>>(*g, n, x, lp, u_new, u_old, F_old, F_new, b_old, b_new, alpha_old, alpha_new,
                                       p_part);
  else {
    gm::set_error("gm_g2p: 3D single-phase is not a supported form");
    return -1;
  }
  GM_LAUNCH_CHECK();
  return 0;
}

// update_GIMP_lp_2d/3d (gimp.py:428-507)
__global__ void k_update_lp(int n, int dim, const double* F_new, double* lp, double lp0, const double* x,
                            double mnx, double mxx, double mny, double mxy, double mnz, double mxz) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double mn[3] = {mnx, mny, mnz}, mx[3] = {mxx, mxy, mxz};
  const double* F = F_new + 9 * p;
  for (int a = 0; a < dim; ++a) {
    double U = sqrt(F[0 * 3 + a] * F[0 * 3 + a] + F[1 * 3 + a] * F[1 * 3 + a] + F[2 * 3 + a] * F[2 * 3 + a]);
    double l = lp0 * U;
    double xa = x[p * dim + a];
    if (xa - l < mn[a]) l = xa - mn[a] - 1e-6;
    if (xa + l > mx[a]) l = mx[a] - xa - 1e-6;
    lp[p * dim + a] = l;
  }
}

extern "C" int gm_update_lp(int n, int dim, const double* F_new, double* lp, double lp0, const double* x,
                            const double* dmin, const double* dmax, hipStream_t s) {
  if (n == 0) return 0;
  k_update_lp<<<gm_blocks(n, 256), 256, 0, s>>>(n, dim, F_new, lp, lp0, x, dmin[0], dmax[0], dmin[1], dmax[1],
                                                dim == 3 ? dmin[2] : 0.0, dim == 3 ? dmax[2] : 0.0);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== standalone entry points
// Per-particle cpGIMP data in the reference's output layout
// (get_GIMP_shape_function_and_gradient_2d/3d, gimp.py:25-112, 197-315, and the
// _avg_ variants :115-193, :318-425): out[p] = (w0_x, w0_y[, w0_z], w1_..., w2_...,
// g0_..., g1_..., g2_...) and avg[p] likewise (weights then gradients), base[p*dim+a] = lo.
__global__ void k_gimp_weights(int n, int dim, const double* x, const double* lp, double h0, double h1, double h2,
                               double* out, double* avg, int* base) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double hs[3] = {h0, h1, h2};
  for (int a = 0; a < dim; ++a) {
    Axis ax = gimp_axis(x[p * dim + a], lp[p * dim + a], hs[a], 1.0 / hs[a]);
    for (int k = 0; k < 3; ++k) {
      out[(long)p * 6 * dim + k * dim + a] = ax.w[k];
      out[(long)p * 6 * dim + 3 * dim + k * dim + a] = ax.g[k];
      if (avg) {
        avg[(long)p * 6 * dim + k * dim + a] = ax.wa[k];
        // the reference's average gradients: -sign(x - x_k)/(4 lp) at k = 0, 2 in the cpGIMP branch,
        // with sign(0) = +1 as Warp's wp.sign defines it
        double gk = 0.0;
        double xa = x[p * dim + a], la = lp[p * dim + a], ih = 1.0 / hs[a];
        bool cp = (int)((xa - la) * ih + 1e-8) != (int)((xa + la) * ih + 1e-8);  // cpGIMP branch
        if (cp && k != 1) {
          double d = xa - (ax.lo + k) * hs[a];
          gk = -1.0 / (4.0 * la) * (d < 0.0 ? -1.0 : 1.0);
        }
        avg[(long)p * 6 * dim + 3 * dim + k * dim + a] = gk;
      }
    }
    if (base) base[p * dim + a] = ax.lo;
  }
}

extern "C" int gm_gimp_weights(int n, int dim, const double* x, const double* lp, const double* h, double* out,
                               double* avg, int* base, hipStream_t s) {
  GM_REQUIRE(h && x && lp && out, "gm_gimp_weights: null argument");
  if (n == 0) return 0;
  k_gimp_weights<<<gm_blocks(n, 256), 256, 0, s>>>(n, dim, x, lp, h[0], h[1], dim == 3 ? h[2] : h[0], out, avg,
                                                   base);
  GM_LAUNCH_CHECK();
  return 0;
}

// Standalone principal return mapping (plasticity.py:8-37 J2; triaxial_DruckerPrager.py:308-412 DP):
// eps_tr, eps (n x 3 principal log strains), dE (n x 9 consistent tangent d eps/d eps_tr, may be null),
// alpha_old/new (J2 hardening, may be null).
__global__ void k_return_map(int n, MatParams m, const double* etr, double* eps, double* dE, const double* ao,
                             double* an) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double e[3], d[9], a_new = 0.0;
  double et[3] = {etr[3 * p], etr[3 * p + 1], etr[3 * p + 2]};
  if (m.kind == MAT_J2) {
    return_j2(et, m, ao ? ao[p] : 0.0, e, d, a_new);
  } else if (m.kind == MAT_DP) {
    return_dp(et, m, e, d);
  } else {
    for (int i = 0; i < 3; ++i) e[i] = et[i];
    for (int i = 0; i < 9; ++i) d[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  for (int i = 0; i < 3; ++i) eps[3 * p + i] = e[i];
  if (dE)
    for (int i = 0; i < 9; ++i) dE[9 * p + i] = d[i];
  if (an) an[p] = a_new;
}

extern "C" int gm_return_map(int n, int mat_kind, int dp_apex, double lam, double mu, double kappa, double H,
                             double phi, double psi, double coh, double shape, const double* eps_tr, double* eps,
                             double* dE, const double* alpha_old, double* alpha_new, hipStream_t s) {
  if (n == 0) return 0;
  MatParams m;
  m.kind = mat_kind;
  m.dp_apex = dp_apex;
  m.lam = lam;
  m.mu = mu;
  m.E = 0.0;
  m.kappa = kappa;
  m.H = H;
  m.phi = phi;
  m.psi = psi;
  m.coh = coh;
  m.shape = shape;
  k_return_map<<<gm_blocks(n, 256), 256, 0, s>>>(n, m, eps_tr, eps, dE, alpha_old, alpha_new);
  GM_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------------------------
// NorSand return mapping (triaxial_NorSand.py:155-400), one particle per lane (gm_device.h
// return_norsand: pressure-dependent hyperelasticity P = p0 exp(-(eps_v - ev0)/kappa), Q = 3 mu
// eps_s; yield Q + eta(P, pi) P; fixed 15-iteration local Newton on (eps_v, eps_s, dlambda, pi_i)
// with hardening h (pi_i - pi*); tension (eps_v > tol) returns zero strain).
// ---------------------------------------------------------------------------------------------
namespace {
__global__ void k_return_map_norsand(int n, gm_norsand_params q, const double* etr, const double* ev_total,
                                     const double* pi_old, double* eps, double* pi_new, double* res_last) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double et[3] = {etr[3 * p], etr[3 * p + 1], etr[3 * p + 2]};
  double e[3], pn, r[4];
  return_norsand(et, ev_total[p], pi_old[p], q, e, pn, r, nullptr, nullptr);
  for (int i = 0; i < 3; ++i) eps[3 * p + i] = e[i];
  pi_new[p] = pn;
  if (res_last)
    for (int i = 0; i < 4; ++i) res_last[4 * p + i] = r[i];
}
}  // namespace

extern "C" int gm_return_map_norsand(int n, const gm_norsand_params* prm, const double* eps_tr,
                                     const double* eps_total_trace, const double* pi_old, double* eps, double* pi_new,
                                     double* res_last, hipStream_t s) {
  GM_REQUIRE(prm, "gm_return_map_norsand: null parameters");
  if (n == 0) return 0;
  GM_REQUIRE(eps_tr && eps_total_trace && pi_old && eps && pi_new, "gm_return_map_norsand: null argument");
  k_return_map_norsand<<<gm_blocks(n, 128), 128, 0, s>>>(n, *prm, eps_tr, eps_total_trace, pi_old, eps, pi_new,
                                                         res_last);
  GM_LAUNCH_CHECK();
  return 0;
}

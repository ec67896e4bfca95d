// Node Jacobian of the 3D u-p form, assembled per particle bin on the matrix cores.
//
// K_IJ = -sum_p L_I(p) D_p M_J(p) (block 4x4; I, J nodes; p particles).  L_I and
// M_J are linear in the particle's shape-function values at the node,
// g = (w, dw/dx, dw/dy, dw/dz, w_avg), so with E_I = L_I D_p (4 x 15) a bin's
// contribution is
//     C_b[(I,r),(J,c)] = sum_{p in b} sum_beta X_p[(I,r),(c,beta)] g_p[J,beta],
// X_p[(I,r),(c,beta)] an entry of E_I (or zero).  All particles of a bin (same
// stencil base node b) reach the same 3x3x3 nodes, so for every column dof c this
// is a dense product (108 rows (I,r)) x (5 n_p) x (27 columns J): the matrix
// cores take it 16x16x4 at a time, FP32 when the tangent is FP32 (mixed form),
// FP64 otherwise.
//   k_jac_bin     a 256-lane workgroup per bin (or a persistent sweep over bins):
//                 E_I for a group of particles in LDS, 14 output tiles of 16x16
//                 over the 4 waves, the bin's 27 x 27 blocks written contiguously
//                 to a workspace (n_nodes x 11664 values).
//   k_jac_gather  one workgroup per block row I: every block (I, J) of the
//                 pattern sums the <= 27 bins b = I - o holding both nodes in the
//                 fixed order (oz, oy, ox); masks, sign and identity rows as in
//                 the node-row kernels (k_jac_gather_staged: the same sums from
//                 chunks staged in LDS).
// Both steps are deterministic (fixed accumulation order).  The workspace is
// written once and read once per assembly, so the pair is bound by HBM on
// 2 x n_bins x 11664 values plus the matrix itself.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gm_common.h"
#include "gm_device.h"
#include "gm_types.h"

using namespace gm;

namespace {

// 3D u-p layout of the particle stress vector Y (17) and unknowns z (15): P (9),
// body/penalty force T (3), mass balance Q, Darcy flux m (3), PPP term S; z =
// grad du (9), du_z, p, pbar, grad p (3).
constexpr int NY = 17, NZ = 15;
constexpr int Y_T = 9, Y_Q = 12, Y_M = 13, Y_S = 16;
constexpr int Z_DU = 9, Z_P = 10, Z_PB = 11, Z_G = 12;
constexpr long CB_BIN = 27L * 27L * 16L;  // values per bin: 27 x 27 blocks of 4 x 4
// chunk of bin `bin`'s blocks (I_loc = o, J_loc = 0..26) in the workspace: bins one
// after another, each its 27 chunks in I_loc order
constexpr long CHUNK = 27L * 16L;
__device__ __forceinline__ long chunk_at(int o, int bin) { return (long)bin * CB_BIN + o * CHUNK; }

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// A (16 x 4): lane l holds A[l % 16][l / 16]; B (4 x 16): B[l / 16][l % 16]; the
// 16 x 16 result: item v of lane l is row 4 (l / 16) + v (FP32) or l / 16 + 4 v
// (FP64) of column l % 16.  An M tile orders its 16 rows (I, r) I-major for FP32
// and r-major for FP64, so that either way lane l ends with the four rows r of
// one node I = l / 16 of the tile.
template <>
struct Mfma<float> {
  typedef f4v acc;
  static constexpr bool r_major = false;
  static __device__ __forceinline__ acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<double> {
  typedef d4v acc;
  static constexpr bool r_major = true;
  static __device__ __forceinline__ acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// z index of E feeding column dof c through shape-function component beta
// (0: w, 1-3: dw/dx_a, 4: w_avg), or ZERO (a column of E that holds zeros) where
// that product vanishes.
constexpr int ZERO = NZ;
__device__ __forceinline__ int zsel(int c, int beta) {
  if (c < 3) {
    if (beta == 0) return c == 2 ? Z_DU : ZERO;
    if (beta == 4) return ZERO;
    return (beta - 1) * 3 + c;
  }
  if (beta == 0) return Z_P;
  if (beta == 4) return Z_PB;
  return Z_G + beta - 1;
}

// shape-function component beta of stencil node J (x fastest) from the axis tables
template <typename TC>
__device__ __forceinline__ TC gcomp(const TC (*ax)[9], int J, int beta) {
  const int jx = J % 3, jy = (J / 3) % 3, jz = J / 9;
  if (beta == 4) return ax[0][6 + jx] * ax[1][6 + jy] * ax[2][6 + jz];
  const TC wx = ax[0][jx], wy = ax[1][jy], wz = ax[2][jz];
  if (beta == 0) return wx * wy * wz;
  if (beta == 1) return ax[0][3 + jx] * wy * wz;
  if (beta == 2) return wx * ax[1][3 + jy] * wz;
  return wx * wy * ax[2][3 + jz];
}

}  // namespace

// Per-bin products.  Workspace cb[b] = 27 x 27 blocks (I_loc, J_loc), I_loc = o_x +
// 3 o_y + 9 o_z with node = b + o.  A workgroup takes bins b = blockIdx.x, +gridDim.x,
// ... (one bin when the grid has one workgroup per bin, a persistent sweep when it is
// smaller) and each bin's particles PG at a time: the E table of a group in LDS, its
// products accumulated in the matrix-core registers, the bin's 27 x 27 blocks stored
// after its last group.  The global loads of the next group (of this bin or of the
// next one) are issued before the products of the current group when PF is set, so
// their latency hides behind the matrix cores (at the cost of the registers that hold
// them), else at the top of the group.
// K order: a K step of the MFMA is one shape-function component beta and four
// particles (lane group l / 16 = particle): the five components of a quad of
// particles, then the next quad, whatever PG is.  After the
// K loop unrolls, every A and B operand address is a per-lane base plus an immediate,
// and the (c, beta) products that vanish (u columns: no w term except the penalty
// on u_z, no w_avg term) are not issued.  Every (PG, OCC, PF, grid) variant adds the
// same products in the same order, so all of them write the same bits.
template <typename TD, typename TC, int PG, int OCC, bool PF>
__global__ void __launch_bounds__(256, OCC) k_jac_bin(gm_grid g, const double* x_s, const double* lp_s,
                                                      const int* cell_start, int n_bins, const TD* D, TC* cb) {
  static_assert(PG % 4 == 0, "particle groups are whole K steps");
  constexpr int KST = 5 * PG / 4;  // K steps of 4 per group
  constexpr int NI = 4 * NZ;       // (z, r) entries of a node row of E_p
  // shape-function tables of a group, two buffers: with PF the next group's are
  // computed (by wave 3, whose last row tile is empty) while the products run
  __shared__ TC AXT_[2][PG][3][9];
  // E_p[I][z][r] (r fastest), node rows 68 entries apart and particles 28 x 68 apart:
  // the 32 (particle, I, r) entries half a wave reads at once fall in 32 distinct banks.
  // Node row I = 27 is zero (the padding rows of the last row tile read it).
  constexpr int EI = 68, EP = 28 * EI;
  __shared__ TC E_[PG * EP];
#define E(p, I, z, r) E_[(p) * EP + (I) * EI + (z) * 4 + (r)]
  // B operand g_p[J][beta], J padded to 32 with zeros, particles 176 entries apart
  constexpr int BB = 32, BP = 5 * BB + 16;
  __shared__ TC BT_[PG * BP];
#define BT(p, beta, J) BT_[(p) * BP + (beta) * BB + (J)]
  int b = blockIdx.x, ks = 0, ke = 0;
  for (; b < n_bins; b += gridDim.x) {  // first non-empty bin (uniform across the workgroup)
    ks = cell_start[b];
    ke = cell_start[b + 1];
    if (ke > ks) break;
  }
  if (b >= n_bins) return;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);  // uniform across the wave
  for (int q = t; q < PG * NI; q += 256) E(q / NI, 27, (q % NI) >> 2, q & 3) = 0;  // never written again
  typedef typename Mfma<TC>::acc ACC;
  ACC acc[4][4];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[s][c] = ACC{0, 0, 0, 0};
  // output tiles tt = wave + 4 s: row tile mt = tt / 2 = wave / 2 + 2 s (16 rows (I, r)),
  // column tile nt = tt % 2 = wave % 2 (16 columns J); waves 2 and 3 have no fourth
  // tile (rows I >= 28)
  const int lg = lane >> 4, rt = lane & 15;
  const int r_l = Mfma<TC>::r_major ? (rt >> 2) : (rt & 3);
  const int i_l = Mfma<TC>::r_major ? (rt & 3) : (rt >> 2);
  int abase[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) abase[s] = lg * EP + min((wave >> 1) * 4 + 8 * s + i_l, 27) * EI + r_l;
  const int bbase = lg * BP + (wave & 1) * 16 + rt;
  constexpr int NQ = (PG * NI + 255) / 256;  // (p, z, r) items per thread
  constexpr int GT0 = PF ? 192 : 0;           // first thread of the shape-function tables
  static_assert(PG * 3 <= 64, "the tables of a group are one wave's work");
  TC dv[NQ][5];
  double xa = 0.0, la = 0.0;
  // tangent entries of this thread's (p, z, r) items, position and domain size of
  // particle (t - GT0) / 3 (axis (t - GT0) % 3) of the group starting at p0 (np particles)
  auto load_group = [&](int p0, int np) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int q = t + 256 * j;
      const int p = q / NI, zr = q % NI, z = zr >> 2, r = zr & 3;
#pragma unroll
      for (int i = 0; i < 5; ++i) dv[j][i] = (TC)0;
      if (q < PG * NI && p < np) {
        const TD* Dp = D + (long)(p0 + p) * NY * NZ;
        if (r < 3) {
          dv[j][0] = (TC)Dp[(0 * 3 + r) * NZ + z];
          dv[j][1] = (TC)Dp[(1 * 3 + r) * NZ + z];
          dv[j][2] = (TC)Dp[(2 * 3 + r) * NZ + z];
          dv[j][3] = (TC)Dp[(Y_T + r) * NZ + z];
        } else {
          const TC dS = (TC)Dp[Y_S * NZ + z];
          dv[j][0] = (TC)Dp[(Y_M + 0) * NZ + z];
          dv[j][1] = (TC)Dp[(Y_M + 1) * NZ + z];
          dv[j][2] = (TC)Dp[(Y_M + 2) * NZ + z];
          dv[j][3] = (TC)Dp[Y_Q * NZ + z] + dS;
          dv[j][4] = -dS;
        }
      }
    }
    xa = 0.0;
    la = 0.0;
    if (t >= GT0 && t < GT0 + PG * 3 && (t - GT0) / 3 < np) {
      xa = x_s[(long)(p0 + (t - GT0) / 3) * 3 + (t - GT0) % 3];
      la = lp_s[(long)(p0 + (t - GT0) / 3) * 3 + (t - GT0) % 3];
    }
  };
  // shape-function tables of the group whose positions are in xa / la (np particles)
  auto axis_tables = [&](int buf, int np) {
    if (t >= GT0 && t < GT0 + PG * 3) {
      const int p = (t - GT0) / 3, a = (t - GT0) % 3;
      if (p < np) {
        const Axis ax = gimp_axis(xa, la, g.h[a], g.inv_h[a]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          AXT_[buf][p][a][i] = (TC)ax.w[i];
          AXT_[buf][p][a][3 + i] = (TC)ax.g[i];
          AXT_[buf][p][a][6 + i] = (TC)ax.wa[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) AXT_[buf][p][a][i] = (TC)0;
      }
    }
  };
  int p0 = ks, cur = 0;
  if (PF) {
    load_group(p0, min(PG, ke - p0));
    axis_tables(0, min(PG, ke - p0));
  }
  while (true) {
    const int np = min(PG, ke - p0);
    if (!PF) load_group(p0, np);
    __syncthreads();  // the previous group's LDS reads are done (and, with PF, this one's tables written)
    if (!PF) {
      axis_tables(cur, np);
      __syncthreads();
    }
    const TC(*AXT)[3][9] = AXT_[cur];
    for (int q = t; q < PG * 5 * BB; q += 256) {
      const int p = q / (5 * BB), beta = (q / BB) % 5, J = q % BB;
      BT(p, beta, J) = (J < 27) ? gcomp<TC>(AXT[p], J, beta) : (TC)0;
    }
    // E_p[I][z][r] = (L_I D_p)[r][z]: u rows sum_b dw/dx_b D[(b,r)] + w D[T_r];
    // the pressure row w D[Q] + sum_b dw/dx_b D[M_b] + (w - w_avg) D[S]
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int q = t + 256 * j;
      if (q >= PG * NI) break;
      const int p = q / NI, zr = q % NI, z = zr >> 2, r = zr & 3;
      const TC d0 = dv[j][0], d1 = dv[j][1], d2 = dv[j][2], dw = dv[j][3], dwa = dv[j][4];
      // sum-factorised over the tensor-product axes: with u = gx d0 + wx dw, v = wx d1,
      // q = wx d2, a = wax dwa per x node, E = wz (wy u + gy v) + gz (wy q) + waz (way a)
      TC wx[3], gx[3], ax[3], wy[3], gy[3], ay[3], wz[3], gz[3], az[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        wx[i] = AXT[p][0][i];
        gx[i] = AXT[p][0][3 + i];
        ax[i] = AXT[p][0][6 + i];
        wy[i] = AXT[p][1][i];
        gy[i] = AXT[p][1][3 + i];
        ay[i] = AXT[p][1][6 + i];
        wz[i] = AXT[p][2][i];
        gz[i] = AXT[p][2][3 + i];
        az[i] = AXT[p][2][6 + i];
      }
#pragma unroll
      for (int ix = 0; ix < 3; ++ix) {
        const TC u = gx[ix] * d0 + wx[ix] * dw, v = wx[ix] * d1, qx = wx[ix] * d2, a = ax[ix] * dwa;
#pragma unroll
        for (int iy = 0; iy < 3; ++iy) {
          const TC U = wy[iy] * u + gy[iy] * v, Q = wy[iy] * qx, A2 = ay[iy] * a;
#pragma unroll
          for (int iz = 0; iz < 3; ++iz) E(p, ix + 3 * iy + 9 * iz, z, r) = wz[iz] * U + gz[iz] * Q + az[iz] * A2;
        }
      }
    }
    // the next group: the rest of this bin, else the first of the next non-empty bin
    int nb = b, nks = ks, nke = ke, np0 = p0 + PG;
    const bool bin_done = np0 >= ke;
    if (bin_done) {
      for (nb = b + gridDim.x; nb < n_bins; nb += gridDim.x) {
        nks = cell_start[nb];
        nke = cell_start[nb + 1];
        if (nke > nks) break;
      }
      np0 = nks;
    }
    const bool more = nb < n_bins;
    if (PF && more) load_group(np0, min(PG, nke - np0));
    __syncthreads();  // E and B tables complete
    // particles past np have all-zero axis tables, hence zero B
#pragma unroll
    for (int kst = 0; kst < KST; ++kst) {
      const int beta = kst % 5, pb = (kst / 5) * 4;  // constants once unrolled
      const TC bk = BT_[bbase + pb * BP + beta * BB];
      // the step's A operands are read before its products are issued
      TC av[4][4];
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int z = zsel(c, beta);
          av[s][c] = (z == ZERO) ? (TC)0 : E_[abase[s] + pb * EP + z * 4];  // z resolved at compile time
        }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (s == 3 && wave >= 2) break;  // wave-uniform: no fourth tile
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (zsel(c, beta) != ZERO) acc[s][c] = Mfma<TC>::step(av[s][c], bk, acc[s][c]);
      }
    }
    if (bin_done) {
      // lane holds rows r = 0..3 of node I_loc = 4 mt + lane / 16 in column J_loc =
      // 16 nt + lane % 16: one whole 4 x 4 block
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int mt = (wave >> 1) + 2 * s, nt = wave & 1;
        const int I = mt * 4 + (lane >> 4), J = nt * 16 + (lane & 15);
        if (I < 27 && J < 27) {
          TC* o = cb + chunk_at(I, b) + J * 16;
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) o[r * 4 + c] = acc[s][c][r];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[s][c] = ACC{0, 0, 0, 0};
      }
    }
    if (!more) break;
    if (PF) {
      cur ^= 1;
      axis_tables(cur, min(PG, nke - np0));  // read after the next barrier
    }
    b = nb;
    ks = nks;
    ke = nke;
    p0 = np0;
  }
}

#undef E
#undef BT

// One workgroup per block row I, four lanes per block (one per block row r).  The
// row's 27 candidate bins (b = I - o, -1 when outside the grid or empty) are read
// once into LDS.
template <typename TC, typename TV>
__global__ void __launch_bounds__(512) k_jac_gather(gm_grid g, const int* cell_start, const TC* cb,
                                                    const uint8_t* freem, const int* row_ptr,
                                                    const uint64_t* row_mask, TV* vals) {
  __shared__ int SLOT_BIT[125];
  __shared__ int BIN[27];
  const int I = blockIdx.x, t = threadIdx.x;
  const uint64_t m0 = row_mask[2 * I], m1 = row_mask[2 * I + 1];
  const int nx1 = g.nx + 1, nxy = (g.nx + 1) * (g.ny + 1);
  const int ci[3] = {I % nx1, (I / nx1) % (g.ny + 1), I / nxy};
  if (t < 125) {
    const uint64_t w = (t < 64) ? m0 : m1;
    const int sh = t & 63;
    if ((w >> sh) & 1ull) {
      const int below = (t < 64) ? __popcll(m0 & ((1ull << sh) - 1ull))
                                 : __popcll(m0) + __popcll(m1 & ((1ull << sh) - 1ull));
      SLOT_BIT[below] = t;
    }
  } else if (t >= 128 && t < 128 + 27) {
    const int o = t - 128, ox = o % 3, oy = (o / 3) % 3, oz = o / 9;
    const int bc[3] = {ci[0] - ox, ci[1] - oy, ci[2] - oz};
    int bn = -1;
    if (bc[0] >= 0 && bc[0] <= g.nx && bc[1] >= 0 && bc[1] <= g.ny && bc[2] >= 0 && bc[2] <= g.nz) {
      bn = bc[0] + bc[1] * nx1 + bc[2] * nxy;
      if (cell_start[bn + 1] == cell_start[bn]) bn = -1;
    }
    BIN[o] = bn;
  }
  __syncthreads();
  const int nnz = __popcll(m0) + __popcll(m1);
  const int slot = t >> 2, r = t & 3;
  if (slot >= nnz) return;
  const int bit = SLOT_BIT[slot];
  const int d[3] = {bit % 5 - 2, (bit / 5) % 5 - 2, bit / 25 - 2};
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    const int kx = d[0] + o % 3, ky = d[1] + (o / 3) % 3, kz = d[2] + o / 9;
    const int bn = BIN[o];
    if (bn >= 0 && kx >= 0 && kx < 3 && ky >= 0 && ky < 3 && kz >= 0 && kz < 3) {
      const TC* src = cb + chunk_at(o, bn) + (kx + 3 * ky + 9 * kz) * 16 + r * 4;
#pragma unroll
      for (int c = 0; c < 4; ++c) s[c] += (double)src[c];
    }
  }
  const int Jn = I + d[0] + d[1] * nx1 + d[2] * nxy;
  const bool diag = (Jn == I);
  const bool fr = freem[I * 4 + r];
  TV* out = vals + (long)(row_ptr[I] + slot) * 16 + r * 4;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bool fc = freem[Jn * 4 + c];
    double v = (fr && fc) ? -s[c] : 0.0;
    if (diag && r == c && !fr) v = 1.0;
    out[c] = (TV)v;
  }
}

// The same sums with coalesced reads: one workgroup per block row I stages the row's
// <= 27 bin chunks (bin I - o, I_loc = o: 27 blocks, contiguous) in LDS
// with whole 16-byte vector loads, OB chunks at a time, and every (block, row r) item
// then adds its <= 27 contributions from LDS in the same order (o = 0..26) as
// k_jac_gather: both write the same bits.
template <typename TC, typename TV, int OB>
__global__ void __launch_bounds__(256, 3) k_jac_gather_staged(gm_grid g, const int* cell_start, const TC* cb,
                                                           const uint8_t* freem, const int* row_ptr,
                                                           const uint64_t* row_mask, TV* vals) {
  static_assert(27 % OB == 0, "chunk rounds cover o = 0..26");
  constexpr int CH = 27 * 16;          // values per chunk
  constexpr int CV = CH * sizeof(TC) / 16;  // 16-byte vectors per chunk
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  __shared__ int SLOT_BIT[125];
  __shared__ int BIN[27];
  __shared__ u4 ST[OB][CV];
  const int I = blockIdx.x, t = threadIdx.x;
  const uint64_t m0 = row_mask[2 * I], m1 = row_mask[2 * I + 1];
  const int nx1 = g.nx + 1, nxy = (g.nx + 1) * (g.ny + 1);
  const int ci[3] = {I % nx1, (I / nx1) % (g.ny + 1), I / nxy};
  if (t < 125) {
    const uint64_t w = (t < 64) ? m0 : m1;
    const int sh = t & 63;
    if ((w >> sh) & 1ull) {
      const int below = (t < 64) ? __popcll(m0 & ((1ull << sh) - 1ull))
                                 : __popcll(m0) + __popcll(m1 & ((1ull << sh) - 1ull));
      SLOT_BIT[below] = t;
    }
  } else if (t >= 128 && t < 128 + 27) {
    const int o = t - 128, ox = o % 3, oy = (o / 3) % 3, oz = o / 9;
    const int bc[3] = {ci[0] - ox, ci[1] - oy, ci[2] - oz};
    int bn = -1;
    if (bc[0] >= 0 && bc[0] <= g.nx && bc[1] >= 0 && bc[1] <= g.ny && bc[2] >= 0 && bc[2] <= g.nz) {
      bn = bc[0] + bc[1] * nx1 + bc[2] * nxy;
      if (cell_start[bn + 1] == cell_start[bn]) bn = -1;
    }
    BIN[o] = bn;
  }
  __syncthreads();
  const int nnz = __popcll(m0) + __popcll(m1);
  // items (slot, r) = (it / 4, it % 4), it = t and t + 256
  int dd[2][3];
  bool live[2];
  double s[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int it = t + 256 * h;
    live[h] = it < 4 * nnz;
    const int bit = live[h] ? SLOT_BIT[it >> 2] : 62;
    dd[h][0] = bit % 5 - 2;
    dd[h][1] = (bit / 5) % 5 - 2;
    dd[h][2] = bit / 25 - 2;
#pragma unroll
    for (int c = 0; c < 4; ++c) s[h][c] = 0.0;
  }
#pragma unroll 1
  for (int o0 = 0; o0 < 27; o0 += OB) {
    if (o0) __syncthreads();  // the previous round's LDS reads are done
    // all of a thread's loads are issued before its first LDS store
    constexpr int NJ = (OB * CV + 255) / 256;
    u4 tmp[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int q = t + 256 * j;
      tmp[j] = u4{0u, 0u, 0u, 0u};
      if (q < OB * CV) {
        const int oo = q / CV, f = q % CV;
        const int bn = BIN[o0 + oo];
        if (bn >= 0) tmp[j] = reinterpret_cast<const u4*>(cb + chunk_at(o0 + oo, bn))[f];
      }
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int q = t + 256 * j;
      if (q < OB * CV) ST[q / CV][q % CV] = tmp[j];
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (!live[h]) continue;
      const int r = (t + 256 * h) & 3;
#pragma unroll 3
      for (int oo = 0; oo < OB; ++oo) {
        const int o = o0 + oo;
        const int kx = dd[h][0] + o % 3, ky = dd[h][1] + (o / 3) % 3, kz = dd[h][2] + o / 9;
        if (BIN[o] >= 0 && kx >= 0 && kx < 3 && ky >= 0 && ky < 3 && kz >= 0 && kz < 3) {
          const TC* src = reinterpret_cast<const TC*>(&ST[oo][0]) + (kx + 3 * ky + 9 * kz) * 16 + r * 4;
#pragma unroll
          for (int c = 0; c < 4; ++c) s[h][c] += (double)src[c];
        }
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (!live[h]) continue;
    const int it = t + 256 * h, slot = it >> 2, r = it & 3;
    const int Jn = I + dd[h][0] + dd[h][1] * nx1 + dd[h][2] * nxy;
    const bool diag = (Jn == I);
    const bool fr = freem[I * 4 + r];
    TV* out = vals + (long)(row_ptr[I] + slot) * 16 + r * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bool fc = freem[Jn * 4 + c];
      double v = (fr && fc) ? -s[h][c] : 0.0;
      if (diag && r == c && !fr) v = 1.0;
      out[c] = (TV)v;
    }
  }
}

// ================================================================== residual (3D u-p)
// r_I = sum_p L_I Y_p per bin (node I_loc of the bin's 3x3x3 box, its particles in
// bin order), then per node the fixed-order sum over the <= 27 bins holding it.
__global__ void __launch_bounds__(64) k_res_bin(gm_grid g, const double* x_s, const double* lp_s,
                                                const int* cell_start, const double* Y, double* cb) {
  constexpr int PB = 16;
  __shared__ double AX[PB][3][9];
  const int b = blockIdx.x;
  const int ks = cell_start[b], ke = cell_start[b + 1];
  if (ks == ke) return;  // uniform across the workgroup
  const int t = threadIdx.x;
  const int ix = t % 3, iy = (t / 3) % 3, iz = t / 9;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  for (int p0 = ks; p0 < ke; p0 += PB) {
    const int np = min(PB, ke - p0);
    __syncthreads();
    if (t < PB * 3) {
      const int p = t / 3, a = t % 3;
      if (p < np) {
        const Axis ax = gimp_axis(x_s[(long)(p0 + p) * 3 + a], lp_s[(long)(p0 + p) * 3 + a], g.h[a], g.inv_h[a]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          AX[p][a][i] = ax.w[i];
          AX[p][a][3 + i] = ax.g[i];
          AX[p][a][6 + i] = ax.wa[i];
        }
      }
    }
    __syncthreads();
    if (t < 27) {
      for (int p = 0; p < np; ++p) {
        const double wx = AX[p][0][ix], wy = AX[p][1][iy], wz = AX[p][2][iz];
        const double w = wx * wy * wz;
        const double G[3] = {AX[p][0][3 + ix] * wy * wz, wx * AX[p][1][3 + iy] * wz, wx * wy * AX[p][2][3 + iz]};
        const double wa = AX[p][0][6 + ix] * AX[p][1][6 + iy] * AX[p][2][6 + iz];
        const double* yp = Y + (long)(p0 + p) * NY;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double sv = 0.0;
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) sv += G[bb] * yp[bb * 3 + c];
          r[c] += sv + w * yp[Y_T + c];
        }
        double sv = w * yp[Y_Q];
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) sv += G[bb] * yp[Y_M + bb];
        sv += (w - wa) * yp[Y_S];
        r[3] += sv;
      }
    }
  }
  if (t < 27)
#pragma unroll
    for (int c = 0; c < 4; ++c) cb[((long)b * 27 + t) * 4 + c] = r[c];
}

__global__ void __launch_bounds__(256) k_res_gather(gm_grid g, const int* cell_start, const double* cb,
                                                    const uint8_t* freem, double* rhs) {
  const int I = blockIdx.x * 64 + threadIdx.x / 4, c = threadIdx.x % 4;
  if (I >= g.n_nodes) return;
  const int nx1 = g.nx + 1, nxy = (g.nx + 1) * (g.ny + 1);
  const int ci[3] = {I % nx1, (I / nx1) % (g.ny + 1), I / nxy};
  double s = 0.0;
  for (int o = 0; o < 27; ++o) {
    const int ox = o % 3, oy = (o / 3) % 3, oz = o / 9;
    const int bc[3] = {ci[0] - ox, ci[1] - oy, ci[2] - oz};
    if (bc[0] < 0 || bc[0] > g.nx || bc[1] < 0 || bc[1] > g.ny || bc[2] < 0 || bc[2] > g.nz) continue;
    const int bn = bc[0] + bc[1] * nx1 + bc[2] * nxy;
    if (cell_start[bn + 1] == cell_start[bn]) continue;
    s += cb[((long)bn * 27 + o) * 4 + c];
  }
  rhs[I * 4 + c] = freem[I * 4 + c] ? s : 0.0;
}

extern "C" size_t gm_residual_work_bytes(const gm_grid* g) {
  return g ? (size_t)g->n_nodes * 27 * 4 * sizeof(double) : 0;
}

extern "C" int gm_node_residual_binned(const gm_grid* g, const gm_bins* b, const double* Y, const uint8_t* freem,
                                       double* rhs, void* work, size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && Y && freem && rhs && work, "gm_node_residual_binned: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_node_residual_binned: the 3D u-p form only");
  GM_REQUIRE(work_bytes >= gm_residual_work_bytes(g),
             "gm_node_residual_binned: workspace smaller than gm_residual_work_bytes");
  const int n = g->n_nodes;
  if (n == 0) return 0;
  k_res_bin<<<n, 64, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, Y, (double*)work);
  GM_LAUNCH_CHECK();
  k_res_gather<<<gm_blocks(n, 64), 256, 0, s>>>(*g, b->cell_start, (const double*)work, freem, rhs);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== consistent mass (3D)
// M_IJ = sum_p m_p w_I w_J in the same node pattern (one value per block slot),
// assembled like the Jacobian: per bin the 27 x 27 products of its particles (in
// bin order), then per row the fixed-order sum over the bins holding both nodes.
__global__ void __launch_bounds__(256) k_mass_bin(gm_grid g, const double* x_s, const double* lp_s,
                                                  const int* cell_start, double p_mass, double* cb) {
  constexpr int PB = 32;
  __shared__ double W[PB][3][3];
  const int b = blockIdx.x;
  const int ks = cell_start[b], ke = cell_start[b + 1];
  if (ks == ke) return;  // uniform across the workgroup
  const int t = threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int p0 = ks; p0 < ke; p0 += PB) {
    const int np = min(PB, ke - p0);
    __syncthreads();
    if (t < PB * 3) {
      const int p = t / 3, a = t % 3;
      if (p < np) {
        const Axis ax = gimp_axis(x_s[(long)(p0 + p) * 3 + a], lp_s[(long)(p0 + p) * 3 + a], g.h[a], g.inv_h[a]);
        W[p][a][0] = ax.w[0];
        W[p][a][1] = ax.w[1];
        W[p][a][2] = ax.w[2];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int q = t + 256 * j;
      if (q < 729) {
        const int I = q / 27, J = q % 27;
        const int ix = I % 3, iy = (I / 3) % 3, iz = I / 9, jx = J % 3, jy = (J / 3) % 3, jz = J / 9;
        for (int p = 0; p < np; ++p) {
          const double wI = W[p][0][ix] * W[p][1][iy] * W[p][2][iz];
          const double wJ = W[p][0][jx] * W[p][1][jy] * W[p][2][jz];
          acc[j] += wI * wJ * p_mass;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int q = t + 256 * j;
    if (q < 729) cb[(long)b * 729 + q] = acc[j];
  }
}

__global__ void __launch_bounds__(128) k_mass_gather(gm_grid g, const int* cell_start, const double* cb,
                                                     const int* row_ptr, const uint64_t* row_mask, double* mvals) {
  __shared__ int SLOT_BIT[125];
  __shared__ int BIN[27];
  const int I = blockIdx.x, t = threadIdx.x;
  const uint64_t m0 = row_mask[2 * I], m1 = row_mask[2 * I + 1];
  const int nx1 = g.nx + 1, nxy = (g.nx + 1) * (g.ny + 1);
  const int ci[3] = {I % nx1, (I / nx1) % (g.ny + 1), I / nxy};
  if (t < 125) {
    const uint64_t w = (t < 64) ? m0 : m1;
    const int sh = t & 63;
    if ((w >> sh) & 1ull) {
      const int below = (t < 64) ? __popcll(m0 & ((1ull << sh) - 1ull))
                                 : __popcll(m0) + __popcll(m1 & ((1ull << sh) - 1ull));
      SLOT_BIT[below] = t;
    }
  }
  if (t < 27) {
    const int ox = t % 3, oy = (t / 3) % 3, oz = t / 9;
    const int bc[3] = {ci[0] - ox, ci[1] - oy, ci[2] - oz};
    int bn = -1;
    if (bc[0] >= 0 && bc[0] <= g.nx && bc[1] >= 0 && bc[1] <= g.ny && bc[2] >= 0 && bc[2] <= g.nz) {
      bn = bc[0] + bc[1] * nx1 + bc[2] * nxy;
      if (cell_start[bn + 1] == cell_start[bn]) bn = -1;
    }
    BIN[t] = bn;
  }
  __syncthreads();
  const int nnz = __popcll(m0) + __popcll(m1);
  if (t >= nnz) return;
  const int bit = SLOT_BIT[t];
  const int d[3] = {bit % 5 - 2, (bit / 5) % 5 - 2, bit / 25 - 2};
  double s = 0.0;
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    const int kx = d[0] + o % 3, ky = d[1] + (o / 3) % 3, kz = d[2] + o / 9;
    const int bn = BIN[o];
    if (bn >= 0 && kx >= 0 && kx < 3 && ky >= 0 && ky < 3 && kz >= 0 && kz < 3)
      s += cb[(long)bn * 729 + o * 27 + (kx + 3 * ky + 9 * kz)];
  }
  mvals[row_ptr[I] + t] = s;
}

extern "C" size_t gm_mass_work_bytes(const gm_grid* g) { return g ? (size_t)g->n_nodes * 729 * sizeof(double) : 0; }

extern "C" int gm_mass_binned(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, double p_mass,
                              double* mvals, void* work, size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && pat && mvals && work, "gm_mass_binned: null argument");
  GM_REQUIRE(g->dim == 3, "gm_mass_binned: 3D grids only");
  GM_REQUIRE(work_bytes >= gm_mass_work_bytes(g), "gm_mass_binned: workspace smaller than gm_mass_work_bytes");
  const int n = g->n_nodes;
  if (n == 0) return 0;
  k_mass_bin<<<n, 256, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, p_mass, (double*)work);
  GM_LAUNCH_CHECK();
  k_mass_gather<<<n, 128, 0, s>>>(*g, b->cell_start, (const double*)work, pat->row_ptr, pat->row_mask, mvals);
  GM_LAUNCH_CHECK();
  return 0;
}

// ================================================================== P2G (3D)
// The activation flags, pressure-boundary flags and consistent-mass right-hand side of
// P2G_coupled_3d (mpm_utils.py:260-386) in two passes: per bin, each of the 27 nodes of
// its 3x3x3 box accumulates its particles' terms in bin order (the right-hand side
// w_I m_p p_p sum_J w_J, activation when w_I > threshold inside the particle's support,
// pressure boundary over all 27 nodes when the node lies at or above a flagged
// particle); per node, a fixed-order pass over the <= 27 bins holding it.
__global__ void __launch_bounds__(64) k_p2g_bin(gm_grid g, const double* x_s, const double* lp_s,
                                                const int* cell_start, const int* sorted, const uint8_t* ext_s,
                                                double threshold, const uint8_t* pbc_flag, double node_h_last,
                                                double p_mass, const double* p_part, double* cb, uint8_t* cf) {
  constexpr int PB = 16;
  __shared__ double W[PB][3][3];
  const int b = blockIdx.x;
  const int ks = cell_start[b], ke = cell_start[b + 1];
  if (ks == ke) return;  // uniform across the workgroup
  const int t = threadIdx.x;
  const int ox = t % 3, oy = (t / 3) % 3, oz = t / 9;
  const int nxy = (g.nx + 1) * (g.ny + 1);
  const double node_z = (double)(b / nxy + oz) * node_h_last;
  double rhs = 0.0;
  bool a_on = false, pb = false;
  for (int p0 = ks; p0 < ke; p0 += PB) {
    const int np = min(PB, ke - p0);
    __syncthreads();
    if (t < PB * 3) {
      const int p = t / 3, a = t % 3;
      if (p < np) {
        const Axis ax = gimp_axis(x_s[(long)(p0 + p) * 3 + a], lp_s[(long)(p0 + p) * 3 + a], g.h[a], g.inv_h[a]);
        W[p][a][0] = ax.w[0];
        W[p][a][1] = ax.w[1];
        W[p][a][2] = ax.w[2];
      }
    }
    __syncthreads();
    if (t < 27) {
      for (int p = 0; p < np; ++p) {
        const int k = p0 + p;
        const int pid = sorted[k];
        if (pbc_flag && pbc_flag[pid] && node_z >= x_s[(long)k * 3 + 2]) pb = true;
        const uint8_t ex = ext_s[k];
        const bool in = ox < ((ex & 1) ? 3 : 2) && oy < ((ex & 2) ? 3 : 2) && oz < ((ex & 4) ? 3 : 2);
        if (!in) continue;
        const double w = W[p][0][ox] * W[p][1][oy] * W[p][2][oz];
        if (threshold < 0.0 || w > threshold) a_on = true;
        if (p_part) {
          double wsum = 1.0;
#pragma unroll
          for (int a = 0; a < 3; ++a) wsum *= W[p][a][0] + W[p][a][1] + W[p][a][2];
          rhs += w * wsum * p_mass * p_part[pid];
        }
      }
    }
  }
  if (t < 27) {
    cb[(long)b * 27 + t] = rhs;
    cf[(long)b * 27 + t] = (uint8_t)((a_on ? 1 : 0) | (pb ? 2 : 0));
  }
}

__global__ void __launch_bounds__(256) k_p2g_gather(gm_grid g, const int* cell_start, const double* cb,
                                                    const uint8_t* cf, uint8_t* act, uint8_t* bflag, double* mrhs) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= g.n_nodes) return;
  const int nx1 = g.nx + 1, nxy = (g.nx + 1) * (g.ny + 1);
  const int ci[3] = {I % nx1, (I / nx1) % (g.ny + 1), I / nxy};
  double s = 0.0;
  int flags = 0;
  for (int o = 0; o < 27; ++o) {
    const int bc[3] = {ci[0] - o % 3, ci[1] - (o / 3) % 3, ci[2] - o / 9};
    if (bc[0] < 0 || bc[1] < 0 || bc[2] < 0) continue;
    const int bn = bc[0] + bc[1] * nx1 + bc[2] * nxy;
    if (cell_start[bn + 1] == cell_start[bn]) continue;
    s += cb[(long)bn * 27 + o];
    flags |= cf[(long)bn * 27 + o];
  }
  if (flags & 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) act[I * 4 + k] = 1;
  if ((flags & 2) && bflag) bflag[I * 4 + 3] = 1;
  if (mrhs) mrhs[I] = s;
}

extern "C" size_t gm_p2g_work_bytes(const gm_grid* g) {
  return g ? (size_t)g->n_nodes * 27 * (sizeof(double) + 1) : 0;
}

extern "C" int gm_p2g_binned(const gm_grid* g, const gm_bins* b, double threshold, const uint8_t* pbc_flag,
                             double node_h_last, double p_mass, const double* p_part, uint8_t* act, uint8_t* bflag,
                             double* mrhs, void* work, size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && act && work, "gm_p2g_binned: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_p2g_binned: the 3D u-p form only");
  GM_REQUIRE(!mrhs || p_part, "gm_p2g_binned: the right-hand side needs the particle pressures");
  GM_REQUIRE(work_bytes >= gm_p2g_work_bytes(g), "gm_p2g_binned: workspace smaller than gm_p2g_work_bytes");
  const int n = g->n_nodes;
  if (n == 0) return 0;
  double* cb = (double*)work;
  uint8_t* cf = (uint8_t*)(cb + (size_t)n * 27);
  k_p2g_bin<<<n, 64, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, b->sorted, b->ext_s, threshold, pbc_flag,
                             node_h_last, p_mass, mrhs ? p_part : nullptr, cb, cf);
  GM_LAUNCH_CHECK();
  k_p2g_gather<<<gm_blocks(n, 256), 256, 0, s>>>(*g, b->cell_start, cb, cf, act, bflag, mrhs);
  GM_LAUNCH_CHECK();
  return 0;
}

extern "C" size_t gm_node_jacobian_work_bytes(const gm_grid* g, int tangent_fp32) {
  if (!g) return 0;
  return (size_t)g->n_nodes * (size_t)CB_BIN * (tangent_fp32 ? sizeof(float) : sizeof(double));
}

// ------------------------------------------------------------ variants and their tuner
// Per-bin kernel variants (particles per group, workgroups per CU, grid):
//   FP32 tangent: 0 (8, 2, one per bin), 1 (8, 2, persistent), 2 (4, 4, persistent), 3 (4, 3, persistent)
//   FP64 tangent: 0 (4, 2, one per bin), 1 (4, 2, persistent), 2 (8, 1, one per bin), 3 (8, 1, persistent)
// "persistent": one workgroup per resident slot (CUs x workgroups per CU), each
// sweeping bins blockIdx.x + k gridDim.x, with the next group's loads issued before
// the current group's products; "one per bin": loads at the top of each group.  Gather
// variants: 0 (per-lane reads from the workspace), 1 (all 27 chunks of a row staged in
// LDS at once; 9 at a time in FP64), 2 (9 at a time; 3 in FP64).  Every variant writes
// the same bits.
namespace {
constexpr int N_BIN_VARIANTS = 4, N_GATHER_VARIANTS = 3;
int g_bin_variant[2] = {0, 0};     // [tangent_fp32]
int g_gather_variant[2] = {0, 0};  // [tangent_fp32]

int cu_count() {
  static int n = 0;
  if (n == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) != hipSuccess) return 0;
    n = pr.multiProcessorCount;
  }
  return n;
}

int launch_bins(const gm_grid* g, const gm_bins* b, const void* D, int fp32, int variant, void* work,
                hipStream_t s) {
  const int n = g->n_nodes;
  const int cus = cu_count();
  GM_REQUIRE(cus > 0, "gm_jacobian_bins: no device properties");
#define GM_BIN(TT, PG, OCC, PERSIST)                                                                     \
  if (PERSIST)                                                                                           \
    k_jac_bin<TT, TT, PG, OCC, true><<<std::min(n, cus * OCC), 256, 0, s>>>(*g, b->x_s, b->lp_s,          \
                                                                        b->cell_start, n, (const TT*)D, \
                                                                        (TT*)work);                     \
  else                                                                                                   \
    k_jac_bin<TT, TT, PG, OCC, false><<<n, 256, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, n, (const TT*)D, \
                                                      (TT*)work)
  if (fp32) {
    switch (variant) {
      case 0: GM_BIN(float, 8, 2, false); break;
      case 1: GM_BIN(float, 8, 2, true); break;
      case 2: GM_BIN(float, 4, 4, true); break;
      default: GM_BIN(float, 4, 3, true); break;
    }
  } else {
    switch (variant) {
      case 0: GM_BIN(double, 4, 2, false); break;
      case 1: GM_BIN(double, 4, 2, true); break;
      case 2: GM_BIN(double, 8, 1, false); break;
      default: GM_BIN(double, 8, 1, true); break;
    }
  }
#undef GM_BIN
  GM_LAUNCH_CHECK();
  return 0;
}

int launch_gather(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, int fp32, int variant,
                  const void* work, const uint8_t* freem, void* vals, int vals_fp32, hipStream_t s) {
  const int n = g->n_nodes;
#define GM_GATHER(TCC, TVV)                                                                                      \
  if (variant == 0)                                                                                              \
    k_jac_gather<TCC, TVV><<<n, 512, 0, s>>>(*g, b->cell_start, (const TCC*)work, freem, pat->row_ptr,         \
                                             pat->row_mask, (TVV*)vals);                                       \
  else if (variant == 1)                                                                                         \
    k_jac_gather_staged<TCC, TVV, sizeof(TCC) == 4 ? 27 : 9><<<n, 256, 0, s>>>(                                 \
        *g, b->cell_start, (const TCC*)work, freem, pat->row_ptr, pat->row_mask, (TVV*)vals);                  \
  else                                                                                                           \
    k_jac_gather_staged<TCC, TVV, sizeof(TCC) == 4 ? 9 : 3><<<n, 256, 0, s>>>(                                  \
        *g, b->cell_start, (const TCC*)work, freem, pat->row_ptr, pat->row_mask, (TVV*)vals)
  if (fp32) {
    if (vals_fp32) GM_GATHER(float, float);
    else GM_GATHER(float, double);
  } else {
    if (vals_fp32) GM_GATHER(double, float);
    else GM_GATHER(double, double);
  }
#undef GM_GATHER
  GM_LAUNCH_CHECK();
  return 0;
}
}  // namespace

extern "C" int gm_jacobian_bins_variant(const gm_grid* g, const gm_bins* b, const void* D, int tangent_fp32,
                                        int variant, void* work, size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && D && work, "gm_jacobian_bins: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_jacobian_bins: the 3D u-p form only");
  GM_REQUIRE(variant >= 0 && variant < N_BIN_VARIANTS, "gm_jacobian_bins: variant out of range");
  GM_REQUIRE(work_bytes >= gm_node_jacobian_work_bytes(g, tangent_fp32),
             "gm_jacobian_bins: workspace smaller than gm_node_jacobian_work_bytes");
  if (g->n_nodes == 0) return 0;
  return launch_bins(g, b, D, tangent_fp32 ? 1 : 0, variant, work, s);
}

extern "C" int gm_jacobian_bins(const gm_grid* g, const gm_bins* b, const void* D, int tangent_fp32, void* work,
                                size_t work_bytes, hipStream_t s) {
  return gm_jacobian_bins_variant(g, b, D, tangent_fp32, g_bin_variant[tangent_fp32 ? 1 : 0], work, work_bytes, s);
}

extern "C" int gm_jacobian_gather_variant(const gm_grid* g, const gm_bins* b, const gm_pattern* pat,
                                          int tangent_fp32, int variant, const void* work, const uint8_t* freem,
                                          void* vals, int vals_fp32, hipStream_t s) {
  GM_REQUIRE(g && b && pat && work && freem && vals, "gm_jacobian_gather: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_jacobian_gather: the 3D u-p form only");
  GM_REQUIRE(variant >= 0 && variant < N_GATHER_VARIANTS, "gm_jacobian_gather: variant out of range");
  if (g->n_nodes == 0) return 0;
  return launch_gather(g, b, pat, tangent_fp32 ? 1 : 0, variant, work, freem, vals, vals_fp32, s);
}

extern "C" int gm_jacobian_gather(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, int tangent_fp32,
                                  const void* work, const uint8_t* freem, void* vals, int vals_fp32, hipStream_t s) {
  return gm_jacobian_gather_variant(g, b, pat, tangent_fp32, g_gather_variant[tangent_fp32 ? 1 : 0], work, freem,
                                    vals, vals_fp32, s);
}

// Times every per-bin and gather variant on the given tangent (`reps` launches each
// after one warm-up, HIP events on the stream), keeps the fastest of each for later
// assemblies in this precision and returns them with the average times (ms[0..3]
// per-bin variants, ms[4..6] gather variants).  The matrix and the workspace are
// overwritten with the assembly.
extern "C" int gm_jacobian_autotune(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, const void* D,
                                    int tangent_fp32, const uint8_t* freem, void* vals, int vals_fp32, void* work,
                                    size_t work_bytes, int reps, int* best_bins, int* best_gather, float* ms,
                                    hipStream_t s) {
  GM_REQUIRE(g && b && pat && D && freem && vals && work && best_bins && best_gather && ms && reps >= 1,
             "gm_jacobian_autotune: bad argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_jacobian_autotune: the 3D u-p form only");
  GM_REQUIRE(work_bytes >= gm_node_jacobian_work_bytes(g, tangent_fp32),
             "gm_jacobian_autotune: workspace smaller than gm_node_jacobian_work_bytes");
  const int fp = tangent_fp32 ? 1 : 0;
  if (g->n_nodes == 0) {
    *best_bins = g_bin_variant[fp];
    *best_gather = g_gather_variant[fp];
    return 0;
  }
  hipEvent_t e0, e1;
  GM_HIP(hipEventCreate(&e0));
  GM_HIP(hipEventCreate(&e1));
  int rc = 0;
  auto time = [&](auto&& launch, float* out) -> int {
    int r = launch();
    if (r) return r;
    GM_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k)
      if ((r = launch())) return r;
    GM_HIP(hipEventRecord(e1, s));
    GM_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    GM_HIP(hipEventElapsedTime(&t, e0, e1));
    *out = t / reps;
    return 0;
  };
  int bb = 0, bg = 0;
  for (int v = 0; v < N_BIN_VARIANTS && !rc; ++v) {
    rc = time([&]() { return launch_bins(g, b, D, fp, v, work, s); }, &ms[v]);
    if (!rc && ms[v] < ms[bb]) bb = v;
  }
  if (!rc) rc = launch_bins(g, b, D, fp, bb, work, s);
  for (int v = 0; v < N_GATHER_VARIANTS && !rc; ++v) {
    rc = time([&]() { return launch_gather(g, b, pat, fp, v, work, freem, vals, vals_fp32, s); },
              &ms[N_BIN_VARIANTS + v]);
    if (!rc && ms[N_BIN_VARIANTS + v] < ms[N_BIN_VARIANTS + bg]) bg = v;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  g_bin_variant[fp] = bb;
  g_gather_variant[fp] = bg;
  *best_bins = bb;
  *best_gather = bg;
  return 0;
}

extern "C" int gm_node_jacobian_binned(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, const void* D,
                                       int tangent_fp32, const uint8_t* freem, void* vals, int vals_fp32,
                                       void* work, size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && pat && D && freem && vals && work, "gm_node_jacobian_binned: null argument");
  int rc = gm_jacobian_bins(g, b, D, tangent_fp32, work, work_bytes, s);
  if (rc) return rc;
  return gm_jacobian_gather(g, b, pat, tangent_fp32, work, freem, vals, vals_fp32, s);
}

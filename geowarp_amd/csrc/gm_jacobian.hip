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
//   k_jac_bin     one 256-lane workgroup per bin: E_I for a group of particles
//                 in LDS, 14 output tiles of 16x16 over the 4 waves, the
//                 bin's 27 x 27 blocks written contiguously to a workspace
//                 (n_nodes x 11664 values).
//   k_jac_gather  one workgroup per block row I: every block (I, J) of the
//                 pattern sums the <= 27 bins b = I - o holding both nodes in the
//                 fixed order (oz, oy, ox); masks, sign and identity rows as in
//                 the node-row kernels.
// Both steps are deterministic (fixed accumulation order).  The workspace is
// written once and read once per assembly, so the pair is bound by HBM on
// 2 x n_bins x 11664 values plus the matrix itself.
#include <hip/hip_runtime.h>

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

// One workgroup per bin (stencil base node b).  Workspace cb[b] = 27 x 27 blocks
// (I_loc, J_loc), I_loc = o_x + 3 o_y + 9 o_z with node = b + o.  The bin's
// particles are taken PG at a time (PG = 8 in FP32, 4 in FP64: the E table of a
// group is 52 KB of LDS either way).
template <typename TD, typename TC>
__global__ void __launch_bounds__(256, 2) k_jac_bin(gm_grid g, const double* x_s, const double* lp_s,
                                                    const int* cell_start, const TD* D, TC* cb) {
  constexpr int PG = sizeof(TC) == 4 ? 8 : 4;
  constexpr int KST = 5 * PG / 4;  // K steps of 4 per group
  __shared__ TC AXT[PG][3][9];
  // E[p][I][r][z] plus a zero node row (I = 27) and a zero column (z = ZERO): every
  // A operand is then an unconditional LDS read.  A node's 4 x 16 entries are padded
  // to 65 words, so that the 16 (I, r) rows an MFMA reads at once fall in distinct banks
  constexpr int ES = 4 * (NZ + 1) + 1;
  __shared__ TC E_[PG][28][ES];
#define E(p, I, r, z) E_[p][I][(r) * (NZ + 1) + (z)]
  __shared__ TC BT[PG][5][32];  // B operand g_p[J][beta], J padded to 32 with zeros
  const int b = blockIdx.x;
  const int ks = cell_start[b], ke = cell_start[b + 1];
  if (ks == ke) return;  // uniform across the workgroup
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int q = t; q < PG * 4 * (NZ + 1); q += 256) E(q / (4 * (NZ + 1)), 27, (q / (NZ + 1)) % 4, q % (NZ + 1)) = 0;
  for (int q = t; q < PG * 27 * 4; q += 256) E(q / 108, (q / 4) % 27, q % 4, ZERO) = 0;
  typedef typename Mfma<TC>::acc ACC;
  ACC acc[4][4];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[s][c] = ACC{0, 0, 0, 0};
  // output tiles tt = wave + 4 s: row tile mt = tt / 2 (16 rows (I, r)), column tile
  // nt = tt % 2 (16 columns J); tiles 14 and 15 (the fourth of waves 2 and 3) have
  // I >= 28, read the zero row and are not stored
  constexpr int NQ = (PG * 4 * NZ + 255) / 256;  // (p, r, z) items per thread
  for (int p0 = ks; p0 < ke; p0 += PG) {
    const int np = min(PG, ke - p0);
    // every global load of the group is issued before the first barrier: the tangent
    // entries of this thread's (p, r, z) items, then the positions and domain sizes
    TC dv[NQ][5];
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int q = t + 256 * j;
      const int p = q / (4 * NZ), rz = q % (4 * NZ), r = rz / NZ, z = rz % NZ;
#pragma unroll
      for (int i = 0; i < 5; ++i) dv[j][i] = (TC)0;
      if (q < PG * 4 * NZ && p < np) {
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
    double xa = 0.0, la = 0.0;
    if (t < PG * 3 && t / 3 < np) {
      xa = x_s[(long)(p0 + t / 3) * 3 + t % 3];
      la = lp_s[(long)(p0 + t / 3) * 3 + t % 3];
    }
    __syncthreads();  // the previous group's LDS reads are done
    if (t < PG * 3) {
      const int p = t / 3, a = t % 3;
      if (p < np) {
        const Axis ax = gimp_axis(xa, la, g.h[a], g.inv_h[a]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          AXT[p][a][i] = (TC)ax.w[i];
          AXT[p][a][3 + i] = (TC)ax.g[i];
          AXT[p][a][6 + i] = (TC)ax.wa[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) AXT[p][a][i] = (TC)0;
      }
    }
    __syncthreads();
    for (int q = t; q < PG * 5 * 32; q += 256) {
      const int p = q / 160, beta = (q / 32) % 5, J = q % 32;
      BT[p][beta][J] = (J < 27) ? gcomp<TC>(AXT[p], J, beta) : (TC)0;
    }
    // E[p][I][r][z] = (L_I D_p)[r][z]: u rows sum_b dw/dx_b D[(b,r)] + w D[T_r];
    // the pressure row w D[Q] + sum_b dw/dx_b D[M_b] + (w - w_avg) D[S]
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int q = t + 256 * j;
      if (q >= PG * 4 * NZ) break;
      const int p = q / (4 * NZ), rz = q % (4 * NZ), r = rz / NZ, z = rz % NZ;
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
          for (int iz = 0; iz < 3; ++iz) E(p, ix + 3 * iy + 9 * iz, r, z) = wz[iz] * U + gz[iz] * Q + az[iz] * A2;
        }
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int kst = 0; kst < KST; ++kst) {
      // B operand (k = (p, beta), column J) of both column tiles
      const int k = kst * 4 + (lane >> 4), p = k / 5, beta = k % 5;
      // particles past np have all-zero axis tables, hence zero B
      const TC b0 = BT[p][beta][lane & 15];
      const TC b1 = BT[p][beta][16 + (lane & 15)];
      int zc[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) zc[c] = zsel(c, beta);
      TC av[4][4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int tt = wave + 4 * s, mt = tt >> 1;
        const int rt = lane & 15;
        const int I = min(mt * 4 + (Mfma<TC>::r_major ? (rt & 3) : (rt >> 2)), 27);
        const int r = Mfma<TC>::r_major ? (rt >> 2) : (rt & 3);
#pragma unroll
        for (int c = 0; c < 4; ++c) av[s][c] = E(p, I, r, zc[c]);
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const TC bk = ((wave + 4 * s) & 1) ? b1 : b0;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[s][c] = Mfma<TC>::step(av[s][c], bk, acc[s][c]);
      }
    }
  }
  // lane holds rows r = 0..3 of node I_loc = 4 mt + lane / 16 in column J_loc =
  // 16 nt + lane % 16: one whole 4 x 4 block
  TC* out = cb + (long)b * CB_BIN;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int tt = wave + 4 * s, mt = tt >> 1, nt = tt & 1;
    const int I = mt * 4 + (lane >> 4), J = nt * 16 + (lane & 15);
    if (I < 27 && J < 27) {
      TC* o = out + (long)(I * 27 + J) * 16;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) o[r * 4 + c] = acc[s][c][r];
    }
  }
}

#undef E

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
      const TC* src = cb + (long)bn * CB_BIN + (long)(o * 27 + (kx + 3 * ky + 9 * kz)) * 16 + r * 4;
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

extern "C" size_t gm_node_jacobian_work_bytes(const gm_grid* g, int tangent_fp32) {
  if (!g) return 0;
  return (size_t)g->n_nodes * (size_t)CB_BIN * (tangent_fp32 ? sizeof(float) : sizeof(double));
}

extern "C" int gm_jacobian_bins(const gm_grid* g, const gm_bins* b, const void* D, int tangent_fp32, void* work,
                                size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && D && work, "gm_jacobian_bins: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_jacobian_bins: the 3D u-p form only");
  GM_REQUIRE(work_bytes >= gm_node_jacobian_work_bytes(g, tangent_fp32),
             "gm_jacobian_bins: workspace smaller than gm_node_jacobian_work_bytes");
  const int n = g->n_nodes;
  if (n == 0) return 0;
  if (tangent_fp32)
    k_jac_bin<float, float><<<n, 256, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, (const float*)D, (float*)work);
  else
    k_jac_bin<double, double><<<n, 256, 0, s>>>(*g, b->x_s, b->lp_s, b->cell_start, (const double*)D,
                                                (double*)work);
  GM_LAUNCH_CHECK();
  return 0;
}

extern "C" int gm_jacobian_gather(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, int tangent_fp32,
                                  const void* work, const uint8_t* freem, void* vals, int vals_fp32, hipStream_t s) {
  GM_REQUIRE(g && b && pat && work && freem && vals, "gm_jacobian_gather: null argument");
  GM_REQUIRE(g->dim == 3 && g->nb == 4, "gm_jacobian_gather: the 3D u-p form only");
  const int n = g->n_nodes;
  if (n == 0) return 0;
#define GM_GATHER(TCC, TVV)                                                                              \
  k_jac_gather<TCC, TVV><<<n, 512, 0, s>>>(*g, b->cell_start, (const TCC*)work, freem, pat->row_ptr, \
                                           pat->row_mask, (TVV*)vals)
  if (tangent_fp32) {
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

extern "C" int gm_node_jacobian_binned(const gm_grid* g, const gm_bins* b, const gm_pattern* pat, const void* D,
                                       int tangent_fp32, const uint8_t* freem, void* vals, int vals_fp32,
                                       void* work, size_t work_bytes, hipStream_t s) {
  GM_REQUIRE(g && b && pat && D && freem && vals && work, "gm_node_jacobian_binned: null argument");
  int rc = gm_jacobian_bins(g, b, D, tangent_fp32, work, work_bytes, s);
  if (rc) return rc;
  return gm_jacobian_gather(g, b, pat, tangent_fp32, work, freem, vals, vals_fp32, s);
}

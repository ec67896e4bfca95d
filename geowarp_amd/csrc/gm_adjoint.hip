// Adjoint building blocks for the 2D plane-strain load step:
//   gm_local_vjp_2d    per-particle vector-Jacobian products of the local map
//                      (residual contributions, next state) by forward-mode dual
//                      numbers, one direction per input component
//   gm_local_eval_2d   the same local map in double (forward values, reaction sums)
//   gm_bsr_transpose   K^T in the same (structurally symmetric) block pattern
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gm_common.h"
#include "gm_types.h"
#include "gm_local2d.h"

namespace {

// Direction layout of gm_local_vjp_2d (GM_ADJ_ND = 25 per particle):
//   0-1 x, 2-3 lp, 4-8 F_old (xx, xy, yx, yy, zz), 9-12 b_old (xx, xy=yx, yy, zz),
//   13 alpha_old, 14-17 H seed (xx, xy, yx, yy), 18-24 E, nu, kappa, H, phi, psi, cohesion
constexpr int ND = 25;

__device__ inline void load_in(const gm_local_args& a, int p, LocalIn<double>& in) {
  for (int c = 0; c < 2; ++c) {
    in.x[c] = a.x[2 * p + c];
    in.lp[c] = a.lp[2 * p + c];
  }
  const double* F = a.F_old + 9 * p;
  in.Fo[0] = F[0];
  in.Fo[1] = F[1];
  in.Fo[2] = F[3];
  in.Fo[3] = F[4];
  in.Fo[4] = F[8];
  const double* b = a.b_old + 9 * p;
  in.bo[0] = b[0];
  in.bo[1] = b[1];
  in.bo[2] = b[4];
  in.bo[3] = b[8];
  in.alpha = a.alpha_old ? a.alpha_old[p] : 0.0;
  for (int i = 0; i < 4; ++i) in.dH[i] = 0.0;
  in.E = a.E;
  in.nu = a.nu;
  in.kappa = a.kappa;
  in.Hh = a.H;
  in.phi = a.p_phi ? a.p_phi[p] : a.phi;
  in.psi = a.psi;
  in.coh = a.p_coh ? a.p_coh[p] : a.coh;
  in.ext = a.ext_flag && a.ext_flag[p];
}

__device__ inline LocalConst make_const(const gm_local_args& a) {
  LocalConst k;
  k.h = a.h;
  k.inv_h = 1.0 / a.h;
  k.nx1 = a.nx1;
  k.mat_kind = a.mat_kind;
  k.dp_apex = a.dp_apex;
  k.shape = a.shape;
  k.p_vol = a.p_vol;
  k.p_rho = a.p_rho;
  for (int c = 0; c < 2; ++c) {
    k.gravity[c] = a.gravity[c];
    k.point_load[c] = a.point_load[c];
    k.dmin[c] = a.dmin[c];
    k.dmax[c] = a.dmax[c];
  }
  k.lp0 = a.lp0;
  return k;
}

__device__ inline Dual seed(double v, int d, int want) { return dmk(v, d == want ? 1.0 : 0.0); }

__global__ void __launch_bounds__(128) k_local_vjp_2d(gm_local_args a, double* g) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_particles) return;
  const LocalConst k = make_const(a);
  LocalIn<double> v;
  load_in(a, p, v);
  // state cotangents of the outputs
  double xb[2] = {0, 0}, lb[2] = {0, 0}, Fb[5] = {0, 0, 0, 0, 0}, bb[4] = {0, 0, 0, 0}, ab = 0.0;
  if (a.xbar)
    for (int c = 0; c < 2; ++c) xb[c] = a.xbar[2 * p + c];
  if (a.lpbar)
    for (int c = 0; c < 2; ++c) lb[c] = a.lpbar[2 * p + c];
  if (a.Fbar)
    for (int c = 0; c < 5; ++c) Fb[c] = a.Fbar[5 * p + c];
  if (a.bbar)
    for (int c = 0; c < 4; ++c) bb[c] = a.bbar[4 * p + c];
  if (a.abar) ab = a.abar[p];
  const double cs = a.cot_scale ? a.cot_scale[p] : 1.0;
  for (int d = 0; d < ND; ++d) {
    double outv = 0.0;
    bool relevant = true;
    if (d == 13 && a.mat_kind != LMAT_J2) relevant = false;                  // alpha
    if ((d == 20 || d == 21) && a.mat_kind != LMAT_J2) relevant = false;     // kappa, H
    if (d >= 22 && a.mat_kind != LMAT_DP) relevant = false;                  // phi, psi, c
    if (relevant) {
      LocalIn<Dual> in;
      for (int c = 0; c < 2; ++c) {
        in.x[c] = seed(v.x[c], d, 0 + c);
        in.lp[c] = seed(v.lp[c], d, 2 + c);
      }
      for (int c = 0; c < 5; ++c) in.Fo[c] = seed(v.Fo[c], d, 4 + c);
      for (int c = 0; c < 4; ++c) in.bo[c] = seed(v.bo[c], d, 9 + c);
      in.alpha = seed(v.alpha, d, 13);
      for (int c = 0; c < 4; ++c) in.dH[c] = seed(0.0, d, 14 + c);
      in.E = seed(v.E, d, 18);
      in.nu = seed(v.nu, d, 19);
      in.kappa = seed(v.kappa, d, 20);
      in.Hh = seed(v.Hh, d, 21);
      in.phi = seed(v.phi, d, 22);
      in.psi = seed(v.psi, d, 23);
      in.coh = seed(v.coh, d, 24);
      in.ext = v.ext;
      LocalOut<Dual> o;
      local_map_2d<Dual>(k, in, a.u, a.cot, o);
      outv = cs * o.acc.d + ab * o.alpha.d;
      for (int c = 0; c < 2; ++c) outv += xb[c] * o.xn[c].d + lb[c] * o.lpn[c].d;
      for (int c = 0; c < 5; ++c) outv += Fb[c] * o.Fn[c].d;
      for (int c = 0; c < 4; ++c) outv += bb[c] * o.bn[c].d;
    }
    g[(long)p * ND + d] = outv;
  }
}

__global__ void __launch_bounds__(128) k_local_eval_2d(gm_local_args a, double* acc, double* state) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_particles) return;
  const LocalConst k = make_const(a);
  LocalIn<double> in;
  load_in(a, p, in);
  LocalOut<double> o;
  local_map_2d<double>(k, in, a.u, a.cot, o);
  if (acc) acc[p] = (a.cot_scale ? a.cot_scale[p] : 1.0) * o.acc;
  if (state) {
    // x_new (2), lp_new (2), F_new (5), b_new (4), alpha_new, sigma (xx, xy, yy, zz)
    double* s = state + (long)p * 18;
    s[0] = o.xn[0];
    s[1] = o.xn[1];
    s[2] = o.lpn[0];
    s[3] = o.lpn[1];
    for (int c = 0; c < 5; ++c) s[4 + c] = o.Fn[c];
    for (int c = 0; c < 4; ++c) s[9 + c] = o.bn[c];
    s[13] = o.alpha;
    for (int c = 0; c < 4; ++c) s[14 + c] = o.sig[c];
  }
}

// Block transpose in a structurally symmetric pattern with ascending column
// indices per row: vt(I,J) = v(J,I)^T.
template <typename TV>
__global__ void k_bsr_transpose(int nb, int n_rows, const int* row_ptr, const int* col_idx, const TV* v, TV* vt) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n_rows) return;
  const int bs = nb * nb;
  for (int q = row_ptr[I]; q < row_ptr[I + 1]; ++q) {
    int J = col_idx[q];
    int lo = row_ptr[J], hi = row_ptr[J + 1] - 1, qt = -1;
    while (lo <= hi) {
      int mid = (lo + hi) >> 1;
      int c = col_idx[mid];
      if (c == I) {
        qt = mid;
        break;
      }
      if (c < I) lo = mid + 1;
      else hi = mid - 1;
    }
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) vt[(long)q * bs + i * nb + j] = (qt >= 0) ? v[(long)qt * bs + j * nb + i] : (TV)0;
  }
}

}  // namespace

extern "C" int gm_local_vjp_2d(const gm_local_args* a, double* g, hipStream_t s) {
  GM_REQUIRE(a && g, "gm_local_vjp_2d: null argument");
  GM_REQUIRE(a->x && a->lp && a->F_old && a->b_old && a->u, "gm_local_vjp_2d: missing particle or nodal array");
  if (a->n_particles == 0) return 0;
  k_local_vjp_2d<<<gm_blocks(a->n_particles, 128), 128, 0, s>>>(*a, g);
  GM_LAUNCH_CHECK();
  return 0;
}

extern "C" int gm_local_eval_2d(const gm_local_args* a, double* acc, double* state, hipStream_t s) {
  GM_REQUIRE(a, "gm_local_eval_2d: null argument");
  GM_REQUIRE(a->x && a->lp && a->F_old && a->b_old && a->u, "gm_local_eval_2d: missing particle or nodal array");
  if (a->n_particles == 0) return 0;
  k_local_eval_2d<<<gm_blocks(a->n_particles, 128), 128, 0, s>>>(*a, acc, state);
  GM_LAUNCH_CHECK();
  return 0;
}

extern "C" int gm_bsr_transpose(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx,
                                const void* vals, void* vals_t, hipStream_t s) {
  GM_REQUIRE(row_ptr && col_idx && vals && vals_t, "gm_bsr_transpose: null argument");
  GM_REQUIRE(vals != vals_t, "gm_bsr_transpose: in-place transpose is not supported");
  if (n_rows == 0) return 0;
  if (fp32)
    k_bsr_transpose<float><<<gm_blocks(n_rows, 128), 128, 0, s>>>(nb, n_rows, row_ptr, col_idx, (const float*)vals,
                                                                   (float*)vals_t);
  else
    k_bsr_transpose<double><<<gm_blocks(n_rows, 128), 128, 0, s>>>(nb, n_rows, row_ptr, col_idx,
                                                                    (const double*)vals, (double*)vals_t);
  GM_LAUNCH_CHECK();
  return 0;
}

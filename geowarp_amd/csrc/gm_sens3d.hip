// Derivative of the 3D u-p Neo-Hookean particle stress vector with respect to
// Young's modulus at fixed nodal unknowns: the load-step sensitivity of the
// indentation problem (dL/dE = dL/dE|_u + lambda^T dR/dE with K^T lambda = dL/du).
//
// At fixed deformation the Neo-Hookean Cauchy stress mu/J (F F^T - I) + lambda/J ln J I
// and the indenter penalty coefficient are proportional to E (mu, lambda and
// pen_par are, at fixed Poisson ratio and penalty factor), the PPP coefficient
// 1/(2 mu) is proportional to 1/E, and the pore-pressure, mass-balance and Darcy
// terms do not depend on E.  So, per particle (bin-sorted row kq, particle p):
//   dP/dE   = -V A^T sigma / E        (A = f^-1 = F_old F_new^-1, V = V0 det F_new)
//   dT_z/dE = pen_z / E = -pen_force / E
//   dS/dE   = -S / E
// and zero elsewhere; the residual's dR/dE is the node gather of these rows.
#include <hip/hip_runtime.h>

#include "gm_common.h"
#include "gm_device.h"
#include "gm_types.h"

using namespace gm;

__global__ void k_dY_dE_3d(int n, const int* sorted, const double* F_old, const double* F_new, const double* sigma,
                           const double* pen_force, const double* Y, double p_vol, double E, double* dY) {
  constexpr int NY = 17, Y_T = 9, Y_S = 16;
  const int kq = blockIdx.x * blockDim.x + threadIdx.x;
  if (kq >= n) return;
  const int p = sorted[kq];
  const M3 Fo = m3_load(F_old + 9 * (long)p);
  const M3 Fn = m3_load(F_new + 9 * (long)p);
  const M3 sg = m3_load(sigma + 9 * (long)p);
  const M3 A = m3_mul(Fo, m3_inv(Fn));
  const double V = p_vol * m3_det(Fn);
  double* out = dY + (long)kq * NY;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) s += A.a[a][b] * sg.a[a][c];
      out[b * 3 + c] = -V * s / E;
    }
  for (int y = Y_T; y < NY; ++y) out[y] = 0.0;
  out[Y_T + 2] = -pen_force[p] / E;
  out[Y_S] = -Y[(long)kq * NY + Y_S] / E;
}

extern "C" int gm_stage_de_3d(int n, const int* sorted, const double* F_old, const double* F_new, const double* sigma,
                           const double* pen_force, const double* Y, double p_vol, double E, double* dY,
                           hipStream_t s) {
  GM_REQUIRE(sorted && F_old && F_new && sigma && pen_force && Y && dY, "gm_stage_de_3d: null argument");
  GM_REQUIRE(E > 0.0, "gm_stage_de_3d: Young's modulus must be positive");
  if (n == 0) return 0;
  k_dY_dE_3d<<<gm_blocks(n, 256), 256, 0, s>>>(n, sorted, F_old, F_new, sigma, pen_force, Y, p_vol, E, dY);
  GM_LAUNCH_CHECK();
  return 0;
}

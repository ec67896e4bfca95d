// Test infrastructure (the oracle): a C++17/OpenMP FP64 restatement of one Newton
// iteration's assembly of the 3D u-p Neo-Hookean indentation residual
// (assemble_residual_coupled_3d_neohookean, mpm_utils.py:1020-1294, as restated in
// oracle/mpm.py residual_coupled_3d) with its exact Jacobian.  Used as the CPU
// baseline of the benchmark and checked against the NumPy oracle by the tests;
// nothing in the product links it.
//
// Per particle the residual contribution to node I is r_I = L(w_I, dw_I, wa_I) Y(z)
// with z = (H (9, H_ab = sum_J dw_J[a] du_J[b], the reference's transposed
// increment gradient), du_z, p, p_avg, grad p (3)) linear in the nodal unknowns.
// The particle map Y(z) (17 values: P (9), T (3), Q, m (3), S) is evaluated once
// with forward-mode dual numbers over the 15 components of z, which gives
// D = dY/dz exactly; the Jacobian is then K_IJ = sum_p L_I D_p M_J gathered per
// node row over the particles whose support holds the row's node.
//
// Weights come from the NumPy oracle's per-step support (idx, W, G, Wa of
// oracle/mpm.py): the port covers what runs per Newton iteration.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

constexpr int NZ = 15, NY = 17;

struct Dual {
  double v;
  double d[NZ];
};

inline Dual cst(double v) {
  Dual r;
  r.v = v;
  for (int i = 0; i < NZ; ++i) r.d[i] = 0.0;
  return r;
}
inline Dual var(double v, int k) {
  Dual r = cst(v);
  r.d[k] = 1.0;
  return r;
}
inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
  for (int i = 0; i < NZ; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
  for (int i = 0; i < NZ; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
inline Dual operator-(const Dual& a) {
  Dual r;
  r.v = -a.v;
  for (int i = 0; i < NZ; ++i) r.d[i] = -a.d[i];
  return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
  for (int i = 0; i < NZ; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
inline Dual operator*(double s, const Dual& a) {
  Dual r;
  r.v = s * a.v;
  for (int i = 0; i < NZ; ++i) r.d[i] = s * a.d[i];
  return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
  for (int i = 0; i < NZ; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
inline Dual dlog(const Dual& a) {
  Dual r;
  r.v = log(a.v);
  for (int i = 0; i < NZ; ++i) r.d[i] = a.d[i] / a.v;
  return r;
}

struct Params {
  double mu, lam;           // Lame constants of the skeleton
  double p_vol, p_rho;      // initial particle volume, solid density
  double phi0, mobility, dt;
  double gravity_z;         // (step+1)/total ramped gravity (z component)
  double pen_par, h_cur;    // indenter penalty coefficient and current footing height
  double fminx, fmaxx, fminy, fmaxy;
};

// Particle map Y(z) with its derivative (oracle/mpm.py residual_coupled_3d terms).
void particle_map(const Params& q, const double* zv, const double* Fo, double ppo_term, double lnJo,
                  const double* x, Dual* Y) {
  Dual z[NZ];
  for (int k = 0; k < NZ; ++k) z[k] = var(zv[k], k);
  // f = I + H, F = f F_old
  Dual f[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) f[a][b] = z[3 * a + b] + cst(a == b ? 1.0 : 0.0);
  Dual F[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      Dual s = cst(0.0);
      for (int k = 0; k < 3; ++k) s = s + Fo[3 * k + b] * f[a][k];
      F[a][b] = s;
    }
  // det and inverse of f (A = f^-1)
  const Dual c00 = f[1][1] * f[2][2] - f[1][2] * f[2][1];
  const Dual c01 = f[1][2] * f[2][0] - f[1][0] * f[2][2];
  const Dual c02 = f[1][0] * f[2][1] - f[1][1] * f[2][0];
  const Dual detf = f[0][0] * c00 + f[0][1] * c01 + f[0][2] * c02;
  Dual A[3][3];
  A[0][0] = c00 / detf;
  A[1][0] = c01 / detf;
  A[2][0] = c02 / detf;
  A[0][1] = (f[0][2] * f[2][1] - f[0][1] * f[2][2]) / detf;
  A[1][1] = (f[0][0] * f[2][2] - f[0][2] * f[2][0]) / detf;
  A[2][1] = (f[0][1] * f[2][0] - f[0][0] * f[2][1]) / detf;
  A[0][2] = (f[0][1] * f[1][2] - f[0][2] * f[1][1]) / detf;
  A[1][2] = (f[0][2] * f[1][0] - f[0][0] * f[1][2]) / detf;
  A[2][2] = (f[0][0] * f[1][1] - f[0][1] * f[1][0]) / detf;
  const Dual J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                 F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                 F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
  const Dual lnJ = dlog(J);
  // Neo-Hookean Cauchy stress sigma = mu/J (F F^T - I) + lambda/J ln J I, minus p I
  const Dual pp = z[10], ppa = z[11];
  Dual s3[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      Dual ff = cst(0.0);
      for (int k = 0; k < 3; ++k) ff = ff + F[a][k] * F[b][k];
      Dual s = (q.mu * (ff - cst(a == b ? 1.0 : 0.0))) / J;
      if (a == b) s = s + (q.lam * lnJ) / J - pp;
      s3[a][b] = s;
    }
  const Dual V = q.p_vol * J;
  // P[b][c] = -V sum_a A[a][b] s3[a][c]
  for (int b = 0; b < 3; ++b)
    for (int c = 0; c < 3; ++c) {
      Dual s = cst(0.0);
      for (int a = 0; a < 3; ++a) s = s + A[a][b] * s3[a][c];
      Y[3 * b + c] = -(s * V);
    }
  // T = rho_mix g V (+ penalty on z)
  const Dual npor = cst(1.0) - cst(1.0 - q.phi0) / J;
  const Dual rho = npor + (cst(1.0) - npor) * cst(q.p_rho);
  Y[9] = cst(0.0);
  Y[10] = cst(0.0);
  Y[11] = (q.gravity_z * rho) * V;
  if (x[2] >= q.h_cur && x[0] > q.fminx && x[0] < q.fmaxx && x[1] > q.fminy && x[1] < q.fmaxy)
    Y[11] = Y[11] + (-q.pen_par) * (cst(q.h_cur - x[2]) - z[9]);
  // Q = (ln J - ln J_old) V ;  m[b] = sum_a A[a][b] mobility gp[a] dt V
  Y[12] = (lnJ - cst(lnJo)) * V;
  for (int b = 0; b < 3; ++b) {
    Dual s = cst(0.0);
    for (int a = 0; a < 3; ++a) s = s + A[a][b] * z[12 + a];
    Y[13 + b] = (q.mobility * q.dt) * (s * V);
  }
  // S = tau ((p - p_avg) - (p_old - p_old_avg)) V, tau = 1/(2 mu)
  Y[16] = (0.5 / q.mu) * ((pp - ppa - cst(ppo_term)) * V);
}

}  // namespace

extern "C" {

// One assembly: rhs (4 n_nodes, dof-major, masked by free & support like the
// oracle) and, when Kwin is given, the Jacobian -dR/du as 125 window blocks per
// node row (Kwin[(I*125 + slot)*16 + r*4 + c], slot = (dx+2) + 5 (dy+2) + 25 (dz+2)
// of column node J - I) with `used` set for every block some particle reaches.
// Masks and identity rows are applied by the caller.  Returns 0, or -1 on a
// particle reaching a node outside the window grid.
int op_assemble_nh3d(int n, const int64_t* idx, const double* W, const double* G, const double* Wa,
                     const uint8_t* support, const double* x, const double* F_old, const double* nodal_old,
                     const double* nodal_new, int nx1, int ny1, int nz1, const Params* q, const uint8_t* freem,
                     double* rhs, double* Kwin, uint8_t* used, int n_threads) {
  const long nn = (long)nx1 * ny1 * nz1;
  std::vector<double> Yall((size_t)n * NY), Dall(Kwin ? (size_t)n * NY * NZ : 0);
#pragma omp parallel for num_threads(n_threads) schedule(static)
  for (int p = 0; p < n; ++p) {
    double zv[NZ] = {0};
    double ppo = 0.0, ppoa = 0.0;
    for (int k = 0; k < 27; ++k) {
      if (!support[p * 27 + k]) continue;
      const long J = idx[p * 27 + k];
      const double w = W[p * 27 + k], wa = Wa[p * 27 + k];
      const double* g = G + (p * 27 + k) * 3;
      double du[3];
      for (int c = 0; c < 3; ++c) du[c] = nodal_new[J + c * nn] - nodal_old[J + c * nn];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) zv[3 * a + b] += g[a] * du[b];
      zv[9] += w * du[2];
      const double pn = nodal_new[J + 3 * nn], po = nodal_old[J + 3 * nn];
      zv[10] += w * pn;
      zv[11] += wa * pn;
      for (int a = 0; a < 3; ++a) zv[12 + a] += g[a] * pn;
      ppo += w * po;
      ppoa += wa * po;
    }
    const double* Fo = F_old + 9 * (long)p;
    const double Jo = Fo[0] * (Fo[4] * Fo[8] - Fo[5] * Fo[7]) - Fo[1] * (Fo[3] * Fo[8] - Fo[5] * Fo[6]) +
                      Fo[2] * (Fo[3] * Fo[7] - Fo[4] * Fo[6]);
    Dual Y[NY];
    particle_map(*q, zv, Fo, ppo - ppoa, log(Jo), x + 3 * (long)p, Y);
    for (int y = 0; y < NY; ++y) {
      Yall[(size_t)p * NY + y] = Y[y].v;
      if (Kwin)
        for (int k = 0; k < NZ; ++k) Dall[((size_t)p * NY + y) * NZ + k] = Y[y].d[k];
    }
  }
  // node -> (particle, stencil slot) lists
  std::vector<int> cnt(nn + 1, 0);
  for (long t = 0; t < 27L * n; ++t)
    if (support[t]) {
      if (idx[t] < 0 || idx[t] >= nn) return -1;
      cnt[idx[t] + 1]++;
    }
  for (long i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> lst(cnt[nn]);
  {
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (long t = 0; t < 27L * n; ++t)
      if (support[t]) lst[pos[idx[t]]++] = (int)t;
  }
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
  for (long I = 0; I < nn; ++I) {
    double r[4] = {0, 0, 0, 0};
    const int ix = (int)(I % nx1), iy = (int)((I / nx1) % ny1), iz = (int)(I / ((long)nx1 * ny1));
    for (int e = cnt[I]; e < cnt[I + 1]; ++e) {
      const int t = lst[e], p = t / 27;
      const double wI = W[t], waI = Wa[t];
      const double* gI = G + (long)t * 3;
      const double* Yp = Yall.data() + (size_t)p * NY;
      // r_I = L_I Y
      for (int c = 0; c < 3; ++c) {
        double s = wI * Yp[9 + c];
        for (int b = 0; b < 3; ++b) s += gI[b] * Yp[3 * b + c];
        r[c] += s;
      }
      {
        double s = wI * Yp[12] + (wI - waI) * Yp[16];
        for (int b = 0; b < 3; ++b) s += gI[b] * Yp[13 + b];
        r[3] += s;
      }
      if (!Kwin) continue;
      // E = L_I D (4 x 15)
      const double* Dp = Dall.data() + (size_t)p * NY * NZ;
      double E[4][NZ];
      for (int k = 0; k < NZ; ++k) {
        for (int c = 0; c < 3; ++c) {
          double s = wI * Dp[(9 + c) * NZ + k];
          for (int b = 0; b < 3; ++b) s += gI[b] * Dp[(3 * b + c) * NZ + k];
          E[c][k] = s;
        }
        double s = wI * Dp[12 * NZ + k] + (wI - waI) * Dp[16 * NZ + k];
        for (int b = 0; b < 3; ++b) s += gI[b] * Dp[(13 + b) * NZ + k];
        E[3][k] = s;
      }
      // K_IJ -= E M_J for every node J of the particle's support
      for (int kk = 0; kk < 27; ++kk) {
        const int tJ = p * 27 + kk;
        if (!support[tJ]) continue;
        const long J = idx[tJ];
        const int jx = (int)(J % nx1), jy = (int)((J / nx1) % ny1), jz = (int)(J / ((long)nx1 * ny1));
        const int dx = jx - ix, dy = jy - iy, dz = jz - iz;
        if (dx < -2 || dx > 2 || dy < -2 || dy > 2 || dz < -2 || dz > 2) continue;
        const int slot = (dx + 2) + 5 * (dy + 2) + 25 * (dz + 2);
        const double wJ = W[tJ], waJ = Wa[tJ];
        const double* gJ = G + (long)tJ * 3;
        double* Kb = Kwin + ((size_t)I * 125 + slot) * 16;
        used[(size_t)I * 125 + slot] = 1;
        for (int rr = 0; rr < 4; ++rr) {
          for (int c = 0; c < 3; ++c) {
            // dz/du_J[c]: H[a][c] += gJ[a], du_z (c == 2) += wJ
            double s = 0.0;
            for (int a = 0; a < 3; ++a) s += E[rr][3 * a + c] * gJ[a];
            if (c == 2) s += E[rr][9] * wJ;
            Kb[rr * 4 + c] -= s;
          }
          // dz/dp_J: p += wJ, p_avg += waJ, grad p += gJ
          double s = E[rr][10] * wJ + E[rr][11] * waJ;
          for (int a = 0; a < 3; ++a) s += E[rr][12 + a] * gJ[a];
          Kb[rr * 4 + 3] -= s;
        }
      }
    }
    for (int c = 0; c < 4; ++c) rhs[I + c * nn] = freem[I + c * nn] ? r[c] : 0.0;
  }
  return 0;
}

int op_params_size(void) { return (int)sizeof(Params); }

}  // extern "C"

// Device-side math for the cpGIMP implicit MPM hot path (gfx950, FP64).
//
// * cpGIMP 1-D weights / gradients / cell averages   (GeoWarp mpm/gimp.py:25-193, 197-425)
// * small dense 3x3 algebra, cyclic-Jacobi symmetric eigen-solver
// * constitutive updates with consistent tangents:
//     Hencky elasticity            (mpm/mpm_utils.py:521-560)
//     J2 radial return             (materials/plasticity.py:8-37) + linear hardening
//     Drucker-Prager local Newton  (problems/material_tests/triaxial_DruckerPrager.py:308-412)
//     Neo-Hookean                  (mpm/mpm_utils.py:915, 1205-1208)
//   Hencky-family tangents use the spectral (Daleckii-Krein) form of d tau / d b.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gm_types.h"

#define GM_DEV __device__ __forceinline__

namespace gm {

// ------------------------------------------------------------------ 3x3 algebra
struct M3 {
  double a[3][3];
};

GM_DEV M3 m3_zero() {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = 0.0;
  return r;
}
GM_DEV M3 m3_eye() {
  M3 r = m3_zero();
  r.a[0][0] = r.a[1][1] = r.a[2][2] = 1.0;
  return r;
}
GM_DEV M3 m3_mul(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j];
  return r;
}
GM_DEV M3 m3_T(const M3& x) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[j][i];
  return r;
}
GM_DEV double m3_det(const M3& m) {
  return m.a[0][0] * (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1]) -
         m.a[0][1] * (m.a[1][0] * m.a[2][2] - m.a[1][2] * m.a[2][0]) +
         m.a[0][2] * (m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0]);
}
GM_DEV M3 m3_inv(const M3& m) {
  M3 r;
  double d = m3_det(m);
  double id = 1.0 / d;
  r.a[0][0] = (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1]) * id;
  r.a[0][1] = (m.a[0][2] * m.a[2][1] - m.a[0][1] * m.a[2][2]) * id;
  r.a[0][2] = (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]) * id;
  r.a[1][0] = (m.a[1][2] * m.a[2][0] - m.a[1][0] * m.a[2][2]) * id;
  r.a[1][1] = (m.a[0][0] * m.a[2][2] - m.a[0][2] * m.a[2][0]) * id;
  r.a[1][2] = (m.a[0][2] * m.a[1][0] - m.a[0][0] * m.a[1][2]) * id;
  r.a[2][0] = (m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0]) * id;
  r.a[2][1] = (m.a[0][1] * m.a[2][0] - m.a[0][0] * m.a[2][1]) * id;
  r.a[2][2] = (m.a[0][0] * m.a[1][1] - m.a[0][1] * m.a[1][0]) * id;
  return r;
}
GM_DEV M3 m3_load(const double* p) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = p[3 * i + j];
  return r;
}
GM_DEV void m3_store(double* p, const M3& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) p[3 * i + j] = m.a[i][j];
}

// ------------------------------------------------------------------ cpGIMP
// Per-axis stencil data of one particle.  gimp.py:33-49 (+1e-8 floor bias,
// C truncation), :67-86 linear/cpGIMP branch, :156-171 cell averages.
struct Axis {
  int lo;
  int ext;  // 2 (domain inside one cell) or 3
  double w[3], g[3], wa[3];
};

GM_DEV Axis gimp_axis(double x, double lp, double h, double inv_h) {
  Axis a;
  int lo = (int)((x - lp) * inv_h + 1e-8);
  int hi = (int)((x + lp) * inv_h + 1e-8);
  int base = (int)(x * inv_h + 1e-8);
  double f = x * inv_h - (double)base;
  a.lo = lo;
  if (lo == hi) {
    a.ext = 2;
    a.w[0] = 1.0 - f;
    a.w[1] = f;
    a.w[2] = 0.0;
    a.g[0] = -1.0 * inv_h;
    a.g[1] = inv_h;
    a.g[2] = 0.0;
    a.wa[0] = 0.5;
    a.wa[1] = 0.5;
    a.wa[2] = 0.0;
  } else {
    a.ext = 3;
    double lof = (double)lo;
    double x0 = lof * h, x1 = (lof + 1.0) * h, x2 = (lof + 2.0) * h;
    double d0 = x - x0, d2 = x - x2;
    double a0 = h + lp - fabs(d0);
    double a2 = h + lp - fabs(d2);
    double hl4 = 4.0 * h * lp, hl2 = 2.0 * h * lp;
    a.w[0] = a0 * a0 / hl4;
    a.g[0] = -a0 / hl2 * d0 / fabs(d0);
    a.w[1] = 1.0 - ((x - x1) * (x - x1) + lp * lp) / hl2;
    a.g[1] = -(x - x1) / (h * lp);
    a.w[2] = a2 * a2 / hl4;
    a.g[2] = -a2 / hl2 * d2 / fabs(d2);
    a.wa[0] = a0 / (4.0 * lp);
    a.wa[1] = 0.5;
    a.wa[2] = a2 / (4.0 * lp);
    // a domain ending exactly on a cell face takes this branch with an exactly
    // zero third node: it contributes nothing and is left out of the support
    if (a.w[2] == 0.0 && a.g[2] == 0.0 && a.wa[2] == 0.0) a.ext = 2;
  }
  return a;
}

// Node weight, material gradient and cell-average weight for local offsets.
// Selects instead of runtime array indexing keep the stencil in registers.
GM_DEV double sel3(const double* a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

template <int DIM>
GM_DEV void node_geom(const Axis* ax, const int* o, double& w, double* G, double& wa) {
  double wx = sel3(ax[0].w, o[0]), wy = sel3(ax[1].w, o[1]);
  double gx = sel3(ax[0].g, o[0]), gy = sel3(ax[1].g, o[1]);
  double ax_ = sel3(ax[0].wa, o[0]), ay_ = sel3(ax[1].wa, o[1]);
  if (DIM == 2) {
    w = wx * wy;
    G[0] = gx * wy;
    G[1] = wx * gy;
    wa = ax_ * ay_;
  } else {
    double wz = sel3(ax[2].w, o[2]), gz = sel3(ax[2].g, o[2]), az_ = sel3(ax[2].wa, o[2]);
    w = wx * wy * wz;
    G[0] = gx * wy * wz;
    G[1] = wx * gy * wz;
    G[2] = wx * wy * gz;
    wa = ax_ * ay_ * az_;
  }
}

// ------------------------------------------------------------------ eigen
// Cyclic Jacobi for a symmetric 3x3 matrix: A = V diag(lam) V^T.
GM_DEV void sym_eig3(const M3& Ain, double* lam, M3& V) {
  double a00 = Ain.a[0][0], a11 = Ain.a[1][1], a22 = Ain.a[2][2];
  double a01 = 0.5 * (Ain.a[0][1] + Ain.a[1][0]);
  double a02 = 0.5 * (Ain.a[0][2] + Ain.a[2][0]);
  double a12 = 0.5 * (Ain.a[1][2] + Ain.a[2][1]);
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = a01 * a01 + a02 * a02 + a12 * a12;
    double dn = a00 * a00 + a11 * a11 + a22 * a22;
    if (off <= 1e-36 * dn || off == 0.0) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      int p = (pq == 2) ? 1 : 0;
      int q = (pq == 0) ? 1 : 2;
      double apq = (pq == 0) ? a01 : (pq == 1 ? a02 : a12);
      if (apq == 0.0) continue;
      double app = (p == 0) ? a00 : a11;
      double aqq = (q == 1) ? a11 : a22;
      double theta = (aqq - app) / (2.0 * apq);
      double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      double c = 1.0 / sqrt(t * t + 1.0);
      double s = t * c;
      // rotate rows/cols p,q of the symmetric matrix
      double r = 3 - p - q;  // the remaining index
      int rr = (int)r;
      double apr = (p == 0) ? (rr == 1 ? a01 : a02) : (rr == 0 ? a01 : a12);
      double aqr = (q == 1) ? (rr == 0 ? a01 : a12) : (rr == 0 ? a02 : a12);
      double napp = app - t * apq;
      double naqq = aqq + t * apq;
      double napr = c * apr - s * aqr;
      double naqr = s * apr + c * aqr;
      // write back
      if (p == 0) a00 = napp; else a11 = napp;
      if (q == 1) a11 = naqq; else a22 = naqq;
      if (pq == 0) a01 = 0.0; else if (pq == 1) a02 = 0.0; else a12 = 0.0;
      // a_pr and a_qr
      if (p == 0 && rr == 1) a01 = napr;
      if (p == 0 && rr == 2) a02 = napr;
      if (p == 1 && rr == 0) a01 = napr;
      if (p == 1 && rr == 2) a12 = napr;
      if (q == 1 && rr == 0) a01 = naqr;
      if (q == 1 && rr == 2) a12 = naqr;
      if (q == 2 && rr == 0) a02 = naqr;
      if (q == 2 && rr == 1) a12 = naqr;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double vkp = v[k][p], vkq = v[k][q];
        v[k][p] = c * vkp - s * vkq;
        v[k][q] = s * vkp + c * vkq;
      }
    }
  }
  lam[0] = a00;
  lam[1] = a11;
  lam[2] = a22;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V.a[i][j] = v[i][j];
}

// ------------------------------------------------------------------ materials
enum { MAT_HENCKY = 0, MAT_J2 = 1, MAT_NEOHOOKEAN = 2, MAT_DP = 3, MAT_NORSAND = 4 };

struct MatParams {
  int kind;
  int dp_apex;  // 0: reference (no return for P_trial >= 0), 1: apex return
  double lam, mu, E;
  double kappa, H;             // J2
  double phi, psi, coh, shape;  // DP (degrees, kPa)
  gm_norsand_params ns;        // NorSand
  double lnJ;                  // NorSand: total volumetric strain ln J (specific volume v0 exp(ln J))
};

// ---- NorSand (triaxial_NorSand.py:155-400)
GM_DEV double ns_eta(double M, double N, double P, double pi) {
  if (fabs(N) < 1e-10) return M * (1.0 + log(pi / P));
  return (M / N) * (1.0 - (1.0 - N) * pow(P / pi, N / (1.0 - N)));
}
GM_DEV double ns_deta_dP(double M, double N, double P, double pi) {
  if (fabs(N) < 1e-10) return -M / P;
  return -M * pow(P / pi, (2.0 * N - 1.0) / (1.0 - N)) / pi;
}
GM_DEV double ns_deta_dpi(double M, double N, double P, double pi) {
  if (fabs(N) < 1e-10) return M / pi;
  return M * pow(P / pi, (2.0 * N - 1.0) / (1.0 - N)) * P / (pi * pi);
}
GM_DEV double ns_eps_s(const double* e) {
  double a = e[0] - e[1], b = e[1] - e[2], c = e[0] - e[2];
  return sqrt(2.0 / 9.0 * (a * a + b * b + c * c));
}

// Local residual R and Jacobian dR/dx of the NorSand return at x = (ev, es, dl, pi) with
// P = P(ev) (triaxial_NorSand.py:338-396); dR_dv (optional): dR/d(specific volume), nonzero
// in the hardening row through pi*(psi_i).
GM_DEV void ns_local(const gm_norsand_params& q, double erv, double ers, double dl, double pi, double ev_tr,
                     double es_tr, double pi_old, double new_v, double r[4], double J[4][4], double* dR_dv) {
  const double M = q.M, N = q.N, alpha_bar = -3.5 / q.beta;
  const double Pt = q.p0 * exp(-(erv - q.ev0) / q.kappa);
  const double dEdP = ns_deta_dP(M, N, Pt, pi), dEdPi = ns_deta_dpi(M, N, Pt, pi), eta = ns_eta(M, N, Pt, pi);
  const double dFdP = (eta - M) / (1.0 - N);
  const double psi_i = new_v - q.vc0 + q.tilde_lambda * log(-pi);
  // pi*: the yield-test tolerance picks the branch for pi*, 1e-10 for its derivatives
  // (triaxial_NorSand.py:356-362 and the derivative helpers)
  const double pstar = fabs(N) < q.tol ? Pt * exp(alpha_bar * psi_i / M)
                                       : Pt * pow(1.0 - alpha_bar * psi_i * N / M, (N - 1.0) / N);
  const double dPs_dP = fabs(N) < 1e-10 ? exp(alpha_bar * psi_i / M)
                                        : pow(1.0 - alpha_bar * psi_i * N / M, (N - 1.0) / N);
  // d pi* / d psi_i
  const double dPs_dpsi = fabs(N) < 1e-10 ? Pt * exp(alpha_bar * psi_i / M) * alpha_bar / M
                                          : Pt * (N - 1.0) / N *
                                                pow(1.0 / (1.0 - alpha_bar * psi_i * N / M), 1.0 / N) *
                                                (-alpha_bar * N / M);
  const double Qt = 3.0 * q.mu * ers;
  r[0] = erv - ev_tr + dl * q.beta * dFdP;
  r[1] = ers - es_tr + dl;
  r[2] = Qt + eta * Pt;
  r[3] = pi - pi_old + q.h * (pi - pstar) * dl;
  const double dPdev = Pt / (-q.kappa);
  J[0][0] = 1.0 + dl * q.beta * dEdP * dPdev / (1.0 - N);
  J[0][1] = 0.0;
  J[0][2] = q.beta * dFdP;
  J[0][3] = dl * q.beta * dEdPi / (1.0 - N);
  J[1][0] = 0.0;
  J[1][1] = 1.0;
  J[1][2] = 1.0;
  J[1][3] = 0.0;
  J[2][0] = (dEdP * Pt + eta) * dPdev;
  J[2][1] = 3.0 * q.mu;
  J[2][2] = 0.0;
  J[2][3] = Pt * dEdPi;
  J[3][0] = -q.h * dl * dPs_dP * dPdev;
  J[3][1] = 0.0;
  J[3][2] = q.h * (pi - pstar);
  J[3][3] = 1.0 + q.h * dl * (1.0 - dPs_dpsi * q.tilde_lambda / pi);
  if (dR_dv) {
    dR_dv[0] = dR_dv[1] = dR_dv[2] = 0.0;
    dR_dv[3] = -q.h * dl * dPs_dpsi;
  }
}

GM_DEV void solve4(double M[4][4], double* r);

// NorSand return of principal trial strains et with total volumetric strain ev_total and image
// pressure pi_old: the reference's tension branch (zero strain when tr et > tol), yield test and
// fixed 15-iteration local Newton.  Optionally the consistent tangent d e / d et (dE, 3x3,
// row-major) and d e / d ln J (dEJ, 3) from the local system at the returned state (implicit
// function theorem), and the last iterate's local residuals (res_last).
GM_DEV void return_norsand(const double* et, double ev_total, double pi_old, const gm_norsand_params& q, double* e,
                           double& pi_new, double* res_last, double* dE, double* dEJ) {
  const double ev = et[0] + et[1] + et[2];
  const double es = ns_eps_s(et);
  double nd[3] = {0.0, 0.0, 0.0};
  if (es > q.tol) {
    const double s23 = sqrt(2.0 / 3.0);
    for (int i = 0; i < 3; ++i) nd[i] = s23 * (et[i] - ev / 3.0) / es;
  }
  const double new_v = exp(ev_total) * q.v0;
  double Pt = q.p0 * exp(-(ev - q.ev0) / q.kappa);
  double Qt = 3.0 * q.mu * es;
  e[0] = et[0];
  e[1] = et[1];
  e[2] = et[2];
  pi_new = pi_old;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (dE)
    for (int i = 0; i < 9; ++i) dE[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (dEJ) dEJ[0] = dEJ[1] = dEJ[2] = 0.0;
  if (ev > q.tol) {
    e[0] = e[1] = e[2] = 0.0;
    if (dE)
      for (int i = 0; i < 9; ++i) dE[i] = 0.0;
  } else if (Qt + ns_eta(q.M, q.N, Pt, pi_old) * Pt >= q.tol) {
    double dl = 0.0, pi = pi_old;
    const double s32 = sqrt(1.5);
    for (int it = 0; it < 15; ++it) {
      const double erv = e[0] + e[1] + e[2];
      const double ers = ns_eps_s(e);
      double J[4][4];
      ns_local(q, erv, ers, dl, pi, ev, es, pi_old, new_v, r, J, nullptr);
      double x[4] = {r[0], r[1], r[2], r[3]};
      solve4(J, x);
      const double nev = erv - x[0];
      const double nes = ers - x[1];
      dl -= x[2];
      pi -= x[3];
      for (int i = 0; i < 3; ++i) e[i] = nev / 3.0 + nes * nd[i] * s32;
    }
    pi_new = pi;
    if (dE || dEJ) {
      // d(ev, es, dl, pi) = J^-1 (d ev_tr e0 + d es_tr e1 - dR/dv dv) at the returned state
      const double erv = e[0] + e[1] + e[2];
      const double ers = ns_eps_s(e);
      double J[4][4], rr[4], dRv[4];
      ns_local(q, erv, ers, dl, pi, ev, es, pi_old, new_v, rr, J, dRv);
      double X[3][4];
      for (int c = 0; c < 3; ++c) {
        double Jc[4][4];
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) Jc[i][j] = J[i][j];
        double b[4] = {c == 0 ? 1.0 : 0.0, c == 1 ? 1.0 : 0.0, 0.0, 0.0};
        if (c == 2)
          for (int i = 0; i < 4; ++i) b[i] = -dRv[i];
        solve4(Jc, b);
        for (int i = 0; i < 4; ++i) X[c][i] = b[i];
      }
      // e_i = ev/3 + sqrt(3/2) es nd_i, nd_i = sqrt(2/3)(et_i - ev_tr/3)/es_tr
      if (dE) {
        const double s23 = sqrt(2.0 / 3.0);
        for (int j = 0; j < 3; ++j) {
          const double des_tr = s23 * nd[j];  // d es_tr / d et_j
          const double dev_j = X[0][0] + X[1][0] * des_tr;
          const double des_j = X[0][1] + X[1][1] * des_tr;
          for (int i = 0; i < 3; ++i) {
            double dnd = (es > q.tol) ? s23 / es * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0 - nd[i] * nd[j]) : 0.0;
            dE[3 * i + j] = dev_j / 3.0 + s32 * (des_j * nd[i] + ers * dnd);
          }
        }
      }
      if (dEJ)
        for (int i = 0; i < 3; ++i) dEJ[i] = (X[2][0] / 3.0 + s32 * X[2][1] * nd[i]) * new_v;
    }
  }
  if (res_last)
    for (int i = 0; i < 4; ++i) res_last[i] = r[i];
}

// Principal Kirchhoff stress of NorSand's pressure-dependent hyperelasticity
// (triaxial_NorSand.py:455-461) and its derivative C = d tau / d e.
GM_DEV void ns_tau(const double* e, const gm_norsand_params& q, double* t, double* C) {
  const double tr = e[0] + e[1] + e[2];
  const double P = q.p0 * exp(-(tr - q.ev0) / q.kappa);
  for (int i = 0; i < 3; ++i) t[i] = P + 2.0 * q.mu * (e[i] - tr / 3.0);
  if (C)
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) C[3 * i + k] = -P / q.kappa + 2.0 * q.mu * ((i == k ? 1.0 : 0.0) - 1.0 / 3.0);
}

// Principal return maps: eps_tr -> eps (3), d eps / d eps_tr (3x3), alpha.
GM_DEV void tau_principal(const double* e, double lam, double mu, double* t) {
  double tr = e[0] + e[1] + e[2];
  t[0] = lam * tr + 2.0 * mu * e[0];
  t[1] = lam * tr + 2.0 * mu * e[1];
  t[2] = lam * tr + 2.0 * mu * e[2];
}

// J2 radial return (plasticity.py:14-37); hardening kappa(alpha) = kappa + sqrt(2/3) H alpha.
GM_DEV void return_j2(const double* etr, const MatParams& m, double alpha_old, double* e, double* dE,
                      double& alpha_new) {
  double t[3];
  tau_principal(etr, m.lam, m.mu, t);
  double P = (t[0] + t[1] + t[2]) / 3.0;
  double S[3] = {t[0] - P, t[1] - P, t[2] - P};
  double Sn = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  const double sq23 = 0.816496580927726;
  double kap = m.kappa + sq23 * m.H * alpha_old;
#pragma unroll
  for (int i = 0; i < 9; ++i) dE[i] = (i % 4 == 0) ? 1.0 : 0.0;
  alpha_new = alpha_old;
  e[0] = etr[0];
  e[1] = etr[1];
  e[2] = etr[2];
  if (Sn > kap) {
    double n[3] = {S[0] / Sn, S[1] / Sn, S[2] / Sn};
    double den = 2.0 * m.mu + 2.0 / 3.0 * m.H;
    double dg = (Sn - kap) / den;
    e[0] = etr[0] - dg * n[0];
    e[1] = etr[1] - dg * n[1];
    e[2] = etr[2] - dg * n[2];
    alpha_new = alpha_old + sq23 * dg;
    // d eps/d eps_tr = I - (2mu/den) n n - (dg/|e_dev|)(P_dev - n n); |e_dev| = Sn/(2mu)
    double c1 = 2.0 * m.mu / den;
    double c2 = dg * 2.0 * m.mu / Sn;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double pd = (i == j ? 1.0 : 0.0) - 1.0 / 3.0;
        dE[3 * i + j] = (i == j ? 1.0 : 0.0) - c1 * n[i] * n[j] - c2 * (pd - n[i] * n[j]);
      }
  }
}

// ---- Drucker-Prager (triaxial_DruckerPrager.py:120-412)
GM_DEV void dp_AB(double angle, double coh, double& A, double& B) {
  const double pi_ref = 3.14159265358979;  // the reference's literal
  double c = cos(angle * pi_ref / 180.0), s = sin(angle * pi_ref / 180.0);
  A = 2.0 * sqrt(6.0) * coh * c / (3.0 - s);
  B = 2.0 * sqrt(6.0) * s / (3.0 - s);
}

GM_DEV void dp_PQS(const double* e, const MatParams& m, double* t, double& P, double* S, double& Q) {
  tau_principal(e, m.lam, m.mu, t);
  P = (t[0] + t[1] + t[2]) / 3.0;
  S[0] = t[0] - P;
  S[1] = t[1] - P;
  S[2] = t[2] - P;
  Q = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]) * sqrt(1.5);
}

GM_DEV double dp_yield(double P, double Q, double A, double B, double shape) {
  double om = 2.0 / 3.0 * Q * Q + shape * shape * A * A;
  return sqrt(om) - (A - B * P);
}

GM_DEV void dp_grad(const double* t, double Q, double A, double B, double shape, double* gr) {
  double om = 2.0 / 3.0 * Q * Q + shape * shape * A * A;
  double dFdQ = 2.0 / 3.0 * Q / sqrt(om);
  double sg = Q < 0.0 ? -1.0 : 1.0;
  double Qt = fabs(Q) < 1e-10 ? 1e-10 * sg : Q;
  gr[0] = B / 3.0 + dFdQ * (0.5 / Qt * (2.0 * t[0] - t[1] - t[2]));
  gr[1] = B / 3.0 + dFdQ * (0.5 / Qt * (2.0 * t[1] - t[0] - t[2]));
  gr[2] = B / 3.0 + dFdQ * (0.5 / Qt * (2.0 * t[2] - t[0] - t[1]));
}

GM_DEV void dp_hess(const double* S, double Q, double A, double shape, double* hg) {
  double chi = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  if (fabs(chi) < 1e-10) chi = 1e-10;
  double c1 = sqrt(1.5) / chi;
  double om = 2.0 / 3.0 * Q * Q + shape * shape * A * A;
  double so = sqrt(om);
  double dFdQ = 2.0 / 3.0 * Q / so;
  double d2 = 2.0 / 3.0 / so - 4.0 / 9.0 * Q * Q / (om * so);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double hq = c1 * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0 - S[i] * S[j] / (chi * chi));
      hg[3 * i + j] = dFdQ * hq + d2 * (c1 * S[i]) * (c1 * S[j]);
    }
}

// Solve 4x4 system M x = r in place (Gaussian elimination, partial pivoting).
GM_DEV void solve4(double M[4][4], double* r) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int piv = k;
    double best = fabs(M[k][k]);
#pragma unroll
    for (int i = k + 1; i < 4; ++i)
      if (fabs(M[i][k]) > best) {
        best = fabs(M[i][k]);
        piv = i;
      }
    if (piv != k) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double tmp = M[k][j];
        M[k][j] = M[piv][j];
        M[piv][j] = tmp;
      }
      double tmp = r[k];
      r[k] = r[piv];
      r[piv] = tmp;
    }
    double inv = 1.0 / M[k][k];
#pragma unroll
    for (int i = k + 1; i < 4; ++i) {
      double f = M[i][k] * inv;
#pragma unroll
      for (int j = k; j < 4; ++j) M[i][j] -= f * M[k][j];
      r[i] -= f * r[k];
    }
  }
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    double s = r[k];
#pragma unroll
    for (int j = k + 1; j < 4; ++j) s -= M[k][j] * r[j];
    r[k] = s / M[k][k];
  }
}

// DP local Jacobian at (e, dl) (triaxial_DruckerPrager.py:379-384).
GM_DEV void dp_local_jac(const double* e, double dl, const MatParams& m, double Af, double Bf, double Ag,
                         double Bg, double J[4][4], double* res, const double* etr) {
  double t[3], S[3], P, Q;
  dp_PQS(e, m, t, P, S, Q);
  double gg[3], gf[3], hg[9];
  dp_grad(t, Q, Ag, Bg, m.shape, gg);
  dp_grad(t, Q, Af, Bf, m.shape, gf);
  dp_hess(S, Q, Ag, m.shape, hg);
  double a[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) a[i][j] = m.lam + (i == j ? 2.0 * m.mu : 0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double he = hg[3 * i + 0] * a[0][j] + hg[3 * i + 1] * a[1][j] + hg[3 * i + 2] * a[2][j];
      J[i][j] = (i == j ? 1.0 : 0.0) + dl * he;
    }
    J[i][3] = gg[i];
    J[3][i] = gf[0] * a[0][i] + gf[1] * a[1][i] + gf[2] * a[2][i];
  }
  J[3][3] = 0.0;
  if (res) {
    res[0] = e[0] - etr[0] + dl * gg[0];
    res[1] = e[1] - etr[1] + dl * gg[1];
    res[2] = e[2] - etr[2] + dl * gg[2];
    res[3] = dp_yield(P, Q, Af, Bf, m.shape);
  }
}

GM_DEV void return_dp(const double* etr, const MatParams& m, double* e, double* dE) {
  double Af, Bf, Ag, Bg;
  dp_AB(m.phi, m.coh, Af, Bf);
  dp_AB(m.psi, m.coh, Ag, Bg);
  double t[3], S[3], P, Q;
  dp_PQS(etr, m, t, P, S, Q);
  double y = dp_yield(P, Q, Af, Bf, m.shape);
  e[0] = etr[0];
  e[1] = etr[1];
  e[2] = etr[2];
#pragma unroll
  for (int i = 0; i < 9; ++i) dE[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (y < 1e-10) return;
  double K = m.lam + 2.0 * m.mu / 3.0;
  bool tip;
  if (m.dp_apex == 0) {
    tip = (P >= 0.0);  // the reference: no return in tension (triaxial_DruckerPrager.py:345-349)
  } else if (m.shape > 0.0) {
    // hyperbolic surface: its tip (Q = 0) sits at P_tip = A (1 - a) / B; a trial state at or
    // beyond it in tension has no return along the flow direction when the flow is deviatoric
    // (dilation 0), and is cut off to the tip (tension cut-off)
    tip = (P >= Af * (1.0 - m.shape) / Bf);
  } else {
    // cone: the closed-form cone return would reverse the deviator -> apex
    double Sn = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
    double dl_cone = y / (2.0 * m.mu + Bf * K * Bg);
    tip = (2.0 * m.mu * dl_cone >= Sn);
  }
  if (tip) {
    if (m.dp_apex == 1) {
      double ev = (Af * (1.0 - m.shape) / Bf) / K;
      e[0] = e[1] = e[2] = ev / 3.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) dE[i] = 0.0;
    }
    return;
  }
  double dl = 0.0;
  for (int it = 0; it < 10; ++it) {
    double J[4][4], r[4];
    dp_local_jac(e, dl, m, Af, Bf, Ag, Bg, J, r, etr);
    solve4(J, r);
    e[0] -= r[0];
    e[1] -= r[1];
    e[2] -= r[2];
    dl -= r[3];
  }
  // consistent tangent: d(e,dl)/d etr = J^{-1} [I; 0] at the converged iterate
  double J[4][4];
  dp_local_jac(e, dl, m, Af, Bf, Ag, Bg, J, nullptr, etr);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double Jc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) Jc[i][j] = J[i][j];
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    r[k] = 1.0;
    solve4(Jc, r);
    dE[0 * 3 + k] = r[0];
    dE[1 * 3 + k] = r[1];
    dE[2 * 3 + k] = r[2];
  }
}

// Hencky-family update from b_tr.  Outputs Kirchhoff tau, b_new, alpha_new and
// (optionally) the spectral tangent as a callable applied to d b.
struct SpectralState {
  M3 V;         // eigenvectors (columns)
  double lam[3];  // eigenvalues of b_tr
  double tau[3];  // principal Kirchhoff
  double T[9];    // d tau_i / d eps_tr_k
  double TJ[3];   // d tau_i / d ln J (NorSand: the specific volume enters the return)
};

GM_DEV void hencky_family(const M3& btr, const MatParams& m, double alpha_old, SpectralState& ss, M3& tau,
                          M3& bnew, double& alpha_new) {
  sym_eig3(btr, ss.lam, ss.V);
  double etr[3], e[3], dE[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) etr[i] = 0.5 * log(fmax(ss.lam[i], 1e-8));
  alpha_new = alpha_old;
  ss.TJ[0] = ss.TJ[1] = ss.TJ[2] = 0.0;
  if (m.kind == MAT_NORSAND) {
    // alpha carries the image pressure pi; elasticity is NorSand's own
    double dEJ[3], C[9];
    return_norsand(etr, m.lnJ, alpha_old, m.ns, e, alpha_new, nullptr, dE, dEJ);
    ns_tau(e, m.ns, ss.tau, C);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double sj = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) s += C[3 * i + j] * dE[3 * j + k];
        ss.T[3 * i + k] = s;
        sj += C[3 * i + k] * dEJ[k];
      }
      ss.TJ[i] = sj;
    }
  } else {
    if (m.kind == MAT_J2) {
      return_j2(etr, m, alpha_old, e, dE, alpha_new);
    } else if (m.kind == MAT_DP) {
      return_dp(etr, m, e, dE);
    } else {
      e[0] = etr[0];
      e[1] = etr[1];
      e[2] = etr[2];
#pragma unroll
      for (int i = 0; i < 9; ++i) dE[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    tau_principal(e, m.lam, m.mu, ss.tau);
    // T = a_el * dE
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) s += (m.lam + (i == j ? 2.0 * m.mu : 0.0)) * dE[3 * j + k];
        ss.T[3 * i + k] = s;
      }
  }
  double ex[3] = {exp(2.0 * e[0]), exp(2.0 * e[1]), exp(2.0 * e[2])};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double st = 0.0, sb = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double nn = ss.V.a[i][k] * ss.V.a[j][k];
        st += ss.tau[k] * nn;
        sb += ex[k] * nn;
      }
      tau.a[i][j] = st;
      bnew.a[i][j] = sb;
    }
}

// d tau for a perturbation d b of b_tr (Daleckii-Krein with divided differences).
GM_DEV M3 spectral_dtau(const SpectralState& ss, const M3& db) {
  // rotate into the eigenbasis
  M3 Vt = m3_T(ss.V);
  M3 dbp = m3_mul(Vt, m3_mul(db, ss.V));
  M3 dtp;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += ss.T[3 * i + k] * dbp.a[k][k] / (2.0 * ss.lam[k]);
    dtp.a[i][i] = s;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (i == j) continue;
      double li = ss.lam[i], lj = ss.lam[j];
      double th;
      if (fabs(li - lj) > 1e-7 * fmax(fabs(li), fabs(lj))) {
        th = (ss.tau[i] - ss.tau[j]) / (li - lj);
      } else {
        th = (ss.T[3 * i + i] - ss.T[3 * i + j]) / (2.0 * li);
      }
      dtp.a[i][j] = th * dbp.a[i][j];
    }
  return m3_mul(ss.V, m3_mul(dtp, Vt));
}

}  // namespace gm

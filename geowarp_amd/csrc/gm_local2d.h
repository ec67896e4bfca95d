// Per-particle local map of one 2D plane-strain load step, generic in the
// scalar type (double or Dual), for the adjoint of quasi_static_solver_2d.
//
// For particle p with state s_p = (x, lp, F_old, b_old, alpha_old), material
// parameters theta = (E, nu, kappa, H, phi, psi, c) and the converged nodal
// increment u, it evaluates
//   * its residual contributions r_I (I over the 3x3 support), contracted with
//     a nodal cotangent:  acc = sum_I c_I . r_I      (assemble_residual_2d_*,
//     mpm_utils.py:416-780; the same expressions as the forward particle stage),
//   * the next state: F_new = f F_old, b_new (return-mapped), alpha_new,
//     x_new = x + sum_I w_I u_I (G2P_2d, mpm_utils.py:1351-1419) and
//     lp_new = lp0 |F_new column| clipped to the domain (update_GIMP_lp_2d,
//     gimp.py:428-462).
// The deformation-increment gradient takes an extra additive seed dH so the
// derivative with respect to H (and through it, u) is available.
//
// Principal values of the in-plane block of b_trial come from the closed-form
// 2x2 eigen-decomposition (the zz entry is its own principal value); the
// forward stage uses a Jacobi sweep, which gives the same spectral functions to
// round-off.
#pragma once
#include "gm_dual.h"

struct LocalConst {
  double h, inv_h;
  int nx1;               // n_grid_x + 1
  int mat_kind, dp_apex;
  double shape;          // Drucker-Prager shape factor
  double p_vol, p_rho;
  double gravity[2], point_load[2];  // already ramped by (step+1)/total
  double lp0, dmin[2], dmax[2];
};

template <class T>
struct LocalIn {
  T x[2], lp[2];
  T Fo[5];  // F_old: xx, xy, yx, yy, zz
  T bo[4];  // b_old: xx, xy (= yx), yy, zz
  T alpha;
  T dH[4];  // additive seed of H: xx, xy, yx, yy
  T E, nu, kappa, Hh, phi, psi, coh;
  bool ext;
};

template <class T>
struct LocalOut {
  T acc;  // sum_I c_I . r_I
  T xn[2], lpn[2], Fn[5], bn[4], alpha;
  T sig[4];  // Cauchy stress xx, xy, yy, zz
};

template <class T>
struct LAxis {
  int lo, ext;
  T w[3], g[3];
};

template <class T>
GM_DD LAxis<T> l_gimp_axis(T x, T lp, double h, double inv_h) {
  LAxis<T> a;
  const double xv = val(x), lv = val(lp);
  int lo = (int)((xv - lv) * inv_h + 1e-8);
  int hi = (int)((xv + lv) * inv_h + 1e-8);
  int base = (int)(xv * inv_h + 1e-8);
  a.lo = lo;
  if (lo == hi) {
    T f = x * inv_h - (double)base;
    a.ext = 2;
    a.w[0] = 1.0 - f;
    a.w[1] = f;
    a.w[2] = T(0.0 * f);
    a.g[0] = T(0.0 * f) + (-1.0 * inv_h);
    a.g[1] = T(0.0 * f) + inv_h;
    a.g[2] = T(0.0 * f);
  } else {
    a.ext = 3;
    double lof = (double)lo;
    double x0 = lof * h, x1 = (lof + 1.0) * h, x2 = (lof + 2.0) * h;
    T d0 = x - x0, d2 = x - x2;
    T a0 = (h + lp) - fabs(d0);
    T a2 = (h + lp) - fabs(d2);
    T hl4 = 4.0 * h * lp, hl2 = 2.0 * h * lp;
    double s0 = val(d0) / fabs(val(d0)), s2 = val(d2) / fabs(val(d2));
    a.w[0] = a0 * a0 / hl4;
    a.g[0] = -a0 / hl2 * s0;
    T dx1 = x - x1;
    a.w[1] = 1.0 - (dx1 * dx1 + lp * lp) / hl2;
    a.g[1] = -dx1 / (h * lp);
    a.w[2] = a2 * a2 / hl4;
    a.g[2] = -a2 / hl2 * s2;
    if (val(a.w[2]) == 0.0 && val(a.g[2]) == 0.0 && val(a2) == 0.0) a.ext = 2;
  }
  return a;
}

template <class T>
GM_DD void l_tau(const T* e, T lam, T mu, T* t) {
  T tr = e[0] + e[1] + e[2];
  for (int i = 0; i < 3; ++i) t[i] = lam * tr + 2.0 * mu * e[i];
}

// ---- J2 radial return with linear hardening (plasticity.py:8-37)
template <class T>
GM_DD void l_return_j2(const T* etr, T lam, T mu, T kappa, T Hh, T alpha_old, T* e, T& alpha_new) {
  T t[3];
  l_tau(etr, lam, mu, t);
  T P = (t[0] + t[1] + t[2]) / 3.0;
  T S[3] = {t[0] - P, t[1] - P, t[2] - P};
  T Sn = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  const double sq23 = 0.816496580927726;
  T kap = kappa + sq23 * Hh * alpha_old;
  for (int i = 0; i < 3; ++i) e[i] = etr[i];
  alpha_new = alpha_old;
  if (val(Sn) > val(kap)) {
    T den = 2.0 * mu + 2.0 / 3.0 * Hh;
    T dg = (Sn - kap) / den;
    for (int i = 0; i < 3; ++i) e[i] = etr[i] - dg * (S[i] / Sn);
    alpha_new = alpha_old + sq23 * dg;
  }
}

// ---- Drucker-Prager (triaxial_DruckerPrager.py:120-412), same iteration as the forward stage
template <class T>
GM_DD void l_dp_AB(T angle, T coh, T& A, T& B) {
  const double pi_ref = 3.14159265358979;
  T c = cos(angle * (pi_ref / 180.0)), s = sin(angle * (pi_ref / 180.0));
  A = 2.0 * sqrt(6.0) * coh * c / (3.0 - s);
  B = 2.0 * sqrt(6.0) * s / (3.0 - s);
}

template <class T>
GM_DD void l_dp_PQS(const T* e, T lam, T mu, T* t, T& P, T* S, T& Q) {
  l_tau(e, lam, mu, t);
  P = (t[0] + t[1] + t[2]) / 3.0;
  for (int i = 0; i < 3; ++i) S[i] = t[i] - P;
  Q = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]) * sqrt(1.5);
}

template <class T>
GM_DD T l_dp_yield(T P, T Q, T A, T B, double shape) {
  T om = 2.0 / 3.0 * Q * Q + shape * shape * A * A;
  return sqrt(om) - (A - B * P);
}

template <class T>
GM_DD void l_dp_grad(const T* t, T Q, T A, T B, double shape, T* gr) {
  T om = 2.0 / 3.0 * Q * Q + shape * shape * A * A;
  T dFdQ = 2.0 / 3.0 * Q / sqrt(om);
  T Qt = Q;
  if (fabs(val(Q)) < 1e-10) Qt = T(0.0 * Q) + (val(Q) < 0.0 ? -1e-10 : 1e-10);
  gr[0] = B / 3.0 + dFdQ * (0.5 / Qt * (2.0 * t[0] - t[1] - t[2]));
  gr[1] = B / 3.0 + dFdQ * (0.5 / Qt * (2.0 * t[1] - t[0] - t[2]));
  gr[2] = B / 3.0 + dFdQ * (0.5 / Qt * (2.0 * t[2] - t[0] - t[1]));
}

template <class T>
GM_DD void l_dp_hess(const T* S, T Q, T A, double shape, T* hg) {
  T chi = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  if (fabs(val(chi)) < 1e-10) chi = T(0.0 * chi) + 1e-10;
  T c1 = sqrt(1.5) / chi;
  T om = 2.0 / 3.0 * Q * Q + shape * shape * A * A;
  T so = sqrt(om);
  T dFdQ = 2.0 / 3.0 * Q / so;
  T d2 = 2.0 / 3.0 / so - 4.0 / 9.0 * Q * Q / (om * so);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T hq = c1 * (((i == j) ? 1.0 : 0.0) - 1.0 / 3.0 - S[i] * S[j] / (chi * chi));
      hg[3 * i + j] = dFdQ * hq + d2 * (c1 * S[i]) * (c1 * S[j]);
    }
}

template <class T>
GM_DD void l_solve4(T M[4][4], T* r) {
  for (int k = 0; k < 4; ++k) {
    int piv = k;
    double best = fabs(val(M[k][k]));
    for (int i = k + 1; i < 4; ++i)
      if (fabs(val(M[i][k])) > best) {
        best = fabs(val(M[i][k]));
        piv = i;
      }
    if (piv != k) {
      for (int j = 0; j < 4; ++j) {
        T tmp = M[k][j];
        M[k][j] = M[piv][j];
        M[piv][j] = tmp;
      }
      T tmp = r[k];
      r[k] = r[piv];
      r[piv] = tmp;
    }
    T inv = 1.0 / M[k][k];
    for (int i = k + 1; i < 4; ++i) {
      T f = M[i][k] * inv;
      for (int j = k; j < 4; ++j) M[i][j] -= f * M[k][j];
      r[i] -= f * r[k];
    }
  }
  for (int k = 3; k >= 0; --k) {
    T s = r[k];
    for (int j = k + 1; j < 4; ++j) s -= M[k][j] * r[j];
    r[k] = s / M[k][k];
  }
}

template <class T>
GM_DD void l_return_dp(const T* etr, T lam, T mu, T phi, T psi, T coh, double shape, int apex, T* e) {
  T Af, Bf, Ag, Bg;
  l_dp_AB(phi, coh, Af, Bf);
  l_dp_AB(psi, coh, Ag, Bg);
  T t[3], S[3], P, Q;
  l_dp_PQS(etr, lam, mu, t, P, S, Q);
  T y = l_dp_yield(P, Q, Af, Bf, shape);
  for (int i = 0; i < 3; ++i) e[i] = etr[i];
  if (val(y) < 1e-10) return;
  T K = lam + 2.0 * mu / 3.0;
  bool tip;
  if (apex == 0) {
    tip = (val(P) >= 0.0);  // the reference: no return in tension
  } else if (shape > 0.0) {
    tip = (val(P) >= val(Af * (1.0 - shape) / Bf));  // at or beyond the hyperbola's tip: cut off
  } else {
    T Sn = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
    T dl_cone = y / (2.0 * mu + Bf * K * Bg);
    tip = (val(2.0 * mu * dl_cone) >= val(Sn));
  }
  if (tip) {
    if (apex == 1) {
      T ev = (Af * (1.0 - shape) / Bf) / K;
      for (int i = 0; i < 3; ++i) e[i] = ev / 3.0;
    }
    return;
  }
  T dl = T(0.0 * y);
  for (int it = 0; it < 10; ++it) {
    T tt[3], SS[3], PP, QQ;
    l_dp_PQS(e, lam, mu, tt, PP, SS, QQ);
    T gg[3], gf[3], hg[9];
    l_dp_grad(tt, QQ, Ag, Bg, shape, gg);
    l_dp_grad(tt, QQ, Af, Bf, shape, gf);
    l_dp_hess(SS, QQ, Ag, shape, hg);
    T J[4][4], r[4];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) {
        T a0 = lam + ((0 == j) ? 2.0 * mu : T(0.0 * mu));
        T a1 = lam + ((1 == j) ? 2.0 * mu : T(0.0 * mu));
        T a2 = lam + ((2 == j) ? 2.0 * mu : T(0.0 * mu));
        T he = hg[3 * i + 0] * a0 + hg[3 * i + 1] * a1 + hg[3 * i + 2] * a2;
        J[i][j] = ((i == j) ? 1.0 : 0.0) + dl * he;
      }
      J[i][3] = gg[i];
    }
    for (int i = 0; i < 3; ++i) {
      T a0 = lam + ((i == 0) ? 2.0 * mu : T(0.0 * mu));
      T a1 = lam + ((i == 1) ? 2.0 * mu : T(0.0 * mu));
      T a2 = lam + ((i == 2) ? 2.0 * mu : T(0.0 * mu));
      J[3][i] = gf[0] * a0 + gf[1] * a1 + gf[2] * a2;
    }
    J[3][3] = T(0.0 * dl);
    for (int i = 0; i < 3; ++i) r[i] = e[i] - etr[i] + dl * gg[i];
    r[3] = l_dp_yield(PP, QQ, Af, Bf, shape);
    l_solve4(J, r);
    for (int i = 0; i < 3; ++i) e[i] -= r[i];
    dl -= r[3];
  }
}

// Material kinds (same numbering as the forward stage)
#define LMAT_HENCKY 0
#define LMAT_J2 1
#define LMAT_NH 2
#define LMAT_DP 3

template <class T>
GM_DD void local_map_2d(const LocalConst& k, const LocalIn<T>& in, const double* u, const double* cot,
                        LocalOut<T>& out) {
  LAxis<T> ax[2];
  for (int a = 0; a < 2; ++a) ax[a] = l_gimp_axis(in.x[a], in.lp[a], k.h, k.inv_h);
  // H_ab = sum_I G_I,a u_I,b (the reference's transposed increment gradient) + seed
  T H[2][2] = {{in.dH[0], in.dH[1]}, {in.dH[2], in.dH[3]}};
  T du[2] = {T(0.0 * in.x[0]), T(0.0 * in.x[0])};
  for (int i = 0; i < ax[0].ext; ++i)
    for (int j = 0; j < ax[1].ext; ++j) {
      int node = (ax[0].lo + i) + (ax[1].lo + j) * k.nx1;
      T w = ax[0].w[i] * ax[1].w[j];
      T G0 = ax[0].g[i] * ax[1].w[j], G1 = ax[0].w[i] * ax[1].g[j];
      double u0 = u[2 * node], u1 = u[2 * node + 1];
      H[0][0] += G0 * u0;
      H[0][1] += G0 * u1;
      H[1][0] += G1 * u0;
      H[1][1] += G1 * u1;
      du[0] += w * u0;
      du[1] += w * u1;
    }
  // kinematics
  T f00 = 1.0 + H[0][0], f01 = H[0][1], f10 = H[1][0], f11 = 1.0 + H[1][1];
  T det = f00 * f11 - f01 * f10;
  T A00 = f11 / det, A01 = -f01 / det, A10 = -f10 / det, A11 = f00 / det;
  const T* Fo = in.Fo;
  T Fn00 = f00 * Fo[0] + f01 * Fo[2], Fn01 = f00 * Fo[1] + f01 * Fo[3];
  T Fn10 = f10 * Fo[0] + f11 * Fo[2], Fn11 = f10 * Fo[1] + f11 * Fo[3];
  T Fn22 = Fo[4];
  T J = Fn22 * (Fn00 * Fn11 - Fn01 * Fn10);
  T lam = in.E * in.nu / ((1.0 + in.nu) * (1.0 - 2.0 * in.nu));
  T mu = in.E / (2.0 * (1.0 + in.nu));
  T s00, s01, s11, s22;  // Cauchy stress
  T bn00, bn01, bn11, bn22;
  out.alpha = in.alpha;
  if (k.mat_kind == LMAT_NH) {
    // sigma = mu/J (F F^T - I) + lambda ln J / J I (mpm_utils.py:915)
    T lnJ = log(J);
    T c0 = lam * lnJ / J;
    s00 = mu / J * (Fn00 * Fn00 + Fn01 * Fn01 - 1.0) + c0;
    s01 = mu / J * (Fn00 * Fn10 + Fn01 * Fn11);
    s11 = mu / J * (Fn10 * Fn10 + Fn11 * Fn11 - 1.0) + c0;
    s22 = mu / J * (Fn22 * Fn22 - 1.0) + c0;
    // b_trial = f b_old f^T
    const T* bo = in.bo;
    T t00 = f00 * bo[0] + f01 * bo[1], t01 = f00 * bo[1] + f01 * bo[2];
    T t10 = f10 * bo[0] + f11 * bo[1], t11 = f10 * bo[1] + f11 * bo[2];
    bn00 = t00 * f00 + t01 * f01;
    bn01 = t00 * f10 + t01 * f11;
    bn11 = t10 * f10 + t11 * f11;
    bn22 = bo[3];
  } else {
    const T* bo = in.bo;
    T t00 = f00 * bo[0] + f01 * bo[1], t01 = f00 * bo[1] + f01 * bo[2];
    T t10 = f10 * bo[0] + f11 * bo[1], t11 = f10 * bo[1] + f11 * bo[2];
    T a = t00 * f00 + t01 * f01;
    T b = 0.5 * ((t00 * f10 + t01 * f11) + (t10 * f00 + t11 * f01));
    T c = t10 * f10 + t11 * f11;
    T mean = 0.5 * (a + c), dif = 0.5 * (a - c);
    T r = sqrt(dif * dif + b * b);
    T l1 = mean + r, l2 = mean - r, l3 = bo[3];
    T th = 0.5 * atan2(b, dif);
    T cs = cos(th), sn = sin(th);
    T etr[3] = {0.5 * log(fmax(l1, 1e-8)), 0.5 * log(fmax(l2, 1e-8)), 0.5 * log(fmax(l3, 1e-8))};
    T e[3];
    if (k.mat_kind == LMAT_J2) {
      l_return_j2(etr, lam, mu, in.kappa, in.Hh, in.alpha, e, out.alpha);
    } else if (k.mat_kind == LMAT_DP) {
      l_return_dp(etr, lam, mu, in.phi, in.psi, in.coh, k.shape, k.dp_apex, e);
    } else {
      for (int i = 0; i < 3; ++i) e[i] = etr[i];
    }
    T t[3];
    l_tau(e, lam, mu, t);
    // v1 = (cs, sn), v2 = (-sn, cs)
    T tau00 = t[0] * cs * cs + t[1] * sn * sn;
    T tau01 = (t[0] - t[1]) * cs * sn;
    T tau11 = t[0] * sn * sn + t[1] * cs * cs;
    s00 = tau00 / J;
    s01 = tau01 / J;
    s11 = tau11 / J;
    s22 = t[2] / J;
    T x1 = exp(2.0 * e[0]), x2 = exp(2.0 * e[1]);
    bn00 = x1 * cs * cs + x2 * sn * sn;
    bn01 = (x1 - x2) * cs * sn;
    bn11 = x1 * sn * sn + x2 * cs * cs;
    bn22 = exp(2.0 * e[2]);
  }
  // residual contributions r_I,c = sum_b G_I,b P_bc + w_I T_c (mpm_utils.py:564-596)
  T V = k.p_vol * J;
  T P00 = -(A00 * s00 + A10 * s01) * V;  // P_bc = -V sum_a A_ab sigma_ac
  T P01 = -(A00 * s01 + A10 * s11) * V;
  T P10 = -(A01 * s00 + A11 * s01) * V;
  T P11 = -(A01 * s01 + A11 * s11) * V;
  T rho = k.p_rho / J;
  T T0 = rho * k.gravity[0] * V + (in.ext ? k.point_load[0] : 0.0);
  T T1 = rho * k.gravity[1] * V + (in.ext ? k.point_load[1] : 0.0);
  T acc = T(0.0 * V);
  if (cot) {
    for (int i = 0; i < ax[0].ext; ++i)
      for (int j = 0; j < ax[1].ext; ++j) {
        int node = (ax[0].lo + i) + (ax[1].lo + j) * k.nx1;
        double c0 = cot[2 * node], c1 = cot[2 * node + 1];
        if (c0 == 0.0 && c1 == 0.0) continue;
        T w = ax[0].w[i] * ax[1].w[j];
        T G0 = ax[0].g[i] * ax[1].w[j], G1 = ax[0].w[i] * ax[1].g[j];
        acc += c0 * (G0 * P00 + G1 * P10 + w * T0) + c1 * (G0 * P01 + G1 * P11 + w * T1);
      }
  }
  out.acc = acc;
  out.sig[0] = s00;
  out.sig[1] = s01;
  out.sig[2] = s11;
  out.sig[3] = s22;
  out.Fn[0] = Fn00;
  out.Fn[1] = Fn01;
  out.Fn[2] = Fn10;
  out.Fn[3] = Fn11;
  out.Fn[4] = Fn22;
  out.bn[0] = bn00;
  out.bn[1] = bn01;
  out.bn[2] = bn11;
  out.bn[3] = bn22;
  // G2P and the lp update (lp from the column norms of F_new, clipped on x_new)
  const T Fc[2][2] = {{Fn00, Fn01}, {Fn10, Fn11}};
  for (int a = 0; a < 2; ++a) {
    out.xn[a] = in.x[a] + du[a];
    T U = sqrt(Fc[0][a] * Fc[0][a] + Fc[1][a] * Fc[1][a]);
    T l = k.lp0 * U;
    if (val(out.xn[a] - l) < k.dmin[a]) l = out.xn[a] - k.dmin[a] - 1e-6;
    if (val(out.xn[a] + l) > k.dmax[a]) l = (k.dmax[a] - out.xn[a]) - 1e-6;
    out.lpn[a] = l;
  }
}

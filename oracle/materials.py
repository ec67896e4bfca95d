"""Constitutive updates (oracle restatement, test infrastructure only).

* Hencky elasticity in principal logarithmic strain: mpm_utils.py:521-560
* J2 radial return (perfectly plastic): materials/plasticity.py:8-37;
  optional linear isotropic hardening (not in the reference, BASELINE c2):
  kappa(alpha) = kappa0 + sqrt(2/3)*H*alpha, dgamma = (|s_tr|-kappa)/(2mu+2H/3)
* Drucker-Prager principal-strain return with a fixed 10-iteration local
  Newton: problems/material_tests/triaxial_DruckerPrager.py:120-412
* NorSand return (Borja & Andrade 2006), fixed 15-iteration local Newton:
  problems/material_tests/triaxial_NorSand.py:155-400
* Neo-Hookean Cauchy stress: mpm_utils.py:915 (2D) and :1205-1208 (3D)

All routines are vectorised over particles; principal values are (N,3).
"""
import numpy as np

SQ23 = np.sqrt(2.0 / 3.0)
PI_REF = 3.14159265358979  # the reference's literal (triaxial_DruckerPrager.py:128)


def spectral(b):
    """Eigen-decomposition of symmetric b (N,3,3): LAPACK eigh (stress is basis-invariant)."""
    lam, vec = np.linalg.eigh(b)
    return lam, vec


def log_strain(lam):
    """0.5*log(max(lambda, 1e-8)) (mpm_utils.py:531-535)."""
    return 0.5 * np.log(np.maximum(lam, 1e-8))


def hencky_principal_tau(eps, lam_l, mu):
    """tau_i = lambda*tr(eps) + 2 mu eps_i (mpm_utils.py:541-542)."""
    tr = eps.sum(axis=1, keepdims=True)
    return lam_l * tr + 2.0 * mu * eps


def assemble_spectral(vals, vec):
    """sum_k vals_k n_k n_k^T (mpm_utils.py:543, :560)."""
    return np.einsum('nk,nik,njk->nij', vals, vec, vec)


def return_mapping_J2(eps_tr, lam_l, mu, kappa, H=0.0, alpha_old=None):
    """plasticity.py:8-37 with optional linear hardening.

    Returns (eps_real, alpha_new)."""
    n = eps_tr.shape[0]
    if alpha_old is None:
        alpha_old = np.zeros(n)
    tau = hencky_principal_tau(eps_tr, lam_l, mu)
    P = tau.sum(axis=1, keepdims=True) / 3.0
    S = tau - P
    Sn = np.sqrt((S ** 2).sum(axis=1))
    kap = kappa + SQ23 * H * alpha_old
    yield_ = Sn > kap
    eps = eps_tr.copy()
    alpha = alpha_old.copy()
    if np.any(yield_):
        nn = S[yield_] / Sn[yield_, None]
        dg = (Sn[yield_] - kap[yield_]) / (2.0 * mu + 2.0 / 3.0 * H)
        eps[yield_] = eps_tr[yield_] - dg[:, None] * nn
        alpha[yield_] = alpha_old[yield_] + SQ23 * dg
    return eps, alpha


# ---------------------------------------------------------------- Drucker-Prager
def _dp_AB(angle_deg, cohesion):
    """A, B of the DP cone (triaxial_DruckerPrager.py:130-134)."""
    c = np.cos(angle_deg * PI_REF / 180.0)
    s = np.sin(angle_deg * PI_REF / 180.0)
    A = 2.0 * np.sqrt(6.0) * cohesion * c / (3.0 - s)
    B = 2.0 * np.sqrt(6.0) * s / (3.0 - s)
    return A, B


def _dp_PQ(eps, lam_l, mu):
    tau = hencky_principal_tau(eps, lam_l, mu)
    P = tau.sum(axis=1) / 3.0
    S = tau - P[:, None]
    Q = np.sqrt((S ** 2).sum(axis=1)) * np.sqrt(1.5)
    return tau, P, S, Q


def dp_yield(P, Q, phi, c, shape):
    """yield_function_DruckerPrager (triaxial_DruckerPrager.py:121-138)."""
    A, B = _dp_AB(phi, c)
    om = 2.0 / 3.0 * Q * Q + shape * shape * A * A
    return np.sqrt(om) - (A - B * P)


def _dp_grad(tau, Q, angle, c, shape):
    """grad_potential / grad_yield (triaxial_DruckerPrager.py:141-226)."""
    A, B = _dp_AB(angle, c)
    om = 2.0 / 3.0 * Q * Q + shape * shape * A * A
    dFdP = B
    dFdQ = 2.0 / 3.0 * Q / np.sqrt(om)
    sgn = np.where(Q < 0.0, -1.0, 1.0)
    Qt = np.where(np.abs(Q) < 1e-10, 1e-10 * sgn, Q)
    t0, t1, t2 = tau[:, 0], tau[:, 1], tau[:, 2]
    gQ = np.stack([0.5 / Qt * (2 * t0 - t1 - t2), 0.5 / Qt * (2 * t1 - t0 - t2),
                   0.5 / Qt * (2 * t2 - t0 - t1)], axis=1)
    return dFdP / 3.0 + dFdQ[:, None] * gQ


def _dp_hess(S, Q, psi, c, shape):
    """hess_potential_DruckerPrager (triaxial_DruckerPrager.py:241-282)."""
    trS2 = (S ** 2).sum(axis=1)
    chi = np.sqrt(trS2)
    chi = np.where(np.abs(chi) < 1e-10, 1e-10, chi)
    c1 = np.sqrt(1.5) / chi
    gQ = c1[:, None] * S
    I3 = np.eye(3)
    hessQ = c1[:, None, None] * (I3[None] - 1.0 / 3.0 - S[:, :, None] * S[:, None, :] / (chi * chi)[:, None, None])
    A, B = _dp_AB(psi, c)
    om = 2.0 / 3.0 * Q * Q + shape * shape * A * A
    dFdQ = 2.0 / 3.0 * Q / np.sqrt(om)
    d2 = 2.0 / 3.0 / np.sqrt(om) - 4.0 / 9.0 * Q * Q / (om * np.sqrt(om))
    return dFdQ[:, None, None] * hessQ + d2[:, None, None] * gQ[:, :, None] * gQ[:, None, :]


def return_mapping_DruckerPrager(eps_tr, lam_l, mu, phi, psi, c, shape=0.0, apex='reference', n_iter=10):
    """return_mapping_DruckerPrager (triaxial_DruckerPrager.py:308-412).

    Fixed 10-iteration local Newton on (eps1, eps2, eps3, dlambda) (n_iter
    other than 10 exists for tests that separate converged returns from
    unconverged ones).  No return
    when f < 1e-10 or when P_trial >= 0 (the reference skips the apex,
    :345-349).  ``apex='return'`` replaces that skip (not in the reference) by:
    on the cone (shape 0), a return to the apex P = A/B when the closed-form
    cone return would reverse the deviator; on the hyperbola (shape a > 0), a
    cut-off to its tip (Q = 0, P = A (1 - a) / B) for trial states at or beyond
    it, which a deviatoric flow (dilation 0) cannot return; other yielding
    states take the local Newton.
    """
    n = eps_tr.shape[0]
    a_el = lam_l * np.ones((3, 3)) + 2.0 * mu * np.eye(3)
    tau, P, S, Q = _dp_PQ(eps_tr, lam_l, mu)
    y = dp_yield(P, Q, phi, c, shape)
    eps = eps_tr.copy()
    plastic = (y >= 1e-10) & (P < 0.0)
    if apex == 'return':
        A, B = _dp_AB(phi, c)
        _, Bg = _dp_AB(psi, c)
        K = lam_l + 2.0 * mu / 3.0
        if shape > 0.0:
            tip = (y >= 1e-10) & (P >= A * (1.0 - shape) / B)
        else:
            Sn = np.sqrt((S ** 2).sum(axis=1))
            dl_cone = y / (2.0 * mu + B * K * Bg)
            tip = (y >= 1e-10) & (2.0 * mu * dl_cone >= Sn)
        if np.any(tip):
            # volumetric strain giving P = P_tip, zero deviator
            eps[tip] = (A * (1.0 - shape) / B) / K / 3.0
        plastic = (y >= 1e-10) & ~tip
    idx = np.nonzero(plastic)[0]
    if idx.size:
        e_tr = eps_tr[idx]
        e = e_tr.copy()
        dl = np.zeros(idx.size)
        t, Pp, Sp, Qp = _dp_PQ(e, lam_l, mu)
        for _ in range(n_iter):
            gg = _dp_grad(t, Qp, psi, c, shape)
            gf = _dp_grad(t, Qp, phi, c, shape)
            gfe = gf @ a_el  # grad_f[a]*a[a,b] (:229-238)
            hg = _dp_hess(Sp, Qp, psi, c, shape)
            hge = hg @ a_el  # (:285-306)
            yv = dp_yield(Pp, Qp, phi, c, shape)
            res = np.concatenate([e - e_tr + dl[:, None] * gg, yv[:, None]], axis=1)
            Jm = np.zeros((idx.size, 4, 4))
            Jm[:, :3, :3] = np.eye(3)[None] + dl[:, None, None] * hge
            Jm[:, :3, 3] = gg
            Jm[:, 3, :3] = gfe
            xd = np.linalg.solve(Jm, res[:, :, None])[:, :, 0]
            e = e - xd[:, :3]
            dl = dl - xd[:, 3]
            t, Pp, Sp, Qp = _dp_PQ(e, lam_l, mu)
        eps[idx] = e
    return eps


def dp_cone_closed_form(eps_tr, lam_l, mu, phi, psi, c):
    """Closed-form DP cone return for shape factor 0 (de Souza Neto et al.,
    Computational Methods for Plasticity, sec. 8.3.2, in the reference's
    P-Q convention).  Used only to pin the local-Newton restatement."""
    A_f, B_f = _dp_AB(phi, c)
    _, B_g = _dp_AB(psi, c)
    tau, P, S, Q = _dp_PQ(eps_tr, lam_l, mu)
    K = lam_l + 2.0 * mu / 3.0
    # f = sqrt(2/3) Q - A + B P ; flow: d eps = dl*(B_g/3 * 1 + sqrt(2/3)*3/2 S/Q)
    # Q decreases by 3 mu sqrt(2/3) dl, P decreases by K B_g dl
    f = SQ23 * Q - A_f + B_f * P
    dl = f / (SQ23 * 3.0 * mu * SQ23 + B_f * K * B_g)
    nhat = S / np.sqrt((S ** 2).sum(axis=1, keepdims=True))
    de = dl[:, None] * (B_g / 3.0 + SQ23 * np.sqrt(1.5) * nhat)
    return eps_tr - de


# ---------------------------------------------------------------- NorSand
def _ns_eta(M, N, P, pi):
    if abs(N) < 1e-10:
        return M * (1.0 + np.log(pi / P))
    return (M / N) * (1.0 - (1.0 - N) * np.power(P / pi, N / (1.0 - N)))


def _ns_deta_dP(M, N, P, pi):
    if abs(N) < 1e-10:
        return M * -1.0 / P
    return M * (-1.0 * np.power(P / pi, (2.0 * N - 1.0) / (1.0 - N)) / pi)


def _ns_deta_dpi(M, N, P, pi):
    if abs(N) < 1e-10:
        return M / pi
    return M * (-1.0 * np.power(P / pi, (2.0 * N - 1.0) / (1.0 - N)) * (-P) / pi ** 2)


def _eps_s(e):
    return np.sqrt(2.0 / 9.0 * ((e[:, 0] - e[:, 1]) ** 2 + (e[:, 1] - e[:, 2]) ** 2 + (e[:, 0] - e[:, 2]) ** 2))


def return_mapping_NorSand(eps_tr, eps_total_trace, pi_old, prm, tol=1e-10):
    """return_mapping_NorSand (triaxial_NorSand.py:243-400).

    prm: dict with kappa, tilde_lambda, M, N, beta, vc0, p0, v0, h, mu, ev0.
    Returns (eps_real, pi_new, local_residuals_of_last_iterate).
    """
    kap, tl, M, N, beta = prm['kappa'], prm['tilde_lambda'], prm['M'], prm['N'], prm['beta']
    vc0, p0, v0, h, mu, ev0 = prm['vc0'], prm['p0'], prm['v0'], prm['h'], prm['mu'], prm['ev0']
    alpha_bar = -3.5 / beta
    n = eps_tr.shape[0]
    eps = eps_tr.copy()
    pi_new = pi_old.copy()
    ev = eps_tr.sum(axis=1)
    es = _eps_s(eps_tr)
    nd = np.zeros((n, 3))
    m = es > tol
    nd[m] = SQ23 * (eps_tr[m] - ev[m, None] / 3.0) / es[m, None]
    P_tr = p0 * np.exp(-(ev - ev0) / kap)
    Q_tr = 3.0 * mu * es
    new_v = np.exp(eps_total_trace) * v0
    res_last = np.zeros((n, 4))
    tension = ev > tol
    eps[tension] = 0.0
    y = Q_tr + _ns_eta(M, N, P_tr, pi_old) * P_tr
    plast = (~tension) & (y >= tol)
    for p in np.nonzero(plast)[0]:
        e = eps_tr[p].copy()
        dl = 0.0
        pi = pi_old[p]
        Pt, Qt = P_tr[p], Q_tr[p]
        for _ in range(15):
            erv = e.sum()
            ers = _eps_s(e[None])[0]
            dEdP = _ns_deta_dP(M, N, Pt, pi)
            dEdPi = _ns_deta_dpi(M, N, Pt, pi)
            eta = _ns_eta(M, N, Pt, pi)
            dFdP = (eta - M) / (1.0 - N)
            dFdQ = 1.0
            psi_i = new_v[p] - vc0 + tl * np.log(-pi)
            if abs(N) < tol:
                pstar = Pt * np.exp(alpha_bar * psi_i / M)
            else:
                pstar = Pt * np.power(1.0 - alpha_bar * psi_i * N / M, (N - 1.0) / N)
            yy = Qt + eta * Pt
            hard = pi - pi_old[p] + h * (pi - pstar) * dl
            r = np.array([erv - ev[p] + dl * beta * dFdP, ers - es[p] + dl * dFdQ, yy, hard])
            res_last[p] = r
            om = -(erv - ev0) / kap
            d2F_dPdev = 1.0 / (1.0 - N) * dEdP * (p0 * np.exp(om) / -kap)
            d2F_dPdpi = 1.0 / (1.0 - N) * dEdPi
            dFdev = (dEdP * Pt + eta) * (p0 * np.exp(om) / (-kap))
            dFdes = 3.0 * mu
            dFdpi = Pt * dEdPi
            if abs(N) < 1e-10:
                dPs_dP = np.exp(alpha_bar * psi_i / M)
            else:
                dPs_dP = np.power(1.0 - alpha_bar * psi_i * N / M, (N - 1.0) / N)
            dPsi_dpi = tl / pi
            if abs(N) < 1e-10:
                dPs_dpi = Pt * np.exp(alpha_bar * psi_i / M) * alpha_bar / M * dPsi_dpi
            else:
                dPs_dpi = Pt * (N - 1.0) / N * np.power(1.0 / (1.0 - alpha_bar * psi_i * N / M), 1.0 / N) \
                    * (-alpha_bar * N / M) * dPsi_dpi
            dHdev = -h * dl * dPs_dP * (p0 * np.exp(om) / (-kap))
            dHdpi = 1.0 + h * dl * (1.0 - dPs_dpi)
            Jm = np.array([[1.0 + dl * beta * d2F_dPdev, 0.0, beta * dFdP, dl * beta * d2F_dPdpi],
                           [0.0, 1.0, dFdQ, 0.0],
                           [dFdev, dFdes, 0.0, dFdpi],
                           [dHdev, 0.0, h * (pi - pstar), dHdpi]])
            xd = np.linalg.solve(Jm, r)
            nev = erv - xd[0]
            nes = ers - xd[1]
            dl = dl - xd[2]
            pi = pi - xd[3]
            e = nev / 3.0 + nes * nd[p] * np.sqrt(1.5)
            erv2 = e.sum()
            ers2 = _eps_s(e[None])[0]
            Pt = p0 * np.exp(-(erv2 - ev0) / kap)
            Qt = 3.0 * mu * ers2
        eps[p] = e
        pi_new[p] = pi
    return eps, pi_new, res_last


# ---------------------------------------------------------------- Neo-Hookean
def neo_hookean_cauchy(F, lam_l, mu):
    """sigma = mu/J (F F^T - I) + lambda lnJ / J I (mpm_utils.py:915; the 3D
    form P F^T / J at :1205-1208 is the same tensor)."""
    J = np.linalg.det(F)
    I = np.eye(3)[None]
    FFt = F @ np.transpose(F, (0, 2, 1))
    return mu / J[:, None, None] * (FFt - I) + (lam_l * np.log(J) / J)[:, None, None] * I, J

"""MPM particle/grid operators (oracle restatement, test infrastructure only).

Follows GeoWarp mpm/mpm_utils.py:
  init_particles_rectangle_2d/3d      :12-51
  initialization                       :53-74
  P2G_2d                               :83-136
  P2G_coupled_2d / P2G_coupled_3d      :138-257 / :259-383
  P2G_set_nodal_solution(_3d)          :387-414
  assemble_residual_2d_hencky / _J2    :416-596 / :598-780
  assemble_residual_coupled_2d_neohookean :782-1017
  assemble_residual_coupled_3d_neohookean :1020-1294
  set_diagnal_component_*              :1299-1341
  G2P_2d / G2P_coupled_2d / _3d        :1351-1590

Vectors are dof-major (dof = node + k*n_nodes), as in the reference.
Material names extend the reference set ('Hencky elasticity', 'J2',
'Neo-Hookean') with 'Drucker-Prager' (triaxial_DruckerPrager.py:308-412 lifted
into the residual) and J2 linear hardening.
"""
import numpy as np
import scipy.sparse as sp

from . import gimp as og
from . import materials as om


def init_particles_rectangle_2d(start_x, start_y, end_x, end_y, dx, n_grid_x, n_grid_y, PPD):
    """mpm_utils.py:12-29 (loop order i, j, p_x, p_y defines particle order)."""
    pts = []
    for i in range(n_grid_x):
        cx = (i + 0.5) * dx
        if not (start_x < cx < end_x):
            continue
        for j in range(n_grid_y):
            cy = (j + 0.5) * dx
            if not (start_y < cy < end_y):
                continue
            for a in range(PPD):
                for b in range(PPD):
                    pts.append((i * dx + (0.5 + a) * dx / PPD, j * dx + (0.5 + b) * dx / PPD))
    return np.array(pts, dtype=np.float64).reshape(-1, 2)


def init_particles_rectangle_3d(start_x, start_y, start_z, end_x, end_y, end_z, dx, nx, ny, nz, PPD):
    """mpm_utils.py:31-51 (loop order i, j, k, p_x, p_y, p_z)."""
    pts = []
    for i in range(nx):
        cx = (i + 0.5) * dx
        if not (start_x < cx < end_x):
            continue
        for j in range(ny):
            cy = (j + 0.5) * dx
            if not (start_y < cy < end_y):
                continue
            for k in range(nz):
                cz = (k + 0.5) * dx
                if not (start_z < cz < end_z):
                    continue
                for a in range(PPD):
                    for b in range(PPD):
                        for c in range(PPD):
                            pts.append((i * dx + (0.5 + a) * dx / PPD, j * dx + (0.5 + b) * dx / PPD,
                                        k * dx + (0.5 + c) * dx / PPD))
    return np.array(pts, dtype=np.float64).reshape(-1, 3)


def initial_F(n):
    """initialization (mpm_utils.py:53-74): F = b = diag(1+1e-7, 1-1e-6, 1)."""
    F = np.zeros((n, 3, 3))
    F[:, 0, 0] = 1.0 + 1e-7
    F[:, 1, 1] = 1.0 - 1e-6
    F[:, 2, 2] = 1.0
    return F


# ------------------------------------------------------------------ P2G
def p2g_2d_activate(st, n_nodes):
    """P2G_2d (mpm_utils.py:83-136): activate x/y dofs of the 2 or 3 nodes per
    axis of the support, without a weight threshold."""
    act = np.zeros(2 * n_nodes, dtype=bool)
    nodes = st.idx[st.support]
    act[nodes] = True
    act[nodes + n_nodes] = True
    return act


def p2g_coupled(st, x, n_nodes, dim, spacing_node, thresh, pbc_flag, boundary, p_mass, p_part):
    """P2G_coupled_2d/3d: activation with weight threshold, pressure Dirichlet
    above flagged particles, consistent mass matrix and rhs.

    2D: mpm_utils.py:209-257 (threshold 1e-6, node y = iy*dy, support-only loops)
    3D: mpm_utils.py:327-383 (threshold 1e-7, node z = iz*dz, all 27 nodes)
    ``spacing_node`` is the spacing used for the pressure-boundary node
    coordinate along the last axis (dy in 2D, dz in 3D).
    Returns (act, boundary, M (csr n_nodes^2), rhs).
    """
    nb = dim + 1
    act = np.zeros(nb * n_nodes, dtype=bool)
    boundary = boundary.copy()
    if dim == 2:
        cand = st.support  # the 2D loops run over i_range x j_range
    else:
        cand = st.valid  # 3D flattened loop over all 27 (in-grid) nodes
    on = cand & (st.W > thresh)
    nodes = st.idx[on]
    for k in range(nb):
        act[nodes + k * n_nodes] = True
    # pressure-boundary nodes (2D: y coordinate of node >= particle y; 3D: z)
    last = dim - 1
    lo_last = st.lo[last]
    off_last = st.offsets[:, last]
    node_coord = (lo_last[:, None] + off_last[None, :]).astype(np.float64) * spacing_node
    pb = cand & pbc_flag[:, None] & (node_coord >= x[:, last][:, None])
    boundary[st.idx[pb] + dim * n_nodes] = True
    # consistent mass: all 3^d x 3^d pairs (mpm_utils.py:238-257, :355-383)
    # pairs outside the support carry exact zeros and are skipped
    K = st.K
    pair = (st.support[:, :, None] & st.support[:, None, :]).ravel()
    rows = np.repeat(st.idx, K, axis=1).ravel()[pair]
    cols = np.tile(st.idx, (1, K)).ravel()[pair]
    vals = (st.W[:, :, None] * st.W[:, None, :] * p_mass).ravel()[pair]
    M = sp.coo_matrix((vals, (rows, cols)), shape=(n_nodes, n_nodes)).tocsr()
    rhs = np.zeros(n_nodes)
    contrib = (st.W[:, :, None] * st.W[:, None, :] * p_mass * p_part[:, None, None]).sum(axis=2)
    np.add.at(rhs, st.idx[st.support], contrib[st.support])
    return act, boundary, M, rhs


def solve_consistent_mass(M, rhs):
    """Solve M p = rhs on the nodes with a nonzero diagonal (the rows the
    reference's Krylov/AMG solve determines); empty rows give p = 0."""
    d = M.diagonal()
    live = np.nonzero(d > 0.0)[0]
    p = np.zeros(M.shape[0])
    if live.size:
        from scipy.sparse.linalg import spsolve
        Ms = M[live][:, live].tocsc()
        p[live] = spsolve(Ms, rhs[live])
    return p


def p2g_set_nodal_solution(boundary, act, n_nodes, dim, old, new, xp2g, require_active):
    """P2G_set_nodal_solution (mpm_utils.py:387-399, 2D: non-Dirichlet p dofs)
    and P2G_set_nodal_solution_3d (:401-414, also requires the dof active)."""
    old = old.copy()
    new = new.copy()
    dof = np.arange(n_nodes) + dim * n_nodes
    m = ~boundary[dof]
    if require_active:
        m &= act[dof]
    old[dof[m]] = xp2g[m]
    new[dof[m]] = xp2g[m]
    return old, new


# ------------------------------------------------------------------ residual
class MaterialParams:
    """Material parameters (constructor dict fields of the reference solvers,
    qs2d:83-90, qc3d:109-114, plus the north-star materials)."""

    def __init__(self, name, E, nu, kappa=0.0, H=0.0, phi=30.0, psi=0.0, cohesion=0.0, shape=0.0,
                 dp_apex='reference', norsand=None):
        self.name = name
        self.E = E
        self.nu = nu
        self.lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        self.mu = E / (2.0 * (1.0 + nu))
        self.kappa = kappa
        self.H = H
        self.phi = phi
        self.psi = psi
        self.cohesion = cohesion
        self.shape = shape
        self.dp_apex = dp_apex
        # NorSand (triaxial_NorSand.py:8-30 parameter names): kappa, tilde_lambda, M, N, beta, vc0, pc,
        # p0, v0, h, G0, ev0 (elastic volumetric strain at which P = p0; 0: the undeformed state)
        self.norsand = norsand_params(norsand) if name == 'NorSand' else None


def norsand_params(d):
    """NorSand parameters with the driver's defaults (triaxial_NorSand.py:8-30)."""
    out = dict(kappa=0.0015, tilde_lambda=0.02, M=1.27, N=0.4, beta=1.0, vc0=1.8911, pc=-715.0, p0=-390.0,
               v0=1.75, h=70.0, G0=4.5e4, ev0=0.0)
    out.update(d or {})
    out['mu'] = out['G0']
    out['pi0'] = out['pc'] * (1.0 - out['N']) ** ((1.0 - out['N']) / out['N'])
    return out


def norsand_tau(e, prm):
    """Principal Kirchhoff stress of NorSand's pressure-dependent hyperelasticity
    (triaxial_NorSand.py:455-461): p0 exp(-(tr e - ev0)/kappa) + 2 G0 (e_i - tr e / 3)."""
    tr = e.sum(axis=1, keepdims=True)
    return prm['p0'] * np.exp(-(tr - prm['ev0']) / prm['kappa']) + 2.0 * prm['mu'] * (e - tr / 3.0)


def constitutive(f3, F_old, b_old, mat, alpha_old):
    """Stress update shared by every residual kernel.

    Hencky family (mpm_utils.py:521-560): b_tr = f b_old f^T, eig, log strain,
    return map (plasticity.py / triaxial_DruckerPrager.py), tau spectral,
    b_new = sum exp(2 eps) n n^T.
    Neo-Hookean (mpm_utils.py:912-915): b_new = f b_old f^T, sigma closed form.
    Returns (F_new, J, sigma3, b_new, alpha_new).
    """
    F_new = f3 @ F_old
    J = np.linalg.det(F_new)
    b_tr = f3 @ b_old @ np.transpose(f3, (0, 2, 1))
    alpha_new = alpha_old.copy()
    if mat.name == 'Neo-Hookean':
        sig, _ = om.neo_hookean_cauchy(F_new, mat.lam, mat.mu)
        return F_new, J, sig, b_tr, alpha_new
    lam, vec = om.spectral(b_tr)
    eps = om.log_strain(lam)
    if mat.name == 'Hencky elasticity':
        er = eps
    elif mat.name == 'J2':
        er, alpha_new = om.return_mapping_J2(eps, mat.lam, mat.mu, mat.kappa, mat.H, alpha_old)
    elif mat.name == 'Drucker-Prager':
        er = om.return_mapping_DruckerPrager(eps, mat.lam, mat.mu, mat.phi, mat.psi, mat.cohesion,
                                             mat.shape, apex=mat.dp_apex)
    elif mat.name == 'NorSand':
        # alpha carries the image pressure pi; the specific volume follows the total volumetric
        # strain ln J (the driver's trace of the total strain, triaxial_NorSand.py:279)
        er, alpha_new, _ = om.return_mapping_NorSand(eps, np.log(J), alpha_old, mat.norsand)
        tau = om.assemble_spectral(norsand_tau(er, mat.norsand), vec)
        return F_new, J, tau / J[:, None, None], om.assemble_spectral(np.exp(2.0 * er), vec), alpha_new
    else:
        raise ValueError("Unimplemented material model: %s" % mat.name)
    tau_p = om.hencky_principal_tau(er, mat.lam, mat.mu)
    tau = om.assemble_spectral(tau_p, vec)
    b_new = om.assemble_spectral(np.exp(2.0 * er), vec)
    return F_new, J, tau / J[:, None, None], b_new, alpha_new


def _embed2(f2):
    n = f2.shape[0]
    f3 = np.zeros((n, 3, 3))
    f3[:, :2, :2] = f2
    f3[:, 2, 2] = 1.0
    return f3


def _free(boundary, act, dofs):
    return (~boundary[dofs]) & act[dofs]


def residual_2d(st, n_nodes, du, F_old, b_old, alpha_old, ext_flag, mat, p_vol, p_rho, gravity,
                point_load, step, total, boundary, act):
    """assemble_residual_2d_hencky / _J2 (mpm_utils.py:416-780); 'Neo-Hookean'
    and 'Drucker-Prager' use the same kinematics.  ``du`` is the dof-major
    displacement increment.  Returns (rhs, F_new, b_new, sigma, alpha_new)."""
    s1 = (step + 1.0) / total
    ux = du[st.idx]
    uy = du[st.idx + n_nodes]
    U = np.stack([ux, uy], axis=2)  # (N,K,2)
    H = np.einsum('nka,nkb->nab', st.G, U)  # delta_u_GRAD = sum outer(GRAD w, du)  (:492-507)
    f2 = np.eye(2)[None] + H
    A = np.linalg.inv(f2)
    F_new, J, sig, b_new, alpha_new = constitutive(_embed2(f2), F_old, b_old, mat, alpha_old)
    V = p_vol * J
    rho = p_rho / J
    g = np.array([0.0, -gravity * s1])
    fext = np.where(ext_flag[:, None], np.array(point_load) * s1, 0.0)
    wg = np.einsum('nab,nkb->nka', A, st.G)  # incr_F_inv @ GRAD w  (:579)
    val = (-np.einsum('nka,nac->nkc', wg, sig[:, :2, :2]) + st.W[:, :, None] * (rho[:, None, None] * g)) \
        * V[:, None, None] + st.W[:, :, None] * fext[:, None, :]
    rhs = np.zeros(2 * n_nodes)
    for c in range(2):
        dofs = st.idx + c * n_nodes
        m = _free(boundary, act, dofs) & st.support
        np.add.at(rhs, dofs[m], val[:, :, c][m])
    return rhs, F_new, b_new, sig, alpha_new


def residual_coupled_2d(st, n_nodes, old, new, F_old, b_old, alpha_old, ext_flag, trac_flag, mat, phi0,
                        mobility, dt, dx, PPD, p_vol, p_rho, gravity, traction, point_load, step, total,
                        boundary, act):
    """assemble_residual_coupled_2d_neohookean (mpm_utils.py:782-1017)."""
    s1 = (step + 1.0) / total
    idx = st.idx
    dU = np.stack([new[idx] - old[idx], new[idx + n_nodes] - old[idx + n_nodes]], axis=2)
    H = np.einsum('nka,nkb->nab', st.G, dU)
    F_old_J = np.linalg.det(F_old)
    f2 = np.eye(2)[None] + H
    A = np.linalg.inv(f2)
    F_new, J, sig, b_new, alpha_new = constitutive(_embed2(f2), F_old, b_old, mat, alpha_old)
    wg = np.einsum('nab,nkb->nka', A, st.G)
    pn = new[idx + 2 * n_nodes]
    po = old[idx + 2 * n_nodes]
    pp = (st.W * pn).sum(1)
    ppo = (st.W * po).sum(1)
    ppa = (st.Wa * pn).sum(1)
    ppoa = (st.Wa * po).sum(1)
    gp = np.einsum('nka,nk->na', wg, pn)  # spatial gradient in 2D (:933, :950)
    V = p_vol * J
    n_por = 1.0 - (1.0 - phi0) / J
    rho_mix = n_por * 1.0 + (1.0 - n_por) * p_rho
    tau = 0.5 / mat.mu
    g = np.array([0.0, -gravity * s1])
    fext = np.where(ext_flag[:, None], np.array(point_load) * s1, 0.0)
    ftr = np.where(trac_flag[:, None], np.array(traction) * dx / PPD, 0.0)
    s2 = sig[:, :2, :2] - pp[:, None, None] * np.eye(2)[None]
    val = (-np.einsum('nka,nac->nkc', wg, s2) + st.W[:, :, None] * (rho_mix[:, None, None] * g)) \
        * V[:, None, None] + st.W[:, :, None] * (fext + ftr)[:, None, :]
    mass = (st.W * (np.log(J) - np.log(F_old_J))[:, None]
            + np.einsum('nka,na->nk', wg, mobility * gp) * dt) * V[:, None]
    ppp = tau * (st.W - st.Wa) * ((pp - ppa) - (ppo - ppoa))[:, None] * V[:, None]
    rhs = np.zeros(3 * n_nodes)
    for c in range(2):
        dofs = idx + c * n_nodes
        m = _free(boundary, act, dofs) & st.support
        np.add.at(rhs, dofs[m], val[:, :, c][m])
    dofs = idx + 2 * n_nodes
    m = _free(boundary, act, dofs) & st.support
    np.add.at(rhs, dofs[m], mass[m])
    np.add.at(rhs, dofs[m], ppp[m])
    return rhs, F_new, b_new, sig, alpha_new


def residual_coupled_3d(st, n_nodes, x, old, new, F_old, b_old, alpha_old, mat, phi0, mobility, dt,
                        dxyz, PPD, p_vol, p_rho, gravity, footing, step, total, boundary, act):
    """assemble_residual_coupled_3d_neohookean (mpm_utils.py:1020-1294).

    Point loads and tractions are computed but never added in the reference's
    3D kernel (:1225-1232 vs :1271); the oracle omits them likewise.
    ``footing`` = (min_x, max_x, min_y, max_y, h0, incr, penalty_factor).
    Returns (rhs, F_new, b_new, sigma, alpha_new, indentation_force).
    """
    s1 = (step + 1.0) / total
    idx = st.idx
    nn = n_nodes
    dU = np.stack([new[idx + c * nn] - old[idx + c * nn] for c in range(3)], axis=2)
    H = np.einsum('nka,nkb->nab', st.G, dU)
    du_p = np.einsum('nk,nkc->nc', st.W, dU)
    pn = new[idx + 3 * nn]
    po = old[idx + 3 * nn]
    pp = (st.W * pn).sum(1)
    ppo = (st.W * po).sum(1)
    ppa = (st.Wa * pn).sum(1)
    ppoa = (st.Wa * po).sum(1)
    gp = np.einsum('nka,nk->na', st.G, pn)  # material gradient in 3D (:1169, :1194)
    Jo = np.linalg.det(F_old)
    f = np.eye(3)[None] + H
    A = np.linalg.inv(f)
    F_new, J, sig, b_new, alpha_new = constitutive(f, F_old, b_old, mat, alpha_old)
    V = p_vol * J
    n_por = 1.0 - (1.0 - phi0) / J
    rho_mix = n_por * 1.0 + (1.0 - n_por) * p_rho
    tau = 0.5 / mat.mu
    g = np.array([0.0, 0.0, -gravity * s1])
    dx, dy, dz = dxyz
    fminx, fmaxx, fminy, fmaxy, h0, incr, pf = footing
    pen_par = mat.E * pf * (dx / PPD) * (dy / PPD)
    hcur = h0 + (step + 1.0) * incr
    inside = (x[:, 2] >= hcur) & (x[:, 0] > fminx) & (x[:, 0] < fmaxx) & (x[:, 1] > fminy) & (x[:, 1] < fmaxy)
    pen_z = np.where(inside, -(hcur - (x[:, 2] + du_p[:, 2])) * pen_par, 0.0)
    force = -pen_z[inside].sum()
    wg = np.einsum('nab,nkb->nka', A, st.G)
    s3 = sig - pp[:, None, None] * np.eye(3)[None]
    val = (-np.einsum('nka,nac->nkc', wg, s3) + st.W[:, :, None] * (rho_mix[:, None, None] * g)) * V[:, None, None]
    mass = (st.W * (np.log(J) - np.log(Jo))[:, None]
            + np.einsum('nka,na->nk', wg, mobility * gp) * dt) * V[:, None]
    ppp = tau * (st.W - st.Wa) * ((pp - ppa) - (ppo - ppoa))[:, None] * V[:, None]
    rhs = np.zeros(4 * nn)
    cand = st.support  # zero-weight, zero-gradient nodes add exact zeros
    for c in range(3):
        dofs = idx + c * nn
        m = _free(boundary, act, dofs) & cand
        v = val[:, :, c]
        if c == 2:
            v = v + pen_z[:, None] * st.W
        np.add.at(rhs, dofs[m], v[m])
    dofs = idx + 3 * nn
    m = _free(boundary, act, dofs) & cand
    np.add.at(rhs, dofs[m], mass[m])
    np.add.at(rhs, dofs[m], ppp[m])
    return rhs, F_new, b_new, sig, alpha_new, force


def g2p(st, x, du_nodes, n_nodes, dim, new_p=None):
    """G2P_2d / G2P_coupled_2d / G2P_coupled_3d (mpm_utils.py:1351-1590):
    x += sum w du, p_p = sum w p_new.  (F_old <- F_new, b_old <- b_new is done
    by the caller.)"""
    x = x.copy()
    for c in range(dim):
        x[:, c] += (st.W * du_nodes[st.idx + c * n_nodes]).sum(1)
    pp = None
    if new_p is not None:
        pp = (st.W * new_p[st.idx]).sum(1)
    return x, pp

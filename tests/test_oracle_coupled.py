"""CPU pins of the oracle's coupled (u-p) functions: consistent-mass projection,
Darcy and PPP terms of the mass balance, the undrained response, Terzaghi
consolidation, and the 3D penalty indenter.

Each check rests on something the oracle does not restate: partition of unity
and linear reproduction of the cpGIMP weights, the divergence theorem for the
Darcy term of a linear pressure field, static equilibrium of a loaded column,
Terzaghi's series solution, and the penalty force written out from the particle
positions.  A plausible slip (a sign, a transposed pair, a factor on the
mobility, a wrong weight in the mass matrix) fails at least one of them; the
last test plants such a slip in the oracle to show the consolidation check
catches it.
"""
import numpy as np
import pytest

import harness as H
import oracle.gimp as og
import oracle.mpm as omp
import oracle.solvers as osv
from geowarp_amd import problems as P


def _terzaghi_mini(ncy=16, total=10, dt=1e-5, n_iter=6):
    cfg = P.terzaghi(n_cells_x=2, n_cells_y=ncy, width=2.0 / ncy, total_steps=total, dt=dt)
    cfg['iteration_dict']['n_iter'] = n_iter
    return cfg


def _indent_mini():
    return P.indent3d(n_cells=(4, 4, 4), dx=0.25, total_steps=10, depth=0.05, footing=0.5, material='Neo-Hookean',
                      tol=1e-10)


# ------------------------------------------------------------------ consistent mass (mpm_utils.py:138-414)
@pytest.mark.parametrize('dim', [2, 3])
def test_consistent_mass_rows_and_linear_pressure_recovery(dim):
    """Row sums of M are the lumped masses sum_p w_ip m_p (sum_j w_jp = 1), and
    the projection M p = rhs returns any linear particle pressure field exactly
    at every node it determines (sum_j w_jp x_j = x_p)."""
    cfg = _terzaghi_mini() if dim == 2 else _indent_mini()
    s = H.make(osv, cfg)
    spacing = (s.dx,) if dim == 2 else (s.dx, s.dy, s.dz)
    st = og.Stencil(s.x_particles, s.lp, dim, spacing, s.n_grid)
    rng = np.random.default_rng(7)
    a, b = 3.0, rng.normal(size=dim)
    p_lin = a + s.x_particles @ b
    thresh = 1e-6 if dim == 2 else 1e-7
    sp_last = s.dy if dim == 2 else s.dz
    none = np.zeros(s.n_particles, dtype=bool)
    act, _, M, rhs = omp.p2g_coupled(st, s.x_particles, s.n_nodes, dim, sp_last, thresh, none, s.boundary.copy(),
                                     s.p_mass, p_lin)
    lumped = np.zeros(s.n_nodes)
    np.add.at(lumped, st.idx[st.support], (st.W * s.p_mass)[st.support])
    assert np.allclose(np.asarray(M.sum(axis=1)).ravel(), lumped, rtol=1e-13, atol=1e-18)
    assert abs(M - M.T).max() == 0.0
    p_nodes = omp.solve_consistent_mass(M, rhs)
    live = M.diagonal() > 0
    coords = osv.node_coords(s.n_grid, dim, s.n_nodes) * np.array(spacing * (dim if dim == 2 else 1))[:dim]
    expect = a + coords @ b
    assert live.sum() > 20
    assert np.allclose(p_nodes[live], expect[live], rtol=1e-9, atol=1e-9)
    assert np.all(p_nodes[~live] == 0.0)


# ------------------------------------------------------------------ mass balance (mpm_utils.py:958-1015)
def _pressure_rows(s, st, f, u):
    """Unmasked residual (no Dirichlet, every dof active) at nodal solution u."""
    s.boundary[:] = False
    s.act[:] = True
    return f(u)


def test_pressure_block_is_symmetric_psd_with_constants_in_its_null_space():
    """dR_p/dp at fixed displacement is the Darcy matrix dt sum_p V k grad w_i .
    grad w_j plus the PPP matrix tau sum_p V (w - wbar)_i (w - wbar)_j: both
    symmetric positive semidefinite, both zero on a constant pressure."""
    cfg = _terzaghi_mini()
    s = H.make(osv, cfg)
    st, f, u0, _ = s.prepare_step(0, cfg['total_steps'])
    n = s.n_nodes
    r0 = _pressure_rows(s, st, f, u0)
    pd = np.arange(n) + 2 * n
    # the pressure rows are affine in the nodal pressures: unit differences are exact
    K = np.zeros((n, n))
    for j in range(n):
        u = u0.copy()
        u[pd[j]] += 1.0
        K[:, j] = _pressure_rows(s, st, f, u)[pd] - r0[pd]
    touched = np.abs(K).sum(1) > 0
    Kt = K[np.ix_(touched, touched)]
    scale = np.abs(Kt).max()
    assert touched.sum() > 30
    assert np.abs(Kt - Kt.T).max() < 1e-9 * scale
    ev = np.linalg.eigvalsh(0.5 * (Kt + Kt.T))
    assert ev.min() > -1e-9 * scale
    assert ev.max() > 0
    assert np.abs(Kt.sum(1)).max() < 1e-9 * scale


def _hat_integral(a, b, center, h):
    """Exact integral over [a, b] of the linear hat of half-width h at `center`."""
    pts = np.unique(np.clip(np.array([a, b, center - h, center, center + h]), a, b))
    vals = np.maximum(0.0, 1.0 - np.abs(pts - center) / h)
    return np.trapezoid(vals, pts)


def test_darcy_term_of_linear_pressure_is_the_boundary_flux():
    """For p = gamma y the Darcy part of the nodal mass balance is
    dt k gamma J int_Omega d_y N_i dV = dt k gamma J [int N_i(x, y_top) dx -
    int N_i(x, y_bot) dx] (cpGIMP gradients are domain averages of the hat
    gradients, so the particle sum is the integral over the particle domains):
    zero at interior nodes and the top/bottom boundary flux at the others."""
    cfg = _terzaghi_mini()
    s = H.make(osv, cfg)
    st, f, u0, _ = s.prepare_step(0, cfg['total_steps'])
    n, h = s.n_nodes, s.dx
    gamma, mob = 7.0, 1e-4
    coords = osv.node_coords(s.n_grid, 2, n) * h
    u = u0.copy()
    u[2 * n:] = gamma * coords[:, 1]
    s.old_solution[:] = 0.0
    s.mobility_constant = mob
    r1 = _pressure_rows(s, st, f, u)
    s.mobility_constant = 0.0
    r0 = _pressure_rows(s, st, f, u)
    darcy = (r1 - r0)[2 * n:]
    x0, x1 = s.start_x, s.end_x
    ybot = (s.x_particles[:, 1] - s.lp[:, 1]).min()
    ytop = (s.x_particles[:, 1] + s.lp[:, 1]).max()
    J = np.linalg.det(s.F_old[0])
    expect = np.zeros(n)
    for i in range(n):
        X, Y = coords[i]
        ix = _hat_integral(x0, x1, X, h)
        expect[i] = ix * (max(0.0, 1 - abs(ytop - Y) / h) - max(0.0, 1 - abs(ybot - Y) / h))
    expect *= s.dt * mob * gamma * J
    scale = np.abs(expect).max()
    assert scale > 0
    # particle domains are 2e-6 narrower than their cells (lp0 = dx/4 - 1e-6, qc2d:81)
    assert np.abs(darcy - expect).max() < 2e-4 * scale
    interior = (coords[:, 1] > ybot + h) & (coords[:, 1] < ytop - h)
    assert np.abs(darcy[interior]).max() < 2e-4 * scale
    assert abs(darcy.sum()) < 1e-12 * scale


def test_undrained_column_carries_the_load_in_pore_pressure():
    """A laterally confined column with incompressible constituents (Biot 1)
    loaded by 10 kPa in a very short step cannot change volume: below the
    loaded boundary layer the pore pressure carries the load (p ~ 10 kPa) and
    the total vertical stress sigma'_yy - p equals the applied -10 kPa."""
    cfg = _terzaghi_mini(dt=1e-7)
    s = H.make(osv, cfg)
    s.advance_one_step(0, cfg['total_steps'])
    y = s.x_particles[:, 1]
    deep = y < 0.5 * y.max()
    assert deep.sum() > 30
    total = s.sigma[deep, 1, 1] - s.pressure[deep]
    assert np.abs(total + 10.0).max() < 1e-3
    assert np.abs(s.pressure[deep] - 10.0).max() < 0.05
    assert np.abs(s.sigma[deep, 1, 1]).max() < 0.05


def _terzaghi_error(mobility_factor=1.0, ncy=16, nsteps=20, Tv_end=0.2):
    """Relative L2 error of the oracle's pore-pressure profile against
    Terzaghi's series at T_v = Tv_end (drained top, impermeable base; step 0
    applies the load undrained, then nsteps equal consolidation steps)."""
    cfg = _terzaghi_mini(ncy=ncy, total=nsteps + 1)
    cfg['porous_media_dict']['mobility_constant'] *= mobility_factor
    s = H.make(osv, cfg)
    Hh = 1.0
    E, nu, mob = 1.0e4, 0.3, 1e-4
    cv = mob * E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    t_end = Tv_end * Hh ** 2 / cv
    dts = np.concatenate([[1e-5], np.full(nsteps, t_end / nsteps)])
    depth = Hh - s.x_particles[:, 1]
    for k in range(nsteps + 1):
        s.dt = dts[k]
        s.advance_one_step(k, nsteps + 1)
        if k == 0:
            p0 = s.pressure.mean()
    M = np.pi * (2 * np.arange(400) + 1) / 2
    series = (2 / M[None] * np.sin(M[None] * depth[:, None] / Hh) * np.exp(-M[None] ** 2 * Tv_end)).sum(1) * p0
    return np.linalg.norm(s.pressure - series) / np.linalg.norm(series)


def test_terzaghi_consolidation_matches_series():
    """16-cell column, 20 backward-Euler steps to T_v = 0.2: within 6 % of the
    series (3.6 % measured; the bound covers space and time discretisation)."""
    assert _terzaghi_error() < 0.06


def test_terzaghi_check_detects_a_mobility_slip():
    """The same check with the mobility planted wrong by a factor of two (in the
    oracle's input, on the CPU) misses the bound by far (33-36 % measured)."""
    assert _terzaghi_error(mobility_factor=2.0) > 0.2
    assert _terzaghi_error(mobility_factor=0.5) > 0.2


# ------------------------------------------------------------------ 3D penalty indenter (mpm_utils.py:1234-1248)
def test_penalty_indenter_force_closed_form():
    """Under a uniform z-increment delta the indentation force is
    sum over particles inside the footing box with z >= h of
    (h - z_p - delta) E pf (dx/PPD)(dy/PPD), and the nodal z-residual changes by
    minus that force in total (partition of unity)."""
    cfg = _indent_mini()
    s = H.make(osv, cfg)
    step, total = 0, cfg['total_steps']
    st, f, u0, _ = s.prepare_step(step, total)
    n = s.n_nodes
    delta = -0.004
    u = u0.copy()
    u[2 * n:3 * n] += delta
    fminx, fmaxx, fminy, fmaxy, h0, incr, pf = s.footing
    h = h0 + (step + 1) * incr
    x = s.x_particles
    inside = (x[:, 2] >= h) & (x[:, 0] > fminx) & (x[:, 0] < fmaxx) & (x[:, 1] > fminy) & (x[:, 1] < fmaxy)
    assert inside.sum() >= 4
    E = cfg['material_dict']['youngs_modulus']
    expect = ((h - x[inside, 2] - delta) * E * pf * (s.dx / s.PPD) * (s.dy / s.PPD)).sum()
    s.boundary[:] = False
    s.act[:] = True
    r_on = f(u, commit=True)
    assert abs(s.indentation_force - expect) < 1e-12 * abs(expect)
    saved = s.footing
    s.footing = saved[:6] + (0.0,)
    r_off = f(u)
    s.footing = saved
    dz = (r_on - r_off)[2 * n:3 * n]
    assert abs(dz.sum() + expect) < 1e-10 * abs(expect)
    other = np.concatenate([(r_on - r_off)[:2 * n], (r_on - r_off)[3 * n:]])
    assert np.abs(other).max() == 0.0

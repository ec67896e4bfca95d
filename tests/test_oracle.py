"""CPU pins of the oracle (no GPU).

Each test checks the oracle against something that does not restate its own
formula: closed forms, brute-force quadrature, invariants, a known answer
printed by a reference driver, or an independent method.
"""
import os

import numpy as np
import pytest

import harness as H
import oracle.gimp as og
import oracle.materials as om
import oracle.mpm as omp
import oracle.solvers as osv


# ------------------------------------------------------------------ GIMP
def _gimp_quadrature(x, lp, h, node_x, n=200001):
    """GIMP weight = average of the linear hat of a node over the particle
    domain [x-lp, x+lp] (Bardenhagen & Kober 2004); brute-force trapezoid."""
    s = np.linspace(x - lp, x + lp, n)
    hat = np.maximum(0.0, 1.0 - np.abs(s - node_x) / h)
    return np.trapezoid(hat, s) / (2.0 * lp)


@pytest.mark.parametrize('x,lp', [(0.37, 0.11), (0.52, 0.2), (0.25, 0.25), (0.9, 0.05), (0.13, 0.1249)])
def test_gimp_weights_equal_domain_average_of_linear_hat(x, lp):
    h = 0.5
    lo, w, g, wa, ext = og.axis_weights(np.array([x]), np.array([lp]), h, 1.0 / h)
    for k in range(3):
        node_x = (lo[0] + k) * h
        ref = _gimp_quadrature(x, lp, h, node_x)
        assert abs(w[0, k] - ref) < 1e-8, (k, w[0, k], ref)


def test_gimp_gradient_matches_fd_and_partition_of_unity():
    rng = np.random.default_rng(1)
    h = 0.1
    x = rng.uniform(0.3, 0.7, 500)
    lp = rng.uniform(0.01, 0.049, 500)
    lo, w, g, wa, ext = og.axis_weights(x, lp, h, 1.0 / h)
    assert np.allclose(w.sum(1), 1.0, atol=1e-14)
    assert np.allclose(g.sum(1), 0.0, atol=1e-10)
    # linear reproduction: sum w_k x_k = x
    xs = (lo[:, None] + np.arange(3)[None]) * h
    assert np.allclose((w * xs).sum(1), x, atol=1e-14)
    # FD of the weights where the branch does not change
    e = 1e-7
    lo2, w2, _, _, ext2 = og.axis_weights(x + e, lp, h, 1.0 / h)
    lo1, w1, _, _, ext1 = og.axis_weights(x - e, lp, h, 1.0 / h)
    same = (lo1 == lo2) & (lo1 == lo) & (ext1 == ext2) & (ext1 == ext)
    fd = (w2 - w1) / (2 * e)
    assert same.sum() > 400
    assert np.allclose(fd[same], g[same], atol=1e-5)


def test_gimp_average_weights_sum_to_one():
    rng = np.random.default_rng(2)
    h = 0.25
    x = rng.uniform(1.0, 2.0, 300)
    lp = rng.uniform(0.01, 0.12, 300)
    _, _, _, wa, _ = og.axis_weights(x, lp, h, 1.0 / h)
    assert np.allclose(wa.sum(1), 1.0, atol=1e-14)


def test_lp_update_clips_to_domain():
    F = np.tile(np.eye(3), (3, 1, 1))
    F[:, 0, 0] = 2.0
    x = np.array([[0.101, 0.5], [0.5, 0.5], [0.899, 0.5]])
    lp = og.update_lp(F, x, 0.01, (0.1, 0.0), (0.9, 1.0))
    assert np.isclose(lp[1, 0], 0.02)
    assert np.isclose(lp[0, 0], 0.001 - 1e-6)
    assert np.isclose(lp[2, 0], 0.001 - 1e-6)


# ------------------------------------------------------------------ materials
def test_j2_return_is_on_the_yield_surface_and_matches_closed_form():
    rng = np.random.default_rng(3)
    lam, mu, kappa, Hh = 5000.0, 3000.0, 2.0, 100.0
    e = rng.normal(scale=2e-3, size=(200, 3))
    alpha = rng.uniform(0, 1e-3, 200)
    er, an = om.return_mapping_J2(e, lam, mu, kappa, Hh, alpha)
    tau = om.hencky_principal_tau(er, lam, mu)
    s = tau - tau.mean(1, keepdims=True)
    sn = np.linalg.norm(s, axis=1)
    kap_new = kappa + np.sqrt(2 / 3) * Hh * an
    yielded = an > alpha
    assert yielded.sum() > 50
    assert np.allclose(sn[yielded], kap_new[yielded], rtol=1e-12)
    # volumetric strain untouched, deviatoric direction preserved
    assert np.allclose(er.sum(1), e.sum(1), atol=1e-15)
    st = om.hencky_principal_tau(e, lam, mu)
    st = st - st.mean(1, keepdims=True)
    cosang = (s * st).sum(1) / (np.linalg.norm(s, axis=1) * np.linalg.norm(st, axis=1))
    assert np.allclose(cosang[yielded], 1.0, atol=1e-12)


def test_drucker_prager_local_newton_matches_closed_form_cone_return():
    rng = np.random.default_rng(4)
    lam, mu = 11538.0, 7692.0
    phi, psi, c = 30.0, 10.0, 5.0
    e = np.column_stack([rng.uniform(-6e-3, -3e-3, 300), rng.uniform(0.0, 1.5e-3, 300), rng.uniform(0.0, 1.5e-3, 300)])
    _, P, S, Q = om._dp_PQ(e, lam, mu)
    f = om.dp_yield(P, Q, phi, c, 0.0)
    Af, Bf = om._dp_AB(phi, c)
    _, Bg = om._dp_AB(psi, c)
    K = lam + 2 * mu / 3
    dl = f / (2 * mu + Bf * K * Bg)
    cone = (f > 1e-10) & (P < 0) & (2 * mu * dl < np.linalg.norm(S, axis=1))
    assert cone.sum() > 50
    er = om.return_mapping_DruckerPrager(e, lam, mu, phi, psi, c)
    ref = om.dp_cone_closed_form(e, lam, mu, phi, psi, c)
    assert np.allclose(er[cone], ref[cone], rtol=1e-10, atol=1e-14)
    _, P2, S2, Q2 = om._dp_PQ(er[cone], lam, mu)
    assert np.abs(om.dp_yield(P2, Q2, phi, c, 0.0)).max() < 1e-8
    el = f < 1e-10
    assert np.array_equal(er[el], e[el])


def test_norsand_converged_state_satisfies_local_equations():
    prm = dict(kappa=0.0015, tilde_lambda=0.02, M=1.27, N=0.4, beta=1.0, vc0=1.8911, p0=-390.0, v0=1.75, h=70.0,
               mu=4.5e4, ev0=0.0)
    pc = -715.0
    pi0 = pc * (1 - 0.4) ** ((1 - 0.4) / 0.4)
    e = np.array([[-1e-4, -1e-4, -8e-3], [-1e-4, -1.5e-4, -6e-3]])
    er, pin, res = om.return_mapping_NorSand(e, e.sum(1), np.full(2, pi0), prm)
    # the local residual of the last iterate (before the final update) is round-off:
    # the fixed 15-iteration Newton (triaxial_NorSand.py:314) has converged here
    assert np.abs(res[:, :3]).max() < 1e-8
    assert np.all(pin != pi0)


def test_neo_hookean_closed_form_uniaxial():
    lam, mu = 1000.0, 500.0
    F = np.tile(np.eye(3), (1, 1, 1))
    F[0, 0, 0] = 1.01
    sig, J = om.neo_hookean_cauchy(F, lam, mu)
    ref = mu / 1.01 * (1.01 ** 2 - 1) + lam * np.log(1.01) / 1.01
    assert np.isclose(sig[0, 0, 0], ref)
    assert np.isclose(sig[0, 1, 1], lam * np.log(1.01) / 1.01)


def test_spectral_hencky_is_isotropic_function():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(50, 3, 3)) * 0.05
    b = np.eye(3)[None] + A @ np.transpose(A, (0, 2, 1))
    lam, vec = om.spectral(b)
    tau = om.assemble_spectral(om.hencky_principal_tau(om.log_strain(lam), 1000.0, 500.0), vec)
    # rotate b -> R b R^T gives R tau R^T
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    lam2, vec2 = om.spectral(R @ b @ R.T)
    tau2 = om.assemble_spectral(om.hencky_principal_tau(om.log_strain(lam2), 1000.0, 500.0), vec2)
    assert np.allclose(tau2, R @ tau @ R.T, atol=1e-10)


# ------------------------------------------------------------------ residual / solver
def _mini(name):
    cfg = H.MINI[name]()
    return cfg, H.make(osv, cfg)


def test_rigid_translation_gives_no_internal_force():
    cfg, s = _mini('bar_hencky')
    s.gravity_mag = 0.0
    st, f, u0, steps = s.prepare_step(0, cfg['total_steps'])
    s.boundary[:] = False
    s.act[:] = True
    u = np.zeros_like(u0)
    u[:s.n_nodes] = 1e-3  # uniform translation
    u[s.n_nodes:] = -2e-3
    r0 = f(np.zeros_like(u0))
    r = f(u)
    # a rigid translation leaves the internal force unchanged (only the tiny
    # stress of the perturbed initial F, mpm_utils.py:64-69, remains)
    assert np.abs(r - r0).max() < 1e-15
    assert np.abs(r0).max() < 1e-4


def test_gravity_column_known_answer():
    """bar_compaction_cpGIMP_J2.py:126: sigma_yy = -rho g (H - (y0 - y_start));
    at a converged step the volume-weighted relative L1 error is a few percent
    or less for this resolution (boundary layers at top and base)."""
    cfg = H.MINI['bar_nh']()
    s = H.make(osv, cfg)
    s.advance_one_step(cfg['total_steps'] - 1, cfg['total_steps'])
    g = cfg['background_grid_dict']
    Hh = g['end_y'] - g['start_y']
    x0 = omp.init_particles_rectangle_2d(g['start_x'], g['start_y'], g['end_x'], g['end_y'], g['dx'],
                                         g['n_grid_x'], g['n_grid_y'], 2)
    ana = -cfg['particles_dict']['p_rho'] * 10.0 * (Hh - (x0[:, 1] - g['start_y']))
    err = np.abs(s.sigma[:, 1, 1] - ana).sum() / np.abs(ana).sum()
    assert err < 0.05


def test_global_force_balance_free_dofs():
    """With no constraints the nodal residual sums to the total external load."""
    cfg, s = _mini('bar_nh')
    st, f, u0, steps = s.prepare_step(0, cfg['total_steps'])
    s.boundary[:] = False
    r = f(np.zeros_like(u0))
    total_y = r[s.n_nodes:].sum()
    W = -s.n_particles * s.p_vol * s.p_rho * 10.0 / cfg['total_steps']
    assert abs(total_y - W) < 1e-12 * abs(W) + 1e-14


def test_fd_jacobian_solves_linearised_system():
    """The FD Jacobian's Newton step reduces the residual quadratically (checked
    on the first two iterations of a golden workload)."""
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'bar_nh.npz'))
    h = g['s0_newton']
    assert h[1] / h[0] < 1e-4 and h[2] / h[1] < 1e-3


def test_golden_fixtures_load():
    gdir = os.path.join(os.path.dirname(__file__), 'golden')
    for name in ['bar_nh', 'bar_hencky', 'bar_j2', 'bar_dp', 'terzaghi', 'indent_nh']:
        g = np.load(os.path.join(gdir, name + '.npz'))
        assert 'jac_val' in g and 's0_x' in g


# ------------------------------------------------------------------ Drucker-Prager apex / tip (dilation 0)
def _dp_psi0_cases(shape):
    """Principal trial strains around the tension region of phi 30, c 5, psi 0."""
    lam, mu = 11538.46, 7692.31
    K = lam + 2 * mu / 3
    A, B = om._dp_AB(30.0, 5.0)
    p_tip = A * (1 - shape) / B
    rng = np.random.default_rng(9)
    P = rng.uniform(-3 * p_tip, 2 * p_tip, 400)
    d = rng.normal(size=(400, 3))
    d -= d.mean(1, keepdims=True)
    d *= (rng.uniform(1e-4, 3e-3, 400) / np.linalg.norm(d, axis=1))[:, None]
    e = P[:, None] / (3 * K) + d  # volumetric part gives P (Hencky: P = K tr e)
    return e, lam, mu, K, A, B, p_tip


@pytest.mark.parametrize('shape', [0.0, 0.1])
def test_dp_dilation_zero_return_keeps_pressure_or_cuts_off_at_the_tip(shape):
    """With dilation 0 the flow is deviatoric: a yielding trial state below the
    tip keeps its pressure P and its deviator direction, and lands on the surface
    at Q = sqrt(1.5 ((A - B P)^2 - a^2 A^2)) (closed form of the yield function);
    a state at or beyond the tip (P >= A (1 - a)/B) is cut off to the tip, Q = 0.
    Tensile states below the tip are returned to the surface, not to the apex."""
    e, lam, mu, K, A, B, p_tip = _dp_psi0_cases(shape)
    _, P, S, Q = om._dp_PQ(e, lam, mu)
    y = om.dp_yield(P, Q, 30.0, 5.0, shape)
    er = om.return_mapping_DruckerPrager(e, lam, mu, 30.0, 0.0, 5.0, shape, apex='return')
    _, P2, S2, Q2 = om._dp_PQ(er, lam, mu)
    plastic = y >= 1e-10
    tip = plastic & (P >= p_tip) if shape > 0 else plastic & (y >= np.linalg.norm(S, axis=1))
    cone = plastic & ~tip
    assert tip.sum() > 20 and cone.sum() > 50 and (cone & (P > 0)).sum() > 10
    assert np.allclose(P2[tip], p_tip, rtol=1e-12) and np.all(np.abs(Q2[tip]) < 1e-9)
    assert np.allclose(P2[cone], P[cone], rtol=1e-9, atol=1e-9)
    q_ref = np.sqrt(1.5 * ((A - B * P[cone]) ** 2 - (shape * A) ** 2))
    assert np.allclose(Q2[cone], q_ref, rtol=1e-8)
    cosang = (S2[cone] * S[cone]).sum(1) / (np.linalg.norm(S2[cone], axis=1) * np.linalg.norm(S[cone], axis=1))
    assert np.allclose(cosang, 1.0, atol=1e-10)
    assert np.array_equal(er[~plastic], e[~plastic])

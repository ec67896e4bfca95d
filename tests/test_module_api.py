"""The reference-named module entry points (geowarp_amd.mpm.gimp,
geowarp_amd.mpm.mpm_utils, geowarp_amd.materials.plasticity,
geowarp_amd.solvers.sparse_differentiation).

CPU tests pin the pure data-format helpers (node picking, seed vectors, COO
scatter) against brute force.  GPU tests call the device entry points and
compare them with the oracle on the same seeded inputs: bit-level agreement is
not expected for floating point, the bars are written in each test (FP64,
1e-12 relative unless the quantity is a finite-difference estimate).
"""
import itertools

import numpy as np
import pytest
import torch

import harness as H
import oracle.gimp as og
import oracle.materials as om
import oracle.mpm as omp
from geowarp_amd.solvers import sparse_differentiation as sd


# ------------------------------------------------------------------ CPU: data formats
def _brute_pick(n_grid, first):
    dims = len(n_grid)
    strides = [1, n_grid[0] + 1, (n_grid[0] + 1) * (n_grid[1] + 1)][:dims]
    start = []
    rem = first
    for a in range(dims - 1, -1, -1):
        start.insert(0, rem // strides[a])
        rem %= strides[a]
    out = []
    ranges = [range(n_grid[a] + 1) for a in range(dims)]
    for c in itertools.product(*reversed(ranges)):  # slowest axis first
        c = c[::-1]
        if all(c[a] >= start[a] and (c[a] - start[a]) % 5 == 0 for a in range(dims)):
            out.append(sum(c[a] * strides[a] for a in range(dims)))
    return np.array(out)


@pytest.mark.parametrize('n_grid', [(7, 11), (4, 4), (12, 5)])
def test_pick_grid_nodes_2d_brute_force(n_grid):
    for first in [0, 1, 3, n_grid[0] + 2, 4 * (n_grid[0] + 1) + 4]:
        np.testing.assert_array_equal(sd.pick_grid_nodes_2d(*n_grid, first), _brute_pick(n_grid, first))


def test_pick_grid_nodes_3d_brute_force():
    n_grid = (6, 8, 7)
    sxy = 7 * 9
    for first in [0, 2, 7 + 3, 2 * sxy + 4 * 7 + 1, 4 * sxy + 4 * 7 + 4]:
        np.testing.assert_array_equal(sd.pick_grid_nodes_3d(*n_grid, first), _brute_pick(n_grid, first))


def test_seed_vectors_partition_the_dofs():
    nx, ny = 9, 6
    n_nodes = (nx + 1) * (ny + 1)
    n = 3 * n_nodes
    sx, sy, sp, ex, ey, ep = sd.precompute_seed_vectors_coupled_2d(nx, ny, n_nodes, n, 40, device='cpu')
    assert len(sx) == 25
    cover = sum(ex) + sum(ey) + sum(ep)
    assert torch.equal(cover, torch.ones(n, dtype=torch.float64))
    for s, e in zip(sy, ey):
        k = int((e > 0).sum())
        assert torch.equal(torch.sort(s[:k]).values, torch.nonzero(e).flatten().to(torch.int32))
        assert torch.all(s[k:] == 0)


def _banded_matrix(n_grid, nb, rng):
    dims = len(n_grid)
    n_nodes = int(np.prod([g + 1 for g in n_grid]))
    coords = np.stack(np.unravel_index(np.arange(n_nodes), [g + 1 for g in n_grid][::-1]), 1)[:, ::-1]
    K = np.zeros((nb * n_nodes, nb * n_nodes))
    for a in range(n_nodes):
        near = np.all(np.abs(coords - coords[a]) <= 2, axis=1)
        for b in np.nonzero(near)[0]:
            K[np.ix_(a + n_nodes * np.arange(nb), b + n_nodes * np.arange(nb))] = rng.normal(size=(nb, nb))
    return K, n_nodes


@pytest.mark.parametrize('n_grid,nb', [((6, 7), 2), ((5, 6), 3), ((5, 4, 6), 4)])
def test_coo_scatter_rebuilds_jacobian_from_row_sweeps(n_grid, nb):
    """Rows picked five nodes apart never share a column within the +-2
    window, so e^T(-K) for each seed holds each selected row exactly; the
    scatter must rebuild K on free rows and columns, with no two triplets on
    one slot."""
    rng = np.random.default_rng(5)
    K, n_nodes = _banded_matrix(n_grid, nb, rng)
    n = K.shape[0]
    boundary = torch.from_numpy(rng.random(n) < 0.15)
    activate = torch.from_numpy(rng.random(n) < 0.9)
    W = 5 ** len(n_grid)
    rows = torch.full((nb * W * n,), -1, dtype=torch.int32)
    cols = torch.full_like(rows, -1)
    vals = torch.zeros(nb * W * n, dtype=torch.float64)
    if len(n_grid) == 2:
        seeds = sd.precompute_seed_vectors_coupled_2d(*n_grid, n_nodes, n, n, device='cpu') if nb == 3 else \
            sd.precompute_seed_vectors_2d(*n_grid, n_nodes, n, n, device='cpu')
    else:
        seeds = sd.precompute_seed_vectors_coupled_3d(*n_grid, n_nodes, n, n, device='cpu')
    sels, es = seeds[:nb], seeds[nb:]
    fn = {2: sd.from_jacobian_to_vector_parallel_2d, 3: sd.from_jacobian_to_vector_parallel_coupled_2d,
          4: sd.from_jacobian_to_vector_parallel_coupled_3d}[nb]
    written = 0
    for k in range(nb):
        for s, e in zip(sels[k], es[k]):
            jac = torch.from_numpy(-(e.numpy() @ K))  # e^T dR/du with K = -dR/du
            fn(jac, rows, cols, vals, *n_grid, n_nodes, n, s, boundary, activate)
    m = rows >= 0
    written = int(m.sum())
    A = np.zeros_like(K)
    np.add.at(A, (rows[m].numpy(), cols[m].numpy()), vals[m].numpy())
    free = (~boundary.numpy()) & activate.numpy()
    ref = K * free[:, None] * free[None, :]
    np.testing.assert_array_equal(A, ref)
    assert written == int(np.count_nonzero(ref))


# ------------------------------------------------------------------ GPU: device entry points
def _dev():
    return torch.device('cuda')


@pytest.mark.gpu
@pytest.mark.parametrize('dim', [2, 3])
def test_gimp_weights_match_oracle(dim):
    from geowarp_amd.mpm import gimp
    rng = np.random.default_rng(11 + dim)
    hs = (0.1, 0.1, 0.1) if dim == 2 else (0.1, 0.125, 0.08)
    n = 4000
    x = rng.uniform(0.5, 1.5, size=(n, dim))
    lp = rng.uniform(0.005, 0.06, size=(n, dim))
    lp[: n // 4] = 0.1 / 4 - 1e-6  # the workloads' initial half-spacing
    xt, lt = torch.from_numpy(x).to(_dev()), torch.from_numpy(lp).to(_dev())
    if dim == 2:
        w = gimp.get_GIMP_shape_function_and_gradient_2d(xt, hs[0], 1 / hs[0], lt[:, 0], lt[:, 1]).cpu().numpy()
        wa = gimp.get_GIMP_shape_function_and_gradient_avg_2d(xt, hs[0], 1 / hs[0], lt[:, 0],
                                                              lt[:, 1]).cpu().numpy()
    else:
        w = gimp.get_GIMP_shape_function_and_gradient_3d(xt, *hs, *[1 / h for h in hs], lt[:, 0], lt[:, 1],
                                                         lt[:, 2]).cpu().numpy()
        wa = gimp.get_GIMP_shape_function_and_gradient_avg_3d(xt, *hs, *[1 / h for h in hs], 8, 8, lt[:, 0],
                                                              lt[:, 1], lt[:, 2]).cpu().numpy()
    for a in range(dim):
        lo, ow, og_, owa, ext = og.axis_weights(x[:, a], lp[:, a], hs[a], 1.0 / hs[a])
        for k in range(3):
            np.testing.assert_allclose(w[:, k * dim + a], ow[:, k], rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(w[:, 3 * dim + k * dim + a], og_[:, k], rtol=1e-12, atol=1e-10)
            np.testing.assert_allclose(wa[:, k * dim + a], owa[:, k], rtol=1e-12, atol=1e-14)
        # average-weight gradient = d(wa)/dx, central FD of the oracle's wa (piecewise linear in x)
        e = 1e-9
        _, _, _, wp_, _ = og.axis_weights(x[:, a] + e, lp[:, a], hs[a], 1.0 / hs[a])
        _, _, _, wm_, _ = og.axis_weights(x[:, a] - e, lp[:, a], hs[a], 1.0 / hs[a])
        same = (og.axis_weights(x[:, a] + e, lp[:, a], hs[a], 1 / hs[a])[0] == lo) & \
               (og.axis_weights(x[:, a] - e, lp[:, a], hs[a], 1 / hs[a])[0] == lo)
        fd = (wp_ - wm_) / (2 * e)
        for k in (0, 2):
            np.testing.assert_allclose(wa[same, 3 * dim + k * dim + a], fd[same, k], rtol=1e-5, atol=1e-3)


def _trial_strains(rng, n, scale):
    e = rng.normal(scale=scale, size=(n, 3))
    e[:, :] -= 0.3 * scale  # mostly compressive, so Drucker-Prager returns to the cone
    return e


@pytest.mark.gpu
def test_return_mapping_j2_matches_oracle():
    from geowarp_amd.materials import plasticity as pl
    rng = np.random.default_rng(3)
    lam, mu, kappa, Hh = 5.0e3, 3.0e3, 20.0, 150.0
    et = _trial_strains(rng, 3000, 0.01)
    ao = rng.uniform(0, 0.01, size=3000)
    ref, aref = om.return_mapping_J2(et, lam, mu, kappa, Hh, ao)
    out, dE, an = pl.return_mapping_J2(torch.from_numpy(np.stack([np.diag(v) for v in et])).to(_dev()), lam, mu,
                                       kappa, Hh, torch.from_numpy(ao).to(_dev()), tangent=True)
    out = torch.diagonal(out, dim1=1, dim2=2).cpu().numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(an.cpu().numpy(), aref, rtol=1e-12, atol=1e-15)
    # consistent tangent d eps / d eps_tr against central FD of the oracle
    h = 1e-7
    fd = np.zeros((3000, 3, 3))
    for j in range(3):
        ep, em = et.copy(), et.copy()
        ep[:, j] += h
        em[:, j] -= h
        fd[:, :, j] = (om.return_mapping_J2(ep, lam, mu, kappa, Hh, ao)[0] -
                       om.return_mapping_J2(em, lam, mu, kappa, Hh, ao)[0]) / (2 * h)
    np.testing.assert_allclose(dE.cpu().numpy().reshape(-1, 3, 3), fd, atol=2e-6)


@pytest.mark.gpu
@pytest.mark.parametrize('apex,shape', [('reference', 0.0), ('return', 0.0), ('return', 0.2)])
def test_return_mapping_dp_matches_oracle(apex, shape):
    from geowarp_amd.materials import plasticity as pl
    rng = np.random.default_rng(4)
    lam, mu, phi, psi, c = 4.0e3, 3.0e3, 30.0, 10.0, 1.0
    et = _trial_strains(rng, 3000, 0.004)
    ref = om.return_mapping_DruckerPrager(et, lam, mu, phi, psi, c, shape, apex)
    out = pl.return_mapping_DruckerPrager(torch.from_numpy(et).to(_dev()), lam, mu, phi, psi, c, shape,
                                          apex=apex).cpu().numpy()
    # the reference's fixed 10 local iterations do not converge for every trial state; where they
    # have not (the result still moves with more iterations) the iterate depends on round-off and
    # is not compared
    ref20 = om.return_mapping_DruckerPrager(et, lam, mu, phi, psi, c, shape, apex, n_iter=20)
    conv = np.abs(ref20 - ref).max(1) <= 1e-13
    plastic = np.abs(ref - et).max(1) > 1e-12
    print('plastic', int(plastic.sum()), 'converged', int((conv & plastic).sum()))
    assert (conv & plastic).sum() >= 0.8 * plastic.sum()
    np.testing.assert_allclose(out[conv], ref[conv], rtol=1e-10, atol=1e-13)


@pytest.mark.gpu
def test_p2g_residual_g2p_match_oracle():
    """P2G_2d activation, assemble_residual_2d_hencky at a random increment and
    G2P_2d, called one by one as a reference driver does, against the oracle."""
    from geowarp_amd.mpm import mpm_utils as mu_
    cfg = H.MINI['bar_nh']()
    cfg['material_dict'] = dict(cfg['material_dict'], material_name='Hencky elasticity')
    import oracle.solvers as osv
    so = H.make(osv, cfg)
    g = cfg['background_grid_dict']
    nx, ny, dx = g['n_grid_x'], g['n_grid_y'], g['dx']
    n_nodes = (nx + 1) * (ny + 1)
    x = so.x_particles.copy()
    n = x.shape[0]
    lp = np.full((n, 2), so.lp[0, 0])
    rng = np.random.default_rng(9)
    x += rng.uniform(-0.2, 0.2, size=x.shape) * dx * 0.1
    st = og.Stencil(x, lp, 2, (dx,), (nx, ny))
    act_ref = omp.p2g_2d_activate(st, n_nodes)
    dev = _dev()
    xt, lt = torch.from_numpy(x).to(dev), torch.from_numpy(lp).to(dev)
    act = torch.zeros(2 * n_nodes, dtype=torch.bool, device=dev)
    mu_.P2G_2d(xt, 1 / dx, dx, nx, n_nodes, act, lt)
    np.testing.assert_array_equal(act.cpu().numpy(), act_ref)

    boundary = so.boundary.copy()
    du = rng.normal(scale=1e-3 * dx, size=2 * n_nodes) * act_ref
    F0 = omp.initial_F(n)
    mat = omp.MaterialParams('Hencky elasticity', 1.0e4, 0.3)
    ext = np.zeros(n, dtype=bool)
    ext[::7] = True
    rref, Fn, bn, sig, _ = omp.residual_2d(st, n_nodes, du, F0, F0.copy(), np.zeros(n), ext, mat, so.p_vol, so.p_rho,
                                           10.0, (0.5, -2.0), 3, 20, boundary, act_ref)
    tens = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    Fo, Fnew, bo, bnew = tens(F0), tens(F0.copy()), tens(F0.copy()), tens(F0.copy())
    sigma = torch.zeros(n, 3, 3, dtype=torch.float64, device=dev)
    rhs = torch.zeros(2 * n_nodes, dtype=torch.float64, device=dev)
    lam = mat.lam
    mu = mat.mu
    mu_.assemble_residual_2d_hencky(xt, dx, 1 / dx, nx, n_nodes, tens(du), Fo, Fnew, bo, bnew, tens(ext), sigma, 10.0,
                                    0.0, 0.0, 0.5, -2.0, 3, 20, lam, mu, so.p_vol, so.p_rho, lt,
                                    tens(boundary), act, rhs)
    r = rhs.cpu().numpy()
    assert H.rel_err(r, rref) < 1e-12
    assert H.rel_err(sigma.cpu().numpy(), sig) < 1e-12
    assert H.rel_err(Fnew.cpu().numpy(), Fn) < 1e-13

    mu_.G2P_2d(lt, xt, dx, 1 / dx, nx, n_nodes, tens(du), Fo, Fnew, bo, bnew)
    xref, _ = omp.g2p(st, x, du, n_nodes, 2)
    assert H.rel_err(xt.cpu().numpy(), xref) < 1e-14
    assert H.rel_err(Fo.cpu().numpy(), Fn) < 1e-13


@pytest.mark.gpu
def test_return_mapping_norsand_matches_oracle():
    """Device NorSand against the oracle on trial states that cover tension,
    elastic and plastic cases; the converged plastic states also satisfy the
    reference's local equations (triaxial_NorSand.py:338-343) to round-off."""
    from geowarp_amd.materials import plasticity as pl
    prm = dict(kappa=0.0015, tilde_lambda=0.02, M=1.27, N=0.4, beta=1.0, vc0=1.8911, p0=-390.0, v0=1.75, h=70.0,
               mu=4.5e4, ev0=0.0)
    pc = -715.0
    pi0 = pc * (1 - 0.4) ** ((1 - 0.4) / 0.4)
    rng = np.random.default_rng(8)
    n = 64
    e = np.stack([rng.uniform(-2e-4, -1e-5, n), rng.uniform(-2e-4, -1e-5, n), rng.uniform(-8e-3, -1e-5, n)], 1)
    e[:4] = np.array([[1e-3, 1e-3, 1e-3], [-1e-6, -1e-6, -1e-6], [-1e-4, -1e-4, -8e-3], [-1e-4, -1.5e-4, -6e-3]])
    tot = e.sum(1) * 1.01
    pi_old = np.full(n, pi0)
    ref_e, ref_pi, ref_res = om.return_mapping_NorSand(e, tot, pi_old, prm)
    dev = _dev()
    tens = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out, pin, res = pl.return_mapping_NorSand(prm['kappa'], prm['tilde_lambda'], prm['M'], prm['N'], prm['beta'],
                                              prm['vc0'], prm['p0'], prm['v0'], prm['h'], prm['mu'], prm['ev0'],
                                              tens(pi_old), tens(e), torch.diag_embed(tens(tot[:, None] / 3 *
                                                                                           np.ones((1, 3)))),
                                              return_residual=True)
    out = out.cpu().numpy()
    pin = pin.cpu().numpy()
    np.testing.assert_allclose(out, ref_e, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(pin, ref_pi, rtol=1e-9)
    assert np.all(out[0] == 0.0)                   # tension: zero strain
    conv = np.abs(ref_res[:, :3]).max(1) < 1e-8     # cases where the fixed 15 iterations converged
    assert conv.sum() >= 2
    assert np.abs(res.cpu().numpy()[conv, :3]).max() < 1e-8


@pytest.mark.gpu
def test_e_grad_matches_fd_of_loss_at_fixed_solution():
    """E_grad of the 3D solver (d loss / d E with the nodal solution held, as
    the reference's tape gives it at qc3d:323-325) against a central FD of the
    loss, where the force comes from the module entry point
    assemble_residual_coupled_3d_neohookean at the converged solution and E +- h."""
    import geowarp_amd.solvers as gs
    from geowarp_amd.mpm import mpm_utils as mu_
    from geowarp_amd.solvers.quasi_static_coupled_solver_3d import get_indentation_loss
    from geowarp_amd import problems as P
    cfg = P.indent3d(n_cells=(4, 4, 4), dx=0.25, total_steps=2, depth=0.01, footing=0.5, material='Neo-Hookean',
                     tol=1e-10, n_iter=12)
    s = H.make(gs, cfg)
    captured = {}
    eng = s.engine
    g2p = eng.g2p

    def capture(*a, **k):
        captured.update(x=s.x_particles.clone(), lp=s.GIMP_lp.clone(), F=s.deformation_gradient_total_old.clone(),
                        old=s.old_solution.clone(), new=s.new_solution.clone(),
                        bnd=s.dofStruct.boundary_flag_array.clone(), act=s.dofStruct.activate_flag_array.clone())
        return g2p(*a, **k)

    eng.g2p = capture
    for k in range(2):
        s.advance_one_step(k, 2)
    assert s.E_grad is not None
    g = cfg['background_grid_dict']
    l = cfg['load_dict']
    dev = s.x_particles.device
    n = s.n_particles

    def force(E):
        F = captured['F'].clone()
        Fn, bo, bn = F.clone(), F.clone(), F.clone()
        sig = torch.zeros(n, 3, 3, dtype=torch.float64, device=dev)
        rhs = torch.zeros(4 * s.n_nodes, dtype=torch.float64, device=dev)
        f = torch.zeros(1, dtype=torch.float64, device=dev)
        fa = torch.zeros(2, dtype=torch.float64, device=dev)
        mu_.assemble_residual_coupled_3d_neohookean(
            captured['x'], g['dx'], g['dy'], g['dz'], 1 / g['dx'], 1 / g['dy'], 1 / g['dz'], s.dt, g['n_grid_x'],
            g['n_grid_y'], s.n_nodes, s.PPD, captured['old'], captured['new'], F, Fn, bo, bn, None, None, sig,
            s.phi_initial, s.mobility_constant, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, l['footing_region_min_x'],
            l['footing_region_max_x'], l['footing_region_min_y'], l['footing_region_max_y'],
            l['footing_initial_height'], l['footing_incr_disp'], l['penalty_factor'], 1, 2,
            torch.tensor([E], dtype=torch.float64), 0.3, s.p_vol, s.p_rho, captured['lp'], captured['bnd'],
            captured['act'], rhs, f, fa)
        return float(f.item())

    E = s.youngs_modulus
    f0 = force(E)
    assert abs(f0 - float(s.indentation_force.item())) <= 1e-6 * abs(f0)
    h = 1e-3 * E
    lp_ = get_indentation_loss(force(E + h), l['footing_incr_disp'], 2.0)
    lm_ = get_indentation_loss(force(E - h), l['footing_incr_disp'], 2.0)
    fd = (lp_ - lm_) / (2 * h)
    assert abs(float(s.E_grad[0]) - fd) <= 1e-6 * abs(fd)


@pytest.mark.gpu
@pytest.mark.parametrize('shape', [0.0, 0.1])
def test_return_mapping_dp_dilation_zero_tension_matches_oracle(shape):
    """Device Drucker-Prager return (phi 30, c 5, psi 0, apex return) on trial states
    around the tension region -- returns to the surface below the tip, cut-off at
    or beyond it -- against the oracle, and its consistent tangent against central
    differences of the oracle away from the regime boundaries."""
    from geowarp_amd.materials import plasticity as pl
    import test_oracle as T
    e, lam, mu, K, A, B, p_tip = T._dp_psi0_cases(shape)
    ref = om.return_mapping_DruckerPrager(e, lam, mu, 30.0, 0.0, 5.0, shape, apex='return', n_iter=30)
    out, dE = pl.return_mapping_DruckerPrager(torch.from_numpy(e).to(_dev()), lam, mu, 30.0, 0.0, 5.0, shape,
                                              apex='return', tangent=True)
    out = out.cpu().numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-10, atol=1e-13)
    h = 1e-9
    fd = np.zeros((e.shape[0], 3, 3))
    for j in range(3):
        ep, em = e.copy(), e.copy()
        ep[:, j] += h
        em[:, j] -= h
        fd[:, :, j] = (om.return_mapping_DruckerPrager(ep, lam, mu, 30.0, 0.0, 5.0, shape, apex='return', n_iter=30)
                       - om.return_mapping_DruckerPrager(em, lam, mu, 30.0, 0.0, 5.0, shape, apex='return',
                                                         n_iter=30)) / (2 * h)
    _, P, S, Q = om._dp_PQ(e, lam, mu)
    y = om.dp_yield(P, Q, 30.0, 5.0, shape)
    # away from the elastic/plastic boundary and from the tip cut-off boundary
    away = (np.abs(y) > 1e-3) & (np.abs(P - p_tip) > 1e-2)
    assert away.sum() > 200
    np.testing.assert_allclose(dE.cpu().numpy().reshape(-1, 3, 3)[away], fd[away], atol=2e-5)

"""MPM entry points with the reference's names and positional arguments
(GeoWarp mpm/mpm_utils.py), over torch tensors on the ROCm device.

Each function runs the same device phases as the solver classes (binning,
particle stage, node gather) on a cached engine for the grid it is given, so
a driver written against the reference's kernel list can call them one by one.
Nodal vectors and flags are dof-major (dof = node + k*n_nodes) as in the
reference; the residual functions ADD into ``rhs`` like the reference's
atomics.  Differences: everything is deterministic; particles whose stencil
leaves the grid raise.
"""
import numpy as np
import torch

from .. import _lib as L
from ..solvers._base import init_particles_rectangle
from ..solvers._engine import Engine, MAT_KIND

_ENGINES = {}


def _stream():
    return L.c_void_p(torch.cuda.current_stream().cuda_stream)


def _engine(dim, coupled, n_grid, h, device):
    key = (dim, bool(coupled), tuple(int(v) for v in n_grid), tuple(float(v) for v in h), str(device))
    if key not in _ENGINES:
        _ENGINES[key] = Engine(dim, coupled, n_grid, h, device=device)
    return _ENGINES[key]


def _grid2d(n_grid_x, n_nodes):
    return (int(n_grid_x), int(n_nodes) // (int(n_grid_x) + 1) - 1)


def _grid3d(n_grid_x, n_grid_y, n_nodes):
    return (int(n_grid_x), int(n_grid_y), int(n_nodes) // ((int(n_grid_x) + 1) * (int(n_grid_y) + 1)) - 1)


def _interleave(eng, v):
    return v.view(eng.nb, eng.n_nodes).t().reshape(-1).contiguous()


def _dofmajor(eng, v):
    return v.view(eng.n_nodes, eng.nb).t().reshape(-1)


def _flags(eng, boundary, activate):
    eng.bflag.copy_(_interleave(eng, boundary.to(torch.uint8)))
    eng.act.copy_(_interleave(eng, activate.to(torch.uint8)))
    L.check(L.lib().gm_free_mask(eng.n_nodes * eng.nb, L.ptr(eng.bflag), L.ptr(eng.act), L.ptr(eng.freem),
                                 _stream()), 'gm_free_mask')


# ------------------------------------------------------------------ set-up
def init_particles_rectangle_2d(start_x, start_y, end_x, end_y, dx, n_grid_x, n_grid_y, PPD, n_particles=None):
    """mpm_utils.py:12-29 (same particle order: i, j, p_x, p_y)."""
    return init_particles_rectangle(2, (start_x, start_y), (end_x, end_y), dx, (n_grid_x, n_grid_y), PPD)


def init_particles_rectangle_3d(start_x, start_y, start_z, end_x, end_y, end_z, dx, n_grid_x, n_grid_y, n_grid_z,
                                PPD, n_particles=None):
    """mpm_utils.py:31-51 (order i, j, k, p_x, p_y, p_z)."""
    return init_particles_rectangle(3, (start_x, start_y, start_z), (end_x, end_y, end_z), dx,
                                    (n_grid_x, n_grid_y, n_grid_z), PPD)


def initialization(deformation_gradient_total_new, deformation_gradient_total_old, left_Cauchy_Green_new,
                   left_Cauchy_Green_old):
    """mpm_utils.py:53-74: F = b_e = diag(1+1e-7, 1-1e-6, 1)."""
    I = torch.diag(torch.tensor([1.0 + 1e-7, 1.0 - 1e-6, 1.0], dtype=torch.float64,
                                device=deformation_gradient_total_new.device))
    for t in (deformation_gradient_total_new, deformation_gradient_total_old, left_Cauchy_Green_new,
              left_Cauchy_Green_old):
        t.copy_(I.expand_as(t))


def initialize_youngs_modulus_diff(youngs_modulus_diff, youngs_modulus):
    """mpm_utils.py:77-81."""
    youngs_modulus_diff[0] = float(youngs_modulus)


# ------------------------------------------------------------------ P2G
def P2G_2d(x_particles, inv_dx, dx, n_grid_x, n_nodes, activate_flag_array, GIMP_lp):
    """mpm_utils.py:83-136: activate x/y dofs of every support node."""
    eng = _engine(2, False, _grid2d(n_grid_x, n_nodes), (dx, dx, 0.0), x_particles.device)
    eng.begin_step(x_particles, GIMP_lp)
    eng.bflag.zero_()
    eng.p2g(x_particles, GIMP_lp, -1.0)
    activate_flag_array |= _dofmajor(eng, eng.act).bool()


def _p2g_coupled(dim, x, h, grid, boundary, activate, lp, pbc, p_mass, p_part, rhs_P2G, rows, cols, vals, thresh,
                 node_h):
    eng = _engine(dim, True, grid, h, x.device)
    eng.begin_step(x, lp)
    eng.bflag.copy_(_interleave(eng, boundary.to(torch.uint8)))
    mrhs = eng.p2g(x, lp, thresh, pbc.to(torch.uint8), node_h, p_mass, p_part, mass=True)
    activate |= _dofmajor(eng, eng.act).bool()
    boundary |= _dofmajor(eng, eng.bflag).bool()
    rhs_P2G += mrhs
    if rows is not None:
        # consistent mass in COO form (the reference's triplet list is per particle
        # (p*81.., p*729..); here one triplet per nonzero of the assembled matrix)
        rp = eng.row_ptr.long()
        nnz = int(rp[-1])
        r = torch.repeat_interleave(torch.arange(eng.n_nodes, device=x.device), rp[1:] - rp[:-1])
        k = min(nnz, rows.numel())
        rows[:k] = r[:k].to(rows.dtype)
        cols[:k] = eng.col_idx[:k].to(cols.dtype)
        vals[:k] = eng.mvals[:k]
    return eng


def P2G_coupled_2d(x_particles, inv_dx, dx, inv_dy, dy, n_grid_x, n_nodes, boundary_flag_array, activate_flag_array,
                   GIMP_lp, particle_pressure_boundary_flag_array, p_mass, particle_pressure_array, rhs_P2G,
                   rows_P2G=None, cols_P2G=None, vals_P2G=None):
    """mpm_utils.py:138-257 (threshold 1e-6, pressure boundary at node y >= particle y)."""
    _p2g_coupled(2, x_particles, (dx, dx, 0.0), _grid2d(n_grid_x, n_nodes), boundary_flag_array,
                 activate_flag_array, GIMP_lp, particle_pressure_boundary_flag_array, p_mass, particle_pressure_array,
                 rhs_P2G, rows_P2G, cols_P2G, vals_P2G, 1e-6, dy)


def P2G_coupled_3d(x_particles, inv_dx, dx, inv_dy, dy, inv_dz, dz, n_grid_x, n_grid_y, n_nodes, boundary_flag_array,
                   activate_flag_array, GIMP_lp, particle_pressure_boundary_flag_array, p_mass,
                   particle_pressure_array, rhs_P2G, rows_P2G=None, cols_P2G=None, vals_P2G=None):
    """mpm_utils.py:259-383 (threshold 1e-7, pressure boundary at node z >= particle z)."""
    _p2g_coupled(3, x_particles, (dx, dy, dz), _grid3d(n_grid_x, n_grid_y, n_nodes), boundary_flag_array,
                 activate_flag_array, GIMP_lp, particle_pressure_boundary_flag_array, p_mass, particle_pressure_array,
                 rhs_P2G, rows_P2G, cols_P2G, vals_P2G, 1e-7, dz)


def P2G_set_nodal_solution(boundary_flag_array, n_nodes, old_solution, new_solution, x_P2G_warp):
    """mpm_utils.py:387-399 (2D: p dofs not Dirichlet)."""
    d = torch.arange(n_nodes, device=old_solution.device) + 2 * n_nodes
    m = ~boundary_flag_array[d]
    old_solution[d[m]] = x_P2G_warp[m]
    new_solution[d[m]] = x_P2G_warp[m]


def P2G_set_nodal_solution_3d(boundary_flag_array, activate_flag_array, n_nodes, old_solution, new_solution,
                              x_P2G_warp):
    """mpm_utils.py:401-414 (3D: p dofs not Dirichlet and active)."""
    d = torch.arange(n_nodes, device=old_solution.device) + 3 * n_nodes
    m = (~boundary_flag_array[d]) & activate_flag_array[d]
    old_solution[d[m]] = x_P2G_warp[m]
    new_solution[d[m]] = x_P2G_warp[m]


# ------------------------------------------------------------------ residuals
def _stage_args(n, coupled, x, lp, Fo, Fn, bo, bn, sig, ext_flag, trac_flag, u_new, u_old, mat, lam, mu, E, kappa,
                p_vol, p_rho, gravity, point_load, traction):
    a = L.gm_stage_args()
    a.n_particles = n
    a.coupled = int(coupled)
    a.x, a.lp = x.data_ptr(), lp.data_ptr()
    a.F_old, a.F_new, a.b_old, a.b_new, a.sigma = Fo.data_ptr(), Fn.data_ptr(), bo.data_ptr(), bn.data_ptr(), \
        sig.data_ptr()
    a.alpha_old = a.alpha_new = None
    a.ext_flag = ext_flag.data_ptr() if ext_flag is not None else None
    a.trac_flag = trac_flag.data_ptr() if trac_flag is not None else None
    a.u_new = u_new.data_ptr()
    a.u_old = u_old.data_ptr() if u_old is not None else None
    a.mat_kind = MAT_KIND[mat]
    a.lam, a.mu, a.E, a.kappa = lam, mu, E, kappa
    a.p_vol, a.p_rho = p_vol, p_rho
    for k in range(3):
        a.gravity[k] = gravity[k]
        a.point_load[k] = point_load[k]
        a.traction[k] = traction[k]
    return a


def _run_residual(eng, a, x, lp, boundary, activate, rhs):
    _flags(eng, boundary, activate)
    eng.stage(a, want_tangent=False)
    eng.residual(x, lp)
    rhs += _dofmajor(eng, eng.rhs)


def _residual_2d(mat, x_particles, dx, n_grid_x, n_nodes, increment_solution, Fo, Fn, bo, bn, ext_flag, sig,
                 gravity_mag, plx, ply, current_step, total_steps, lam, mu, kappa, p_vol, p_rho, GIMP_lp, boundary,
                 activate, rhs):
    eng = _engine(2, False, _grid2d(n_grid_x, n_nodes), (dx, dx, 0.0), x_particles.device)
    eng.begin_step(x_particles, GIMP_lp, pattern=False)
    u = _interleave(eng, increment_solution)
    s1 = (float(current_step) + 1.0) / float(total_steps)
    ef = ext_flag.to(torch.uint8)
    a = _stage_args(x_particles.shape[0], False, x_particles, GIMP_lp, Fo, Fn, bo, bn, sig, ef, None, u, None, mat,
                    lam, mu, 0.0, kappa, p_vol, p_rho, (0.0, -gravity_mag * s1, 0.0), (plx * s1, ply * s1, 0.0),
                    (0.0, 0.0, 0.0))
    _run_residual(eng, a, x_particles, GIMP_lp, boundary, activate, rhs)


def assemble_residual_2d_hencky(x_particles, dx, inv_dx, n_grid_x, n_nodes, increment_solution,
                                deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
                                left_Cauchy_Green_new, particle_external_flag_array, particle_Cauchy_stress_array,
                                gravity_mag, traction_value_x, traction_value_y, point_load_value_x,
                                point_load_value_y, current_step, total_steps, lame_lambda, lame_mu, p_vol, p_rho,
                                GIMP_lp, boundary_flag_array, activate_flag_array, rhs):
    """mpm_utils.py:416-596."""
    _residual_2d('Hencky elasticity', x_particles, dx, n_grid_x, n_nodes, increment_solution,
                 deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
                 left_Cauchy_Green_new, particle_external_flag_array, particle_Cauchy_stress_array, gravity_mag,
                 point_load_value_x, point_load_value_y, current_step, total_steps, lame_lambda, lame_mu, 0.0, p_vol,
                 p_rho, GIMP_lp, boundary_flag_array, activate_flag_array, rhs)


def assemble_residual_2d_J2(x_particles, dx, inv_dx, n_grid_x, n_nodes, increment_solution,
                            deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
                            left_Cauchy_Green_new, particle_external_flag_array, particle_Cauchy_stress_array,
                            gravity_mag, traction_value_x, traction_value_y, point_load_value_x, point_load_value_y,
                            current_step, total_steps, lame_lambda, lame_mu, kappa, p_vol, p_rho, GIMP_lp,
                            boundary_flag_array, activate_flag_array, rhs):
    """mpm_utils.py:598-780 (perfectly plastic J2)."""
    _residual_2d('J2', x_particles, dx, n_grid_x, n_nodes, increment_solution, deformation_gradient_total_old,
                 deformation_gradient_total_new, left_Cauchy_Green_old, left_Cauchy_Green_new,
                 particle_external_flag_array, particle_Cauchy_stress_array, gravity_mag, point_load_value_x,
                 point_load_value_y, current_step, total_steps, lame_lambda, lame_mu, kappa, p_vol, p_rho, GIMP_lp,
                 boundary_flag_array, activate_flag_array, rhs)


def assemble_residual_coupled_2d_neohookean(x_particles, dx, inv_dx, dt, n_grid_x, n_nodes, PPD, old_solution,
                                            new_solution, deformation_gradient_total_old,
                                            deformation_gradient_total_new, left_Cauchy_Green_old,
                                            left_Cauchy_Green_new, particle_external_flag_array,
                                            particle_traction_flag_array, particle_Cauchy_stress_array, phi_initial,
                                            mobility_constant, gravity_mag, traction_value_x, traction_value_y,
                                            point_load_value_x, point_load_value_y, current_step, total_steps,
                                            lame_lambda, lame_mu, p_vol, p_rho, GIMP_lp, boundary_flag_array,
                                            activate_flag_array, rhs):
    """mpm_utils.py:782-1017."""
    eng = _engine(2, True, _grid2d(n_grid_x, n_nodes), (dx, dx, 0.0), x_particles.device)
    eng.begin_step(x_particles, GIMP_lp, pattern=False)
    un, uo = _interleave(eng, new_solution), _interleave(eng, old_solution)
    s1 = (float(current_step) + 1.0) / float(total_steps)
    ef = particle_external_flag_array.to(torch.uint8)
    tf = particle_traction_flag_array.to(torch.uint8)
    a = _stage_args(x_particles.shape[0], True, x_particles, GIMP_lp, deformation_gradient_total_old,
                    deformation_gradient_total_new, left_Cauchy_Green_old, left_Cauchy_Green_new,
                    particle_Cauchy_stress_array, ef, tf, un, uo, 'Neo-Hookean', lame_lambda, lame_mu, 0.0, 0.0,
                    p_vol, p_rho, (0.0, -gravity_mag * s1, 0.0),
                    (point_load_value_x * s1, point_load_value_y * s1, 0.0),
                    (traction_value_x * dx / PPD, traction_value_y * dx / PPD, 0.0))
    a.phi0, a.mobility, a.dt, a.PPD = phi_initial, mobility_constant, dt, float(PPD)
    _run_residual(eng, a, x_particles, GIMP_lp, boundary_flag_array, activate_flag_array, rhs)


def assemble_residual_coupled_3d_neohookean(x_particles, dx, dy, dz, inv_dx, inv_dy, inv_dz, dt, n_grid_x, n_grid_y,
                                            n_nodes, PPD, old_solution, new_solution, deformation_gradient_total_old,
                                            deformation_gradient_total_new, left_Cauchy_Green_old,
                                            left_Cauchy_Green_new, particle_external_flag_array,
                                            particle_traction_flag_array, particle_Cauchy_stress_array, phi_initial,
                                            mobility_constant, gravity_mag, traction_value_x, traction_value_y,
                                            traction_value_z, point_load_value_x, point_load_value_y,
                                            point_load_value_z, footing_region_min_x, footing_region_max_x,
                                            footing_region_min_y, footing_region_max_y, footing_initial_height,
                                            footing_incr_disp, penalty_factor, current_step, total_steps,
                                            youngs_modulus_diff, poisson_ratio, p_vol, p_rho, GIMP_lp,
                                            boundary_flag_array, activate_flag_array, rhs, indentation_force,
                                            indentation_force_array):
    """mpm_utils.py:1020-1294 (point loads and tractions are not applied, as in the reference)."""
    eng = _engine(3, True, _grid3d(n_grid_x, n_grid_y, n_nodes), (dx, dy, dz), x_particles.device)
    eng.begin_step(x_particles, GIMP_lp, pattern=False)
    un, uo = _interleave(eng, new_solution), _interleave(eng, old_solution)
    E = float(youngs_modulus_diff[0])
    nu = float(poisson_ratio)
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    s1 = (float(current_step) + 1.0) / float(total_steps)
    a = _stage_args(x_particles.shape[0], True, x_particles, GIMP_lp, deformation_gradient_total_old,
                    deformation_gradient_total_new, left_Cauchy_Green_old, left_Cauchy_Green_new,
                    particle_Cauchy_stress_array, None, None, un, uo, 'Neo-Hookean', lam, mu, E, 0.0, p_vol, p_rho,
                    (0.0, 0.0, -gravity_mag * s1), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    a.phi0, a.mobility, a.dt, a.PPD = phi_initial, mobility_constant, dt, float(PPD)
    a.footing[0], a.footing[1] = footing_region_min_x, footing_region_max_x
    a.footing[2], a.footing[3] = footing_region_min_y, footing_region_max_y
    a.footing_height = footing_initial_height + (float(current_step) + 1.0) * footing_incr_disp
    a.pen_par = E * penalty_factor * (dx / PPD) * (dy / PPD)
    a.penalty = 1
    a.pen_force = eng.pen_force.data_ptr()
    _run_residual(eng, a, x_particles, GIMP_lp, boundary_flag_array, activate_flag_array, rhs)
    eng.sum_into(eng.pen_force, 1)
    f = eng.scal[1]
    indentation_force[0] += f
    indentation_force_array[int(current_step)] += f


# ------------------------------------------------------------------ matrix bookkeeping
def _diag(boundary, activate, rows, cols, vals, n_matrix_size, slot):
    d = torch.nonzero(boundary | ~activate).flatten()
    rows[slot * n_matrix_size + d] = d.to(rows.dtype)
    cols[slot * n_matrix_size + d] = d.to(cols.dtype)
    vals[slot * n_matrix_size + d] = 1.0


def set_diagnal_component_for_boundary_and_deactivated_dofs_2d(boundary_flag_array, activate_flag_array, rows, cols,
                                                               vals, n_matrix_size):
    """mpm_utils.py:1299-1312 (slot 50*n + dof)."""
    _diag(boundary_flag_array, activate_flag_array, rows, cols, vals, n_matrix_size, 50)


def set_diagnal_component_for_boundary_and_deactivated_dofs_coupled_2d(boundary_flag_array, activate_flag_array, rows,
                                                                       cols, vals, n_matrix_size):
    """mpm_utils.py:1314-1327 (slot 75*n + dof)."""
    _diag(boundary_flag_array, activate_flag_array, rows, cols, vals, n_matrix_size, 75)


def set_diagnal_component_for_boundary_and_deactivated_dofs_coupled_3d(boundary_flag_array, activate_flag_array, rows,
                                                                       cols, vals, n_matrix_size):
    """mpm_utils.py:1329-1341 (slot 500*n + dof)."""
    _diag(boundary_flag_array, activate_flag_array, rows, cols, vals, n_matrix_size, 500)


def from_increment_to_solution(increment_iteration, increment_solution):
    """mpm_utils.py:1343-1348: increment_solution += increment_iteration."""
    L.check(L.lib().gm_axpy(increment_solution.numel(), 1.0, L.ptr(increment_iteration),
                            int(increment_iteration.dtype == torch.float32), L.ptr(increment_solution), _stream()),
            'gm_axpy')


# ------------------------------------------------------------------ G2P
def _g2p(dim, coupled, grid, h, GIMP_lp, x, u_new, u_old, Fo, Fn, bo, bn, p_part):
    eng = _engine(dim, coupled, grid, h, x.device)
    un = _interleave(eng, u_new)
    uo = _interleave(eng, u_old) if u_old is not None else None
    eng.g2p(x, GIMP_lp, un, uo, Fo, Fn, bo, bn, None, None, p_part)


def G2P_2d(GIMP_lp, x_particles, dx, inv_dx, n_grid_x, n_nodes, increment_solution, deformation_gradient_total_old,
           deformation_gradient_total_new, left_Cauchy_Green_old, left_Cauchy_Green_new):
    """mpm_utils.py:1351-1419."""
    _g2p(2, False, _grid2d(n_grid_x, n_nodes), (dx, dx, 0.0), GIMP_lp, x_particles, increment_solution, None,
         deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
         left_Cauchy_Green_new, None)


def G2P_coupled_2d(GIMP_lp, x_particles, dx, inv_dx, n_grid_x, n_nodes, old_solution, new_solution,
                   deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
                   left_Cauchy_Green_new, particle_pressure_array):
    """mpm_utils.py:1422-1499."""
    _g2p(2, True, _grid2d(n_grid_x, n_nodes), (dx, dx, 0.0), GIMP_lp, x_particles, new_solution, old_solution,
         deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
         left_Cauchy_Green_new, particle_pressure_array)


def G2P_coupled_3d(GIMP_lp, x_particles, dx, dy, dz, inv_dx, inv_dy, inv_dz, n_grid_x, n_grid_y, n_nodes,
                   old_solution, new_solution, deformation_gradient_total_old, deformation_gradient_total_new,
                   left_Cauchy_Green_old, left_Cauchy_Green_new, particle_pressure_array):
    """mpm_utils.py:1502-1590."""
    _g2p(3, True, _grid3d(n_grid_x, n_grid_y, n_nodes), (dx, dy, dz), GIMP_lp, x_particles, new_solution,
         old_solution, deformation_gradient_total_old, deformation_gradient_total_new, left_Cauchy_Green_old,
         left_Cauchy_Green_new, particle_pressure_array)

"""Shared set-up for parity tests: build a workload for the oracle (numpy) or the
device solvers (torch), apply its boundary callback / particle flags, run steps,
and collect comparable fields as numpy arrays."""
import numpy as np

from geowarp_amd import problems as P

# Reduced workloads used by the golden fixtures (the oracle runs them in seconds).
MINI = {
    'bar_nh': lambda: P.bar_nh(n_cells_x=4, n_cells_y=16, total_steps=20, tol=1e-9),
    'bar_hencky': lambda: _with_material(P.bar_nh(n_cells_x=4, n_cells_y=16, total_steps=20, tol=1e-9),
                                         dict(material_name='Hencky elasticity')),
    'bar_j2': lambda: _j2_mini(),
    'bar_dp': lambda: _dp_mini(),
    'bar_j2_platen': lambda: P.bar_j2(n_cells_x=4, n_cells_y=16, dx=0.0125, total_steps=10, tol=1e-9,
                                      platen_strain=0.02),
    'terzaghi': lambda: P.terzaghi(n_cells_x=2, n_cells_y=16, width=0.125, total_steps=10, dt=1.0),
    # the full 2D Neo-Hookean bar of 8 x 80 cells (2560 particles), all 20 load steps
    'c1_bar_nh': lambda: P.bar_nh(tol=1e-10),
    # 16-cell column: the 10 kPa load undrained (dt 1e-5 s), then 7 log-spaced consolidation steps
    'terzaghi_log': lambda: dict(P.terzaghi(n_cells_x=2, n_cells_y=16, width=0.125, total_steps=8, dt=1e-5),
                                 dts=[1e-5] + list(np.diff(np.logspace(-3, np.log10(0.15), 8)))),
    'indent_nh': lambda: P.indent3d(n_cells=(4, 4, 4), dx=0.25, total_steps=10, depth=0.05, footing=0.5,
                                    material='Neo-Hookean', tol=1e-10),
    'cantilever': lambda: P.cantilever(PPD=2, total_steps=50, tol=1e-10),
    # NorSand (the reference driver's material) in a rigid box under a platen, 4 x 8 cells.  The
    # platen moves 0.1 % of the height per step (the first iterate strains the top cells by 0.8 %,
    # where the fixed 15-iteration local Newton still converges, so its tangent is the derivative
    # of the stress); the particles start pre-strained by -0.2 % volumetric, away from the
    # return's tension cut at zero volumetric strain.
    'bar_norsand': lambda: P.norsand_box(total_steps=10, strain=0.01, tol=1e-10),
    # Drucker-Prager phi 30, c 5, psi 0 (the 3D indentation's material): the hyperbolic surface
    # (shape factor 0.1) with its tension cut-off under the reference's penalty, and the cone with
    # its apex return under the textbook penalty; the indenter starts at the top particle row
    'indent_dp': lambda: P.indent3d(n_cells=(4, 4, 4), dx=0.25, total_steps=10, depth=0.01, footing=0.5,
                                    material='Drucker-Prager', tol=1e-12, n_iter=20, dp_apex='return',
                                    shape_factor=0.1, footing_start='surface'),
    'indent_dp_cone': lambda: P.indent3d(n_cells=(4, 4, 4), dx=0.25, total_steps=10, depth=0.01, footing=0.5,
                                         material='Drucker-Prager', tol=1e-12, n_iter=20, dp_apex='return',
                                         footing_start='surface', penalty_form='physical'),
}


def _with_material(cfg, upd):
    cfg['material_dict'] = dict(cfg['material_dict'], **upd)
    return cfg


def _j2_mini():
    cfg = P.bar_nh(n_cells_x=4, n_cells_y=16, total_steps=5)
    cfg['particles_dict']['p_rho'] = 2.0
    cfg['material_dict'] = dict(youngs_modulus=1.0e4, poisson_ratio=0.3, material_name='J2', kappa=0.5, H=100.0)
    return cfg


def _dp_mini():
    cfg = P.bar_nh(n_cells_x=4, n_cells_y=16, total_steps=5)
    cfg['particles_dict']['p_rho'] = 2.0
    cfg['material_dict'] = dict(youngs_modulus=1.0e4, poisson_ratio=0.05, material_name='Drucker-Prager',
                                friction_angle=20.0, dilation_angle=5.0, cohesion=0.2)
    return cfg


def make(module, cfg, **kw):
    """Build the solver and apply the workload's boundary / particle flags."""
    s = P.build_solver(module, cfg, **kw)
    is_oracle = module.__name__.startswith('oracle')
    if is_oracle and 'fd_rel' in cfg:
        s.fd_rel = cfg['fd_rel']
    if cfg['solver'] == 'quasi_static_solver_2d':
        flags = s.boundary if is_oracle else s.dofStruct.boundary_flag_array
        cfg['set_boundary_dofs'](flags, s.n_grid_x, s.n_nodes)
    if 'external_flags' in cfg:
        cfg['external_flags'](s.x_particles, s.external_flag if is_oracle else s.particle_external_flag_array)
    if 'particle_flags' in cfg:
        if is_oracle:
            tf = getattr(s, 'traction_flag', np.zeros(s.n_particles, dtype=bool))
            cfg['particle_flags'](s.x_particles, tf, s.pressure_boundary_flag)
            if hasattr(s, 'traction_flag'):
                s.traction_flag = tf
        else:
            cfg['particle_flags'](s.x_particles, s.particle_traction_flag_array,
                                  s.particle_pressure_boundary_flag_array)
    return s


def fields(s):
    """Comparable state of either implementation as numpy arrays."""
    if hasattr(s, 'engine'):
        out = dict(x=s.x_particles.cpu().numpy(), sigma=s.particle_Cauchy_stress_array.cpu().numpy(),
                   F=s.deformation_gradient_total_old.cpu().numpy(), b=s.left_Cauchy_Green_old.cpu().numpy())
        if hasattr(s, 'particle_pressure_array'):
            out['p'] = s.particle_pressure_array.cpu().numpy()
        if hasattr(s, 'indentation_force'):
            out['force'] = np.array([float(s.indentation_force.item())])
        out['alpha'] = s.plastic_strain_old.cpu().numpy()
        return out
    out = dict(x=s.x_particles.copy(), sigma=s.sigma.copy(), F=s.F_old.copy(), b=s.b_old.copy(),
               alpha=s.alpha_old.copy())
    if hasattr(s, 'pressure'):
        out['p'] = s.pressure.copy()
    if hasattr(s, 'indentation_force'):
        out['force'] = np.array([float(s.indentation_force)])
    return out


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d

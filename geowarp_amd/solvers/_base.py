"""Shared implementation of the three quasi-static solver classes.

The public classes (quasi_static_solver_2d, quasi_static_coupled_solver_2d,
quasi_static_coupled_solver_3d) keep the reference constructor signatures,
dict keys, method names and state attributes (solvers/quasi_static_*.py);
their arrays are torch tensors on the ROCm device, FP64, in the reference
shapes, with dof-major nodal vectors.  Every numerical phase runs in
libgeowarp_amd.so; this module only sequences the calls.

Extensions over the reference (all optional, defaults reproduce it):
  iteration_dict['tol_mode']      'absolute' (reference, qs2d:230) or 'relative'
  iteration_dict['precision']     'fp64' (default) or 'mixed' (FP32 Jacobian +
                                  FP32 Krylov, FP64 residual/refinement)
  iteration_dict['linear_tol'], ['linear_maxiter'], ['ir_steps'], ['forcing']
                                  ('fixed' | 'ew': Eisenstat-Walker inner tolerance),
  ['final_forcing']               'ew' inner tolerance of the last correction, the one applied
                                  after the residual has passed the Newton test (1e-1; for J2,
                                  Drucker-Prager and NorSand the forcing term, at most 1e-3)
  ['forcing_max']                 cap of the 'ew' inner tolerance (1e-3; 1e-4 for J2,
                                  Drucker-Prager and NorSand; a hundred times tighter after each Newton
                                  iteration that reduces the residual by less than half),
  ['forcing_gamma'], ['forcing_alpha']  'ew' eta_k = gamma (|r_k|/|r_k-1|)^alpha (0.9; alpha 1,
                                  2 for J2, Drucker-Prager and NorSand); once |r_k| is within
                                  1/forcing_max of the Newton target, eta_k = half the target / |r_k|
  ['line_search'], ['autotune'],
  ['krylov'] ('mg' default: multigrid-preconditioned flexible GMRES |
  'gmres': block-Jacobi GMRES | 'bicgstab'), ['krylov_restart'],
  ['mg_nu'] (both), ['mg_nu_pre'], ['mg_nu_post'] (smoothing sweeps, default 1 and 2),
  ['mg_omega'] (damping, 0.4), ['mg_coarse_sweeps'], ['mg_ublock'], ['mg_min_cells']  V-cycle options
  ['mg_lag']                      coarse operators rebuilt at every setup (0), kept within a load
                                  step (1) or also across load steps (2: the mixed form's default
                                  for elastic materials; 1 otherwise)
  ['mg_lowprec']                  mixed form: FP16 (row-scaled) smoother matrices (default True)
  ['final_chord']                 solve the last correction (applied after the residual passes the
                                  Newton test) on the previous iteration's Jacobian (default: the
                                  mixed form with an elastic material)
  ['mixed_fallback']              mixed form: re-solve a correction in FP64 (FP64 tangent, Jacobian
                                  and hierarchy) when the FP32 solve leaves relres > 0.1 (default True)
  ['mixed_assembly']              mixed form: 'fp32' tangent and Jacobian (default) or 'fp64'
                                  (FP64 assembly, FP32 copy for the Krylov solve, refinement
                                  residual against the FP64 Jacobian)
  ['abs_floor']                   the Newton test also passes once the residual norm is below it
  ['mass_tol'] (coupled)          relative tolerance of the consistent-mass projection (1e-3 x a
                                  relative Newton tolerance within [1e-13, 1e-11]; else 1e-13)
  ['substeps'] (3D u-p)           levels of load-step halving after a failed Newton iteration
  iteration_dict['verbose']       print residual norms like the reference
  material_dict['H']              J2 linear isotropic hardening modulus
  material_dict['friction_angle'], ['dilation_angle'], ['cohesion'],
  ['dp_apex']                     Drucker-Prager ('reference' or 'return')
"""
import math

import numpy as np
import torch

from .. import _lib as L
from ._engine import Engine, MAT_KIND


def norsand_params(overrides=None):
    """NorSand material parameters: the reference driver's values
    (triaxial_NorSand.py:8-30: kappa, tilde_lambda, M, N, beta, vc0, pc, p0, v0, h,
    G0) updated by `overrides`, plus ev0 (the elastic volumetric strain at which
    P = p0; 0 puts the undeformed particle at p0) and the initial image pressure
    pi0 = pc (1 - N)^((1 - N)/N)."""
    out = dict(kappa=0.0015, tilde_lambda=0.02, M=1.27, N=0.4, beta=1.0, vc0=1.8911, pc=-715.0, p0=-390.0, v0=1.75,
               h=70.0, G0=4.5e4, ev0=0.0)
    out.update(overrides or {})
    out['pi0'] = out['pc'] * (1.0 - out['N']) ** ((1.0 - out['N']) / out['N'])
    return out


class DofStruct:
    """DofStruct (qs2d:40-43): activation and Dirichlet flags, dof-major bool."""

    def __init__(self, n, device):
        self.activate_flag_array = torch.zeros(n, dtype=torch.bool, device=device)
        self.boundary_flag_array = torch.zeros(n, dtype=torch.bool, device=device)


def init_particles_rectangle(dim, start, end, dx, n_grid, PPD):
    """init_particles_rectangle_2d/3d (mpm_utils.py:12-51) without Python loops:
    cells whose centre lies strictly inside (start, end), PPD^d points per cell,
    ordered by (i, j[, k], a, b[, c])."""
    axes = []
    for a in range(dim):
        c = (np.arange(n_grid[a]) + 0.5) * dx
        axes.append(np.nonzero((start[a] < c) & (c < end[a]))[0])
    sub = (0.5 + np.arange(PPD)) * dx / PPD
    if dim == 2:
        I, J, A, B = np.meshgrid(axes[0], axes[1], np.arange(PPD), np.arange(PPD), indexing='ij')
        x = np.stack([I * dx + sub[A], J * dx + sub[B]], axis=-1)
    else:
        I, J, K, A, B, C = np.meshgrid(axes[0], axes[1], axes[2], np.arange(PPD), np.arange(PPD), np.arange(PPD),
                                       indexing='ij')
        x = np.stack([I * dx + sub[A], J * dx + sub[B], K * dx + sub[C]], axis=-1)
    return x.reshape(-1, dim).astype(np.float64)


class _QuasiStaticBase:
    dim = 2
    coupled = False

    # ------------------------------------------------------------------ construction
    def _common_init(self, background_grid_dict, particles_dict, iteration_dict, material_dict, load_dict,
                     device, x_particles):
        g = background_grid_dict
        self.n_grid_x = g['n_grid_x']
        self.n_grid_y = g['n_grid_y']
        self.dx = g['dx']
        self.inv_dx = 1.0 / self.dx
        self.max_x = g.get('max_x')
        self.start_x, self.end_x = g['start_x'], g['end_x']
        self.start_y, self.end_y = g['start_y'], g['end_y']
        self.domain_min_x, self.domain_max_x = g['domain_min_x'], g['domain_max_x']
        self.domain_min_y, self.domain_max_y = g['domain_min_y'], g['domain_max_y']
        if self.dim == 3:
            self.n_grid_z = g['n_grid_z']
            self.n_nodes = (self.n_grid_x + 1) * (self.n_grid_y + 1) * (self.n_grid_z + 1)
        else:
            self.n_nodes = (self.n_grid_x + 1) * (self.n_grid_y + 1)
        nb = self.dim + (1 if self.coupled else 0)
        self.n_matrix_size = nb * self.n_nodes
        p = particles_dict
        self.PPD = p['PPD']
        self.p_vol = p['p_vol']
        self.p_rho = p['p_rho']
        it = iteration_dict
        self.n_iter = it['n_iter']
        self.tol = it['tol']
        self.solver_name = it.get('solver_name', 'warp')
        self.tol_mode = it.get('tol_mode', 'absolute')
        self.precision = it.get('precision', 'fp64')
        self.linear_tol = it.get('linear_tol', 1e-10 if self.precision == 'fp64' else 1e-6)
        self.linear_maxiter = it.get('linear_maxiter', 20000)
        # FP64 iterative-refinement sweeps around the FP32 Krylov solve (mixed form); a sweep
        # ends the loop once the FP64 residual of the correction meets the inner tolerance
        self.ir_steps = it.get('ir_steps', 1 if self.precision == 'fp64' else 4)
        # linear solver: flexible GMRES(krylov_restart) preconditioned by geometric-multigrid V-cycles
        # ('mg', default), block-Jacobi GMRES ('gmres') or block-Jacobi BiCGStab ('bicgstab')
        self.krylov_method = it.get('krylov', 'mg')
        # row equilibration of the Newton system (u-p: displacement and pressure rows differ by orders)
        self.equilibrate = it.get('equilibrate', False)
        # backtracking line search on the Newton correction (the reference takes full steps)
        self.line_search = it.get('line_search', False)
        # time the SpMV workgroup sizes on the first Jacobian and keep the fastest
        self.autotune = it.get('autotune', True)
        self.line_search_steps = []
        # multigrid V-cycle options (krylov='mg'): smoothing sweeps, damping, coarsest-level sweeps
        mg_opts = {k: it[k2] for k, k2 in (('nu_pre', 'mg_nu'), ('nu_post', 'mg_nu'), ('nu_pre', 'mg_nu_pre'),
                                           ('nu_post', 'mg_nu_post'), ('omega', 'mg_omega'),
                                           ('coarse_sweeps', 'mg_coarse_sweeps'), ('ublock', 'mg_ublock'),
                                           ('lowprec', 'mg_lowprec'), ('min_cells', 'mg_min_cells'))
                   if k2 in it}
        self.krylov_restart = it.get('krylov_restart', 250 if self.krylov_method == 'mg' else 40)
        # inner tolerance: 'fixed' (linear_tol) or 'ew' (Eisenstat-Walker choice 2)
        self.forcing = it.get('forcing', 'fixed' if self.precision == 'fp64' else 'ew')
        self.verbose = it.get('verbose', False)
        m = material_dict
        self.material_name = m['material_name']
        # materials with a return mapping: their tangent changes with the active set, which
        # sets several solver defaults below apart from those of elastic skeletons
        plastic = self.material_name in ('J2', 'Drucker-Prager', 'NorSand')
        # cap of the Eisenstat-Walker inner tolerance.  Materials with a return mapping get a
        # tighter cap: a loose Newton correction moves particles across the yield surface, and
        # the iteration then stalls on the changed active set (a 256 x 1024-cell J2 bar stalls
        # at relative residual 3e-2 with a 1e-2 cap in FP64 as in FP32, and converges
        # quadratically with 1e-4).  Elastic skeletons take 1e-3: on the 64 x 64 x 32-cell 3D
        # indentation it gives four Newton iterations per load step instead of five and 15 %
        # less time than 1e-2 (1.60 s against 1.87 s for load steps 5-24 on an MI355X)
        self.eta_max = it.get('forcing_max', 1e-4 if plastic else 1e-3)
        # mixed form: solve a Newton correction again in FP64 when the FP32 solve fails
        self.mixed_fallback = it.get('mixed_fallback', True)
        # Eisenstat-Walker choice 2 parameters: eta_k = gamma (|r_k| / |r_k-1|)^alpha.  Elastic
        # skeletons take alpha = 1: their Newton iteration gains two to three decades per
        # correction rather than squaring the residual, so alpha = 2 asks the second correction
        # for 1e-6 where 1e-3 leaves the next residual the same (the 64 x 64 x 32 indentation:
        # 9 GMRES iterations instead of 16, same Newton residuals)
        self.ew_gamma = it.get('forcing_gamma', 0.9)
        self.ew_alpha = it.get('forcing_alpha', 2.0 if plastic else 1.0)
        # inner tolerance of the last correction (the one applied after the residual has
        # passed the Newton test) under Eisenstat-Walker forcing; None: the forcing term of
        # that iteration, at most 1e-3.  Elastic skeletons take 1e-1: the accepted state then
        # lies ten times inside the residual the test passed (the indentation: 5 GMRES
        # iterations instead of 13, fields against the FP64 form unchanged at 5e-8)
        self.final_forcing = it.get('final_forcing', None if plastic else 1e-1)
        # absolute floor of the Newton test (None: the tolerance alone)
        self.abs_floor = it.get('abs_floor')
        # coupled forms: relative tolerance of the consistent-mass projection, a thousand times
        # tighter than a relative Newton tolerance, within [1e-13, 1e-11] (1e-13 under an
        # absolute test); the reference solves it to 1e-8 (qc2d:293, qc3d:412)
        self.mass_tol = float(it.get('mass_tol', max(1e-13, min(1e-11, 1e-3 * self.tol))
                                     if self.tol_mode == 'relative' else 1e-13))
        # the last correction (its residual already passes the Newton test) solved on the
        # previous iteration's Jacobian, a chord step: the mixed form's default for elastic
        # materials, whose Jacobian barely moves over the last Newton correction
        self.final_chord = bool(it.get('final_chord', self.precision == 'mixed' and not plastic))
        if self.material_name not in MAT_KIND:
            raise ValueError("Unimplemented material model: %s" % self.material_name)
        self.youngs_modulus = m['youngs_modulus']
        self.poisson_ratio = m['poisson_ratio']
        self._set_lame()
        self.kappa = m.get('kappa', 0.0)
        self.hardening = m.get('H', 0.0)
        self.friction_angle = m.get('friction_angle', 30.0)
        self.dilation_angle = m.get('dilation_angle', 0.0)
        self.cohesion = m.get('cohesion', 0.0)
        self.shape_factor = m.get('shape_factor', 0.0)
        self.dp_apex = {'reference': 0, 'return': 1}[m.get('dp_apex', 'reference')]
        self.norsand = norsand_params(m.get('norsand')) if self.material_name == 'NorSand' else None
        self.gravity_mag = load_dict['gravity_mag']
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        spacing = (self.dx, self.dx, 0.0) if self.dim == 2 else (self.dx, self.dy, self.dz)
        n_grid = (self.n_grid_x, self.n_grid_y) if self.dim == 2 else (self.n_grid_x, self.n_grid_y, self.n_grid_z)
        self.engine = Engine(self.dim, self.coupled, n_grid, spacing, device=self.device, precision=self.precision)
        self.engine.mixed_assembly = it.get('mixed_assembly', 'fp32')
        # coarse multigrid operators: 1 kept within a load step, 2 also across load steps (the
        # mixed form's default for elastic skeletons: its GMRES iterations are cheap, so trading
        # a few more of them for the Galerkin set-up pays -- 1.51 s against 1.61 s on load steps
        # 5-24 of the 64 x 64 x 32 indentation -- while the FP64 form's cost more than the
        # set-up: 6.80 s against 6.37 s.  Plastic skeletons keep 1: yielding between load steps
        # changes the tangent, and on the 256 x 1024-cell J2 bar a correction of load step 1
        # then takes 435 iterations on step 0's coarse operators against 149 on fresh ones)
        self.engine.mg_lag = int(it.get('mg_lag', 2 if self.precision == 'mixed' and not plastic else 1))
        self.engine.mg_options = mg_opts
        # coupled forms: start each consistent-mass projection from the last nodal pressures
        self.warm_mass = bool(it.get('warm_mass', True))
        dev = self.device
        self.dofStruct = DofStruct(self.n_matrix_size, dev)
        if x_particles is None:
            start = (self.start_x, self.start_y) + ((self.start_z,) if self.dim == 3 else ())
            end = (self.end_x, self.end_y) + ((self.end_z,) if self.dim == 3 else ())
            x_particles = init_particles_rectangle(self.dim, start, end, self.dx, n_grid, self.PPD)
        x_np = np.asarray(x_particles, dtype=np.float64).reshape(-1, self.dim)
        self.n_particles = x_np.shape[0]
        if 'n_particles' in p and p['n_particles'] != self.n_particles:
            raise ValueError('particles_dict n_particles=%d but the lattice holds %d' % (p['n_particles'],
                                                                                         self.n_particles))
        f64 = dict(dtype=torch.float64, device=dev)
        n = self.n_particles
        self.x_particles = torch.tensor(x_np, **f64)
        self.x0 = self.x_particles.clone()
        self.deformation_gradient_total_new = torch.zeros(n, 3, 3, **f64)
        self.deformation_gradient_total_old = torch.zeros(n, 3, 3, **f64)
        self.left_Cauchy_Green_new = torch.zeros(n, 3, 3, **f64)
        self.left_Cauchy_Green_old = torch.zeros(n, 3, 3, **f64)
        self.particle_Cauchy_stress_array = torch.zeros(n, 3, 3, **f64)
        self.plastic_strain_new = torch.zeros(n, **f64)
        self.plastic_strain_old = torch.zeros(n, **f64)
        self.GIMP_lp = torch.zeros(n, self.dim, **f64)
        self.particle_external_flag_array = torch.zeros(n, dtype=torch.bool, device=dev)
        self.residual_history = []
        self.newton_iterations = 0
        self._initialization()

    def _set_lame(self):
        E, nu = self.youngs_modulus, self.poisson_ratio
        self.lame_lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        self.lame_mu = E / (2.0 * (1.0 + nu))

    def _initialization(self):
        """initialization (mpm_utils.py:53-74) + initialize_GIMP_lp (gimp.py:8-22)."""
        I = torch.diag(torch.tensor([1.0 + 1e-7, 1.0 - 1e-6, 1.0], dtype=torch.float64, device=self.device))
        for t in (self.deformation_gradient_total_new, self.deformation_gradient_total_old,
                  self.left_Cauchy_Green_new, self.left_Cauchy_Green_old):
            t.copy_(I.expand_as(t))
        # hardening variable: J2's equivalent plastic strain (0), or NorSand's image pressure pi_0
        a0 = self.norsand['pi0'] if self.norsand is not None else 0.0
        self.plastic_strain_new.fill_(a0)
        self.plastic_strain_old.fill_(a0)
        self.GIMP_lp.fill_(self.GIMP_lp_initial)

    # ------------------------------------------------------------------ stage arguments
    def _stage_args(self, current_step, total_steps, u_new, u_old):
        a = L.gm_stage_args()
        a.n_particles = self.n_particles
        a.coupled = int(self.coupled)
        a.x = self.x_particles.data_ptr()
        a.lp = self.GIMP_lp.data_ptr()
        a.F_old = self.deformation_gradient_total_old.data_ptr()
        a.F_new = self.deformation_gradient_total_new.data_ptr()
        a.b_old = self.left_Cauchy_Green_old.data_ptr()
        a.b_new = self.left_Cauchy_Green_new.data_ptr()
        a.sigma = self.particle_Cauchy_stress_array.data_ptr()
        a.alpha_old = self.plastic_strain_old.data_ptr()
        a.alpha_new = self.plastic_strain_new.data_ptr()
        self._ext_u8 = self.particle_external_flag_array.to(torch.uint8)
        a.ext_flag = self._ext_u8.data_ptr()
        a.trac_flag = None
        a.u_new = u_new.data_ptr()
        a.u_old = u_old.data_ptr() if u_old is not None else None
        a.pen_force = self.engine.pen_force.data_ptr() if self.engine.n_particles > 0 else None
        a.mat_kind = MAT_KIND[self.material_name]
        a.dp_apex = self.dp_apex
        a.lam, a.mu, a.E = self.lame_lambda, self.lame_mu, self.youngs_modulus
        a.kappa, a.H = self.kappa, self.hardening
        a.phi, a.psi, a.coh, a.shape = self.friction_angle, self.dilation_angle, self.cohesion, self.shape_factor
        # optional per-particle Drucker-Prager parameters (batched specimens)
        pa = getattr(self, 'friction_angle_array', None)
        ca = getattr(self, 'cohesion_array', None)
        a.p_phi = pa.data_ptr() if pa is not None else None
        a.p_coh = ca.data_ptr() if ca is not None else None
        if self.norsand is not None:
            q = self.norsand
            a.ns = L.gm_norsand_params(q['kappa'], q['tilde_lambda'], q['M'], q['N'], q['beta'], q['vc0'], q['p0'],
                                       q['v0'], q['h'], q['G0'], q['ev0'], 1e-10)
        a.p_vol, a.p_rho = self.p_vol, self.p_rho
        a.PPD = float(self.PPD)
        s1 = (float(current_step) + 1.0) / float(total_steps)
        g = [0.0, 0.0, 0.0]
        g[self.dim - 1] = -self.gravity_mag * s1
        for k in range(3):
            a.gravity[k] = g[k]
        return a

    # ------------------------------------------------------------------ Newton loop
    def _newton(self, current_step, total_steps, u_new, u_old, args, after_residual=None):
        """newton_iter of the reference (qs2d:151-231): per iteration residual,
        Jacobian, solve, u += delta, then the test on the pre-update norm."""
        eng = self.engine
        x, lp = self.x_particles, self.GIMP_lp
        self.residual_history = []
        r0 = None
        it = 0
        tm = eng.timers
        eta = None
        nr_prev = None
        self._ls_ready = None  # norm of an accepted line-search trial (full step)
        self._ls_tangent = True  # whether that trial's stage computed the tangent
        jac_fresh = False  # the engine holds this load step's last Newton Jacobian (FP32 in the mixed form)
        for it in range(self.n_iter):
            if tm is not None:
                tm.mark('start')
            reuse = self._ls_ready
            has_tangent = reuse is None or self._ls_tangent
            self._ls_ready = None
            self._aims_to_pass = False
            if reuse is None:
                eng.stage(args, want_tangent=True)
                if tm is not None:
                    tm.mark('particle_stage')
                eng.residual(x, lp)
                eng.norm2(eng.rhs, 0)
            # else: the accepted line-search trial left this iterate's tangent, residual and norm
            if after_residual is not None:
                after_residual()
            if tm is not None:
                tm.mark('node_residual')
            nr = reuse
            if nr is None and r0 is not None and self.final_chord:
                nr = math.sqrt(max(eng.scal[0].item(), 0.0))  # decides whether the Jacobian is assembled
            if self.final_chord and jac_fresh and nr is not None and r0 is not None and self._passes(nr, r0):
                # the correction applied after the passing residual: on the last Jacobian
                eng.stats['chord_steps'] = eng.stats.get('chord_steps', 0) + 1
            else:
                if not has_tangent:
                    # the line search skipped the tangent, expecting this residual to pass
                    eng.stage(args, want_tangent=True)
                eng.jacobian(x, lp)
                jac_fresh = True
                if self.autotune and eng.stats.get('jacobian_tune') is None and eng.step_id >= 2:
                    # tuned from the second load step on: the undeformed first step has a sparser block pattern
                    eng.autotune_jacobian()
                if self.autotune and eng.stats.get('spmv_tune') is None:
                    eng.autotune_spmv()
            if tm is not None:
                tm.mark('node_jacobian')
            if nr is None:
                nr = math.sqrt(max(eng.scal[0].item(), 0.0))
            tol_lin = self.linear_tol
            if self.forcing == 'ew':
                if nr_prev is None:
                    self._eta_cap = self.eta_max
                elif nr > 0.5 * nr_prev:
                    # the iteration stalls: later corrections are solved a hundred times tighter
                    self._eta_cap = max(self.linear_tol, 1e-2 * self._eta_cap)
                eta = self._ew_forcing(nr, nr_prev, eta, r0 if r0 is not None else nr)
                tol_lin = eta
                if r0 is not None and self._passes(nr, r0):
                    # this residual passes the test, so its correction is the last one applied
                    # (reference ordering): solve it to final_forcing, so that the accepted state
                    # lies well inside the tolerance rather than at it
                    tol_lin = min(tol_lin, 1e-3) if self.final_forcing is None else self.final_forcing
            eng.solve(tol_lin, self.linear_maxiter, ir_steps=self.ir_steps, method=self.krylov_method,
                      restart=self.krylov_restart, equilibrate=self.equilibrate)
            if (eng.precision == 'mixed' and self.mixed_fallback and not self.equilibrate
                    and not (eng.stats['krylov_relres'][-1] <= 0.1)):
                # the FP32 solve did not even reduce the residual tenfold: assemble the tangent and
                # the Jacobian in FP64 at the same iterate and solve this correction in FP64
                prev = eng.mixed_assembly
                eng.mixed_assembly = 'fp64'
                jac_fresh = False  # the matrices now hold the FP64 assembly
                try:
                    eng.stage(args, want_tangent=True)
                    eng.jacobian(x, lp)
                finally:
                    eng.mixed_assembly = prev
                eng.solve_fp64(tol_lin, self.linear_maxiter, method=self.krylov_method, restart=self.krylov_restart)
            if tm is not None:
                tm.mark('krylov')
            converged_here = self._passes(nr, r0)
            if self.line_search and not converged_here and it + 1 < self.n_iter:
                self._line_search(args, u_new, nr, after_residual)
            else:
                eng.axpy(1.0, eng.delta, u_new)
            if tm is not None:
                tm.mark('line_search')
            nr_prev = nr
            self.residual_history.append(nr)
            if self.verbose:
                print('residual.norm:', nr, 'krylov iterations:', eng.stats['krylov_iters'][-1],
                      'relres: %.2e' % eng.stats['krylov_relres'][-1], flush=True)
            if not math.isfinite(nr):
                raise L.GMError('Newton residual is not finite at iteration %d' % it)
            if r0 is None:
                r0 = nr
            passed = self._passes(nr, r0)
            if passed:
                break
        self.newton_iterations = it + 1
        self.newton_converged = bool(self.residual_history) and passed

    def _passes(self, nr, r0):
        """Newton test on the residual norm nr (r0: the first of the load step, None
        before it is known): absolute (reference, qs2d:230) or relative to r0, and
        with iteration_dict['abs_floor'] also passed once nr is below that floor (a
        relative test cannot pass when r0 itself is near the residual's round-off
        level)."""
        if self.abs_floor is not None and nr < self.abs_floor:
            return True
        if self.tol_mode == 'absolute':
            return nr < self.tol
        if r0 is None:
            return False
        return (nr / r0 if r0 > 0 else 0.0) < self.tol

    def _line_search(self, args, u_new, nr, after_residual, max_halvings=6):
        """Backtracking on the Newton correction (not in the reference, which
        always takes the full step): accept u + a*delta for the first
        a = 1, 1/2, ... whose residual norm is finite and below (1 - 1e-4 a) nr.
        The full step is evaluated with the particle tangent: when it is accepted,
        the next Newton iteration starts from its tangent, residual and norm
        (self._ls_ready) instead of evaluating them again."""
        eng = self.engine
        base = u_new.clone()
        a = 1.0
        finite = False
        accepted = False
        # the full step's tangent serves the next Newton Jacobian; none is assembled after a
        # correction solved to pass the Newton test when the last correction is a chord step
        full_tangent = not (self.final_chord and getattr(self, '_aims_to_pass', False))
        for _ in range(max_halvings + 1):
            u_new.copy_(base)
            eng.axpy(a, eng.delta, u_new)
            eng.stage(args, want_tangent=(a == 1.0 and full_tangent))
            eng.residual(self.x_particles, self.GIMP_lp)
            eng.norm2(eng.rhs, 0)
            trial = math.sqrt(max(eng.scal[0].item(), 0.0))
            finite = math.isfinite(trial)
            if finite and trial <= (1.0 - 1e-4 * a) * nr:
                accepted = True
                break
            if _ < max_halvings:
                a *= 0.5
        if not finite:
            u_new.copy_(base)
            raise L.GMError('line search: no finite residual along the Newton correction (down to step %g)' % a)
        if accepted and a == 1.0:
            self._ls_ready = trial
            self._ls_tangent = full_tangent
        self.line_search_steps.append(a)

    def _ew_forcing(self, nr, nr_prev, eta_prev, r0):
        """Eisenstat & Walker (1996) choice 2 forcing term with their safeguards,
        clamped to [linear_tol, forcing_max], and not tighter than needed to
        reach the Newton tolerance."""
        gamma, alpha = self.ew_gamma, self.ew_alpha
        if nr_prev is None or eta_prev is None:
            eta = min(1e-2, self.eta_max)
        else:
            eta = gamma * (nr / nr_prev) ** alpha
            sg = gamma * eta_prev ** alpha
            if sg > 0.1:
                eta = max(eta, sg)
        target = self.tol * r0 if self.tol_mode == 'relative' else self.tol
        cap = getattr(self, '_eta_cap', self.eta_max)
        self._aims_to_pass = False
        if nr > 0:
            eta = max(eta, 0.5 * target / nr)
            if nr * cap < target:
                # the residual is within 1/cap of the target: solve to half the target, so that
                # the next residual can pass the test
                eta = min(eta, 0.5 * target / nr)
                self._aims_to_pass = 0.5 * target / nr >= self.linear_tol
        return min(max(eta, self.linear_tol), cap)

    # ------------------------------------------------------------------ checkpoint / restart
    _STATE = ('x_particles', 'GIMP_lp', 'deformation_gradient_total_old', 'deformation_gradient_total_new',
              'left_Cauchy_Green_old', 'left_Cauchy_Green_new', 'particle_Cauchy_stress_array',
              'plastic_strain_old', 'plastic_strain_new')
    _STATE_EXTRA = ()

    def state_dict(self):
        """Device copies of the particle state between load steps (positions, lp,
        F and b_e old/new, stress, hardening variable and the form's extra
        arrays, e.g. pore pressures): enough to restart the load path from here
        with load_state_dict (the nodal state is rebuilt by every load step's
        P2G)."""
        return {k: getattr(self, k).detach().clone() for k in self._STATE + self._STATE_EXTRA}

    def load_state_dict(self, state):
        for k in self._STATE + self._STATE_EXTRA:
            getattr(self, k).copy_(state[k])

    # ------------------------------------------------------------------ dof-major views
    def _sync_flags_out(self):
        eng = self.engine
        self.dofStruct.activate_flag_array.copy_(eng.dofmajor(eng.act).bool())
        self.dofStruct.boundary_flag_array.copy_(eng.dofmajor(eng.bflag).bool())

    @property
    def rhs(self):
        return self.engine.dofmajor(self.engine.rhs)

    @property
    def increment_iteration(self):
        return self.engine.dofmajor(self.engine.delta)

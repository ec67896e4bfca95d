"""quasi_static_coupled_solver_3d: drop-in for GeoWarp
solvers/quasi_static_coupled_solver_3d.py:56-420.

3D u-p (4 dofs per node) implicit cpGIMP MPM with a penalty rigid indenter.
``set_boundary_dofs`` is a Python callable
``f(boundary_flag_array, n_grid_x, n_grid_y, n_nodes, dx, dy, end_x, end_y)``
(the reference launches a Warp kernel over (nx+1, ny+1, nz+1), qc3d:400-402).
G2P and the lp update run at the end of ``newton_iter`` as in the reference
(qc3d:378-386).  ``indentation_force_array`` grows to the number of load steps
(the reference fixes it at 25 entries, qc3d:162).
"""
import numpy as np
import torch

from .. import _lib as L
from ._base import _QuasiStaticBase, init_particles_rectangle
from ._engine import _stream

TARGET_FORCE = 13567.27337196  # qc3d:52, "from E = 1e4 kPa"


def get_indentation_loss(indentation_force, footing_incr_disp, total_steps, target=TARGET_FORCE):
    """get_indentation_loss (qc3d:45-52)."""
    return (indentation_force / (abs(footing_incr_disp) * total_steps) - target) ** 2


class quasi_static_coupled_solver_3d(_QuasiStaticBase):
    dim = 3
    coupled = True
    _STATE_EXTRA = ('particle_pressure_array',)

    def __init__(self, background_grid_dict, particles_dict, iteration_dict, material_dict, porous_media_dict,
                 load_dict, set_boundary_dofs, device=None, x_particles=None):
        g = background_grid_dict
        self.max_y, self.max_z = g.get('max_y'), g.get('max_z')
        self.dy, self.dz = g['dy'], g['dz']
        self.inv_dy, self.inv_dz = 1.0 / self.dy, 1.0 / self.dz
        self.start_z, self.end_z = g['start_z'], g['end_z']
        self.domain_min_z, self.domain_max_z = g['domain_min_z'], g['domain_max_z']
        self.GIMP_lp_initial = g['dx'] / particles_dict['PPD'] / 2.0  # qc3d:100 (no -1e-6 in 3D)
        self._common_init(background_grid_dict, particles_dict, iteration_dict, material_dict, load_dict, device,
                          x_particles)
        self.p_mass = self.p_vol * self.p_rho
        self.dt = iteration_dict['dt']
        self.phi_initial = porous_media_dict['phi_initial']
        self.mobility_constant = porous_media_dict['mobility_constant']
        l = load_dict
        self.footing_region_min_x = l['footing_region_min_x']
        self.footing_region_max_x = l['footing_region_max_x']
        self.footing_region_min_y = l['footing_region_min_y']
        self.footing_region_max_y = l['footing_region_max_y']
        self.footing_initial_height = l['footing_initial_height']
        self.footing_incr_disp = l['footing_incr_disp']
        self.penalty_factor = l['penalty_factor']
        # 'reference': r_z += -(h - (z + du_z)) k w_I (mpm_utils.py:1240-1246), a force pushing a
        # penetrating particle up and a negative stiffness -k w_I w_J in the Newton matrix;
        # 'physical': the textbook penalty r_z += (h - (z + du_z)) k w_I (push down, stiffness +k)
        self.penalty_form = l.get('penalty_form', 'reference')
        if self.penalty_form not in ('reference', 'physical'):
            raise ValueError("penalty_form must be 'reference' or 'physical'")
        self.set_boundary_dofs = set_boundary_dofs
        # multigrid hierarchy from the Newton matrix with the indenter penalty's sign reversed
        # (the Krylov operator stays the exact matrix); False builds it from the matrix itself
        self.engine.penalty_pc = iteration_dict.get('penalty_pc', True)
        # adaptive load-step splitting on Newton failure (0: off, the reference's behaviour)
        self.max_substeps = int(iteration_dict.get('substeps', 0))
        # d loss / d E of every converged load step at its fixed initial state (the adjoint of
        # the load step), in E_grad_step; the explicit part alone is E_grad_step_explicit
        self.step_sensitivity = bool(iteration_dict.get('step_sensitivity', False))
        self.E_grad_step = None
        self.E_grad_step_explicit = None
        self.substep_count = 0
        self.step_newton_iterations = 0
        dev = self.device
        n = self.n_particles
        self.particle_pressure_array = torch.zeros(n, dtype=torch.float64, device=dev)
        self.particle_traction_flag_array = torch.zeros(n, dtype=torch.bool, device=dev)
        self.particle_pressure_boundary_flag_array = torch.zeros(n, dtype=torch.bool, device=dev)
        self._old = torch.zeros(self.n_matrix_size, dtype=torch.float64, device=dev)
        self._new = torch.zeros(self.n_matrix_size, dtype=torch.float64, device=dev)
        self.youngs_modulus_diff = torch.tensor([self.youngs_modulus], dtype=torch.float64, device=dev)
        self.indentation_force = torch.zeros(1, dtype=torch.float64, device=dev)
        self.indentation_force_array = torch.zeros(25, dtype=torch.float64, device=dev)
        self.indentation_force_list = np.empty(0)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)
        self.E_grad = None

    @property
    def old_solution(self):
        return self.engine.dofmajor(self._old)

    @property
    def new_solution(self):
        return self.engine.dofmajor(self._new)

    # ---------------------------------------------------------------- reference API
    def reinit_everything(self):
        """qc3d:206-221."""
        start = (self.start_x, self.start_y, self.start_z)
        end = (self.end_x, self.end_y, self.end_z)
        x = init_particles_rectangle(3, start, end, self.dx, (self.n_grid_x, self.n_grid_y, self.n_grid_z), self.PPD)
        self.x_particles.copy_(torch.tensor(x, dtype=torch.float64, device=self.device))
        self.E_grad = None
        self.indentation_force_list = np.empty(0)
        self.particle_pressure_array.zero_()
        self._initialization()

    def reset_grid(self, dx, dy, dz):
        """qc3d:223-230."""
        self.dx, self.dy, self.dz = dx, dy, dz
        self.inv_dx, self.inv_dy, self.inv_dz = 1.0 / dx, 1.0 / dy, 1.0 / dz

    def reset_youngs_modulus(self, youngs_modulus):
        """qc3d:232-239."""
        self.youngs_modulus = youngs_modulus
        self._set_lame()
        self.youngs_modulus_diff.fill_(youngs_modulus)

    def reset(self):
        """qc3d:242-255."""
        self._old.zero_()
        self._new.zero_()
        self.engine.delta.zero_()
        self.dofStruct.boundary_flag_array.zero_()
        self.dofStruct.activate_flag_array.zero_()

    def _sync_grid(self):
        # drivers may move the mesh between steps (3d_indentation_coupled_cpGIMP.py:202-207)
        g = self.engine.grid
        hs = (self.dx, self.dy, self.dz)
        for a in range(3):
            g.h[a] = hs[a]
            g.inv_h[a] = 1.0 / hs[a]

    def _args(self, current_step, total_steps):
        a = self._stage_args(current_step, total_steps, self._new, self._old)
        a.phi0 = self.phi_initial
        a.mobility = self.mobility_constant
        a.dt = self.dt
        a.footing[0] = self.footing_region_min_x
        a.footing[1] = self.footing_region_max_x
        a.footing[2] = self.footing_region_min_y
        a.footing[3] = self.footing_region_max_y
        a.footing_height = self.footing_initial_height + (float(current_step) + 1.0) * self.footing_incr_disp
        sign = 1.0 if self.penalty_form == 'reference' else -1.0
        a.pen_par = sign * self.youngs_modulus * self.penalty_factor * (self.dx / self.PPD) * (self.dy / self.PPD)
        a.penalty = 1
        return a

    def newton_iter(self, current_step, total_steps, force_index=None):
        """qc3d:258-392 (Newton loop, then G2P and the lp update).  force_index is
        the indentation_force_array entry (default int(current_step); a sub-step of
        load step s writes entry s)."""
        eng = self.engine
        step = int(current_step) if force_index is None else int(force_index)
        if step >= self.indentation_force_array.numel():
            grown = torch.zeros(step + 1, dtype=torch.float64, device=self.device)
            grown[:self.indentation_force_array.numel()] = self.indentation_force_array
            self.indentation_force_array = grown
        last = float(current_step) == float(total_steps) - 1

        def after_residual():
            eng.sum_into(eng.pen_force, 1)
            f = eng.scal[1:2]
            self.indentation_force.copy_(f)
            self.indentation_force_array[step] = f[0]
            self.loss.copy_(get_indentation_loss(f, self.footing_incr_disp, float(total_steps)))
            if last:
                # explicit d loss / d E at fixed u: the force is linear in E through the penalty (qc3d:323-325)
                F = f[0]
                S = abs(self.footing_incr_disp) * float(total_steps)
                self.E_grad = (2.0 * (F / S - TARGET_FORCE) * F / (self.youngs_modulus * S)).reshape(1).cpu().numpy()

        args = self._args(current_step, total_steps)
        eng.penalty = dict(footing=(self.footing_region_min_x, self.footing_region_max_x, self.footing_region_min_y,
                                    self.footing_region_max_y), height=args.footing_height, pen_par=args.pen_par,
                           lp_max=float(self.GIMP_lp.max()))
        self._newton(current_step, total_steps, self._new, self._old, args, after_residual=after_residual)
        if self.step_sensitivity:
            self._step_E_sensitivity(total_steps, args)
        self._update_particles()
        self.indentation_force_list = np.append(self.indentation_force_list, float(self.indentation_force.item()))

    def _step_E_sensitivity(self, total_steps, args, tol=1e-12):
        """Adjoint of the converged load step for the indentation loss L = (F/S - target)^2
        (qc3d:45-52) with respect to Young's modulus, at the step's fixed initial state:
        dL/dE = dL/dE|_u + lambda . dR/dE with K^T lambda = dL/du (K = -dR/du on the free
        dofs).  dL/dE|_u is the closed form of the explicit gradient (the force is linear in
        E through the penalty, qc3d:323-325); dL/du = dL/dF dF/du with dF/du_{J,z} =
        -pen_par sum_p w_J over the penalised particles; dR/dE gathers the particle rows of
        gm_stage_de_3d.  Both vectors are gathered by the residual kernel itself with a
        substituted particle stress vector.  FP64 form and Neo-Hookean skeleton only (there
        dR/dE follows from the stress being proportional to E at fixed deformation)."""
        eng = self.engine
        if eng.precision != 'fp64' or self.material_name != 'Neo-Hookean':
            raise L.GMError('the load-step sensitivity needs the FP64 form and the Neo-Hookean skeleton')
        x, lp = self.x_particles, self.GIMP_lp
        n, NY = self.n_particles, eng.NY
        # stage, residual and Jacobian at the converged unknowns (after the last update)
        eng.stage(args, want_tangent=True)
        eng.residual(x, lp)
        eng.sum_into(eng.pen_force, 1)
        F = float(eng.scal[1].item())
        eng.jacobian(x, lp)
        S = abs(self.footing_incr_disp) * float(total_steps)
        dLdF = 2.0 * (F / S - TARGET_FORCE) / S
        E = float(self.youngs_modulus)
        h = float(args.footing_height)
        inside = ((x[:, 2] >= h) & (x[:, 0] > self.footing_region_min_x) & (x[:, 0] < self.footing_region_max_x)
                  & (x[:, 1] > self.footing_region_min_y) & (x[:, 1] < self.footing_region_max_y))
        order = eng.sorted[:n].long()
        Y = eng.Y
        try:
            yg = torch.zeros(n * NY, dtype=torch.float64, device=self.device)
            yg.view(n, NY)[:, 11] = inside[order].to(torch.float64) * (-float(args.pen_par) * dLdF)
            eng.Y = yg
            eng.residual(x, lp)
            g = eng.rhs.clone()
            dY = torch.empty(n * NY, dtype=torch.float64, device=self.device)
            L.check(L.lib().gm_stage_de_3d(n, L.ptr(eng.sorted), L.ptr(self.deformation_gradient_total_old),
                                        L.ptr(self.deformation_gradient_total_new),
                                        L.ptr(self.particle_Cauchy_stress_array), L.ptr(eng.pen_force), L.ptr(Y),
                                        float(self.p_vol), E, L.ptr(dY), _stream()), 'gm_stage_de_3d')
            eng.Y = dY
            eng.residual(x, lp)
            rE = eng.rhs.clone()
        finally:
            eng.Y = Y
        # K^T lambda = g on the FP64 Jacobian (block transpose; a full hierarchy for K^T)
        nb = eng.nb
        vt = torch.empty(eng.nnzb * nb * nb, dtype=torch.float64, device=self.device)
        L.check(L.lib().gm_bsr_transpose(nb, 0, eng.n_nodes, L.ptr(eng.row_ptr), L.ptr(eng.col_idx), L.ptr(eng.vals),
                                         L.ptr(vt), _stream()), 'gm_bsr_transpose')
        vals, lag = eng.vals, eng.mg_lag
        eng.vals, eng.mg_lag = vt, False
        eng.jac_version += 1
        lam = torch.zeros_like(g)
        try:
            its, rel = eng.krylov(g, lam, False, tol, self.linear_maxiter, self.krylov_method if
                                  self.krylov_method in ('mg', 'gmres') else 'gmres', self.krylov_restart)
        finally:
            eng.vals, eng.mg_lag = vals, lag
            eng.jac_version += 1
        self.E_grad_step_explicit = dLdF * F / E
        self.E_grad_step = self.E_grad_step_explicit + float(torch.dot(lam, rE).item())
        self.sensitivity_stats = dict(krylov_iters=its, relres=rel, force=F)

    def _set_E(self, E):
        self.youngs_modulus = float(E)
        self._set_lame()

    def _update_particles(self):
        """G2P and the particle-domain update at the current nodal solution (the end of
        newton_iter, qc3d:378-392)."""
        eng = self.engine
        eng.g2p(self.x_particles, self.GIMP_lp, self._new, self._old, self.deformation_gradient_total_old,
                self.deformation_gradient_total_new, self.left_Cauchy_Green_old, self.left_Cauchy_Green_new,
                self.plastic_strain_old, self.plastic_strain_new, self.particle_pressure_array)
        eng.update_lp(self.deformation_gradient_total_new, self.GIMP_lp, self.GIMP_lp_initial, self.x_particles,
                      (self.domain_min_x, self.domain_min_y, self.domain_min_z),
                      (self.domain_max_x, self.domain_max_y, self.domain_max_z))

    def path_E_gradient(self, n_steps, total_steps, h_rel=1e-6, tol=1e-12):
        """Total derivative of the indentation loss after load steps 0 .. n_steps-1 with
        respect to Young's modulus through the whole load path, i.e. including the
        particle state (positions, domains, deformation gradients, stresses, pore
        pressures) every load step hands to the next; the reference's tape covers only
        the last step's residual (qc3d:322-325, 3d_indentation_coupled_cpGIMP.py:195-232).

        Forward (tangent-linear) sensitivity along the forward run, from the current
        state: for every load step k with initial state s and converged nodal solution u,
            K du/dE = dR/ds . ds/dE + dR/dE          (K = -dR/du, converged FP64 Jacobian)
            ds'/dE = dG/du . du/dE + dG/ds . ds/dE + dG/dE
        with G the load step's particle update (P2G projection, stage, G2P, domain
        update).  The products with dR/ds, dR/dE and with G's derivatives are central
        differences of the device residual and update in the tangent's own direction
        (E + t, s + t ds/dE, u + t du/dE, t = h_rel E): one pair of evaluations each, at
        the cost of O(t^2) truncation.  At the last step the indentation force is
        differentiated the same way.  FP64 form; each step must converge in one part.
        Returns dL/dE (also in self.E_grad_path, the force derivative in
        self.path_force_derivative) and leaves the solver at the end of the path."""
        eng = self.engine
        if eng.precision != 'fp64':
            raise L.GMError('the load-path sensitivity needs the FP64 form')
        E = float(self.youngs_modulus)
        t = h_rel * E
        keys = [k for k in self._STATE + self._STATE_EXTRA if getattr(self, k).dtype == torch.float64]
        ds = {k: torch.zeros_like(getattr(self, k)) for k in keys}
        x, lp = self.x_particles, self.GIMP_lp

        def shifted(state, sg):
            return {k: (v + (sg * t) * ds[k]) if k in ds else v for k, v in state.items()}

        def at(k, state, sg, u):
            # perturbed initial state and modulus, the step's P2G projection, nodal solution u
            self._set_E(E + sg * t)
            self.load_state_dict(shifted(state, sg))
            self.prepare_step(k, total_steps)
            self._new.copy_(u)
            return self._args(k, total_steps)

        dF = 0.0
        its_total = 0
        warm = self.warm_mass
        for k in range(n_steps):
            s0 = self.state_dict()
            self.advance_one_step(k, total_steps)
            if not self.newton_converged or self.substep_count != 1:
                raise L.GMError('load-path sensitivity: load step %d did not converge in one part' % k)
            u = self._new.clone()
            s1 = self.state_dict()
            force = float(self.indentation_force.item())
            self.warm_mass = False  # every evaluation below projects from x0 = 0, the same way
            try:
                # dR/ds . ds/dE + dR/dE at u
                r = []
                for sg in (1.0, -1.0):
                    args = at(k, s0, sg, u)
                    eng.stage(args, want_tangent=False)
                    eng.residual(x, lp)
                    r.append(eng.rhs.clone())
                rhs = (r[0] - r[1]) / (2.0 * t)
                # the converged Jacobian at the unperturbed state
                args = at(k, s0, 0.0, u)
                eng.stage(args, want_tangent=True)
                eng.residual(x, lp)
                eng.jacobian(x, lp)
                du = torch.zeros_like(rhs)
                its, rel = eng.krylov(rhs, du, False, tol, self.linear_maxiter,
                                      self.krylov_method if self.krylov_method in ('mg', 'gmres') else 'gmres',
                                      self.krylov_restart)
                its_total += its
                # the state the step hands on, differentiated along (E, s, u)
                g = []
                for sg in (1.0, -1.0):
                    args = at(k, s0, sg, u + (sg * t) * du)
                    eng.stage(args, want_tangent=False)
                    if k == n_steps - 1:
                        eng.residual(x, lp)
                        eng.sum_into(eng.pen_force, 1)
                        g.append(float(eng.scal[1].item()))
                    self._update_particles()
                    g.append({kk: getattr(self, kk).clone() for kk in keys})
                if k == n_steps - 1:
                    dF = (g[0] - g[2]) / (2.0 * t)
                    g = [g[1], g[3]]
                ds = {kk: (g[0][kk] - g[1][kk]) / (2.0 * t) for kk in keys}
            finally:
                self._set_E(E)
                self.load_state_dict(s1)
                self.warm_mass = warm
        S = abs(self.footing_incr_disp) * float(total_steps)
        self.path_force_derivative = dF
        self.E_grad_path = 2.0 * (force / S - TARGET_FORCE) / S * dF
        self.path_stats = dict(krylov_iters=its_total, force=force)
        return self.E_grad_path

    def prepare_step(self, current_step, total_steps):
        """reset, boundary callback, P2G and consistent-mass projection (qc3d:397-417).
        The projection starts from the nodal pressures the solver last held (the previous
        load step's, from which the particle pressures were interpolated)."""
        x0 = self._new.view(-1, self.engine.nb)[:, -1].clone() if self.warm_mass else None
        self.reset()
        self._sync_grid()
        eng = self.engine
        self.set_boundary_dofs(self.dofStruct.boundary_flag_array, self.n_grid_x, self.n_grid_y, self.n_nodes,
                               self.dx, self.dy, self.end_x, self.end_y)
        eng.begin_step(self.x_particles, self.GIMP_lp)
        eng.set_boundary_from_dofmajor(self.dofStruct.boundary_flag_array)
        pbc = self.particle_pressure_boundary_flag_array.to(torch.uint8)
        mrhs = eng.p2g(self.x_particles, self.GIMP_lp, 1e-7, pbc, self.dz, self.p_mass, self.particle_pressure_array,
                       mass=True)
        self._sync_flags_out()
        xp = eng.solve_mass(mrhs, tol=self.mass_tol, x0=x0)
        eng.set_nodal_pressure(xp, self._old, self._new, require_active=True)

    def advance_one_step(self, current_step, total_steps):
        """qc3d:396-420.  With iteration_dict['substeps'] = d > 0 (not in the
        reference, which moves on after n_iter iterations), a load step whose
        Newton iteration does not converge is restarted from its initial state and
        split into two halves, recursively down to 2^-d of the step: each part is a
        full load step (P2G, Newton, G2P) to the load of its end point (the load
        factor (s + f)/total and the indenter height h0 + (s + f) * incr are those
        of a fractional current_step s - 1 + f) with the time step dt scaled by
        the part's length.  substep_count and step_newton_iterations report the
        parts and Newton iterations the step took."""
        self.substep_count = 0
        self.step_newton_iterations = 0
        if not self.max_substeps:
            self.prepare_step(current_step, total_steps)
            self.newton_iter(current_step, total_steps)
            self.substep_count, self.step_newton_iterations = 1, self.newton_iterations
            return
        s = float(current_step)
        self._advance_range(s - 1.0, s, total_steps, 0, int(current_step))

    def _advance_range(self, a, b, total_steps, depth, step_index):
        snap = self.state_dict() if depth < self.max_substeps else None
        dt_full = self.dt
        self.dt = dt_full * (b - a)
        try:
            self.prepare_step(b, total_steps)
            self.newton_iter(b, total_steps, force_index=step_index)
        finally:
            self.dt = dt_full
        self.substep_count += 1
        self.step_newton_iterations += self.newton_iterations
        if self.newton_converged or snap is None:
            return
        self.load_state_dict(snap)
        m = 0.5 * (a + b)
        self._advance_range(a, m, total_steps, depth + 1, step_index)
        if self.newton_converged:
            self._advance_range(m, b, total_steps, depth + 1, step_index)

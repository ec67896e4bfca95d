"""quasi_static_coupled_solver_2d: drop-in for GeoWarp
solvers/quasi_static_coupled_solver_2d.py:45-313.

2D plane-strain u-p (3 dofs per node) implicit cpGIMP MPM with PPP-stabilised
mass balance and consistent-mass pressure projection at every load step.
``set_boundary_dofs`` is a Python callable
``f(boundary_flag_array, n_grid_x, n_nodes)`` that sets dof-major Dirichlet
flags on a torch bool tensor (the reference launches a Warp kernel with the same
inputs over (nx+1, ny+1), qc2d:278-280).
"""
import torch

from ._base import _QuasiStaticBase


class quasi_static_coupled_solver_2d(_QuasiStaticBase):
    dim = 2
    coupled = True
    _STATE_EXTRA = ('particle_pressure_array',)

    def __init__(self, background_grid_dict, particles_dict, iteration_dict, material_dict, porous_media_dict,
                 load_dict, set_boundary_dofs, device=None, x_particles=None):
        g = background_grid_dict
        self.max_y = g.get('max_y')
        self.dy = g['dy']
        self.inv_dy = 1.0 / self.dy
        self.GIMP_lp_initial = g['dx'] / particles_dict['PPD'] / 2.0 - 1e-6  # qc2d:81
        self._common_init(background_grid_dict, particles_dict, iteration_dict, material_dict, load_dict, device,
                          x_particles)
        self.p_mass = self.p_vol * self.p_rho
        self.dt = iteration_dict['dt']
        self.phi_initial = porous_media_dict['phi_initial']
        self.mobility_constant = porous_media_dict['mobility_constant']
        l = load_dict
        self.traction_value_x = l.get('traction_value_x', 0.0)
        self.traction_value_y = l.get('traction_value_y', 0.0)
        self.point_load_value_x = l.get('point_load_value_x', 0.0)
        self.point_load_value_y = l.get('point_load_value_y', 0.0)
        self.set_boundary_dofs = set_boundary_dofs
        dev = self.device
        n = self.n_particles
        self.particle_pressure_array = torch.zeros(n, dtype=torch.float64, device=dev)
        self.particle_traction_flag_array = torch.zeros(n, dtype=torch.bool, device=dev)
        self.particle_pressure_boundary_flag_array = torch.zeros(n, dtype=torch.bool, device=dev)
        self._old = torch.zeros(self.n_matrix_size, dtype=torch.float64, device=dev)
        self._new = torch.zeros(self.n_matrix_size, dtype=torch.float64, device=dev)

    @property
    def old_solution(self):
        return self.engine.dofmajor(self._old)

    @property
    def new_solution(self):
        return self.engine.dofmajor(self._new)

    def reset(self):
        """qc2d:164-175."""
        self._old.zero_()
        self._new.zero_()
        self.engine.delta.zero_()
        self.dofStruct.boundary_flag_array.zero_()
        self.dofStruct.activate_flag_array.zero_()

    def _args(self, current_step, total_steps):
        a = self._stage_args(current_step, total_steps, self._new, self._old)
        s1 = (float(current_step) + 1.0) / float(total_steps)
        self._trac_u8 = self.particle_traction_flag_array.to(torch.uint8)
        a.trac_flag = self._trac_u8.data_ptr()
        a.point_load[0] = self.point_load_value_x * s1
        a.point_load[1] = self.point_load_value_y * s1
        # traction is a per-particle line load t*dx/PPD, not ramped (mpm_utils.py:972-974)
        a.traction[0] = self.traction_value_x * self.dx / self.PPD
        a.traction[1] = self.traction_value_y * self.dx / self.PPD
        a.phi0 = self.phi_initial
        a.mobility = self.mobility_constant
        a.dt = self.dt
        return a

    def newton_iter(self, current_step, total_steps):
        """qc2d:179-266."""
        self._newton(current_step, total_steps, self._new, self._old, self._args(current_step, total_steps))

    def prepare_step(self, current_step, total_steps):
        """reset, boundary callback, P2G and consistent-mass projection (qc2d:275-300).
        The projection starts from the nodal pressures the solver last held (the previous
        load step's, from which the particle pressures were interpolated)."""
        x0 = self._new.view(-1, self.engine.nb)[:, -1].clone() if self.warm_mass else None
        self.reset()
        eng = self.engine
        self.set_boundary_dofs(self.dofStruct.boundary_flag_array, self.n_grid_x, self.n_nodes)
        eng.begin_step(self.x_particles, self.GIMP_lp)
        eng.set_boundary_from_dofmajor(self.dofStruct.boundary_flag_array)
        pbc = self.particle_pressure_boundary_flag_array.to(torch.uint8)
        mrhs = eng.p2g(self.x_particles, self.GIMP_lp, 1e-6, pbc, self.dy, self.p_mass, self.particle_pressure_array,
                       mass=True)
        self._sync_flags_out()
        xp = eng.solve_mass(mrhs, tol=self.mass_tol, x0=x0)
        eng.set_nodal_pressure(xp, self._old, self._new, require_active=False)

    def advance_one_step(self, current_step, total_steps):
        """qc2d:274-313."""
        self.prepare_step(current_step, total_steps)
        eng = self.engine
        self.newton_iter(current_step, total_steps)
        eng.g2p(self.x_particles, self.GIMP_lp, self._new, self._old, self.deformation_gradient_total_old,
                self.deformation_gradient_total_new, self.left_Cauchy_Green_old, self.left_Cauchy_Green_new,
                self.plastic_strain_old, self.plastic_strain_new, self.particle_pressure_array)
        eng.update_lp(self.deformation_gradient_total_new, self.GIMP_lp, self.GIMP_lp_initial, self.x_particles,
                      (self.domain_min_x, self.domain_min_y), (self.domain_max_x, self.domain_max_y))

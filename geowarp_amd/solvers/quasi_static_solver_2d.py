"""quasi_static_solver_2d: drop-in for GeoWarp solvers/quasi_static_solver_2d.py:45-258.

2D plane-strain single-phase implicit cpGIMP MPM.  Unknown: nodal displacement
increment since the start of the load step.  Materials: 'Hencky elasticity',
'J2' (reference) and 'Neo-Hookean', 'Drucker-Prager' (extensions).
"""
import torch

from ._base import _QuasiStaticBase


class quasi_static_solver_2d(_QuasiStaticBase):
    dim = 2
    coupled = False

    def __init__(self, background_grid_dict, particles_dict, iteration_dict, material_dict, load_dict,
                 device=None, x_particles=None):
        self.GIMP_lp_initial = background_grid_dict['dx'] / particles_dict['PPD'] / 2.0 - 1e-6  # qs2d:75
        self._common_init(background_grid_dict, particles_dict, iteration_dict, material_dict, load_dict, device,
                          x_particles)
        l = load_dict
        self.traction_value_x = l.get('traction_value_x', 0.0)
        self.traction_value_y = l.get('traction_value_y', 0.0)
        self.point_load_value_x = l.get('point_load_value_x', 0.0)
        self.point_load_value_y = l.get('point_load_value_y', 0.0)
        # optional driver hook for non-zero Dirichlet increments (e.g. a displacement-controlled platen):
        # f(x_particles, GIMP_lp, boundary_flag_array, prescribed, n_grid_x, n_nodes, dx, current_step,
        #   total_steps) sets boundary flags and writes prescribed increments into the dof-major `prescribed`
        self.prescribed_increment = l.get('prescribed_increment')
        self._inc = torch.zeros(self.n_matrix_size, dtype=torch.float64, device=self.device)  # node-interleaved

    @property
    def increment_solution(self):
        return self.engine.dofmajor(self._inc)

    def reset(self):
        """qs2d:144-147: zero the increment and the activation flags (boundary flags are kept)."""
        self._inc.zero_()
        self.dofStruct.activate_flag_array.zero_()

    def _args(self, current_step, total_steps):
        a = self._stage_args(current_step, total_steps, self._inc, None)
        s1 = (float(current_step) + 1.0) / float(total_steps)
        a.point_load[0] = self.point_load_value_x * s1
        a.point_load[1] = self.point_load_value_y * s1
        return a

    def newton_iter(self, current_step, total_steps):
        """qs2d:151-231."""
        self._newton(current_step, total_steps, self._inc, None, self._args(current_step, total_steps))

    def prepare_step(self, current_step, total_steps):
        """reset + P2G of advance_one_step (qs2d:240-245)."""
        self.reset()
        eng = self.engine
        pres = None
        if self.prescribed_increment is not None:
            pres = torch.zeros(self.n_matrix_size, dtype=torch.float64, device=self.device)
            self.prescribed_increment(self.x_particles, self.GIMP_lp, self.dofStruct.boundary_flag_array, pres,
                                      self.n_grid_x, self.n_nodes, self.dx, current_step, total_steps)
        eng.begin_step(self.x_particles, self.GIMP_lp)
        eng.set_boundary_from_dofmajor(self.dofStruct.boundary_flag_array)
        eng.p2g(self.x_particles, self.GIMP_lp, -1.0)  # P2G_2d: every support node, no threshold
        self._prescribed = pres
        if pres is not None:
            self._inc.copy_(eng.from_dofmajor(pres * self.dofStruct.boundary_flag_array))
        self._sync_flags_out()

    def advance_one_step(self, current_step, total_steps):
        """qs2d:239-258: reset, P2G activation, Newton, G2P, lp update."""
        self.prepare_step(current_step, total_steps)
        self.newton_iter(current_step, total_steps)
        self.finish_step()

    def finish_step(self):
        """G2P and the lp update that close a load step (qs2d:247-258)."""
        eng = self.engine
        eng.g2p(self.x_particles, self.GIMP_lp, self._inc, None, self.deformation_gradient_total_old,
                self.deformation_gradient_total_new, self.left_Cauchy_Green_old, self.left_Cauchy_Green_new,
                self.plastic_strain_old, self.plastic_strain_new, None)
        eng.update_lp(self.deformation_gradient_total_new, self.GIMP_lp, self.GIMP_lp_initial, self.x_particles,
                      (self.domain_min_x, self.domain_min_y), (self.domain_max_x, self.domain_max_y))

"""Load-step Newton loops (oracle restatement, test infrastructure only).

Mirrors the control flow of the reference solver classes:
  quasi_static_solver_2d          solvers/quasi_static_solver_2d.py:45-258
  quasi_static_coupled_solver_2d  solvers/quasi_static_coupled_solver_2d.py:45-313
  quasi_static_coupled_solver_3d  solvers/quasi_static_coupled_solver_3d.py:56-420

The Jacobian is built by coloured central finite differences with the
reference's distance-5 node colouring (sparse_differentiation.py:8-210) and its
sign/masking semantics (vals = -dR/du, entries only between free & active dofs,
identity on the rest: sparse_differentiation.py:213-401, mpm_utils.py:1299-1341).
The linear solve is scipy.sparse.linalg.spsolve (qs2d:214-220).

Newton ordering follows qs2d:151-231: evaluate r_k, solve, apply delta, then
test the *pre-update* norm.  ``tol_mode='relative'`` tests ||r_k||/||r_0||.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from . import gimp as og
from . import mpm as om


def colour_columns(n_grid, dim):
    """Node colour classes (ix%5, iy%5[, iz%5]) -- pick_grid_nodes_2d/3d
    (sparse_differentiation.py:8-56) started from every node of the first 5^d
    block (precompute_seed_vectors_*, :59-210)."""
    nx, ny = n_grid[0], n_grid[1]
    if dim == 2:
        ix, iy = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing='xy')
        node = ix + iy * (nx + 1)
        cls = (ix % 5) + 5 * (iy % 5)
        return node.ravel(), cls.ravel(), 25
    nz = n_grid[2]
    iz, iy, ix = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing='ij')
    node = ix + iy * (nx + 1) + iz * (nx + 1) * (ny + 1)
    cls = (ix % 5) + 5 * (iy % 5) + 25 * (iz % 5)
    return node.ravel(), cls.ravel(), 125


def node_coords(n_grid, dim, n_nodes):
    nx, ny = n_grid[0], n_grid[1]
    nodes = np.arange(n_nodes)
    ix = nodes % (nx + 1)
    if dim == 2:
        iy = nodes // (nx + 1)
        return np.stack([ix, iy], 1)
    iy = (nodes // (nx + 1)) % (ny + 1)
    iz = nodes // ((nx + 1) * (ny + 1))
    return np.stack([ix, iy, iz], 1)


def fd_jacobian(resfun, u, n_grid, dim, nb, n_nodes, boundary, act, steps):
    """Coloured central-FD Jacobian of resfun at u, returned as the reference's
    matrix: -dR/du between free & active dofs plus identity on the others."""
    nodes, cls, ncls = colour_columns(n_grid, dim)
    free = (~boundary) & act
    coords = node_coords(n_grid, dim, n_nodes)
    rows_all, cols_all, vals_all = [], [], []
    for k in range(nb):
        h = steps[k]
        for c in range(ncls):
            sel = nodes[cls == c]
            cdofs = sel + k * n_nodes
            cdofs = cdofs[free[cdofs]]
            if cdofs.size == 0:
                continue
            e = np.zeros_like(u)
            e[cdofs] = h
            dr = (resfun(u + e) - resfun(u - e)) / (2.0 * h)
            # every row sees at most one column of this colour within +-2 nodes
            rdofs = np.nonzero(free & (dr != 0.0))[0]
            if rdofs.size == 0:
                continue
            rnode = rdofs % n_nodes
            rc = coords[rnode]
            cc = np.zeros_like(rc)
            for a in range(dim):
                target = (c // (5 ** a)) % 5
                cc[:, a] = rc[:, a] - 2 + ((target - (rc[:, a] - 2)) % 5)
            if dim == 2:
                cnode = cc[:, 0] + cc[:, 1] * (n_grid[0] + 1)
            else:
                cnode = cc[:, 0] + cc[:, 1] * (n_grid[0] + 1) + cc[:, 2] * (n_grid[0] + 1) * (n_grid[1] + 1)
            cdof = cnode + k * n_nodes
            rows_all.append(rdofs)
            cols_all.append(cdof)
            vals_all.append(-dr[rdofs])
    n = nb * n_nodes
    fix = np.nonzero(~free)[0]
    rows_all.append(fix)
    cols_all.append(fix)
    vals_all.append(np.ones(fix.size))
    return sp.coo_matrix((np.concatenate(vals_all), (np.concatenate(rows_all), np.concatenate(cols_all))),
                         shape=(n, n)).tocsr()


class _NewtonMixin:
    tol_mode = 'absolute'
    # central-difference step of the coloured Jacobian, relative to dx (displacement dofs)
    fd_rel = 1e-6

    def first_jacobian(self, current_step, total_steps):
        """Reference-convention matrix (-dR/du, masked, identity rows) and
        residual at the first Newton iterate of a load step (state is left
        as after the P2G of that step)."""
        st, f, u0, steps = self.prepare_step(current_step, total_steps)
        r = f(u0, commit=False)
        K = fd_jacobian(lambda v: f(v, commit=False), u0, self.n_grid, self.dim, self.nb, self.n_nodes,
                        self.boundary, self.act, steps)
        return K, r, u0

    def _newton(self, resfun, commit, u, steps):
        self.residual_history = []
        r0 = None
        for it in range(self.n_iter):
            r = resfun(u, commit=True)
            nr = np.linalg.norm(r)
            self.residual_history.append(nr)
            K = fd_jacobian(lambda v: resfun(v, commit=False), u, self.n_grid, self.dim, self.nb,
                            self.n_nodes, self.boundary, self.act, steps)
            delta = spsolve(K.tocsc(), r)
            u = u + delta
            if r0 is None:
                r0 = nr
            crit = nr if self.tol_mode == 'absolute' else (nr / r0 if r0 > 0 else 0.0)
            if crit < self.tol:
                break
        self.newton_iterations = it + 1
        return u


class quasi_static_solver_2d(_NewtonMixin):
    """Oracle of quasi_static_solver_2d (qs2d:45-258). Adds 'Neo-Hookean',
    'Drucker-Prager' and J2 hardening ('H' in material_dict)."""
    dim = 2
    nb = 2

    def __init__(self, background_grid_dict, particles_dict, iteration_dict, material_dict, load_dict,
                 x_particles=None):
        g = background_grid_dict
        self.n_grid_x, self.n_grid_y = g['n_grid_x'], g['n_grid_y']
        self.n_grid = (self.n_grid_x, self.n_grid_y)
        self.n_nodes = (self.n_grid_x + 1) * (self.n_grid_y + 1)
        self.n_matrix_size = 2 * self.n_nodes
        self.dx = g['dx']
        self.start_x, self.end_x, self.start_y, self.end_y = g['start_x'], g['end_x'], g['start_y'], g['end_y']
        self.dmin = (g['domain_min_x'], g['domain_min_y'])
        self.dmax = (g['domain_max_x'], g['domain_max_y'])
        p = particles_dict
        self.PPD, self.p_vol, self.p_rho = p['PPD'], p['p_vol'], p['p_rho']
        self.GIMP_lp_initial = self.dx / self.PPD / 2.0 - 1e-6  # qs2d:75
        it = iteration_dict
        self.n_iter, self.tol = it['n_iter'], it['tol']
        self.tol_mode = it.get('tol_mode', 'absolute')
        m = material_dict
        self.mat = om.MaterialParams(m['material_name'], m['youngs_modulus'], m['poisson_ratio'],
                                     kappa=m.get('kappa', 0.0), H=m.get('H', 0.0), phi=m.get('friction_angle', 30.0),
                                     psi=m.get('dilation_angle', 0.0), cohesion=m.get('cohesion', 0.0),
                                     shape=m.get('shape_factor', 0.0), dp_apex=m.get('dp_apex', 'reference'),
                                     norsand=m.get('norsand'))
        l = load_dict
        self.gravity_mag = l['gravity_mag']
        self.point_load = (l['point_load_value_x'], l['point_load_value_y'])
        self.prescribed_increment = l.get('prescribed_increment')  # driver hook, see prepare_step
        if x_particles is None:
            x_particles = om.init_particles_rectangle_2d(self.start_x, self.start_y, self.end_x, self.end_y,
                                                         self.dx, self.n_grid_x, self.n_grid_y, self.PPD)
        self.x_particles = np.array(x_particles, dtype=np.float64)
        self.n_particles = self.x_particles.shape[0]
        self.F_old = om.initial_F(self.n_particles)
        self.F_new = self.F_old.copy()
        self.b_old = self.F_old.copy()
        self.b_new = self.F_old.copy()
        # hardening variable: J2's equivalent plastic strain, or NorSand's image pressure pi
        a0 = self.mat.norsand['pi0'] if self.mat.norsand is not None else 0.0
        self.alpha_old = np.full(self.n_particles, a0)
        self.alpha_new = self.alpha_old.copy()
        self.sigma = np.zeros((self.n_particles, 3, 3))
        self.lp = og.init_lp(self.n_particles, self.GIMP_lp_initial, 2)
        self.external_flag = np.zeros(self.n_particles, dtype=bool)
        self.boundary = np.zeros(self.n_matrix_size, dtype=bool)
        self.act = np.zeros(self.n_matrix_size, dtype=bool)
        self.increment_solution = np.zeros(self.n_matrix_size)
        self.reaction_history = []

    def reset(self):
        """qs2d:144-147 (boundary flags are kept)."""
        self.increment_solution[:] = 0.0
        self.act[:] = False

    def _res(self, st, step, total):
        def f(u, commit=False):
            r, Fn, bn, sig, an = om.residual_2d(st, self.n_nodes, u, self.F_old, self.b_old, self.alpha_old,
                                                self.external_flag, self.mat, self.p_vol, self.p_rho,
                                                self.gravity_mag, self.point_load, step, total,
                                                self.boundary, self.act)
            if commit:
                self.F_new, self.b_new, self.sigma, self.alpha_new = Fn, bn, sig, an
            return r
        return f

    def prepare_step(self, current_step, total_steps):
        """reset + P2G (qs2d:240-245); returns (stencil, residual fn, u0, FD steps)."""
        self.reset()
        if self.prescribed_increment is not None:
            # non-zero Dirichlet increments set by the driver (fixed dofs keep them: their rhs is 0)
            pres = np.zeros(self.n_matrix_size)
            self.prescribed_increment(self.x_particles, self.lp, self.boundary, pres, self.n_grid_x, self.n_nodes,
                                      self.dx, current_step, total_steps)
            self.increment_solution = pres * self.boundary
            self._prescribed = pres
        st = og.Stencil(self.x_particles, self.lp, 2, (self.dx,), self.n_grid)
        self.act = om.p2g_2d_activate(st, self.n_nodes)
        steps = (self.fd_rel * self.dx, self.fd_rel * self.dx)
        return st, self._res(st, current_step, total_steps), self.increment_solution.copy(), steps

    def advance_one_step(self, current_step, total_steps):
        """qs2d:239-258."""
        st, f, u0, steps = self.prepare_step(current_step, total_steps)
        u = self._newton(f, None, u0, steps)
        self.increment_solution = u
        if self.prescribed_increment is not None:
            # reaction of the displacement-controlled dofs: the unmasked residual summed over them
            r, _, _, _, _ = om.residual_2d(st, self.n_nodes, u, self.F_old, self.b_old, self.alpha_old,
                                           self.external_flag, self.mat, self.p_vol, self.p_rho, self.gravity_mag,
                                           self.point_load, current_step, total_steps,
                                           np.zeros(self.n_matrix_size, dtype=bool),
                                           np.ones(self.n_matrix_size, dtype=bool))
            m = (self._prescribed != 0) & self.boundary
            self.reaction_history.append(float(r[m].sum()))
        self.x_particles, _ = om.g2p(st, self.x_particles, u, self.n_nodes, 2)
        self.F_old = self.F_new.copy()
        self.b_old = self.b_new.copy()
        self.alpha_old = self.alpha_new.copy()
        self.lp = og.update_lp(self.F_new, self.x_particles, self.GIMP_lp_initial, self.dmin, self.dmax)


class quasi_static_coupled_solver_2d(_NewtonMixin):
    """Oracle of quasi_static_coupled_solver_2d (qc2d:45-313)."""
    dim = 2
    nb = 3

    def __init__(self, background_grid_dict, particles_dict, iteration_dict, material_dict, porous_media_dict,
                 load_dict, set_boundary_dofs, x_particles=None):
        g = background_grid_dict
        self.n_grid_x, self.n_grid_y = g['n_grid_x'], g['n_grid_y']
        self.n_grid = (self.n_grid_x, self.n_grid_y)
        self.n_nodes = (self.n_grid_x + 1) * (self.n_grid_y + 1)
        self.n_matrix_size = 3 * self.n_nodes
        self.dx, self.dy = g['dx'], g['dy']
        self.start_x, self.end_x, self.start_y, self.end_y = g['start_x'], g['end_x'], g['start_y'], g['end_y']
        self.dmin = (g['domain_min_x'], g['domain_min_y'])
        self.dmax = (g['domain_max_x'], g['domain_max_y'])
        p = particles_dict
        self.PPD, self.p_vol, self.p_rho = p['PPD'], p['p_vol'], p['p_rho']
        self.p_mass = self.p_vol * self.p_rho
        self.GIMP_lp_initial = self.dx / self.PPD / 2.0 - 1e-6  # qc2d:81
        it = iteration_dict
        self.n_iter, self.tol, self.dt = it['n_iter'], it['tol'], it['dt']
        self.tol_mode = it.get('tol_mode', 'absolute')
        m = material_dict
        self.mat = om.MaterialParams(m['material_name'], m['youngs_modulus'], m['poisson_ratio'],
                                     kappa=m.get('kappa', 0.0), H=m.get('H', 0.0), phi=m.get('friction_angle', 30.0),
                                     psi=m.get('dilation_angle', 0.0), cohesion=m.get('cohesion', 0.0),
                                     shape=m.get('shape_factor', 0.0), dp_apex=m.get('dp_apex', 'reference'),
                                     norsand=m.get('norsand'))
        self.phi_initial = porous_media_dict['phi_initial']
        self.mobility_constant = porous_media_dict['mobility_constant']
        l = load_dict
        self.gravity_mag = l['gravity_mag']
        self.traction = (l['traction_value_x'], l['traction_value_y'])
        self.point_load = (l['point_load_value_x'], l['point_load_value_y'])
        self.set_boundary_dofs = set_boundary_dofs
        if x_particles is None:
            x_particles = om.init_particles_rectangle_2d(self.start_x, self.start_y, self.end_x, self.end_y,
                                                         self.dx, self.n_grid_x, self.n_grid_y, self.PPD)
        self.x_particles = np.array(x_particles, dtype=np.float64)
        self.n_particles = self.x_particles.shape[0]
        self.F_old = om.initial_F(self.n_particles)
        self.F_new = self.F_old.copy()
        self.b_old = self.F_old.copy()
        self.b_new = self.F_old.copy()
        # hardening variable: J2's equivalent plastic strain, or NorSand's image pressure pi
        a0 = self.mat.norsand['pi0'] if self.mat.norsand is not None else 0.0
        self.alpha_old = np.full(self.n_particles, a0)
        self.alpha_new = self.alpha_old.copy()
        self.sigma = np.zeros((self.n_particles, 3, 3))
        self.pressure = np.zeros(self.n_particles)
        self.lp = og.init_lp(self.n_particles, self.GIMP_lp_initial, 2)
        self.external_flag = np.zeros(self.n_particles, dtype=bool)
        self.traction_flag = np.zeros(self.n_particles, dtype=bool)
        self.pressure_boundary_flag = np.zeros(self.n_particles, dtype=bool)
        self.boundary = np.zeros(self.n_matrix_size, dtype=bool)
        self.act = np.zeros(self.n_matrix_size, dtype=bool)
        self.old_solution = np.zeros(self.n_matrix_size)
        self.new_solution = np.zeros(self.n_matrix_size)

    def reset(self):
        """qc2d:164-175."""
        self.old_solution[:] = 0.0
        self.new_solution[:] = 0.0
        self.boundary[:] = False
        self.act[:] = False

    def _res(self, st, step, total):
        def f(u, commit=False):
            r, Fn, bn, sig, an = om.residual_coupled_2d(
                st, self.n_nodes, self.old_solution, u, self.F_old, self.b_old, self.alpha_old, self.external_flag,
                self.traction_flag, self.mat, self.phi_initial, self.mobility_constant, self.dt, self.dx, self.PPD,
                self.p_vol, self.p_rho, self.gravity_mag, self.traction, self.point_load, step, total,
                self.boundary, self.act)
            if commit:
                self.F_new, self.b_new, self.sigma, self.alpha_new = Fn, bn, sig, an
            return r
        return f

    def prepare_step(self, current_step, total_steps):
        """qc2d:275-300; returns (stencil, residual fn, u0, FD steps)."""
        self.reset()
        self.set_boundary_dofs(self.boundary, self.n_grid_x, self.n_nodes)
        st = og.Stencil(self.x_particles, self.lp, 2, (self.dx,), self.n_grid)
        self.act, self.boundary, M, rhs = om.p2g_coupled(st, self.x_particles, self.n_nodes, 2, self.dy, 1e-6,
                                                         self.pressure_boundary_flag, self.boundary, self.p_mass,
                                                         self.pressure)
        xp = om.solve_consistent_mass(M, rhs)
        self.old_solution, self.new_solution = om.p2g_set_nodal_solution(
            self.boundary, self.act, self.n_nodes, 2, self.old_solution, self.new_solution, xp, False)
        steps = (self.fd_rel * self.dx, self.fd_rel * self.dx, 1.0)
        return st, self._res(st, current_step, total_steps), self.new_solution.copy(), steps

    def advance_one_step(self, current_step, total_steps):
        """qc2d:274-313."""
        st, f, u0, steps = self.prepare_step(current_step, total_steps)
        u = self._newton(f, None, u0, steps)
        self.new_solution = u
        du = u - self.old_solution
        self.x_particles, self.pressure = om.g2p(st, self.x_particles, du, self.n_nodes, 2,
                                                 new_p=u[2 * self.n_nodes:])
        self.F_old = self.F_new.copy()
        self.b_old = self.b_new.copy()
        self.alpha_old = self.alpha_new.copy()
        self.lp = og.update_lp(self.F_new, self.x_particles, self.GIMP_lp_initial, self.dmin, self.dmax)


class quasi_static_coupled_solver_3d(_NewtonMixin):
    """Oracle of quasi_static_coupled_solver_3d (qc3d:56-420)."""
    dim = 3
    nb = 4

    def __init__(self, background_grid_dict, particles_dict, iteration_dict, material_dict, porous_media_dict,
                 load_dict, set_boundary_dofs, x_particles=None, total_steps_hint=25):
        g = background_grid_dict
        self.n_grid_x, self.n_grid_y, self.n_grid_z = g['n_grid_x'], g['n_grid_y'], g['n_grid_z']
        self.n_grid = (self.n_grid_x, self.n_grid_y, self.n_grid_z)
        self.n_nodes = (self.n_grid_x + 1) * (self.n_grid_y + 1) * (self.n_grid_z + 1)
        self.n_matrix_size = 4 * self.n_nodes
        self.dx, self.dy, self.dz = g['dx'], g['dy'], g['dz']
        self.start = (g['start_x'], g['start_y'], g['start_z'])
        self.end = (g['end_x'], g['end_y'], g['end_z'])
        self.end_x, self.end_y = g['end_x'], g['end_y']
        self.dmin = (g['domain_min_x'], g['domain_min_y'], g['domain_min_z'])
        self.dmax = (g['domain_max_x'], g['domain_max_y'], g['domain_max_z'])
        p = particles_dict
        self.PPD, self.p_vol, self.p_rho = p['PPD'], p['p_vol'], p['p_rho']
        self.p_mass = self.p_vol * self.p_rho
        self.GIMP_lp_initial = self.dx / self.PPD / 2.0  # qc3d:100 (no -1e-6)
        it = iteration_dict
        self.n_iter, self.tol, self.dt = it['n_iter'], it['tol'], it['dt']
        self.tol_mode = it.get('tol_mode', 'absolute')
        m = material_dict
        self.mat = om.MaterialParams(m['material_name'], m['youngs_modulus'], m['poisson_ratio'],
                                     kappa=m.get('kappa', 0.0), H=m.get('H', 0.0), phi=m.get('friction_angle', 30.0),
                                     psi=m.get('dilation_angle', 0.0), cohesion=m.get('cohesion', 0.0),
                                     shape=m.get('shape_factor', 0.0), dp_apex=m.get('dp_apex', 'reference'),
                                     norsand=m.get('norsand'))
        self.phi_initial = porous_media_dict['phi_initial']
        self.mobility_constant = porous_media_dict['mobility_constant']
        l = load_dict
        self.gravity_mag = l['gravity_mag']
        self.footing = (l['footing_region_min_x'], l['footing_region_max_x'], l['footing_region_min_y'],
                        l['footing_region_max_y'], l['footing_initial_height'], l['footing_incr_disp'],
                        # the reference's sign (mpm_utils.py:1240-1246), or a negative factor for
                        # the textbook penalty that pushes a penetrating particle down
                        l['penalty_factor'] * (1.0 if l.get('penalty_form', 'reference') == 'reference' else -1.0))
        self.set_boundary_dofs = set_boundary_dofs
        if x_particles is None:
            x_particles = om.init_particles_rectangle_3d(*self.start, *self.end, self.dx, *self.n_grid, self.PPD)
        self.x_particles = np.array(x_particles, dtype=np.float64)
        self.n_particles = self.x_particles.shape[0]
        self.F_old = om.initial_F(self.n_particles)
        self.F_new = self.F_old.copy()
        self.b_old = self.F_old.copy()
        self.b_new = self.F_old.copy()
        # hardening variable: J2's equivalent plastic strain, or NorSand's image pressure pi
        a0 = self.mat.norsand['pi0'] if self.mat.norsand is not None else 0.0
        self.alpha_old = np.full(self.n_particles, a0)
        self.alpha_new = self.alpha_old.copy()
        self.sigma = np.zeros((self.n_particles, 3, 3))
        self.pressure = np.zeros(self.n_particles)
        self.lp = og.init_lp(self.n_particles, self.GIMP_lp_initial, 3)
        self.pressure_boundary_flag = np.zeros(self.n_particles, dtype=bool)
        self.boundary = np.zeros(self.n_matrix_size, dtype=bool)
        self.act = np.zeros(self.n_matrix_size, dtype=bool)
        self.old_solution = np.zeros(self.n_matrix_size)
        self.new_solution = np.zeros(self.n_matrix_size)
        self.indentation_force = 0.0
        self.indentation_force_array = np.zeros(total_steps_hint)

    def reset(self):
        """qc3d:242-255."""
        self.old_solution[:] = 0.0
        self.new_solution[:] = 0.0
        self.boundary[:] = False
        self.act[:] = False

    def _res(self, st, step, total):
        def f(u, commit=False):
            r, Fn, bn, sig, an, force = om.residual_coupled_3d(
                st, self.n_nodes, self.x_particles, self.old_solution, u, self.F_old, self.b_old, self.alpha_old,
                self.mat, self.phi_initial, self.mobility_constant, self.dt, (self.dx, self.dy, self.dz), self.PPD,
                self.p_vol, self.p_rho, self.gravity_mag, self.footing, step, total, self.boundary, self.act)
            if commit:
                self.F_new, self.b_new, self.sigma, self.alpha_new = Fn, bn, sig, an
                self.indentation_force = force
                if step < self.indentation_force_array.size:
                    self.indentation_force_array[int(step)] = force
            return r
        return f

    def prepare_step(self, current_step, total_steps):
        """qc3d:397-417; returns (stencil, residual fn, u0, FD steps)."""
        self.reset()
        self.set_boundary_dofs(self.boundary, self.n_grid_x, self.n_grid_y, self.n_nodes, self.dx, self.dy,
                               self.end_x, self.end_y)
        st = og.Stencil(self.x_particles, self.lp, 3, (self.dx, self.dy, self.dz), self.n_grid)
        self.act, self.boundary, M, rhs = om.p2g_coupled(st, self.x_particles, self.n_nodes, 3, self.dz, 1e-7,
                                                         self.pressure_boundary_flag, self.boundary, self.p_mass,
                                                         self.pressure)
        xp = om.solve_consistent_mass(M, rhs)
        self.old_solution, self.new_solution = om.p2g_set_nodal_solution(
            self.boundary, self.act, self.n_nodes, 3, self.old_solution, self.new_solution, xp, True)
        steps = (self.fd_rel * self.dx, self.fd_rel * self.dx, self.fd_rel * self.dx, 1.0)
        return st, self._res(st, current_step, total_steps), self.new_solution.copy(), steps

    def advance_one_step(self, current_step, total_steps):
        """qc3d:396-420 (G2P and the lp update run at the end of newton_iter, :378-386)."""
        st, f, u0, steps = self.prepare_step(current_step, total_steps)
        u = self._newton(f, None, u0, steps)
        self.new_solution = u
        du = u - self.old_solution
        self.x_particles, self.pressure = om.g2p(st, self.x_particles, du, self.n_nodes, 3,
                                                 new_p=u[3 * self.n_nodes:])
        self.F_old = self.F_new.copy()
        self.b_old = self.b_new.copy()
        self.alpha_old = self.alpha_new.copy()
        self.lp = og.update_lp(self.F_new, self.x_particles, self.GIMP_lp_initial, self.dmin, self.dmax)

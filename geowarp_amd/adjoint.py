"""Hand-written adjoint of a quasi_static_solver_2d load path.

Observable: the reaction force of the displacement-controlled dofs (the dofs the
solver's ``prescribed_increment`` hook gives a non-zero increment, e.g. a
platen), per load step and per particle group (one group per specimen when
several specimens share the background grid):

    force[k, g] = sum over particles p in group g of sum over prescribed dofs I
                  of the residual contribution r_I^p at the converged increment
                  of step k (assemble_residual_2d_*, mpm_utils.py:416-780, before
                  the Dirichlet mask).

Reverse sweep for step k, with state S_k = (x, lp, F_old, b_old, alpha_old) per
particle, converged increment u_k, K = -dR/du (the solver's Jacobian) and the
state cotangent Sbar_{k+1} carried back from step k+1:

  1. local VJP of the step's particle map with the reaction cotangent and
     Sbar_{k+1}:  gives Hbar (cotangent of the increment gradient), the direct
     terms of Sbar_k and of the parameters;
  2. ubar = gather over particles of Hbar and of xbar_{k+1} (the transposes of
     the gradient and the G2P interpolation), masked to free dofs;
  3. transposed solve K^T lambda = ubar (block transpose of K; GMRES with the
     solver's preconditioner: multigrid on K^T or block Jacobi);
  4. local VJP with the residual cotangent lambda: Sbar_k and parameter terms
     through R (du/dS = K^-1 dR/dS).

The per-particle VJPs are forward-mode dual-number evaluations of the particle
map, one direction per input (gm_local_vjp_2d).  Load steps are recomputed
from checkpoints taken every ``checkpoint_every`` steps, so memory is one
particle-state checkpoint per segment plus one segment of states.  The forward
run also keeps each step's converged nodal increment (store_increments, default
on: n_nodes x 2 doubles per step), so a recomputation replays the step (P2G,
G2P) without its Newton solve; the replayed states differ from the forward
ones only by the last Newton correction (the forward commits the stage of the
iterate before it).

Discrete choices (active nodes, particle supports, the platen rows) are held
fixed: the gradient is that of the piecewise-smooth branch the forward run
takes.  ``LoadPathFunction`` wraps forward + adjoint as a torch.autograd.Function
of the per-group friction angle and cohesion (Drucker-Prager).
"""

import torch

from . import _lib as L
from .solvers._engine import MAT_KIND

ND = 25
DIR_X, DIR_LP, DIR_F, DIR_B, DIR_ALPHA, DIR_H = slice(0, 2), slice(2, 4), slice(4, 9), slice(9, 13), 13, \
    slice(14, 18)
DIR_THETA = {'E': 18, 'nu': 19, 'kappa': 20, 'H': 21, 'phi': 22, 'psi': 23, 'coh': 24}


def _stream():
    return L.c_void_p(torch.cuda.current_stream().cuda_stream)


class LoadPathAdjoint:
    """Forward load path with reaction-force records and its reverse sweep.

    solver: a quasi_static_solver_2d (fresh: its current state is step 0's
      initial state) whose load_dict has a ``prescribed_increment`` hook.
    groups: int64 tensor (n_particles,) of group ids (default: one group).
    """

    def __init__(self, solver, total_steps, n_steps=None, groups=None, checkpoint_every=5, adjoint_tol=1e-12,
                 initial=None, store_increments=True):
        if solver.dim != 2 or solver.coupled:
            raise ValueError('the adjoint covers quasi_static_solver_2d')
        if solver.prescribed_increment is None:
            raise ValueError("the reaction observable needs load_dict['prescribed_increment']")
        if solver.precision != 'fp64':
            raise ValueError('the adjoint runs on the FP64 solver (precision=\'fp64\')')
        self.s = solver
        self.total_steps = int(total_steps)
        self.n_steps = int(n_steps if n_steps is not None else total_steps)
        dev = solver.device
        n = solver.n_particles
        self.groups = groups.to(dev).long() if groups is not None else torch.zeros(n, dtype=torch.long, device=dev)
        self.n_groups = int(self.groups.max().item()) + 1 if n else 1
        self.k = max(1, int(checkpoint_every))
        self.adjoint_tol = adjoint_tol
        # initial state of the load path (default: the solver's current state)
        self.initial = initial if initial is not None else self._snapshot()
        # keep every step's converged nodal increment (n_nodes x nb doubles, small next to a particle
        # state): recomputing a segment from its checkpoint is then a replay (P2G, G2P) with no Newton
        # solve; without them the segment's Newton solves run again
        self.store_increments = bool(store_increments)
        self.increments = {}
        self.stats = dict(forward_steps=0, adjoint_steps=0, replayed_steps=0, krylov_iters=[],
                          unconverged_steps=[])

    # ------------------------------------------------------------------ state
    def _snapshot(self):
        s = self.s
        # the boundary flags are state too: a prescribed-increment hook may change them every step
        return dict(x=s.x_particles.clone(), lp=s.GIMP_lp.clone(), F=s.deformation_gradient_total_old.clone(),
                    b=s.left_Cauchy_Green_old.clone(), a=s.plastic_strain_old.clone(),
                    bc=s.dofStruct.boundary_flag_array.clone())

    def _restore(self, st):
        s = self.s
        s.x_particles.copy_(st['x'])
        s.GIMP_lp.copy_(st['lp'])
        s.deformation_gradient_total_old.copy_(st['F'])
        s.left_Cauchy_Green_old.copy_(st['b'])
        s.plastic_strain_old.copy_(st['a'])
        s.dofStruct.boundary_flag_array.copy_(st['bc'])

    def _local_args(self, step, u, cot=None, cot_scale=None, bars=None):
        s = self.s
        a = L.gm_local_args()
        a.n_particles = s.n_particles
        a.nx1 = s.n_grid_x + 1
        a.mat_kind = MAT_KIND[s.material_name]
        a.dp_apex = s.dp_apex
        a.h = s.dx
        a.shape = s.shape_factor
        a.p_vol, a.p_rho = s.p_vol, s.p_rho
        s1 = (float(step) + 1.0) / float(self.total_steps)
        a.gravity[0], a.gravity[1] = 0.0, -s.gravity_mag * s1
        a.point_load[0], a.point_load[1] = s.point_load_value_x * s1, s.point_load_value_y * s1
        a.lp0 = s.GIMP_lp_initial
        a.dmin[0], a.dmin[1] = s.domain_min_x, s.domain_min_y
        a.dmax[0], a.dmax[1] = s.domain_max_x, s.domain_max_y
        a.E, a.nu, a.kappa, a.H = s.youngs_modulus, s.poisson_ratio, s.kappa, s.hardening
        a.phi, a.psi, a.coh = s.friction_angle, s.dilation_angle, s.cohesion
        pa = getattr(s, 'friction_angle_array', None)
        ca = getattr(s, 'cohesion_array', None)
        a.p_phi = pa.data_ptr() if pa is not None else None
        a.p_coh = ca.data_ptr() if ca is not None else None
        a.x, a.lp = s.x_particles.data_ptr(), s.GIMP_lp.data_ptr()
        a.F_old = s.deformation_gradient_total_old.data_ptr()
        a.b_old = s.left_Cauchy_Green_old.data_ptr()
        a.alpha_old = s.plastic_strain_old.data_ptr()
        self._ext = s.particle_external_flag_array.to(torch.uint8)
        a.ext_flag = self._ext.data_ptr()
        a.u = u.data_ptr()
        a.cot = cot.data_ptr() if cot is not None else None
        a.cot_scale = cot_scale.data_ptr() if cot_scale is not None else None
        for name in ('xbar', 'lpbar', 'Fbar', 'bbar', 'abar'):
            setattr(a, name, bars[name].data_ptr() if bars is not None else None)
        return a

    def _reaction_cot(self):
        """Node-interleaved indicator of the prescribed (displacement-controlled) dofs."""
        s = self.s
        pres = s._prescribed
        m = ((pres != 0) & s.dofStruct.boundary_flag_array).to(torch.float64)
        return s.engine.from_dofmajor(m)

    def _reaction(self, step):
        s = self.s
        cot = self._reaction_cot()
        acc = torch.empty(s.n_particles, dtype=torch.float64, device=s.device)
        a = self._local_args(step, s._inc, cot=cot)
        L.check(L.lib().gm_local_eval_2d(L.ctypes.byref(a), L.ptr(acc), None, _stream()), 'gm_local_eval_2d')
        f = torch.zeros(self.n_groups, dtype=torch.float64, device=s.device)
        f.index_add_(0, self.groups, acc)
        return f

    def _step(self, step):
        """One forward load step; returns (u_k, reaction forces)."""
        s = self.s
        s.prepare_step(step, self.total_steps)
        s.newton_iter(step, self.total_steps)
        if not getattr(s, 'newton_converged', True):
            # the sensitivities assume R(u_k) = 0; record the steps where Newton stopped short
            self.stats['unconverged_steps'].append((step, s.residual_history[-1]))
        u = s._inc.clone()
        f = self._reaction(step)
        s.finish_step()
        self.stats['forward_steps'] += 1
        return u, f

    def _replay(self, step, u):
        """Advance one load step with its stored converged increment (no Newton solve):
        the particle stage at u_k, then G2P.  (The Newton loop commits the stage of its
        last iterate before the final correction, which differs from u_k by that
        correction; the replay commits the state at u_k itself, the state the reverse
        sweep differentiates.)"""
        s = self.s
        s.prepare_step(step, self.total_steps)
        s._inc.copy_(u)
        s.engine.stage(s._args(step, self.total_steps), want_tangent=False)
        s.finish_step()
        self.stats['replayed_steps'] += 1

    # ------------------------------------------------------------------ forward
    def forward(self):
        """Run the load path from the initial state; returns forces (n_steps, n_groups)."""
        self._restore(self.initial)
        self.checkpoints = {}
        self.increments = {}
        forces = torch.zeros(self.n_steps, self.n_groups, dtype=torch.float64, device=self.s.device)
        for k in range(self.n_steps):
            if k % self.k == 0:
                self.checkpoints[k] = self._snapshot()
            u, forces[k] = self._step(k)
            if self.store_increments:
                self.increments[k] = u
        return forces

    # ------------------------------------------------------------------ reverse
    def _vjp(self, step, u, cot=None, cot_scale=None, bars=None):
        s = self.s
        g = torch.empty(s.n_particles, ND, dtype=torch.float64, device=s.device)
        a = self._local_args(step, u, cot=cot, cot_scale=cot_scale, bars=bars)
        L.check(L.lib().gm_local_vjp_2d(L.ctypes.byref(a), L.ptr(g), _stream()), 'gm_local_vjp_2d')
        return g

    def _transposed_solve(self, ubar):
        """K^T lambda = ubar on the current step's Jacobian (FP64 GMRES, block-Jacobi)."""
        eng = self.s.engine
        nb = eng.nb
        nnz = eng.nnzb * nb * nb
        vt = torch.empty(nnz, dtype=torch.float64, device=eng.dev)
        L.check(L.lib().gm_bsr_transpose(nb, 0, eng.n_nodes, L.ptr(eng.row_ptr), L.ptr(eng.col_idx),
                                         L.ptr(eng.vals), L.ptr(vt), _stream()), 'gm_bsr_transpose')
        vals = eng.vals
        eng.vals = vt
        eng.jac_version += 1  # a multigrid hierarchy is rebuilt for K^T (and again for K afterwards)
        lam = torch.zeros_like(ubar)
        method = 'mg' if self.s.krylov_method == 'mg' else 'gmres'
        try:
            its, rel = eng.krylov(ubar, lam, False, self.adjoint_tol, self.s.linear_maxiter, method,
                                  self.s.krylov_restart)
        finally:
            eng.vals = vals
            eng.jac_version += 1
        self.stats['krylov_iters'].append(its)
        return lam

    def _reverse_step(self, step, state, u, fbar_k, Sbar):
        s = self.s
        eng = s.engine
        self._restore(state)
        s.prepare_step(step, self.total_steps)
        s._inc.copy_(u)
        args = s._args(step, self.total_steps)
        eng.stage(args, want_tangent=True)
        eng.jacobian(s.x_particles, s.GIMP_lp)
        # 1. reaction cotangent and the carried state cotangent
        cot = self._reaction_cot()
        cs = fbar_k[self.groups].contiguous()
        g1 = self._vjp(step, u, cot=cot, cot_scale=cs, bars=Sbar)
        # 2. ubar = gather(Hbar, xbar_{k+1}) over the particles, free dofs only
        Yb = torch.cat([g1[:, DIR_H], Sbar['xbar']], 1)[eng.sorted.long()].contiguous()
        ubar = torch.zeros(eng.n_nodes * eng.nb, dtype=torch.float64, device=eng.dev)
        L.check(L.lib().gm_node_residual(L.ctypes.byref(eng.grid), L.ctypes.byref(eng.bins), L.ptr(Yb),
                                         L.ptr(eng.freem), L.ptr(ubar), 0, _stream()), 'gm_node_residual')
        # 3. transposed solve
        lam = self._transposed_solve(ubar)
        # 4. residual cotangent lambda
        g2 = self._vjp(step, u, cot=lam)
        g = g1 + g2
        new_bar = dict(xbar=g[:, DIR_X].contiguous(), lpbar=g[:, DIR_LP].contiguous(),
                       Fbar=g[:, DIR_F].contiguous(), bbar=g[:, DIR_B].contiguous(),
                       abar=g[:, DIR_ALPHA].contiguous())
        self.stats['adjoint_steps'] += 1
        return new_bar, g[:, 18:]

    def backward(self, fbar):
        """fbar: (n_steps, n_groups) cotangent of the forces.  Returns a dict of
        per-group gradients: 'phi', 'coh' (sum over the group's particles) and
        global 'E', 'nu', 'kappa', 'H', 'psi'."""
        s = self.s
        dev = s.device
        n = s.n_particles
        fbar = fbar.to(dev, torch.float64)
        Sbar = dict(xbar=torch.zeros(n, 2, dtype=torch.float64, device=dev),
                    lpbar=torch.zeros(n, 2, dtype=torch.float64, device=dev),
                    Fbar=torch.zeros(n, 5, dtype=torch.float64, device=dev),
                    bbar=torch.zeros(n, 4, dtype=torch.float64, device=dev),
                    abar=torch.zeros(n, dtype=torch.float64, device=dev))
        theta = torch.zeros(n, 7, dtype=torch.float64, device=dev)
        starts = sorted(self.checkpoints)
        for c in reversed(starts):
            end = min(c + self.k, self.n_steps)
            # recompute the segment's states and converged increments
            self._restore(self.checkpoints[c])
            seg = []
            for k in range(c, end):
                st = self._snapshot()
                if k in self.increments:
                    u = self.increments[k]
                    self._replay(k, u)
                else:
                    u, _ = self._step(k)
                seg.append((k, st, u))
            for k, st, u in reversed(seg):
                Sbar, th = self._reverse_step(k, st, u, fbar[k], Sbar)
                theta += th
        out = {}
        for name, col in (('E', 0), ('nu', 1), ('kappa', 2), ('H', 3), ('psi', 5)):
            out[name] = theta[:, col].sum()
        for name, col in (('phi', 4), ('coh', 6)):
            gsum = torch.zeros(self.n_groups, dtype=torch.float64, device=dev)
            gsum.index_add_(0, self.groups, theta[:, col])
            out[name] = gsum
        # leave the solver at the initial state of the load path
        self._restore(self.initial)
        return out


class LoadPathFunction(torch.autograd.Function):
    """forces = LoadPathFunction.apply(phi, coh, adjoint): per-group friction
    angle and cohesion (tensors of n_groups) -> reaction forces (n_steps,
    n_groups); backward runs the adjoint sweep."""

    @staticmethod
    def forward(ctx, phi, coh, adj):
        s = adj.s
        g = adj.groups
        s.friction_angle_array = phi.detach().to(s.device, torch.float64)[g].contiguous()
        s.cohesion_array = coh.detach().to(s.device, torch.float64)[g].contiguous()
        ctx.adj = adj
        return adj.forward()

    @staticmethod
    def backward(ctx, fbar):
        out = ctx.adj.backward(fbar)
        return out['phi'], out['coh'], None

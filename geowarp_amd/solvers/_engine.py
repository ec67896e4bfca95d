"""Device engine behind the three solver classes.

Owns every device buffer of one background grid + particle set and runs the
load-step phases through the C ABI (include/geowarp_amd.h):

  begin_step      bin particles by stencil base, build the per-step block-CSR
                  pattern (weights are fixed inside a load step)
  p2g_*           activation / pressure-Dirichlet flags / consistent mass
  newton          residual (FP64) -> Jacobian (FP64 or FP32) -> Krylov -> update
  g2p, update_lp

Nodal vectors are node-interleaved internally (value k of node n at n*nb+k);
the solver classes expose the reference's dof-major views.
"""
import math
import os
import time

import torch

from .. import _lib as L
from ._multigrid import Multigrid

MAT_KIND = {'Hencky elasticity': 0, 'J2': 1, 'Neo-Hookean': 2, 'Drucker-Prager': 3, 'NorSand': 4}


def _stream():
    return L.c_void_p(torch.cuda.current_stream().cuda_stream)


def _layout(dim, coupled):
    nh = dim * dim
    if not coupled:
        return nh + dim, nh  # NY, NZ
    nz = nh + (1 if dim == 3 else 0) + 2 + dim
    ny = nh + dim + 1 + dim + 1
    return ny, nz


class Engine:
    def __init__(self, dim, coupled, n_grid, h, device=None, precision='fp64'):
        L.lib()
        if device is None:
            if not torch.cuda.is_available():
                raise L.GMError('geowarp_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
            device = torch.device('cuda', torch.cuda.current_device())
        self.dev = torch.device(device)
        if self.dev.type != 'cuda':
            raise L.GMError('geowarp_amd runs on a ROCm device only')
        if precision not in ('fp64', 'mixed'):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.dim = dim
        self.coupled = bool(coupled)
        self.nb = dim + (1 if coupled else 0)
        nx, ny = n_grid[0], n_grid[1]
        nz = n_grid[2] if dim == 3 else 0
        self.n_nodes = (nx + 1) * (ny + 1) * ((nz + 1) if dim == 3 else 1)
        g = L.gm_grid()
        g.dim, g.nx, g.ny, g.nz, g.n_nodes, g.nb = dim, nx, ny, nz, self.n_nodes, self.nb
        hs = (h[0], h[0], 0.0) if dim == 2 else tuple(h)
        for a in range(3):
            g.h[a] = hs[a]
            g.inv_h[a] = (1.0 / hs[a]) if hs[a] != 0.0 else 0.0
        self.grid = g
        self.NY, self.NZ = _layout(dim, coupled)
        dev = self.dev
        n = self.n_nodes
        i32 = dict(dtype=torch.int32, device=dev)
        self.cell_start = torch.zeros(n + 1, **i32)
        self.cell_count = torch.zeros(n + 1, **i32)
        self.scan_tmp = torch.zeros(1100, **i32)
        self.err = torch.zeros(1, **i32)
        self.row_ptr = torch.zeros(n + 1, **i32)
        self.row_nnzb = torch.zeros(n + 1, **i32)
        self.row_mask = torch.zeros(2 * n, dtype=torch.int64, device=dev)
        self.cap = int(os.environ.get('GM_INITIAL_BLOCKS_PER_NODE', 12 if dim == 2 else 32)) * n
        self.col_idx = torch.zeros(self.cap, **i32)
        nbv = n * self.nb
        f64 = dict(dtype=torch.float64, device=dev)
        self.rhs = torch.zeros(nbv, **f64)
        self.delta = torch.zeros(nbv, **f64)
        self.bflag = torch.zeros(nbv, dtype=torch.uint8, device=dev)
        self.act = torch.zeros(nbv, dtype=torch.uint8, device=dev)
        self.freem = torch.zeros(nbv, dtype=torch.uint8, device=dev)
        self.scal = torch.zeros(8, **f64)
        self.part = torch.zeros(2 * max(n // 32 + 2, 2048), **f64)
        self._alloc_vals()
        self._ws = {}
        self.n_particles = -1
        self.timers = None
        # CG on the consistent mass: diagonal preconditioner (True: the lumped masses, its row
        # sums; on the 3D indentation both take 72 iterations to 1e-13)
        self.mass_lumped_pc = False
        self.stats = dict(krylov_iters=[], krylov_relres=[])
        self.jac_version = 0
        self.mg_options = {}
        self.mixed_inner_maxiter = None  # FP32 Krylov iterations per refinement sweep (None: maxiter)
        # mixed form: 'fp32' assembles the tangent and the Jacobian in FP32; 'fp64' assembles
        # them in FP64, solves with an FP32 copy and refines against the FP64 Jacobian
        self.mixed_assembly = 'fp32'
        # 3D indenter: dict(footing=(min_x, max_x, min_y, max_y), height, pen_par, lp_max) of the
        # current load step, set by the solver; with penalty_pc the multigrid hierarchy is built
        # from a copy of the Newton matrix whose penalty term has the opposite sign
        self.penalty = None
        self.penalty_pc = True
        self.vals_pc = None
        # node Jacobian and residual of the 3D u-p form: 'bins' (per-bin sums -- the Jacobian's
        # as products on the matrix cores -- into a workspace of up to n_nodes x 11664 values,
        # then a fixed-order gather) or 'rows' (one workgroup or thread per node row, no
        # workspace)
        self.jacobian_kernel = os.environ.get('GM_JACOBIAN_KERNEL', 'bins')
        # 3D u-p P2G (flags and consistent-mass right-hand side): per particle bin, then a
        # fixed-order gather per node (True), or one thread per node over its candidates
        self.binned_p2g = True
        # multigrid: rebuild the coarse Galerkin operators at the first solve of a load step
        # only (later Newton corrections refresh level 0 and keep them), unless a solve then
        # needs three times the iterations of the step's first
        self.mg_lag = False
        self.step_id = 0

    # ------------------------------------------------------------------ buffers
    def _alloc_vals(self):
        jdt = torch.float32 if self.precision == 'mixed' else torch.float64
        self.vals = torch.zeros(self.cap * self.nb * self.nb, dtype=jdt, device=self.dev)
        self.vals64 = None  # FP64 Jacobian of the mixed form with FP64 assembly (allocated on first use)
        self.mvals = None

    def _tangent_fp32(self):
        return self.precision == 'mixed' and getattr(self, 'mixed_assembly', 'fp32') == 'fp32'

    def _grow(self, nnzb):
        self.cap = int(nnzb * 1.15) + 16
        self.col_idx = torch.zeros(self.cap, dtype=torch.int32, device=self.dev)
        self._alloc_vals()

    def ensure_particles(self, n):
        if n == self.n_particles:
            return
        dev = self.dev
        self.n_particles = n
        self.sorted = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.pbase = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.ext = torch.zeros(max(n, 1), dtype=torch.uint8, device=dev)
        self.ext_s = torch.zeros(max(n, 1), dtype=torch.uint8, device=dev)
        self.x_s = torch.zeros(max(n, 1) * self.dim, dtype=torch.float64, device=dev)
        self.lp_s = torch.zeros(max(n, 1) * self.dim, dtype=torch.float64, device=dev)
        self.Y = torch.zeros(max(n, 1) * self.NY, dtype=torch.float64, device=dev)
        ddt = torch.float32 if self._tangent_fp32() else torch.float64
        self.D = torch.zeros(max(n, 1) * self.NY * self.NZ, dtype=ddt, device=dev)
        self.pen_force = torch.zeros(max(n, 1), dtype=torch.float64, device=dev)

    def workspace(self, nb, fp32):
        key = (nb, fp32)
        if key not in self._ws:
            dt = torch.float32 if fp32 else torch.float64
            n = self.n_nodes * nb
            vecs = [torch.zeros(n, dtype=dt, device=self.dev) for _ in range(8)]
            dinv = torch.zeros(self.n_nodes * nb * nb, dtype=dt, device=self.dev)
            part = torch.zeros(2 * max(self.n_nodes // 16 + 2, 2048), dtype=torch.float64, device=self.dev)
            sc = torch.zeros(16, dtype=torch.float64, device=self.dev)
            ws = L.gm_krylov_ws(*[v.data_ptr() for v in vecs], dinv.data_ptr(), part.data_ptr(), sc.data_ptr())
            self._ws[key] = (ws, vecs, dinv, part, sc)
        return self._ws[key][0]

    def _bins(self):
        b = L.gm_bins()
        b.n_particles = self.n_particles
        b.cell_start = self.cell_start.data_ptr()
        b.cell_count = self.cell_count.data_ptr()
        b.sorted = self.sorted.data_ptr()
        b.ext = self.ext.data_ptr()
        b.scan_tmp = self.scan_tmp.data_ptr()
        b.ext_s = self.ext_s.data_ptr()
        b.x_s = self.x_s.data_ptr()
        b.lp_s = self.lp_s.data_ptr()
        return b

    def _pattern(self):
        p = L.gm_pattern()
        p.row_ptr = self.row_ptr.data_ptr()
        p.col_idx = self.col_idx.data_ptr()
        p.row_mask = self.row_mask.data_ptr()
        p.nnzb = self.nnzb
        return p

    # ------------------------------------------------------------------ step phases
    def begin_step(self, x, lp, pattern=True):
        """Bin particles by stencil base node and build the block-CSR pattern."""
        lib = L.lib()
        self.step_id += 1
        self.ensure_particles(x.shape[0])
        self.bins = self._bins()
        L.check(lib.gm_bin_particles(L.ctypes.byref(self.grid), L.ptr(x), L.ptr(lp), L.ctypes.byref(self.bins),
                                     L.ptr(self.pbase), L.ptr(self.err), _stream()), 'gm_bin_particles')
        nbad = int(self.err.item())
        if nbad:
            raise L.GMError('%d particle stencils leave the background grid (or are not finite)' % nbad)
        self.nnzb = 0
        if not pattern:
            return
        for _ in range(3):
            pat = self._pattern()
            rc = lib.gm_bsr_pattern(L.ctypes.byref(self.grid), L.ctypes.byref(self.bins), L.ctypes.byref(pat),
                                    L.ptr(self.row_nnzb), self.cap, _stream())
            if rc == -4:
                self._grow(pat.nnzb)
                continue
            L.check(rc, 'gm_bsr_pattern')
            self.nnzb = pat.nnzb
            break
        self.pattern = self._pattern()

    def set_boundary_from_dofmajor(self, bflag_dofmajor):
        nb = self.nb
        self.bflag.copy_(bflag_dofmajor.view(nb, self.n_nodes).t().reshape(-1).to(torch.uint8))

    def dofmajor(self, v):
        return v.view(self.n_nodes, self.nb).t().reshape(-1)

    def from_dofmajor(self, v):
        return v.view(self.nb, self.n_nodes).t().reshape(-1).contiguous()

    def p2g(self, x, lp, threshold, pbc_flag=None, node_h_last=0.0, p_mass=0.0, p_part=None, mass=False):
        lib = L.lib()
        self.act.zero_()
        mrhs = None
        if mass:
            if self.mvals is None or self.mvals.numel() < self.cap:
                self.mvals = torch.zeros(self.cap, dtype=torch.float64, device=self.dev)
            mrhs = torch.zeros(self.n_nodes, dtype=torch.float64, device=self.dev)
        binned_mass = mass and self.dim == 3
        if self.dim == 3 and self.nb == 4 and self.binned_p2g:
            # flags and right-hand side per particle bin, then a fixed-order gather per node
            g = L.ctypes.byref(self.grid)
            need = int(lib.gm_p2g_work_bytes(g))
            w = self._bins_workspace(need)
            L.check(lib.gm_p2g_binned(g, L.ctypes.byref(self.bins), float(threshold), L.ptr(pbc_flag),
                                      float(node_h_last), float(p_mass), L.ptr(p_part), L.ptr(self.act),
                                      L.ptr(self.bflag), L.ptr(mrhs), L.ptr(w), need, _stream()), 'gm_p2g_binned')
        else:
            L.check(lib.gm_p2g(L.ctypes.byref(self.grid), L.ctypes.byref(self.bins), L.ctypes.byref(self.pattern),
                               float(threshold), L.ptr(pbc_flag), float(node_h_last),
                               float(p_mass), L.ptr(p_part), L.ptr(self.act), L.ptr(self.bflag),
                               L.ptr(self.mvals) if mass and not binned_mass else None, L.ptr(mrhs), _stream()),
                    'gm_p2g')
        if binned_mass:
            g = L.ctypes.byref(self.grid)
            need = int(lib.gm_mass_work_bytes(g))
            w = self._bins_workspace(need)
            L.check(lib.gm_mass_binned(g, L.ctypes.byref(self.bins), L.ctypes.byref(self.pattern), float(p_mass),
                                       L.ptr(self.mvals), L.ptr(w), need, _stream()), 'gm_mass_binned')
        L.check(lib.gm_free_mask(self.n_nodes * self.nb, L.ptr(self.bflag), L.ptr(self.act), L.ptr(self.freem),
                                 _stream()), 'gm_free_mask')
        return mrhs

    def solve_mass(self, mrhs, tol=1e-13, maxiter=20000, x0=None):
        """Consistent-mass projection M p = rhs (the reference solves it with
        pyamg SA-BiCGStab (qc2d:288-293) or Warp Jacobi-BiCGStab (qc3d:410-412)
        at 1e-8; here to `tol`): conjugate gradients, the matrix being symmetric
        positive definite, preconditioned by its diagonal (mass_lumped_pc: by the
        lumped masses, its row sums), from x0 when given (e.g. the previous load
        step's nodal pressures), with Jacobi-BiCGStab as the fallback."""
        lib = L.lib()
        L.check(lib.gm_fix_empty_diag(self.n_nodes, L.ptr(self.row_ptr), L.ptr(self.col_idx), L.ptr(self.mvals),
                                      L.ptr(mrhs), _stream()), 'gm_fix_empty_diag')
        if x0 is not None:
            xp = x0.to(torch.float64).clone()
        else:
            xp = torch.zeros(self.n_nodes, dtype=torch.float64, device=self.dev)
        ws = self.workspace(1, False)
        it = L.c_int(0)
        rr = L.c_double(0.0)
        head = (1, 0, self.n_nodes, L.ptr(self.row_ptr), L.ptr(self.col_idx), L.ptr(self.mvals), L.ptr(mrhs),
                L.ptr(xp), L.ctypes.byref(ws), tol, maxiter)
        tail = (L.ctypes.byref(it), L.ctypes.byref(rr), _stream())
        # CG stops on the device; the host reads its state every 16 iterations
        L.check(lib.gm_pcg(*head, 16, int(self.mass_lumped_pc) | (2 if x0 is not None else 0), *tail),
                'gm_pcg(mass)')
        self.mass_solver = 'pcg'
        if not (rr.value <= tol):
            its = it.value
            L.check(lib.gm_bicgstab(*head, 4, *tail), 'gm_bicgstab(mass)')  # from x0 = 0
            it.value += its
            self.mass_solver = 'bicgstab'
        self.mass_iters = it.value
        self.mass_relres = rr.value
        return xp

    def set_nodal_pressure(self, xp, u_old, u_new, require_active):
        L.check(L.lib().gm_set_nodal_pressure(self.n_nodes, self.nb, L.ptr(self.bflag), L.ptr(self.act),
                                              int(require_active), L.ptr(xp), L.ptr(u_old), L.ptr(u_new),
                                              _stream()), 'gm_set_nodal_pressure')

    def stage(self, args, want_tangent):
        t32 = self._tangent_fp32()
        if self.D.dtype != (torch.float32 if t32 else torch.float64):
            self.D = torch.zeros(self.D.numel(), dtype=torch.float32 if t32 else torch.float64, device=self.dev)
        args.want_tangent = int(want_tangent)
        args.tangent_fp32 = int(t32)
        args.sorted = self.sorted.data_ptr()
        args.Y = self.Y.data_ptr()
        args.D = self.D.data_ptr()
        L.check(L.lib().gm_particle_stage(L.ctypes.byref(self.grid), L.ctypes.byref(args), _stream()),
                'gm_particle_stage')

    def residual(self, x, lp):
        g = L.ctypes.byref(self.grid)
        if self.dim == 3 and self.coupled and self.jacobian_kernel == 'bins':
            # per-bin sums into the shared bin workspace, then a fixed-order gather per node
            need = int(L.lib().gm_residual_work_bytes(g))
            w = self._bins_workspace(need)
            L.check(L.lib().gm_node_residual_binned(g, L.ctypes.byref(self.bins), L.ptr(self.Y), L.ptr(self.freem),
                                                    L.ptr(self.rhs), L.ptr(w), need, _stream()),
                    'gm_node_residual_binned')
            return
        L.check(L.lib().gm_node_residual(g, L.ctypes.byref(self.bins), L.ptr(self.Y),
                                         L.ptr(self.freem), L.ptr(self.rhs), int(self.coupled), _stream()),
                'gm_node_residual')

    def norm2(self, v, slot=0):
        L.check(L.lib().gm_norm2(v.numel(), L.ptr(v), L.ptr(self.part), L.c_void_p(self.scal.data_ptr() + 8 * slot),
                                 _stream()), 'gm_norm2')

    def multigrid(self, fp32):
        key = ('mg', bool(fp32))
        if key not in self._ws:
            # [hierarchy, Jacobian version it was set up for, load step of its last full
            # setup, GMRES iterations of that step's first solve, rebuild at the next setup]
            self._ws[key] = [Multigrid(self, fp32, **self.mg_options), -1, -1, 0, False]
        return self._ws[key]

    def autotune_spmv(self, reps=10):
        """Pick the SpMV workgroup size on the current Jacobian (once per engine
        and precision); the timings are kept in stats['spmv_tune']."""
        fp32 = self.precision == 'mixed'
        n = self.n_nodes * self.nb
        dt = torch.float32 if fp32 else torch.float64
        x = torch.ones(n, dtype=dt, device=self.dev)
        y = torch.empty(n, dtype=dt, device=self.dev)
        best = L.c_int(0)
        ms = (L.ctypes.c_float * 4)()
        L.check(L.lib().gm_spmv_autotune(self.nb, int(fp32), self.n_nodes, L.ptr(self.row_ptr), L.ptr(self.col_idx),
                                         L.ptr(self.vals), L.ptr(x), L.ptr(y), reps, L.ctypes.byref(best), ms,
                                         _stream()), 'gm_spmv_autotune')
        self.stats['spmv_tune'] = dict(threads=best.value, ms={t: round(v, 5) for t, v in
                                                               zip((64, 128, 256, 512), list(ms))}, nnzb=self.nnzb)
        return best.value

    def autotune_jacobian(self, reps=3):
        """Pick the per-bin and gather kernel variants of the binned 3D u-p Jacobian
        on the current tangent (once per engine and precision; every variant writes
        the same bits); the timings are kept in stats['jacobian_tune']."""
        if not (self.dim == 3 and self.coupled and self.jacobian_kernel == 'bins'):
            self.stats['jacobian_tune'] = False  # nothing to choose for this form
            return None
        fp32 = self.precision == 'mixed'
        t32 = fp32 and self._tangent_fp32()
        vals, v32 = (self.vals64, False) if (fp32 and not t32) else (self.vals, fp32)
        g, b, pat = L.ctypes.byref(self.grid), L.ctypes.byref(self.bins), L.ctypes.byref(self.pattern)
        need = int(L.lib().gm_node_jacobian_work_bytes(g, int(t32)))
        w = self._bins_workspace(need)
        bb, bg = L.c_int(0), L.c_int(0)
        ms = (L.ctypes.c_float * 7)()
        L.check(L.lib().gm_jacobian_autotune(g, b, pat, L.ptr(self.D), int(t32), L.ptr(self.freem), L.ptr(vals),
                                             int(v32), L.ptr(w), need, reps, L.ctypes.byref(bb),
                                             L.ctypes.byref(bg), ms, _stream()), 'gm_jacobian_autotune')
        if fp32 and not t32:
            L.check(L.lib().gm_d2f(self.nnzb * self.nb * self.nb, L.ptr(self.vals64), L.ptr(self.vals), _stream()),
                    'gm_d2f')
        self.stats['jacobian_tune'] = dict(bins_variant=bb.value, gather_variant=bg.value,
                                           bins_ms=[round(v, 4) for v in list(ms)[:4]],
                                           gather_ms=[round(v, 4) for v in list(ms)[4:]])
        return bb.value, bg.value

    def jacobian(self, x, lp):
        fp32 = self.precision == 'mixed'
        self.jac_version += 1
        if fp32 and not self._tangent_fp32():
            # FP64 assembly, FP32 copy for the Krylov solve
            if self.vals64 is None or self.vals64.numel() != self.vals.numel():
                self.vals64 = torch.zeros(self.vals.numel(), dtype=torch.float64, device=self.dev)
            self._assemble(False, self.vals64, False)
            L.check(L.lib().gm_d2f(self.nnzb * self.nb * self.nb, L.ptr(self.vals64), L.ptr(self.vals), _stream()),
                    'gm_d2f')
            return
        self._assemble(fp32, self.vals, fp32)

    def _bins_workspace(self, nbytes):
        """Per-bin workspace of the binned 3D assemblies (Jacobian, consistent mass),
        grown on demand; the assemblies run in stream order, so they share it."""
        w = self._ws.get('bins')
        if w is None or w.numel() < nbytes:
            self._ws.pop('bins', None)
            w = self._ws['bins'] = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        return w

    def _assemble(self, tangent_fp32, vals, vals_fp32):
        g, b, pat = L.ctypes.byref(self.grid), L.ctypes.byref(self.bins), L.ctypes.byref(self.pattern)
        if self.dim == 3 and self.coupled and self.jacobian_kernel == 'bins':
            need = int(L.lib().gm_node_jacobian_work_bytes(g, int(tangent_fp32)))
            w = self._bins_workspace(need)
            L.check(L.lib().gm_node_jacobian_binned(g, b, pat, L.ptr(self.D), int(tangent_fp32), L.ptr(self.freem),
                                                    L.ptr(vals), int(vals_fp32), L.ptr(w), need, _stream()),
                    'gm_node_jacobian_binned')
            return
        L.check(L.lib().gm_node_jacobian(g, b, pat, L.ptr(self.D), int(tangent_fp32), L.ptr(self.freem), L.ptr(vals),
                                         int(vals_fp32), int(self.coupled), _stream()),
                'gm_node_jacobian')

    def gmres_workspace(self, nb, fp32, m):
        key = ('gmres', nb, fp32, m)
        if key not in self._ws:
            dt = torch.float32 if fp32 else torch.float64
            n = self.n_nodes * nb
            V = torch.zeros((m + 1) * n, dtype=dt, device=self.dev)
            z = torch.zeros(n, dtype=dt, device=self.dev)
            w = torch.zeros(n, dtype=dt, device=self.dev)
            dinv = torch.zeros(self.n_nodes * nb * nb, dtype=dt, device=self.dev)
            part = torch.zeros(max(64 * (m + 1), 2048), dtype=torch.float64, device=self.dev)
            dh = torch.zeros(3 * (m + 2), dtype=torch.float64, device=self.dev)
            self._ws[key] = (V, z, w, dinv, part, dh)
        return self._ws[key]

    def krylov(self, b, x, fp32, tol, maxiter, method='gmres', restart=40, check_every=8):
        """Krylov solve with a fallback: when the multigrid-preconditioned
        solve fails (non-finite Arnoldi vectors or a residual that grows) the
        system is solved again with block-Jacobi GMRES."""
        try:
            its, rel = self._krylov(b, x, fp32, tol, maxiter, method, restart, check_every)
        except L.GMError:
            if method != 'mg':
                raise
            its, rel = maxiter, float('inf')
        if not bool(torch.isfinite(x).all()):
            x.zero_()
            if method != 'mg':
                raise L.GMError('%s: the Krylov solution is not finite' % method)
            rel = float('inf')
        if method == 'mg' and not (rel <= 1.0):
            self.stats.setdefault('fallbacks', 0)
            self.stats['fallbacks'] += 1
            its2, rel = self._krylov(b, x, fp32, tol, maxiter, 'gmres', min(restart, 100), check_every)
            its += its2
        return its, rel

    def _hierarchy_matrix(self):
        """Level-0 matrix of the multigrid hierarchy: the Newton matrix itself, or
        (3D indenter, penalty_pc) a copy whose penalty term -pen_par w_I w_J on the
        contact z-dofs is replaced by +pen_par w_I w_J (gm_penalty_pc).  Returns
        (tensor, operator) with operator the Krylov matrix when it differs."""
        pen = self.penalty
        if pen is None or not self.penalty_pc or not (self.dim == 3 and self.coupled) or pen['pen_par'] <= 0.0:
            return self.vals, None  # (a textbook penalty, pen_par < 0, already adds +k w_I w_J)
        n = self.nnzb * self.nb * self.nb
        if self.vals_pc is None or self.vals_pc.dtype != self.vals.dtype or self.vals_pc.numel() < n:
            self.vals_pc = torch.zeros(self.vals.numel(), dtype=self.vals.dtype, device=self.dev)
        self.vals_pc[:n].copy_(self.vals[:n])
        g = self.grid
        h = (g.h[0], g.h[1], g.h[2])
        fx0, fx1, fy0, fy1 = pen['footing']
        m = 3 + int(math.ceil(pen.get('lp_max', 0.5 * h[0]) / h[0]))
        lo = [max(0, int(math.floor(fx0 / h[0])) - m), max(0, int(math.floor(fy0 / h[1])) - m),
              max(0, int(math.floor(pen['height'] / h[2])) - m)]
        hi = [min(g.nx, int(math.ceil(fx1 / h[0])) + m), min(g.ny, int(math.ceil(fy1 / h[1])) + m), g.nz]
        cnt = [max(0, hi[a] - lo[a] + 1) for a in range(3)]
        L.check(L.lib().gm_penalty_pc(L.ctypes.byref(self.grid), L.ctypes.byref(self.bins),
                                      L.ctypes.byref(self.pattern), L.ptr(self.freem), L.ptr(self.vals_pc),
                                      int(self.vals.dtype == torch.float32), (L.c_int * 3)(*lo),
                                      (L.c_int * 3)(*cnt), (L.c_double * 4)(fx0, fx1, fy0, fy1),
                                      float(pen['height']), 2.0 * float(pen['pen_par']), _stream()),
                'gm_penalty_pc')
        return self.vals_pc, self.vals

    def _krylov(self, b, x, fp32, tol, maxiter, method='gmres', restart=40, check_every=8):
        """One Krylov solve K x = b on the current Jacobian: GMRES(restart)
        preconditioned by geometric-multigrid V-cycles ('mg') or block Jacobi
        ('gmres'), or block-Jacobi BiCGStab ('bicgstab'); returns (iterations, relres)."""
        lib = L.lib()
        it = L.c_int(0)
        rr = L.c_double(0.0)
        if method == 'mg':
            entry = self.multigrid(fp32)
            mgh = entry[0]
            full = False  # coarse operators built for this solve (not for a refinement sweep or a
            # later solve on the same Jacobian, which reuse the set-up as it stands)
            if entry[1] != self.jac_version:
                # mg_lag 1: coarse operators kept within a load step; 2: also across load steps
                # (the coarse grids and their patterns do not change; level 0 is refreshed)
                # (entry[2] < 0: the coarse operators have never been built)
                full = (not self.mg_lag or entry[4] or mgh.ublock or entry[2] < 0
                        or (entry[2] != self.step_id and self.mg_lag < 2))
                hv, op = self._hierarchy_matrix() if not mgh.ublock else (self.vals, None)
                mgh.setup(_stream(), vals=hv, op_vals=op, coarse=full)
                entry[1] = self.jac_version
                if full:
                    entry[2], entry[4] = self.step_id, False
            V, z, w, dinv, part, dh = self.gmres_workspace(self.nb, fp32, restart)
            # a solve on kept coarse operators gets at most max(20x, 1000) the iterations of
            # the solve after their setup (a safety net: operators that no longer fit the
            # matrix); past that they are rebuilt and the system solved again.  Ordinary
            # drift (more than 3x + 20) marks them for a rebuild at the next setup instead.
            cap = maxiter if full else min(maxiter, max(20 * entry[3], 1000))
            its, rel = mgh.gmres(b, x, V, z, w, part, dh, restart, tol, cap, _stream())
            if full:
                entry[3] = its
            elif not (rel <= tol) and cap < maxiter:
                self.stats['mg_rebuilds'] = self.stats.get('mg_rebuilds', 0) + 1
                hv, op = self._hierarchy_matrix()
                mgh.setup(_stream(), vals=hv, op_vals=op, coarse=True)
                entry[2], entry[4] = self.step_id, False
                its2, rel = mgh.gmres(b, x, V, z, w, part, dh, restart, tol, maxiter, _stream())
                entry[3] = its2
                its += its2
            elif its > 3 * entry[3] + 20:
                entry[4] = True
            return its, rel
        if method == 'gmres':
            V, z, w, dinv, part, dh = self.gmres_workspace(self.nb, fp32, restart)
            L.check(lib.gm_gmres(self.nb, int(fp32), self.n_nodes, L.ptr(self.row_ptr), L.ptr(self.col_idx),
                                 L.ptr(self.vals), L.ptr(b), L.ptr(x), L.ptr(V), L.ptr(z), L.ptr(w), L.ptr(dinv),
                                 L.ptr(part), L.ptr(dh), restart, tol, maxiter, L.ctypes.byref(it),
                                 L.ctypes.byref(rr), _stream()), 'gm_gmres')
        else:
            ws = self.workspace(self.nb, fp32)
            L.check(lib.gm_bicgstab(self.nb, int(fp32), self.n_nodes, L.ptr(self.row_ptr), L.ptr(self.col_idx),
                                    L.ptr(self.vals), L.ptr(b), L.ptr(x), L.ctypes.byref(ws), tol, maxiter,
                                    check_every, L.ctypes.byref(it), L.ctypes.byref(rr), _stream()), 'gm_bicgstab')
        return it.value, rr.value

    def solve(self, tol, maxiter, check_every=8, ir_steps=1, method='gmres', restart=40, equilibrate=False):
        """Solve K delta = rhs (K = -dR/du).  FP64: Krylov in FP64.  Mixed:
        Krylov in FP32 on the FP32 Jacobian, wrapped in `ir_steps` sweeps of
        iterative refinement whose residual rhs - K delta is formed in FP64.
        equilibrate: scale rows by 1/|K_ii| first (K is overwritten; the
        residual vector is not)."""
        lib = L.lib()
        n = self.n_nodes * self.nb
        b = self.rhs
        if equilibrate:
            if getattr(self, '_rhs_s', None) is None or self._rhs_s.numel() != n:
                self._rhs_s = torch.zeros(n, dtype=torch.float64, device=self.dev)
            L.check(lib.gm_row_scale(self.nb, int(self.precision == 'mixed'), self.n_nodes, L.ptr(self.row_ptr),
                                     L.ptr(self.col_idx), L.ptr(self.vals), L.ptr(self.rhs), L.ptr(self._rhs_s),
                                     _stream()), 'gm_row_scale')
            self.jac_version += 1
            b = self._rhs_s
        if self.precision == 'fp64':
            its, rel = self.krylov(b, self.delta, False, tol, maxiter, method, restart, check_every)
            self.stats['krylov_iters'].append(its)
            self.stats['krylov_relres'].append(rel)
            return
        if not hasattr(self, '_b32'):
            self._b32 = torch.zeros(n, dtype=torch.float32, device=self.dev)
            self._x32 = torch.zeros(n, dtype=torch.float32, device=self.dev)
            self._res64 = torch.zeros(n, dtype=torch.float64, device=self.dev)
            self._ax64 = torch.zeros(n, dtype=torch.float64, device=self.dev)
        self.delta.zero_()
        total = 0
        bnorm = None
        inner_max = maxiter if self.mixed_inner_maxiter is None else min(maxiter, self.mixed_inner_maxiter)
        rel_in = float('inf')
        for k in range(ir_steps):
            if k == 0:
                src = b
            else:
                if rel_in <= tol:
                    break  # the FP32 solve met its tolerance: no further sweep
                # FP64 refinement residual b - K delta; stop once it meets the tolerance
                k64 = not self._tangent_fp32() and not equilibrate  # refine against the FP64 Jacobian
                L.check(lib.gm_residual_refine(self.nb, 0 if k64 else 1, self.n_nodes, L.ptr(self.row_ptr),
                                               L.ptr(self.col_idx), L.ptr(self.vals64 if k64 else self.vals),
                                               L.ptr(b), L.ptr(self.delta), L.ptr(self._ax64), L.ptr(self._res64),
                                               _stream()), 'gm_residual_refine')
                src = self._res64
                if bnorm is None:
                    bnorm = float(torch.linalg.norm(b))
                rel = float(torch.linalg.norm(self._res64)) / bnorm if bnorm > 0 else 0.0
                if rel <= tol:
                    break
            L.check(lib.gm_d2f(n, L.ptr(src), L.ptr(self._b32), _stream()), 'gm_d2f')
            its, rel_in = self.krylov(self._b32, self._x32, True, tol, inner_max, method, restart, check_every)
            total += its
            rel = rel_in
            L.check(lib.gm_axpy(n, 1.0, L.ptr(self._x32), 1, L.ptr(self.delta), _stream()), 'gm_axpy')
        self.stats['krylov_iters'].append(total)
        self.stats['krylov_relres'].append(rel)

    def solve_fp64(self, tol, maxiter, method='mg', restart=250, check_every=8):
        """Mixed form, fallback: solve K delta = rhs in FP64 on the FP64 Jacobian
        (vals64, from an FP64 assembly) with an FP64 hierarchy; replaces the last
        solve's statistics."""
        if self.vals64 is None:
            raise L.GMError('solve_fp64: no FP64 Jacobian assembled')
        saved = self.vals
        self.vals = self.vals64
        self.jac_version += 1  # hierarchies are rebuilt for the FP64 matrix (and again after)
        try:
            self.delta.zero_()
            its, rel = self.krylov(self.rhs, self.delta, False, tol, maxiter, method, restart, check_every)
        finally:
            self.vals = saved
            self.jac_version += 1
        self.stats['krylov_iters'][-1] += its
        self.stats['krylov_relres'][-1] = rel
        self.stats['fp64_fallbacks'] = self.stats.get('fp64_fallbacks', 0) + 1
        return its, rel

    def axpy(self, a, x, y):
        L.check(L.lib().gm_axpy(x.numel(), float(a), L.ptr(x), int(x.dtype == torch.float32), L.ptr(y), _stream()),
                'gm_axpy')

    def sum_into(self, v, slot):
        L.check(L.lib().gm_sum(v.numel(), L.ptr(v), L.ptr(self.part), L.c_void_p(self.scal.data_ptr() + 8 * slot),
                               _stream()), 'gm_sum')

    def g2p(self, x, lp, u_new, u_old, F_old, F_new, b_old, b_new, alpha_old, alpha_new, p_part):
        L.check(L.lib().gm_g2p(L.ctypes.byref(self.grid), x.shape[0], L.ptr(x), L.ptr(lp), L.ptr(u_new),
                               L.ptr(u_old), L.ptr(F_old), L.ptr(F_new), L.ptr(b_old), L.ptr(b_new), L.ptr(alpha_old),
                               L.ptr(alpha_new), L.ptr(p_part), int(self.coupled), _stream()), 'gm_g2p')

    def update_lp(self, F_new, lp, lp0, x, dmin, dmax):
        dmn = (L.c_double * 3)(*(list(dmin) + [0.0] * (3 - len(dmin))))
        dmx = (L.c_double * 3)(*(list(dmax) + [0.0] * (3 - len(dmax))))
        L.check(L.lib().gm_update_lp(x.shape[0], self.dim, L.ptr(F_new), L.ptr(lp), float(lp0), L.ptr(x), dmn, dmx,
                                     _stream()), 'gm_update_lp')


def bsr_to_scipy_dofmajor(row_ptr, col_idx, vals, n_nodes, nb):
    """Host copy of a node-block CSR matrix as a dof-major scipy CSR matrix
    (dof = node + k*n_nodes), for tests and diagnostics."""
    import numpy as np
    import scipy.sparse as sp
    rp = row_ptr.cpu().numpy().astype(np.int64)
    ci = col_idx[:rp[-1]].cpu().numpy().astype(np.int64)
    v = vals[:rp[-1] * nb * nb].double().cpu().numpy().reshape(-1, nb, nb)
    rows_blk = np.repeat(np.arange(n_nodes), np.diff(rp))
    r = (rows_blk[:, None, None] + n_nodes * np.arange(nb)[None, :, None]) * np.ones((1, 1, nb), dtype=np.int64)
    c = (ci[:, None, None] + n_nodes * np.arange(nb)[None, None, :]) * np.ones((1, nb, 1), dtype=np.int64)
    m = sp.coo_matrix((v.ravel(), (r.ravel(), c.ravel())), shape=(n_nodes * nb, n_nodes * nb)).tocsr()
    m.eliminate_zeros()
    return m


class PhaseTimer:
    """Per-phase HIP-event timer on the current stream: mark('start') opens an
    interval, every later mark(name) closes the interval since the previous mark
    and charges it to `name` (milliseconds)."""

    def __init__(self):
        self.events = []

    def mark(self, name):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        self.events.append((name, e))

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for (n0, e0), (n1, e1) in zip(self.events[:-1], self.events[1:]):
            if n1 == 'start':
                continue
            out[n1] = out.get(n1, 0.0) + e0.elapsed_time(e1)
        return out

"""Geometric multigrid hierarchy over the background grid (device buffers and
the calls of gm_mg_setup / gm_gmres_mg).

Level 0 is the Newton Jacobian in its per-step block-CSR pattern; level l+1
keeps every second node of level l per axis (ceil(n/2) cells).  Coarse levels
use the full 5^dim window pattern clipped to the grid, which depends on the
grid only, so their patterns are built once.
"""
import numpy as np
import torch

from .. import _lib as L


def window_pattern(dim, cells):
    """row_ptr, col_idx and window bitmasks (2 x uint64 per row, as int64) of
    the full 5^dim window clipped to a grid of `cells` cells per axis."""
    n1 = [c + 1 for c in cells[:dim]]
    grids = np.meshgrid(*[np.arange(k) for k in reversed(n1)], indexing='ij')
    coords = np.stack([g.reshape(-1) for g in reversed(grids)], 1)  # x fastest
    W = 5 ** dim
    bits = np.arange(W)
    offs = np.stack([bits % 5 - 2, (bits // 5) % 5 - 2] + ([bits // 25 - 2] if dim == 3 else []), 1)
    nbc = coords[:, None, :] + offs[None, :, :]
    valid = np.all((nbc >= 0) & (nbc < np.array(n1)[None, None, :]), axis=2)
    strides = np.array([1, n1[0], n1[0] * (n1[1] if dim == 3 else 1)])[:dim]
    nb_idx = (nbc * strides[None, None, :]).sum(2)
    counts = valid.sum(1)
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    col_idx = nb_idx[valid].astype(np.int32)
    m = np.zeros((coords.shape[0], 2), dtype=np.uint64)
    for b in range(W):
        word, sh = divmod(b, 64)
        m[:, word] |= valid[:, b].astype(np.uint64) << np.uint64(sh)
    return row_ptr, col_idx, m.view(np.int64)


def _robust_inverse(A):
    """Inverse of the small coarsest operator: LU when it is regular, else a
    pseudo-inverse from the (slower, more robust) QR-iteration SVD driver."""
    try:
        inv = np.linalg.inv(A)
        if np.all(np.isfinite(inv)):
            return inv
    except np.linalg.LinAlgError:
        pass
    import scipy.linalg
    u, sv, vt = scipy.linalg.svd(A, full_matrices=False, lapack_driver='gesvd')
    keep = sv > sv.max() * max(A.shape) * np.finfo(np.float64).eps
    return (vt[keep].T / sv[keep]) @ u[:, keep].T


class Multigrid:
    """Hierarchy for the whole system (single phase) or, with ``ublock=True``,
    for the displacement block of a u-p system (nb - 1 dofs per node) inside the
    block-triangular u-p preconditioner (gm_up_setup / gm_gmres_up)."""

    def __init__(self, eng, fp32, max_levels=8, min_cells=2, nu_pre=1, nu_post=2, coarse_sweeps=20, omega=0.4,
                 max_dense=3000, ublock=False, gamma=1, lowprec=None, stop_dofs=2000):
        self.eng = eng
        self.fp32 = bool(fp32)
        # FP16 copies of the level matrices for the smoother (FP32 hierarchies; default on)
        self.lowprec = self.fp32 and (True if lowprec is None else bool(lowprec))
        self.dim = eng.dim
        self.ublock = bool(ublock) and eng.coupled
        self.nb = eng.nb - 1 if self.ublock else eng.nb
        self.nu_pre, self.nu_post, self.coarse_sweeps, self.omega = nu_pre, nu_post, coarse_sweeps, omega
        g = eng.grid
        dims = [(g.nx, g.ny, g.nz if self.dim == 3 else 0)]
        while len(dims) < max_levels:
            nx, ny, nz = dims[-1]
            # at least two levels; then stop at the first level small enough for the dense
            # coarsest solve (stop_dofs): the levels below it are a handful of short launches
            # each per cycle (on the 64 x 64 x 32-cell indentation stopping at 8 x 8 x 4 cells,
            # 1620 dofs, takes 3 % off both forms' time)
            if len(dims) >= 2 and (nx + 1) * (ny + 1) * ((nz + 1) if self.dim == 3 else 1) * self.nb <= stop_dofs:
                break
            c = ((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2 if self.dim == 3 else 0)
            if min(c[:self.dim]) < min_cells:
                break
            dims.append(c)
        self.dims = dims
        dev = eng.dev
        dt = torch.float32 if self.fp32 else torch.float64
        nb2 = self.nb * self.nb
        self._keep = []
        self.levels = (L.gm_mg_level * len(dims))()
        for l, (nx, ny, nz) in enumerate(dims):
            n = (nx + 1) * (ny + 1) * ((nz + 1) if self.dim == 3 else 1)
            lv = self.levels[l]
            lv.nx, lv.ny, lv.nz, lv.n_nodes = nx, ny, nz, n
            lv.gamma = int(gamma)
            bufs = dict(dinv=torch.zeros(n * nb2, dtype=dt, device=dev),
                        r=torch.zeros(n * self.nb, dtype=dt, device=dev))
            if l > 0:
                rp, ci, mk = window_pattern(self.dim, (nx, ny, nz))
                bufs.update(rp=torch.from_numpy(rp).to(dev), ci=torch.from_numpy(ci).to(dev),
                            mask=torch.from_numpy(mk.reshape(-1).copy()).to(dev),
                            vals=torch.zeros(len(ci) * nb2, dtype=torch.float64, device=dev),
                            freem=torch.zeros(n * self.nb, dtype=torch.uint8, device=dev),
                            x=torch.zeros(n * self.nb, dtype=dt, device=dev),
                            b=torch.zeros(n * self.nb, dtype=dt, device=dev))
                lv.nnzb = len(ci)
                lv.row_ptr, lv.col_idx, lv.row_mask = (bufs['rp'].data_ptr(), bufs['ci'].data_ptr(),
                                                       bufs['mask'].data_ptr())
                lv.vals, lv.freem = bufs['vals'].data_ptr(), bufs['freem'].data_ptr()
                lv.x, lv.b = bufs['x'].data_ptr(), bufs['b'].data_ptr()
                if self.lowprec:
                    bufs.update(lo=torch.zeros(len(ci) * nb2, dtype=torch.float16, device=dev),
                                scale=torch.zeros(n * nb2, dtype=torch.float32, device=dev))
                    lv.vals_lo, lv.row_scale = bufs['lo'].data_ptr(), bufs['scale'].data_ptr()
            lv.dinv, lv.r = bufs['dinv'].data_ptr(), bufs['r'].data_ptr()
            self._keep.append(bufs)
        n0 = self.levels[0].n_nodes
        # Galerkin workspace, coarse operators and the coarsest inverse are FP64 in both precisions
        self.ap = torch.zeros(n0 * (4 ** self.dim) * nb2, dtype=torch.float64, device=dev) if len(dims) > 1 else None
        self.dt = dt
        if self.ublock:
            # level 0 holds the extracted displacement block; the Schur diagonal and split vectors
            self._keep[0].update(vals=torch.zeros(0, dtype=dt, device=dev),
                                 freem=torch.zeros(n0 * self.nb, dtype=torch.uint8, device=dev))
            self.levels[0].freem = self._keep[0]['freem'].data_ptr()
            self.sinv = torch.zeros(n0, dtype=dt, device=dev)
            self.vu = torch.zeros(n0 * self.nb, dtype=dt, device=dev)
            self.zu = torch.zeros(n0 * self.nb, dtype=dt, device=dev)
        # the coarsest level is solved exactly through a dense inverse when it is small enough
        nc = self.levels[len(dims) - 1].n_nodes * self.nb
        self.dense = len(dims) > 1 and nc <= max_dense
        self.cinv = torch.zeros(nc * nc, dtype=torch.float64, device=dev) if self.dense else None
        self.levels[len(dims) - 1].dense_inv = self.cinv.data_ptr() if self.dense else None

    @property
    def n_levels(self):
        return len(self.dims)

    def setup(self, stream, vals=None, op_vals=None, coarse=True):
        """Point level 0 at the current Jacobian (or at `vals`, a matrix in its
        pattern the hierarchy is built from, with op_vals then the Krylov
        operator) and build the coarse operators; with coarse=False only level 0
        (its block-diagonal inverses and FP16 copy) is refreshed and the coarse
        operators of the last full setup are kept (same pattern and free dofs)."""
        eng = self.eng
        lv = self.levels[0]
        self.hier_vals = eng.vals if vals is None else vals
        self.op_vals = op_vals
        lv.nnzb = eng.nnzb
        lv.row_ptr, lv.col_idx, lv.row_mask = eng.row_ptr.data_ptr(), eng.col_idx.data_ptr(), \
            eng.row_mask.data_ptr()
        if self.lowprec:
            k0 = self._keep[0]
            need = max(eng.nnzb, 1) * self.nb * self.nb
            if k0.get('lo') is None or k0['lo'].numel() < need:
                k0['lo'] = torch.zeros(int(need * 1.15) + 16, dtype=torch.float16, device=eng.dev)
            if k0.get('scale') is None:
                k0['scale'] = torch.zeros(lv.n_nodes * self.nb * self.nb, dtype=torch.float32, device=eng.dev)
            lv.vals_lo, lv.row_scale = k0['lo'].data_ptr(), k0['scale'].data_ptr()
        if self.ublock:
            k0 = self._keep[0]
            need = max(eng.nnzb, 1) * self.nb * self.nb
            if k0['vals'].numel() < need:
                k0['vals'] = torch.zeros(int(need * 1.15) + 16, dtype=self.dt, device=eng.dev)
            lv.vals = k0['vals'].data_ptr()
            L.check(L.lib().gm_up_setup(self.dim, eng.nb, int(self.fp32), eng.n_nodes, eng.nnzb,
                                        L.ptr(eng.row_ptr), L.ptr(eng.col_idx), L.ptr(eng.vals), L.ptr(eng.freem),
                                        self.n_levels, self.levels, L.ptr(self.ap), L.ptr(self.sinv), stream),
                    'gm_up_setup')
        else:
            lv.vals, lv.freem = self.hier_vals.data_ptr(), eng.freem.data_ptr()
            L.check(L.lib().gm_mg_setup(self.dim, self.nb, int(self.fp32), self.n_levels if coarse else 1,
                                        self.levels, L.ptr(self.ap), stream), 'gm_mg_setup')
            if not coarse:
                return
        if self.dense:
            # exact coarsest solve: pseudo-inverse of the (small) coarsest operator, formed on the host
            k = self._keep[-1]
            lv = self.levels[self.n_levels - 1]
            nb, n = self.nb, lv.n_nodes
            rp = k['rp'].cpu().numpy().astype(np.int64)
            ci = k['ci'].cpu().numpy().astype(np.int64)
            v = k['vals'].double().cpu().numpy().reshape(-1, nb, nb)
            A = np.zeros((n * nb, n * nb))
            brow = np.repeat(np.arange(n), np.diff(rp))
            for r in range(nb):
                for c in range(nb):
                    A[brow * nb + r, ci * nb + c] += v[:, r, c]
            self.cinv.copy_(torch.from_numpy(_robust_inverse(A).reshape(-1)))

    def gmres(self, b, x, V, z, w, part, dh, m, tol, maxiter, stream):
        eng = self.eng
        n = self.levels[0].n_nodes * eng.nb
        if getattr(self, '_Z', None) is None or self._Z.numel() < m * n:
            self._Z = torch.zeros(m * n, dtype=V.dtype, device=V.device)
        it = L.c_int(0)
        rr = L.c_double(0.0)
        if self.ublock:
            L.check(L.lib().gm_gmres_up(self.dim, eng.nb, int(self.fp32), eng.n_nodes, L.ptr(eng.row_ptr),
                                        L.ptr(eng.col_idx), L.ptr(eng.vals), L.ptr(self.sinv), self.n_levels,
                                        self.levels, self.nu_pre, self.nu_post, self.coarse_sweeps, self.omega,
                                        L.ptr(self.vu), L.ptr(self.zu), L.ptr(b), L.ptr(x), L.ptr(V),
                                        L.ptr(self._Z), L.ptr(z), L.ptr(w), L.ptr(part), L.ptr(dh), m, tol,
                                        maxiter, L.ctypes.byref(it), L.ctypes.byref(rr), stream), 'gm_gmres_up')
            return it.value, rr.value
        L.check(L.lib().gm_gmres_mg(self.dim, self.nb, int(self.fp32), self.n_levels, self.levels,
                                    L.ptr(getattr(self, 'op_vals', None)), self.nu_pre,
                                    self.nu_post, self.coarse_sweeps, self.omega, L.ptr(b), L.ptr(x), L.ptr(V),
                                    L.ptr(self._Z), L.ptr(z), L.ptr(w), L.ptr(part), L.ptr(dh), m, tol, maxiter,
                                    L.ctypes.byref(it), L.ctypes.byref(rr), stream), 'gm_gmres_mg')
        return it.value, rr.value

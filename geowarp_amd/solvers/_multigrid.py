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


class Multigrid:
    def __init__(self, eng, fp32, max_levels=8, min_cells=2, nu_pre=2, nu_post=2, coarse_sweeps=20, omega=0.7):
        self.eng = eng
        self.fp32 = bool(fp32)
        self.dim = eng.dim
        self.nb = eng.nb
        self.nu_pre, self.nu_post, self.coarse_sweeps, self.omega = nu_pre, nu_post, coarse_sweeps, omega
        g = eng.grid
        dims = [(g.nx, g.ny, g.nz if self.dim == 3 else 0)]
        while len(dims) < max_levels:
            nx, ny, nz = dims[-1]
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
            bufs = dict(dinv=torch.zeros(n * nb2, dtype=dt, device=dev),
                        r=torch.zeros(n * self.nb, dtype=dt, device=dev))
            if l > 0:
                rp, ci, mk = window_pattern(self.dim, (nx, ny, nz))
                bufs.update(rp=torch.from_numpy(rp).to(dev), ci=torch.from_numpy(ci).to(dev),
                            mask=torch.from_numpy(mk.reshape(-1).copy()).to(dev),
                            vals=torch.zeros(len(ci) * nb2, dtype=dt, device=dev),
                            freem=torch.zeros(n * self.nb, dtype=torch.uint8, device=dev),
                            x=torch.zeros(n * self.nb, dtype=dt, device=dev),
                            b=torch.zeros(n * self.nb, dtype=dt, device=dev))
                lv.nnzb = len(ci)
                lv.row_ptr, lv.col_idx, lv.row_mask = (bufs['rp'].data_ptr(), bufs['ci'].data_ptr(),
                                                       bufs['mask'].data_ptr())
                lv.vals, lv.freem = bufs['vals'].data_ptr(), bufs['freem'].data_ptr()
                lv.x, lv.b = bufs['x'].data_ptr(), bufs['b'].data_ptr()
            lv.dinv, lv.r = bufs['dinv'].data_ptr(), bufs['r'].data_ptr()
            self._keep.append(bufs)
        n0 = self.levels[0].n_nodes
        self.ap = torch.zeros(n0 * (4 ** self.dim) * nb2, dtype=dt, device=dev) if len(dims) > 1 else None

    @property
    def n_levels(self):
        return len(self.dims)

    def setup(self, stream):
        """Point level 0 at the current Jacobian and build the coarse operators."""
        eng = self.eng
        lv = self.levels[0]
        lv.nnzb = eng.nnzb
        lv.row_ptr, lv.col_idx, lv.row_mask = eng.row_ptr.data_ptr(), eng.col_idx.data_ptr(), \
            eng.row_mask.data_ptr()
        lv.vals, lv.freem = eng.vals.data_ptr(), eng.freem.data_ptr()
        L.check(L.lib().gm_mg_setup(self.dim, self.nb, int(self.fp32), self.n_levels, self.levels,
                                    L.ptr(self.ap), stream), 'gm_mg_setup')

    def gmres(self, b, x, V, z, w, part, dh, m, tol, maxiter, stream):
        it = L.c_int(0)
        rr = L.c_double(0.0)
        L.check(L.lib().gm_gmres_mg(self.dim, self.nb, int(self.fp32), self.n_levels, self.levels, self.nu_pre,
                                    self.nu_post, self.coarse_sweeps, self.omega, L.ptr(b), L.ptr(x), L.ptr(V),
                                    L.ptr(z), L.ptr(w), L.ptr(part), L.ptr(dh), m, tol, maxiter,
                                    L.ctypes.byref(it), L.ctypes.byref(rr), stream), 'gm_gmres_mg')
        return it.value, rr.value

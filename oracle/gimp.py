"""cpGIMP shape functions (oracle restatement, test infrastructure only).

Follows GeoWarp mpm/gimp.py:
  * 1-D cpGIMP weights/gradients:   get_GIMP_shape_function_and_gradient_2d  gimp.py:25-112
                                    get_GIMP_shape_function_and_gradient_3d  gimp.py:197-315
  * cell-average (PPP) weights:      get_GIMP_shape_function_and_gradient_avg_2d gimp.py:115-193
                                    get_GIMP_shape_function_and_gradient_avg_3d gimp.py:318-425
  * particle half-length update:     update_GIMP_lp_2d/3d gimp.py:428-507

The reference evaluates one axis at a time with the same formulae, so the
oracle exposes a vectorised per-axis routine ``axis_weights`` and builds the
tensor-product stencil from it.
"""
import numpy as np


def c_int(v):
    """C-style float->int conversion (truncation toward zero), as wp.int()."""
    return np.trunc(v).astype(np.int64)


def axis_weights(x, lp, h, inv_h):
    """Per-axis cpGIMP data for arrays of particle coordinates.

    gimp.py:33-49 (corner/base floors with the +1e-8 bias), :67-86 (branch on
    whether the particle domain lies in one cell), :156-171 (average weights).

    Returns (lo, w, g, wa, ext):
      lo  int64 (N,)   first stencil node index along this axis
      w   (N,3)        weights of nodes lo, lo+1, lo+2
      g   (N,3)        d w / d x
      wa  (N,3)        cell-average weights used by the PPP stabilisation
      ext (N,)         2 when the particle lies in one cell, else 3
    """
    x = np.asarray(x, dtype=np.float64)
    lp = np.asarray(lp, dtype=np.float64)
    lo = c_int((x - lp) * inv_h + 1e-8)
    hi = c_int((x + lp) * inv_h + 1e-8)
    base = c_int(x * inv_h + 1e-8)
    f = x * inv_h - base.astype(np.float64)
    n = x.shape[0]
    w = np.zeros((n, 3))
    g = np.zeros((n, 3))
    wa = np.zeros((n, 3))
    one = lo == hi
    # particle domain inside one cell: linear hat (gimp.py:67-71, avg :156-158)
    w[one, 0] = 1.0 - f[one]
    w[one, 1] = f[one]
    g[one, 0] = -inv_h
    g[one, 1] = inv_h
    wa[one, 0] = 0.5
    wa[one, 1] = 0.5
    two = ~one
    if np.any(two):
        xs = x[two]
        l = lp[two]
        lof = lo[two].astype(np.float64)
        x0 = lof * h
        x1 = (lof + 1.0) * h
        x2 = (lof + 2.0) * h
        d0 = xs - x0
        d2 = xs - x2
        a0 = h + l - np.abs(d0)
        a2 = h + l - np.abs(d2)
        # Coombs et al. section 2.4.2 (gimp.py:76-86)
        w[two, 0] = a0 ** 2 / (4.0 * h * l)
        g[two, 0] = -a0 / (2.0 * h * l) * d0 / np.abs(d0)
        w[two, 1] = 1.0 - ((xs - x1) ** 2 + l ** 2) / (2.0 * h * l)
        g[two, 1] = -(xs - x1) / (h * l)
        w[two, 2] = a2 ** 2 / (4.0 * h * l)
        g[two, 2] = -a2 / (2.0 * h * l) * d2 / np.abs(d2)
        # cell averages (gimp.py:163-171)
        wa[two, 0] = a0 / (4.0 * l)
        wa[two, 1] = 0.5
        wa[two, 2] = a2 / (4.0 * l)
    ext = np.where(one, 2, 3)
    # a domain ending exactly on a cell face takes the cpGIMP branch with an
    # exactly zero third node; it adds exact zeros and is left out of the support
    zero3 = (~one) & (w[:, 2] == 0.0) & (g[:, 2] == 0.0) & (wa[:, 2] == 0.0)
    ext = np.where(zero3, 2, ext)
    return lo, w, g, wa, ext


class Stencil:
    """Tensor-product stencil of every particle: 3^d nodes in reference order.

    2D order is (i, j) with j fastest only for storage; node index is
    ix + iy*(nx+1) (mpm_utils:132).  3D index is ix + iy*(nx+1) + iz*(nx+1)(ny+1)
    (mpm_utils:336).  ``spacing`` is (dx,) for 2D — the reference 2D shape
    functions use dx on both axes (gimp.py:26-31) — or (dx, dy, dz) for 3D.
    """

    def __init__(self, x, lp, dim, spacing, n_grid):
        x = np.asarray(x, dtype=np.float64)
        lp = np.asarray(lp, dtype=np.float64)
        self.dim = dim
        self.n = x.shape[0]
        if dim == 2:
            hs = (spacing[0], spacing[0])
        else:
            hs = tuple(spacing)
        self.lo = []
        self.w = []
        self.g = []
        self.wa = []
        self.ext = []
        for a in range(dim):
            lo, w, g, wa, ext = axis_weights(x[:, a], lp[:, a], hs[a], 1.0 / hs[a])
            self.lo.append(lo)
            self.w.append(w)
            self.g.append(g)
            self.wa.append(wa)
            self.ext.append(ext)
        nx = n_grid[0]
        ny = n_grid[1]
        K = 3 ** dim
        self.K = K
        self.idx = np.zeros((self.n, K), dtype=np.int64)
        self.W = np.zeros((self.n, K))
        self.Wa = np.zeros((self.n, K))
        self.G = np.zeros((self.n, K, dim))
        self.offsets = []
        k = 0
        if dim == 2:
            for i in range(3):
                for j in range(3):
                    ix = self.lo[0] + i
                    iy = self.lo[1] + j
                    self.idx[:, k] = ix + iy * (nx + 1)
                    self.W[:, k] = self.w[0][:, i] * self.w[1][:, j]
                    self.Wa[:, k] = self.wa[0][:, i] * self.wa[1][:, j]
                    self.G[:, k, 0] = self.g[0][:, i] * self.w[1][:, j]
                    self.G[:, k, 1] = self.w[0][:, i] * self.g[1][:, j]
                    self.offsets.append((i, j))
                    k += 1
        else:
            for i in range(3):
                for j in range(3):
                    for l in range(3):
                        ix = self.lo[0] + i
                        iy = self.lo[1] + j
                        iz = self.lo[2] + l
                        self.idx[:, k] = ix + iy * (nx + 1) + iz * (nx + 1) * (ny + 1)
                        wx, wy, wz = self.w[0][:, i], self.w[1][:, j], self.w[2][:, l]
                        gx, gy, gz = self.g[0][:, i], self.g[1][:, j], self.g[2][:, l]
                        self.W[:, k] = wx * wy * wz
                        self.Wa[:, k] = self.wa[0][:, i] * self.wa[1][:, j] * self.wa[2][:, l]
                        self.G[:, k, 0] = gx * wy * wz
                        self.G[:, k, 1] = wx * gy * wz
                        self.G[:, k, 2] = wx * wy * gz
                        self.offsets.append((i, j, l))
                        k += 1
        self.offsets = np.array(self.offsets)
        # in-support mask: node offset below the per-axis extent
        self.support = np.ones((self.n, K), dtype=bool)
        for a in range(dim):
            self.support &= self.offsets[None, :, a] < self.ext[a][:, None]
        # node inside the background grid (the reference does not check; an
        # out-of-grid node only ever carries a zero weight in valid set-ups)
        self.valid = np.ones((self.n, K), dtype=bool)
        for a in range(dim):
            c = self.lo[a][:, None] + self.offsets[None, :, a]
            self.valid &= (c >= 0) & (c <= n_grid[a])
        if np.any(self.support & ~self.valid):
            raise IndexError("particle support leaves the background grid")
        # off-grid slots carry zero weight and gradient; point them at node 0
        self.idx = np.where(self.valid, self.idx, 0)


def init_lp(n, lp0, dim):
    """initialize_GIMP_lp_2d/3d (gimp.py:8-22)."""
    return np.full((n, dim), lp0, dtype=np.float64)


def update_lp(F_new, x, lp0, dmin, dmax):
    """update_GIMP_lp_2d/3d (gimp.py:428-507): lp = lp0*||F col a||, clipped to
    keep x +- lp inside the domain box with a 1e-6 margin."""
    dim = x.shape[1]
    lp = np.zeros_like(x)
    for a in range(dim):
        U = np.sqrt(F_new[:, 0, a] ** 2 + F_new[:, 1, a] ** 2 + F_new[:, 2, a] ** 2)
        l = lp0 * U
        m = x[:, a] - l < dmin[a]
        l = np.where(m, x[:, a] - dmin[a] - 1e-6, l)
        m = x[:, a] + l > dmax[a]
        l = np.where(m, dmax[a] - x[:, a] - 1e-6, l)
        lp[:, a] = l
    return lp

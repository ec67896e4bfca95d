"""cpGIMP entry points with the reference's names and positional arguments
(GeoWarp mpm/gimp.py).  Arrays are torch tensors on the ROCm device; the work
runs in libgeowarp_amd.so.

The reference's ``get_GIMP_shape_function_and_gradient_*`` are per-particle
``wp.func``s; here they take tensors of particles and return one row per
particle in the reference's component order.
"""
import torch

from .. import _lib as L


def _stream():
    return L.c_void_p(torch.cuda.current_stream().cuda_stream)


def initialize_GIMP_lp_2d(GIMP_lp, GIMP_lp_initial):
    """gimp.py:8-14."""
    GIMP_lp.fill_(float(GIMP_lp_initial))


def initialize_GIMP_lp_3d(GIMP_lp, GIMP_lp_initial):
    """gimp.py:16-22."""
    GIMP_lp.fill_(float(GIMP_lp_initial))


def _weights(xp, lp, hs, avg):
    n, dim = xp.shape
    xp = xp.contiguous()
    lp = lp.contiguous()
    out = torch.empty(n, 6 * dim, dtype=torch.float64, device=xp.device)
    av = torch.empty(n, 6 * dim, dtype=torch.float64, device=xp.device) if avg else None
    h = (L.c_double * 3)(*(list(hs) + [0.0] * (3 - len(hs))))
    L.check(L.lib().gm_gimp_weights(n, dim, L.ptr(xp), L.ptr(lp), h, L.ptr(out), L.ptr(av), None, _stream()),
            'gm_gimp_weights')
    return av if avg else out


def get_GIMP_shape_function_and_gradient_2d(xp, dx, inv_dx, lpx, lpy):
    """gimp.py:25-112: rows (wx0, wy0, wx1, wy1, wx2, wy2, gwx0, gwy0, gwx1, gwy1, gwx2, gwy2);
    dx is used on both axes as in the reference."""
    return _weights(xp, torch.stack([lpx, lpy], 1), (dx, dx), False)


def get_GIMP_shape_function_and_gradient_avg_2d(xp, dx, inv_dx, lpx, lpy):
    """gimp.py:115-193 (cell-average weights and their gradients)."""
    return _weights(xp, torch.stack([lpx, lpy], 1), (dx, dx), True)


def get_GIMP_shape_function_and_gradient_3d(xp, dx, dy, dz, inv_dx, inv_dy, inv_dz, lpx, lpy, lpz):
    """gimp.py:197-315: rows (w0 xyz, w1 xyz, w2 xyz, g0 xyz, g1 xyz, g2 xyz)."""
    return _weights(xp, torch.stack([lpx, lpy, lpz], 1), (dx, dy, dz), False)


def get_GIMP_shape_function_and_gradient_avg_3d(xp, dx, dy, dz, inv_dx, inv_dy, inv_dz, n_grid_x, n_grid_y, lpx, lpy,
                                                lpz):
    """gimp.py:318-425."""
    return _weights(xp, torch.stack([lpx, lpy, lpz], 1), (dx, dy, dz), True)


def _update(F, GIMP_lp, lp0, x, dmin, dmax):
    dim = x.shape[1]
    dmn = (L.c_double * 3)(*(list(dmin) + [0.0] * (3 - len(dmin))))
    dmx = (L.c_double * 3)(*(list(dmax) + [0.0] * (3 - len(dmax))))
    L.check(L.lib().gm_update_lp(x.shape[0], dim, L.ptr(F), L.ptr(GIMP_lp), float(lp0), L.ptr(x), dmn, dmx,
                                 _stream()), 'gm_update_lp')


def update_GIMP_lp_2d(deformation_gradient_total_new, GIMP_lp, GIMP_lp_initial, x_particles, min_x, max_x, min_y,
                      max_y):
    """gimp.py:428-462."""
    _update(deformation_gradient_total_new, GIMP_lp, GIMP_lp_initial, x_particles, (min_x, min_y), (max_x, max_y))


def update_GIMP_lp_3d(deformation_gradient_total_new, GIMP_lp, GIMP_lp_initial, x_particles, min_x, max_x, min_y,
                      max_y, min_z, max_z):
    """gimp.py:465-507."""
    _update(deformation_gradient_total_new, GIMP_lp, GIMP_lp_initial, x_particles, (min_x, min_y, min_z),
            (max_x, max_y, max_z))

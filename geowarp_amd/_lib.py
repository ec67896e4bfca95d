"""ctypes boundary to libgeowarp_amd.so (the C ABI declared in include/geowarp_amd.h).

The product path has no CPU fallback: importing the solvers without the built
library, or calling them without a ROCm device, raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libgeowarp_amd.so')

c_int = ctypes.c_int
c_long = ctypes.c_long
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p


class gm_grid(ctypes.Structure):
    _fields_ = [('dim', c_int), ('nx', c_int), ('ny', c_int), ('nz', c_int), ('n_nodes', c_int), ('nb', c_int),
                ('h', c_double * 3), ('inv_h', c_double * 3)]


class gm_bins(ctypes.Structure):
    _fields_ = [('n_particles', c_int), ('cell_start', c_void_p), ('cell_count', c_void_p), ('sorted', c_void_p),
                ('ext', c_void_p), ('scan_tmp', c_void_p), ('ext_s', c_void_p), ('x_s', c_void_p),
                ('lp_s', c_void_p)]


class gm_pattern(ctypes.Structure):
    _fields_ = [('row_ptr', c_void_p), ('col_idx', c_void_p), ('row_mask', c_void_p), ('nnzb', c_int)]


class gm_norsand_params(ctypes.Structure):
    _fields_ = [(k, c_double) for k in ('kappa', 'tilde_lambda', 'M', 'N', 'beta', 'vc0', 'p0', 'v0', 'h', 'mu',
                                        'ev0', 'tol')]


class gm_stage_args(ctypes.Structure):
    _fields_ = [('n_particles', c_int), ('coupled', c_int), ('want_tangent', c_int), ('tangent_fp32', c_int),
                ('x', c_void_p), ('lp', c_void_p), ('F_old', c_void_p), ('F_new', c_void_p), ('b_old', c_void_p),
                ('b_new', c_void_p), ('sigma', c_void_p), ('alpha_old', c_void_p), ('alpha_new', c_void_p),
                ('ext_flag', c_void_p), ('trac_flag', c_void_p), ('u_new', c_void_p), ('u_old', c_void_p),
                ('sorted', c_void_p), ('Y', c_void_p), ('D', c_void_p), ('pen_force', c_void_p),
                ('mat_kind', c_int), ('dp_apex', c_int),
                ('lam', c_double), ('mu', c_double), ('E', c_double), ('kappa', c_double), ('H', c_double),
                ('phi', c_double), ('psi', c_double), ('coh', c_double), ('shape', c_double),
                ('p_vol', c_double), ('p_rho', c_double), ('phi0', c_double), ('mobility', c_double),
                ('dt', c_double), ('PPD', c_double),
                ('gravity', c_double * 3), ('point_load', c_double * 3), ('traction', c_double * 3),
                ('footing', c_double * 4), ('footing_height', c_double), ('pen_par', c_double), ('penalty', c_int),
                ('p_phi', c_void_p), ('p_coh', c_void_p), ('ns', gm_norsand_params)]




class gm_local_args(ctypes.Structure):
    _fields_ = [('n_particles', c_int), ('nx1', c_int), ('mat_kind', c_int), ('dp_apex', c_int),
                ('h', c_double), ('shape', c_double), ('p_vol', c_double), ('p_rho', c_double),
                ('gravity', c_double * 2), ('point_load', c_double * 2), ('lp0', c_double),
                ('dmin', c_double * 2), ('dmax', c_double * 2)] + \
        [(k, c_double) for k in ('E', 'nu', 'kappa', 'H', 'phi', 'psi', 'coh')] + \
        [(k, c_void_p) for k in ('p_phi', 'p_coh', 'x', 'lp', 'F_old', 'b_old', 'alpha_old', 'ext_flag', 'u',
                                 'cot', 'cot_scale', 'xbar', 'lpbar', 'Fbar', 'bbar', 'abar')]


class gm_mg_level(ctypes.Structure):
    _fields_ = [('nx', c_int), ('ny', c_int), ('nz', c_int), ('n_nodes', c_int), ('nnzb', c_int),
                ('row_ptr', c_void_p), ('col_idx', c_void_p), ('row_mask', c_void_p), ('vals', c_void_p),
                ('dinv', c_void_p), ('freem', c_void_p), ('x', c_void_p), ('b', c_void_p), ('r', c_void_p),
                ('dense_inv', c_void_p), ('gamma', c_int), ('vals_lo', c_void_p), ('row_scale', c_void_p)]


class gm_krylov_ws(ctypes.Structure):
    _fields_ = [('r', c_void_p), ('rhat', c_void_p), ('p', c_void_p), ('v', c_void_p), ('s', c_void_p),
                ('t', c_void_p), ('y', c_void_p), ('z', c_void_p), ('dinv', c_void_p), ('part', c_void_p),
                ('sc', c_void_p)]


_P = c_void_p
_SIGS = {
    'gm_version': (c_int, []),
    'gm_last_error': (ctypes.c_char_p, []),
    'gm_exclusive_scan': (c_int, [_P, _P, c_int, _P, _P]),
    'gm_bin_particles': (c_int, [ctypes.POINTER(gm_grid), _P, _P, ctypes.POINTER(gm_bins), _P, _P, _P]),
    'gm_bsr_pattern': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern), _P,
                               c_int, _P]),
    'gm_particle_stage': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_stage_args), _P]),
    'gm_node_residual': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), _P, _P, _P, c_int, _P]),
    'gm_node_jacobian': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern), _P,
                                 c_int, _P, _P, c_int, c_int, _P]),
    'gm_node_jacobian_work_bytes': (ctypes.c_size_t, [ctypes.POINTER(gm_grid), c_int]),
    'gm_jacobian_bins': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), _P, c_int, _P, ctypes.c_size_t,
                                 _P]),
    'gm_jacobian_gather': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern),
                                   c_int, _P, _P, _P, c_int, _P]),
    'gm_jacobian_bins_variant': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), _P, c_int, c_int, _P,
                                         ctypes.c_size_t, _P]),
    'gm_jacobian_gather_variant': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins),
                                           ctypes.POINTER(gm_pattern), c_int, c_int, _P, _P, _P, c_int, _P]),
    'gm_jacobian_autotune': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern),
                                     _P, c_int, _P, _P, c_int, _P, ctypes.c_size_t, c_int, _P, _P, _P, _P]),
    'gm_mass_work_bytes': (ctypes.c_size_t, [ctypes.POINTER(gm_grid)]),
    'gm_stage_de_3d': (c_int, [c_int, _P, _P, _P, _P, _P, _P, c_double, c_double, _P, _P]),
    'gm_residual_work_bytes': (ctypes.c_size_t, [ctypes.POINTER(gm_grid)]),
    'gm_node_residual_binned': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), _P, _P, _P, _P,
                                        ctypes.c_size_t, _P]),
    'gm_mass_binned': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern),
                               c_double, _P, _P, ctypes.c_size_t, _P]),
    'gm_node_jacobian_binned': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins),
                                        ctypes.POINTER(gm_pattern), _P, c_int, _P, _P, c_int, _P, ctypes.c_size_t,
                                        _P]),
    'gm_p2g': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern),
                       c_double, _P, c_double, c_double, _P, _P, _P, _P, _P, _P]),
    'gm_p2g_work_bytes': (ctypes.c_size_t, [ctypes.POINTER(gm_grid)]),
    'gm_pattern_kernel': (c_int, [c_int]),
    'gm_p2g_binned': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), c_double, _P, c_double, c_double,
                              _P, _P, _P, _P, _P, ctypes.c_size_t, _P]),
    'gm_free_mask': (c_int, [c_int, _P, _P, _P, _P]),
    'gm_set_nodal_pressure': (c_int, [c_int, c_int, _P, _P, c_int, _P, _P, _P, _P]),
    'gm_g2p': (c_int, [ctypes.POINTER(gm_grid), c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_int, _P]),
    'gm_update_lp': (c_int, [c_int, c_int, _P, _P, c_double, _P, ctypes.POINTER(c_double),
                             ctypes.POINTER(c_double), _P]),
    'gm_bicgstab': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, ctypes.POINTER(gm_krylov_ws), c_double, c_int,
                            c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_double), _P]),
    'gm_pcg': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, ctypes.POINTER(gm_krylov_ws), c_double, c_int,
                       c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_double), _P]),
    'gm_gmres': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_int, c_double, c_int,
                         ctypes.POINTER(c_int), ctypes.POINTER(c_double), _P]),
    'gm_bsr_spmv': (c_int, [c_int, c_int, c_int, c_int, _P, _P, _P, _P, _P, _P]),
    'gm_norm2': (c_int, [c_long, _P, _P, _P, _P]),
    'gm_sum': (c_int, [c_long, _P, _P, _P, _P]),
    'gm_norm2_noalloc': (c_int, [c_long, _P, _P, _P]),
    'gm_axpy': (c_int, [c_long, c_double, _P, c_int, _P, _P]),
    'gm_residual_refine': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    'gm_d2f': (c_int, [c_long, _P, _P, _P]),
    'gm_fix_empty_diag': (c_int, [c_int, _P, _P, _P, _P, _P]),
    'gm_gimp_weights': (c_int, [c_int, c_int, _P, _P, ctypes.POINTER(c_double), _P, _P, _P, _P]),
    'gm_return_map': (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_double,
                              c_double, c_double, _P, _P, _P, _P, _P, _P]),
    'gm_return_map_norsand': (c_int, [c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    'gm_local_eval_2d': (c_int, [_P, _P, _P, _P]),
    'gm_local_vjp_2d': (c_int, [_P, _P, _P]),
    'gm_bsr_transpose': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P]),
    'gm_mg_setup': (c_int, [c_int, c_int, c_int, c_int, _P, _P, _P]),
    'gm_mg_vcycle': (c_int, [c_int, c_int, c_int, c_int, _P, c_int, c_int, c_int, c_double, _P, _P, _P]),
    'gm_mg_level_op': (c_int, [c_int, c_int, _P, c_int, c_int, _P, _P, c_double, _P, _P]),
    'gm_gmres_mg': (c_int, [c_int, c_int, c_int, c_int, _P, _P, c_int, c_int, c_int, c_double, _P, _P, _P, _P,
                            _P, _P, _P, _P, c_int, c_double, c_int, _P, _P, _P]),
    'gm_penalty_pc': (c_int, [ctypes.POINTER(gm_grid), ctypes.POINTER(gm_bins), ctypes.POINTER(gm_pattern), _P, _P,
                              c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_double), c_double,
                              c_double, _P]),
    'gm_krylov_graphs': (c_int, [c_int]),
    'gm_spmv_autotune': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, c_int, _P, _P, _P]),
    'gm_row_scale': (c_int, [c_int, c_int, c_int, _P, _P, _P, _P, _P, _P]),
    'gm_up_setup': (c_int, [c_int, c_int, c_int, c_int, c_int, _P, _P, _P, _P, c_int, _P, _P, _P, _P]),
    'gm_gmres_up': (c_int, [c_int, c_int, c_int, c_int, _P, _P, _P, _P, c_int, _P, c_int, c_int, c_int, c_double,
                            _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_int, c_double, c_int, _P, _P, _P]),
}

_lib = None


class GMError(RuntimeError):
    pass


def lib():
    """Load libgeowarp_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GMError('libgeowarp_amd.so is not built (expected at %s); run __graft_entry__.build()' % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return sorted(_SIGS.keys())


def check(rc, what=''):
    if rc != 0:
        msg = lib().gm_last_error()
        msg = msg.decode() if msg else ''
        raise GMError('%s failed (rc=%d): %s' % (what, rc, msg))


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return c_void_p(t.data_ptr())

"""Return mappings with the reference's names (GeoWarp materials/plasticity.py and
the Drucker-Prager and NorSand functions of problems/material_tests/triaxial_*.py),
one particle per lane on the device.

Inputs follow the reference: trial strains as (N, 3, 3) principal (diagonal)
log-strain matrices; outputs likewise.
"""
import torch

from .. import _lib as L


def _stream():
    return L.c_void_p(torch.cuda.current_stream().cuda_stream)


def _principal(e):
    if e.dim() == 3:
        return torch.diagonal(e, dim1=1, dim2=2).contiguous()
    return e.contiguous()


def _call(kind, e_trial, lam, mu, kappa=0.0, H=0.0, phi=0.0, psi=0.0, c=0.0, shape=0.0, apex=0, alpha_old=None,
          tangent=False):
    et = _principal(e_trial).to(torch.float64)
    n = et.shape[0]
    eps = torch.empty_like(et)
    dE = torch.empty(n, 9, dtype=torch.float64, device=et.device) if tangent else None
    an = torch.empty(n, dtype=torch.float64, device=et.device) if alpha_old is not None else None
    L.check(L.lib().gm_return_map(n, kind, apex, float(lam), float(mu), float(kappa), float(H), float(phi),
                                  float(psi), float(c), float(shape), L.ptr(et), L.ptr(eps), L.ptr(dE),
                                  L.ptr(alpha_old), L.ptr(an), _stream()), 'gm_return_map')
    out = torch.diag_embed(eps) if e_trial.dim() == 3 else eps
    return out, dE, an


def return_mapping_J2(e_trial, lame_lambda, lame_mu, kappa, H=0.0, alpha_old=None, tangent=False):
    """plasticity.py:8-37 (radial return, principal log strains); H adds linear
    hardening.  Returns e_real (and, with tangent=True, (e_real, d e/d e_tr))."""
    out, dE, an = _call(1, e_trial, lame_lambda, lame_mu, kappa=kappa, H=H, alpha_old=alpha_old, tangent=tangent)
    if tangent or alpha_old is not None:
        return out, dE, an
    return out


def return_mapping_DruckerPrager(trial_strain, lame_lambda, lame_mu, friction_angle, dilation_angle, cohesion,
                                 shape_factor=0.0, tol=1e-10, real_strain_history=None, delta_lambda_history=None,
                                 apex='reference', tangent=False):
    """triaxial_DruckerPrager.py:308-412.  The history arrays of the reference
    exist only for its tape and are ignored."""
    out, dE, _ = _call(3, trial_strain, lame_lambda, lame_mu, phi=friction_angle, psi=dilation_angle, c=cohesion,
                       shape=shape_factor, apex={'reference': 0, 'return': 1}[apex], tangent=tangent)
    return (out, dE) if tangent else out


def return_mapping_NorSand(kappa, tilde_lambda, M, N, beta, vc0, p0, v0, h, lame_mu, initial_volumetric_strain,
                           old_pi_array, trial_strain, total_strain, tol=1e-10, return_residual=False):
    """triaxial_NorSand.py:243-400 for tensors of particles: trial_strain and
    total_strain as (N, 3, 3) principal matrices (or (N, 3)), old_pi_array (N,).
    Returns (real_strain, new_pi_array) in the input's form, plus the local
    residual of the last Newton iterate (N, 4) with return_residual=True (the
    reference keeps it as saved_local_residual)."""
    et = _principal(trial_strain).to(torch.float64)
    tt = _principal(total_strain).to(torch.float64)
    n = et.shape[0]
    ev_total = tt.sum(1).contiguous()
    pi_old = old_pi_array.to(torch.float64).contiguous()
    eps = torch.empty_like(et)
    pi_new = torch.empty_like(pi_old)
    res = torch.empty(n, 4, dtype=torch.float64, device=et.device) if return_residual else None
    prm = L.gm_norsand_params(kappa, tilde_lambda, M, N, beta, vc0, p0, v0, h, lame_mu, initial_volumetric_strain,
                              tol)
    L.check(L.lib().gm_return_map_norsand(n, L.ctypes.byref(prm), L.ptr(et), L.ptr(ev_total), L.ptr(pi_old),
                                          L.ptr(eps), L.ptr(pi_new), L.ptr(res), _stream()),
            'gm_return_map_norsand')
    out = torch.diag_embed(eps) if trial_strain.dim() == 3 else eps
    return (out, pi_new, res) if return_residual else (out, pi_new)

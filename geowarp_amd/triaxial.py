"""Single-point strain-driven triaxial tests on the device return maps.

Drop-in for the reference's material-test drivers
(problems/material_tests/triaxial_DruckerPrager.py:497-611,
problems/material_tests/triaxial_NorSand.py:497-628): the axial principal
strain is decreased by final_axial_strain / n_steps per step, and a Newton
iteration on the two lateral principal strains holds the lateral stresses at
their targets.  Per iteration (as the drivers do): evaluate the return map at
the trial strain, residual = target - stress (xx, yy), solve with the tangent,
add the correction to the trial and total strains, then stop when the
pre-update residual norm is below `tol` (absolute) or after n_iter iterations.
After the step the trial strain becomes the returned strain of the last
evaluation.  The initial strain is the linear elastic strain of the target
stresses (the drivers' 6 x 6 stiffness solve).  Reported per step: P =
-tr(sigma)/3, Q = -sigma_zz + sigma_xx of the last evaluation (the drivers'
printed invariants), the axial strain and the volumetric strain.

Drucker-Prager tests run as a batch, one material point per lane of
gm_return_map, with per-point target stresses (for instance several confining
pressures in one launch); the Newton tangent is the return map's consistent
tangent C dE.  NorSand runs gm_return_map_norsand, and its Newton tangent is a
central difference of the device map over the two lateral strains (the five
evaluations share one launch); the reference differentiates its tape instead,
which changes the iterates but not the converged state.
"""
import numpy as np
import torch

from .materials import plasticity as pl


def _lame(E, nu):
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def elastic_strain(lam, mu, stress):
    """Principal strains e with lam tr(e) + 2 mu e_i = stress_i (rows)."""
    C = lam * np.ones((3, 3)) + 2.0 * mu * np.eye(3)
    return np.linalg.solve(C, np.atleast_2d(np.asarray(stress, dtype=np.float64)).T).T


def _device(device):
    return torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())


def triaxial_drucker_prager(youngs_modulus=25000.0, poisson_ratio=0.3, friction_angle=35.0, dilation_angle=5.0,
                            cohesion=0.0, shape_factor=0.0, final_axial_strain=0.1, n_steps=100,
                            target_stress=(-150.0, -150.0, -150.0), tol=1e-10, n_iter=10, apex='reference',
                            device=None):
    """Drained triaxial compression of Drucker-Prager material points
    (defaults: the driver's material and loading, triaxial_DruckerPrager.py:10-31).

    target_stress: (3,) or (m, 3) initial principal stresses (xx, yy, zz) in kPa,
    compression negative; xx and yy are held during the test.  Returns numpy
    arrays P, Q, volumetric_strain, newton_iterations (n_steps, m) and
    axial_strain (n_steps,)."""
    dev = _device(device)
    lam, mu = _lame(youngs_modulus, poisson_ratio)
    tgt = np.atleast_2d(np.asarray(target_stress, dtype=np.float64))
    m = tgt.shape[0]
    f64 = dict(dtype=torch.float64, device=dev)
    trial = torch.tensor(elastic_strain(lam, mu, tgt), **f64)
    total = trial.clone()
    ev0 = total.sum(1).clone()
    target = torch.tensor(tgt, **f64)
    C = torch.tensor(lam * np.ones((3, 3)) + 2.0 * mu * np.eye(3), **f64)
    rate = final_axial_strain / n_steps
    out = dict(P=np.zeros((n_steps, m)), Q=np.zeros((n_steps, m)), volumetric_strain=np.zeros((n_steps, m)),
               newton_iterations=np.zeros((n_steps, m), dtype=np.int64),
               axial_strain=rate * np.arange(1, n_steps + 1), initial_P=-tgt.sum(1) / 3.0,
               initial_Q=tgt[:, 0] - tgt[:, 2])
    for step in range(n_steps):
        trial[:, 2] -= rate
        total[:, 2] -= rate
        active = torch.ones(m, dtype=torch.bool, device=dev)
        iters = torch.zeros(m, dtype=torch.int64, device=dev)
        e_keep = torch.zeros_like(trial)
        s_keep = torch.zeros_like(trial)
        for _ in range(n_iter):
            e, dE = pl.return_mapping_DruckerPrager(trial, lam, mu, friction_angle, dilation_angle, cohesion,
                                                    shape_factor, apex=apex, tangent=True)
            sig = e @ C
            r = target[:, :2] - sig[:, :2]
            T = C @ dE.view(m, 3, 3)  # d sigma / d trial strain
            upd = torch.linalg.solve(T[:, :2, :2], r.unsqueeze(-1)).squeeze(-1)
            a = active[:, None]
            e_keep = torch.where(a, e, e_keep)
            s_keep = torch.where(a, sig, s_keep)
            trial[:, :2] += torch.where(a, upd, torch.zeros_like(upd))
            total[:, :2] += torch.where(a, upd, torch.zeros_like(upd))
            iters += active.long()
            active = active & ~(torch.linalg.norm(r, dim=1) < tol)
            if not bool(active.any()):
                break
        trial.copy_(e_keep)
        out['P'][step] = (-s_keep.sum(1) / 3.0).cpu().numpy()
        out['Q'][step] = (s_keep[:, 0] - s_keep[:, 2]).cpu().numpy()
        out['volumetric_strain'][step] = (-(total.sum(1) - ev0)).cpu().numpy()
        out['newton_iterations'][step] = iters.cpu().numpy()
    return out


def norsand_stress(e, ev0, p0, kappa, mu):
    """Pressure-dependent hyperelastic principal stress of the NorSand driver
    (triaxial_NorSand.py:455-461): p0 exp(-(tr e - ev0)/kappa) + 2 mu (e_i - tr e / 3)."""
    tr = e.sum(1, keepdim=True)
    return p0 * torch.exp(-(tr - ev0) / kappa) + 2.0 * mu * (e - tr / 3.0)


def triaxial_norsand(G0=4.5e4, kappa=0.0015, tilde_lambda=0.02, M=1.27, N=0.4, beta=1.0, vc0=1.8911, pc=-715.0,
                     p0=-390.0, v0=1.75, h=70.0, K0_consolidation=0.45, final_axial_strain=0.1, n_steps=500,
                     tol=1e-10, n_iter=10, fd_step=1e-8, device=None):
    """Drained triaxial compression of one NorSand material point from a K0
    state (defaults: the driver's, triaxial_NorSand.py:8-45).  Returns numpy
    arrays P, Q, volumetric_strain (percent), newton_iterations (n_steps,) and
    axial_strain (percent, the driver's printed abscissa)."""
    dev = _device(device)
    K0 = p0 / (-kappa)
    lam, mu = K0 - 2.0 / 3.0 * G0, G0
    slope = (1.0 - K0_consolidation) / (1.0 / 3.0 * (1.0 + 2.0 * K0_consolidation))
    sxx = (3.0 - slope) / 3.0 * p0
    szz = (3.0 + 2.0 * slope) / 3.0 * p0
    tgt = np.array([sxx, sxx, szz])
    f64 = dict(dtype=torch.float64, device=dev)
    trial = torch.tensor(elastic_strain(lam, mu, tgt)[0], **f64)
    total = trial.clone()
    ev0 = float(total.sum())
    pi_old = torch.tensor([pc * (1.0 - N) ** ((1.0 - N) / N)], **f64)
    target = torch.tensor(tgt[:2], **f64)
    rate = final_axial_strain / n_steps
    out = dict(P=np.zeros(n_steps), Q=np.zeros(n_steps), volumetric_strain=np.zeros(n_steps),
               newton_iterations=np.zeros(n_steps, dtype=np.int64),
               axial_strain=100.0 * rate * np.arange(1, n_steps + 1), initial_Q=-tgt[2] + tgt[0])
    # lanes: base, +h / -h on e_xx, +h / -h on e_yy
    pert = torch.zeros(5, 3, **f64)
    pert[1, 0], pert[2, 0], pert[3, 1], pert[4, 1] = fd_step, -fd_step, fd_step, -fd_step
    for step in range(n_steps):
        trial[2] -= rate
        total[2] -= rate
        it = 0
        for it in range(1, n_iter + 1):
            tr5 = trial[None, :] + pert  # the tangent is d stress / d trial strain (total held)
            to5 = total[None, :].expand(5, 3).contiguous()
            e5, pi5 = pl.return_mapping_NorSand(kappa, tilde_lambda, M, N, beta, vc0, p0, v0, h, mu, ev0,
                                                pi_old.expand(5).contiguous(), tr5, to5)
            s5 = norsand_stress(e5, ev0, p0, kappa, mu)
            sig, e_last, pi_last = s5[0], e5[0].clone(), pi5[0].clone()
            r = target - sig[:2]
            T = torch.stack([(s5[1, :2] - s5[2, :2]), (s5[3, :2] - s5[4, :2])], 1) / (2.0 * fd_step)
            upd = torch.linalg.solve(T, r)
            trial[:2] += upd
            total[:2] += upd
            if float(torch.linalg.norm(r)) < tol:
                break
        trial.copy_(e_last)
        pi_old = pi_last.reshape(1)
        out['P'][step] = float(-sig.sum() / 3.0)
        out['Q'][step] = float(-sig[2] + sig[0])
        out['volumetric_strain'][step] = float(-total.sum() * 100.0 - (-ev0 * 100.0))
        out['newton_iterations'][step] = it
    return out

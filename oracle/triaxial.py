"""Single-point triaxial drivers (oracle restatement, test infrastructure only).

Follows problems/material_tests/triaxial_DruckerPrager.py:497-611 and
problems/material_tests/triaxial_NorSand.py:497-628: elastic initial strain of
the target stresses, then per step the axial strain decreases by
final_axial_strain / n_steps and a Newton iteration on the two lateral strains
holds the lateral stresses (update applied, then the pre-update residual norm
tested against `tol`); the trial strain becomes the returned strain of the
last evaluation.  The Newton tangent is a central difference of the return map
(the reference differentiates its tape; the converged states agree).
"""
import numpy as np

from . import materials as om


def _lame(E, nu):
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def _run(stress_of, e0, target, rate, n_steps, tol, n_iter, fd, state0=None, commit=None):
    trial = e0.copy()
    total = e0.copy()
    ev0 = total.sum()
    state = state0
    rows = []
    for _ in range(n_steps):
        trial[2] -= rate
        total[2] -= rate
        for it in range(1, n_iter + 1):
            sig, e_last, st_last = stress_of(trial, total, state)
            r = target[:2] - sig[:2]
            T = np.zeros((2, 2))
            for j in range(2):
                dp = np.zeros(3)
                dp[j] = fd
                T[:, j] = (stress_of(trial + dp, total, state)[0][:2] - stress_of(trial - dp, total, state)[0][:2]) \
                    / (2.0 * fd)
            upd = np.linalg.solve(T, r)
            trial[:2] += upd
            total[:2] += upd
            if np.linalg.norm(r) < tol:
                break
        trial = e_last.copy()
        state = st_last
        rows.append((-sig.sum() / 3.0, sig[0] - sig[2], -(total.sum() - ev0), it))
    return np.array(rows)


def triaxial_drucker_prager(E=25000.0, nu=0.3, phi=35.0, psi=5.0, c=0.0, shape=0.0, final_axial_strain=0.1,
                            n_steps=100, target=(-150.0, -150.0, -150.0), tol=1e-10, n_iter=10, apex='reference'):
    """Returns an (n_steps, 4) array of (P, Q, volumetric strain, Newton iterations)."""
    lam, mu = _lame(E, nu)
    C = lam * np.ones((3, 3)) + 2.0 * mu * np.eye(3)
    target = np.asarray(target, dtype=np.float64)
    e0 = np.linalg.solve(C, target)

    def stress_of(trial, total, state):
        e = om.return_mapping_DruckerPrager(trial[None], lam, mu, phi, psi, c, shape, apex=apex)[0]
        return C @ e, e, None

    return _run(stress_of, e0, target, final_axial_strain / n_steps, n_steps, tol, n_iter, 1e-8)


def triaxial_norsand(G0=4.5e4, kappa=0.0015, tilde_lambda=0.02, M=1.27, N=0.4, beta=1.0, vc0=1.8911, pc=-715.0,
                     p0=-390.0, v0=1.75, h=70.0, K0_consolidation=0.45, final_axial_strain=0.1, n_steps=500,
                     tol=1e-10, n_iter=10):
    """Returns an (n_steps, 4) array of (P, Q, volumetric strain, Newton iterations)."""
    K0 = p0 / (-kappa)
    lam, mu = K0 - 2.0 / 3.0 * G0, G0
    slope = (1.0 - K0_consolidation) / (1.0 / 3.0 * (1.0 + 2.0 * K0_consolidation))
    sxx = (3.0 - slope) / 3.0 * p0
    target = np.array([sxx, sxx, (3.0 + 2.0 * slope) / 3.0 * p0])
    C = lam * np.ones((3, 3)) + 2.0 * mu * np.eye(3)
    e0 = np.linalg.solve(C, target)
    ev0 = e0.sum()
    prm = dict(kappa=kappa, tilde_lambda=tilde_lambda, M=M, N=N, beta=beta, vc0=vc0, p0=p0, v0=v0, h=h, mu=mu,
               ev0=ev0)

    def stress_of(trial, total, pi_old):
        e, pi_new, _ = om.return_mapping_NorSand(trial[None], np.array([total.sum()]), np.array([pi_old]), prm)
        e = e[0]
        tr = e.sum()
        sig = p0 * np.exp(-(tr - ev0) / kappa) + 2.0 * mu * (e - tr / 3.0)
        return sig, e, float(pi_new[0])

    pi0 = pc * (1.0 - N) ** ((1.0 - N) / N)
    return _run(stress_of, e0, target, final_axial_strain / n_steps, n_steps, tol, n_iter, 1e-8, state0=pi0)

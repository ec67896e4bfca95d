"""Single-point triaxial harness (triaxial_DruckerPrager.py:497-611,
triaxial_NorSand.py:497-628).

CPU: the oracle harness against textbook answers -- the elastic slope dq/de_a = E
of a triaxial test at constant lateral stress, and the perfectly plastic
plateau q_f = sigma_3 (N_phi - 1), N_phi = (1 + sin phi) / (1 - sin phi), the
Mohr-Coulomb triaxial-compression strength the reference's Drucker-Prager
constants (compression-cone fit) reproduce.
GPU: the device harness (batched Drucker-Prager over confining pressures,
NorSand with a difference tangent) against the oracle harness.
"""
import numpy as np
import pytest

import oracle.triaxial as ot


def _mc_strength(phi_deg, sigma3):
    s = np.sin(np.radians(phi_deg))
    return sigma3 * ((1 + s) / (1 - s) - 1)


@pytest.mark.parametrize('sigma3', [50.0, 150.0])
def test_oracle_dp_triaxial_elastic_slope_and_mohr_coulomb_plateau(sigma3):
    r = ot.triaxial_drucker_prager(target=(-sigma3, -sigma3, -sigma3), n_steps=100)
    q = r[:, 1]
    # elastic branch: the first steps stay below the yield stress
    assert abs((q[1] - q[0]) / 1e-3 - 25000.0) < 1e-6 * 25000.0
    assert abs(q[-1] - _mc_strength(35.0, sigma3)) < 1e-8 * q[-1]
    assert np.all(r[:, 3] <= 10)
    # dilation 5 deg: the specimen dilates once it flows
    assert r[-1, 2] < r[-10, 2] < 0.0


def test_oracle_norsand_triaxial_converges_and_hardens_to_a_peak():
    r = ot.triaxial_norsand(n_steps=200)
    assert np.all(r[:, 3] < 10)
    eta = r[:, 1] / r[:, 0]
    assert eta.max() > 1.27  # dense state: peak stress ratio above the critical-state slope M
    assert r[0, 1] > 300.0 and np.all(np.isfinite(r))


@pytest.mark.gpu
def test_device_dp_triaxial_batch_matches_oracle():
    from geowarp_amd import triaxial as T
    s3 = np.array([50.0, 150.0, 300.0])
    tgt = -np.repeat(s3[:, None], 3, axis=1)
    out = T.triaxial_drucker_prager(target_stress=tgt, n_steps=100, device='cuda:0')
    for k, s in enumerate(s3):
        ref = ot.triaxial_drucker_prager(target=(-s, -s, -s), n_steps=100)
        for col, key in ((0, 'P'), (1, 'Q'), (2, 'volumetric_strain')):
            a, b = out[key][:, k], ref[:, col]
            err = np.abs(a - b).max() / np.abs(b).max()
            print('sigma3 %g %s max rel err %.2e' % (s, key, err))
            assert err < 1e-9, (s, key, err)
        assert abs(out['Q'][-1, k] - _mc_strength(35.0, s)) < 1e-8 * out['Q'][-1, k]


@pytest.mark.gpu
def test_device_norsand_triaxial_matches_oracle():
    from geowarp_amd import triaxial as T
    out = T.triaxial_norsand(n_steps=200, device='cuda:0')
    ref = ot.triaxial_norsand(n_steps=200)
    for col, key in ((0, 'P'), (1, 'Q')):
        err = np.abs(out[key] - ref[:, col]).max() / np.abs(ref[:, col]).max()
        print('NorSand %s max rel err %.2e' % (key, err))
        assert err < 1e-8
    dv = np.abs(out['volumetric_strain'] - 100.0 * ref[:, 2]).max()
    print('NorSand volumetric strain (percent) max abs err %.2e' % dv)
    assert dv < 1e-6

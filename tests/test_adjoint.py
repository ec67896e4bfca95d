"""Adjoint of the 2D load path (geowarp_amd.adjoint) on the GPU.

* The double evaluation of the per-particle local map reproduces the forward
  particle stage, residual, G2P and lp update (FP64, 1e-12 relative).
* Reaction forces match the oracle's and the adjoint's sensitivities match the
  oracle's central finite differences (tests/golden/adjoint_dp_specimen.npz) to
  1e-4 relative.
* Batched specimens: torch.autograd through LoadPathFunction gives per-specimen
  gradients equal to central finite differences of the device forward
  (h = 1e-6 deg / kPa) to 1e-4 relative, and neither the checkpoint interval
  nor replaying the stored increments (instead of re-solving the segments)
  changes them.
* A load path whose platen rows move repeats exactly from its initial state.
"""
import os

import numpy as np
import pytest
import torch

import harness as H
from geowarp_amd import problems as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _solver(cfg):
    import geowarp_amd.solvers as gs
    s = H.make(gs, cfg)
    dev = s.device
    g = torch.as_tensor(cfg['groups'], device=dev)
    s.friction_angle_array = torch.as_tensor(cfg['phi'], device=dev)[g].contiguous()
    s.cohesion_array = torch.as_tensor(cfg['coh'], device=dev)[g].contiguous()
    return s, g


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.gpu
def test_local_map_matches_forward_stage():
    from geowarp_amd import _lib as L
    from geowarp_amd.adjoint import LoadPathAdjoint, _stream
    cfg = P.specimens2d([3, 4], [28.0, 36.0], [4.0, 12.0], total_steps=4, strain=0.06, tol=1e-12)
    s, g = _solver(cfg)
    adj = LoadPathAdjoint(s, 4, groups=g)
    s.prepare_step(0, 4)
    s.newton_iter(0, 4)
    eng = s.engine
    # forward stage and residual at the converged increment plus a perturbation of the free dofs
    # (at the converged increment itself the residual is round-off)
    rng = np.random.default_rng(2)
    fm = eng.freem.to(torch.float64)
    s._inc.add_(torch.from_numpy(rng.normal(size=eng.n_nodes * 2)).to(s.device) * fm * (0.01 * s.dx))
    eng.stage(s._args(0, 4), want_tangent=False)
    eng.residual(s.x_particles, s.GIMP_lp)
    cot = torch.from_numpy(rng.normal(size=eng.n_nodes * 2)).to(s.device) * fm
    a = adj._local_args(0, s._inc, cot=cot)
    acc = torch.empty(s.n_particles, dtype=torch.float64, device=s.device)
    st = torch.empty(s.n_particles, 18, dtype=torch.float64, device=s.device)
    L.check(L.lib().gm_local_eval_2d(L.ctypes.byref(a), L.ptr(acc), L.ptr(st), _stream()), 'gm_local_eval_2d')
    ref_dot = float((cot * eng.rhs).sum())
    scale = float((cot.abs() * eng.rhs.abs()).sum())
    got = float(acc.sum())
    print('cot.r local', got, 'forward', ref_dot, 'scale', scale)
    assert scale > 0.0 and abs(got - ref_dot) <= 1e-12 * scale
    Fn = s.deformation_gradient_total_new
    F5 = torch.stack([Fn[:, 0, 0], Fn[:, 0, 1], Fn[:, 1, 0], Fn[:, 1, 1], Fn[:, 2, 2]], 1)
    bn = s.left_Cauchy_Green_new
    b4 = torch.stack([bn[:, 0, 0], bn[:, 0, 1], bn[:, 1, 1], bn[:, 2, 2]], 1)
    sg = s.particle_Cauchy_stress_array
    s4 = torch.stack([sg[:, 0, 0], sg[:, 0, 1], sg[:, 1, 1], sg[:, 2, 2]], 1)
    assert _rel(st[:, 4:9].cpu(), F5.cpu()) < 1e-13
    assert _rel(st[:, 9:13].cpu(), b4.cpu()) < 1e-12
    assert _rel(st[:, 14:18].cpu(), s4.cpu()) < 1e-12
    s.finish_step()
    assert _rel(st[:, 0:2].cpu(), s.x_particles.cpu()) < 1e-15
    assert _rel(st[:, 2:4].cpu(), s.GIMP_lp.cpu()) < 1e-14


@pytest.mark.gpu
def test_adjoint_matches_oracle_finite_differences():
    from geowarp_amd.adjoint import LoadPathAdjoint
    gold = np.load(os.path.join(GOLD, 'adjoint_dp_specimen.npz'))
    cfg = P.specimens2d([3], [30.0], [5.0], total_steps=4, strain=0.06, tol=1e-12)
    s, g = _solver(cfg)
    adj = LoadPathAdjoint(s, 4, groups=g, checkpoint_every=2)
    f = adj.forward()[:, 0].cpu().numpy()
    print('forces', f, 'oracle', gold['forces'])
    assert _rel(f, gold['forces']) < 1e-8
    for w in (np.ones(4), np.array([0.3, -1.0, 2.0, 0.5])):
        out = adj.backward(torch.from_numpy(w[:, None]))
        for name, key in (('phi', 'dphi'), ('coh', 'dcoh'), ('E', 'dE')):
            got = float(out[name].sum())
            ref = float(w @ gold[key])
            print(name, got, ref)
            assert abs(got - ref) <= 1e-4 * abs(ref), (name, got, ref)


@pytest.mark.gpu
def test_batched_gradients_match_device_fd_and_checkpointing():
    from geowarp_amd.adjoint import LoadPathAdjoint, LoadPathFunction
    widths = [2, 3, 4]
    phi0 = np.array([27.0, 33.0, 38.0])
    coh0 = np.array([2.0, 9.0, 15.0])
    cfg = P.specimens2d(widths, phi0, coh0, total_steps=4, strain=0.06, tol=1e-12)
    s, g = _solver(cfg)
    # observed curves from other parameters
    adj_obs = LoadPathAdjoint(s, 4, groups=g)
    dev = s.device
    with torch.no_grad():
        f_obs = LoadPathFunction.apply(torch.tensor(phi0 + 1.5, device=dev), torch.tensor(coh0 + 2.0, device=dev),
                                       adj_obs)
    grads = []
    for every, store in ((10, True), (2, True), (2, False)):
        adj = LoadPathAdjoint(s, 4, groups=g, checkpoint_every=every, initial=adj_obs.initial,
                              store_increments=store)
        phi = torch.tensor(phi0, device=dev, requires_grad=True)
        coh = torch.tensor(coh0, device=dev, requires_grad=True)
        f = LoadPathFunction.apply(phi, coh, adj)
        loss = ((f - f_obs) ** 2).sum()
        loss.backward()
        grads.append((phi.grad.cpu().numpy(), coh.grad.cpu().numpy()))
    np.testing.assert_allclose(grads[0][0], grads[1][0], rtol=1e-10)
    np.testing.assert_allclose(grads[0][1], grads[1][1], rtol=1e-10)
    # replaying stored increments or re-solving the segments gives the same gradients (the two
    # recomputations differ by the last Newton correction, below 1e-12 relative here)
    np.testing.assert_allclose(grads[1][0], grads[2][0], rtol=1e-9)
    np.testing.assert_allclose(grads[1][1], grads[2][1], rtol=1e-9)

    def loss_at(ph, co):
        adj = LoadPathAdjoint(s, 4, groups=g, initial=adj_obs.initial)
        with torch.no_grad():
            f = LoadPathFunction.apply(torch.tensor(ph, device=dev), torch.tensor(co, device=dev), adj)
        return float(((f - f_obs) ** 2).sum())

    h = 1e-6
    for i in range(len(widths)):
        e = np.zeros(len(widths))
        e[i] = h
        fd_phi = (loss_at(phi0 + e, coh0) - loss_at(phi0 - e, coh0)) / (2 * h)
        fd_coh = (loss_at(phi0, coh0 + e) - loss_at(phi0, coh0 - e)) / (2 * h)
        print(i, 'phi', grads[0][0][i], fd_phi, 'coh', grads[0][1][i], fd_coh)
        assert abs(grads[0][0][i] - fd_phi) <= 1e-4 * abs(fd_phi)
        assert abs(grads[0][1][i] - fd_coh) <= 1e-4 * abs(fd_coh)


@pytest.mark.gpu
def test_forward_repeats_from_the_initial_state():
    """A load path whose platen rows move down (20 % strain) gives the same forces
    when it is run again after a run with other parameters: the snapshot of the
    initial state includes the boundary flags that the platen hook rewrites."""
    from geowarp_amd.adjoint import LoadPathAdjoint, LoadPathFunction
    cfg = P.specimens2d([3, 4], [30.0, 36.0], [5.0, 12.0], total_steps=8, strain=0.2, tol=1e-12)
    s, g = _solver(cfg)
    adj = LoadPathAdjoint(s, 8, groups=g, checkpoint_every=3)
    dev = s.device
    bc0 = s.dofStruct.boundary_flag_array.clone()
    runs = []
    for dphi in (0.0, 3.0, 0.0):
        with torch.no_grad():
            f = LoadPathFunction.apply(torch.tensor([30.0 + dphi, 36.0], device=dev),
                                       torch.tensor([5.0, 12.0], device=dev), adj)
        runs.append(f.cpu().numpy())
        print('dphi', dphi, 'forces', runs[-1][:, 0], 'unconverged', adj.stats['unconverged_steps'])
    assert adj.stats['unconverged_steps'] == []
    assert not np.array_equal(runs[0], runs[1])
    np.testing.assert_array_equal(runs[0], runs[2])
    # the platen rows did move: the flags at the end of the path differ from the initial ones
    assert not torch.equal(s.dofStruct.boundary_flag_array, bc0)

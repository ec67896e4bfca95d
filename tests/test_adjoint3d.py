"""GPU: the adjoint of a converged 3D u-p load step (quasi_static_coupled_solver_3d with
iteration_dict['step_sensitivity']) against central finite differences of the device
forward in Young's modulus, and its explicit term against the closed form of the
explicit gradient (qc3d:323-325).

The load step's initial state is held fixed: for load step 0 (the undeformed block)
that is the total derivative; for a later step the finite differences restart both
perturbed runs from the same snapshot."""
import numpy as np
import pytest

import harness as H

REL = 1e-4  # gradients within 1e-4 of central differences


def _cfg(E=None, sens=False):
    cfg = H.MINI['indent_nh']()
    cfg['iteration_dict']['tol'] = 1e-13
    cfg['iteration_dict']['n_iter'] = 20
    if sens:
        cfg['iteration_dict']['step_sensitivity'] = True
    if E is not None:
        cfg['material_dict'] = dict(cfg['material_dict'], youngs_modulus=E)
    return cfg


def _force_and_loss(s):
    return float(s.indentation_force.item()), float(s.loss.item())


@pytest.mark.gpu
@pytest.mark.parametrize('step', [0, 1])
def test_load_step_E_sensitivity_matches_central_fd(step):
    import geowarp_amd.solvers as gs
    from geowarp_amd.solvers.quasi_static_coupled_solver_3d import TARGET_FORCE
    cfg = _cfg(sens=True)
    total = cfg['total_steps']
    E = cfg['material_dict']['youngs_modulus']
    s = H.make(gs, cfg)
    snap = None
    for k in range(step + 1):
        if k == step:
            snap = s.state_dict()
        s.advance_one_step(k, total)
    adj, expl = s.E_grad_step, s.E_grad_step_explicit
    F = s.sensitivity_stats['force']
    S = abs(s.footing_incr_disp) * total
    dLdF = 2.0 * (F / S - TARGET_FORCE) / S
    closed = 2.0 * (F / S - TARGET_FORCE) * F / (E * S)
    assert abs(expl - closed) <= 1e-12 * abs(closed)
    # central differences of the forward load step from the same initial state
    h = 1e-5 * E
    out = []
    for Ep in (E + h, E - h):
        c = _cfg(E=Ep)
        t = H.make(gs, c)
        if step:
            t.load_state_dict(snap)
        t.advance_one_step(step, total)
        assert t.newton_converged
        out.append(_force_and_loss(t))
    dF_fd = (out[0][0] - out[1][0]) / (2 * h)
    dL_fd = (out[0][1] - out[1][1]) / (2 * h)
    dF_adj = adj / dLdF
    print('step', step, 'dF/dE adjoint %.9e fd %.9e' % (dF_adj, dF_fd), 'dL/dE adjoint %.9e fd %.9e' % (adj, dL_fd),
          'explicit %.3e' % expl, s.sensitivity_stats)
    assert abs(dF_fd) > 0
    assert abs(dF_adj - dF_fd) <= REL * abs(dF_fd)
    assert abs(adj - dL_fd) <= REL * abs(dL_fd)


@pytest.mark.gpu
def test_step_sensitivity_needs_fp64_and_neo_hookean():
    import geowarp_amd.solvers as gs
    from geowarp_amd._lib import GMError
    cfg = _cfg(sens=True)
    cfg['iteration_dict']['precision'] = 'mixed'
    s = H.make(gs, cfg)
    with pytest.raises(GMError):
        s.advance_one_step(0, cfg['total_steps'])


@pytest.mark.gpu
def test_load_path_E_gradient_matches_central_fd():
    """The total dL/dE through three load steps (path_E_gradient: the state each step
    hands to the next is differentiated too) against central differences of the whole
    forward run in E; the single-step derivative of the last step, which holds that
    step's initial state fixed, differs from it."""
    import geowarp_amd.solvers as gs
    n = 3
    cfg = _cfg()
    total = cfg['total_steps']
    E = cfg['material_dict']['youngs_modulus']
    s = H.make(gs, cfg)
    grad = s.path_E_gradient(n, total)
    dF = s.path_force_derivative
    h = 1e-5 * E
    out = []
    for Ep in (E + h, E - h):
        t = H.make(gs, _cfg(E=Ep))
        for k in range(n):
            t.advance_one_step(k, total)
            assert t.newton_converged
        out.append(_force_and_loss(t))
    dF_fd = (out[0][0] - out[1][0]) / (2 * h)
    dL_fd = (out[0][1] - out[1][1]) / (2 * h)
    # the last step alone, from the same state at its start
    s2 = H.make(gs, _cfg(sens=True))
    for k in range(n):
        s2.advance_one_step(k, total)
    print('path dF/dE %.9e fd %.9e | dL/dE path %.9e fd %.9e | last step alone %.9e' %
          (dF, dF_fd, grad, dL_fd, s2.E_grad_step), s.path_stats)
    assert abs(dF_fd) > 0
    assert abs(dF - dF_fd) <= REL * abs(dF_fd)
    assert abs(grad - dL_fd) <= REL * abs(dL_fd)
    assert abs(s2.E_grad_step - dL_fd) > 10 * REL * abs(dL_fd)

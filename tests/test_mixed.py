"""Mixed-precision form (FP32 Jacobian, multigrid and GMRES; FP64 residual,
norm test and iterative refinement) against the FP64 oracle fixtures.

The mixed form converges the same FP64 residual to the fixtures' relative
tolerance, so its fields must agree with the oracle to within what that
tolerance allows: 1e-6 relative (the FP64 form is held to 1e-8).  The same bar
holds with the FP64 assembly option and when every correction falls back to an
FP64 solve.
"""
import os

import numpy as np
import pytest

import harness as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CASES = ['bar_nh', 'bar_j2', 'bar_j2_platen', 'bar_dp', 'terzaghi', 'indent_nh', 'indent_dp', 'indent_dp_cone',
         'cantilever', 'terzaghi_log', 'bar_norsand']


def _run_against_fixture(name, extra, engine_setup=None):
    import geowarp_amd.solvers as gs
    g = np.load(os.path.join(GOLD, name + '.npz'))
    cfg = H.MINI[name]()
    cfg['iteration_dict'] = dict(cfg['iteration_dict'], precision='mixed', **extra)
    s = H.make(gs, cfg)
    if engine_setup is not None:
        engine_setup(s.engine)
    total = int(g['total_steps'][0])
    x0 = s.x0.cpu().numpy()
    for k in range(int(g['n_steps'][0])):
        if cfg.get('dts') is not None:
            s.dt = cfg['dts'][k]
        s.advance_one_step(k, total)
        f = H.fields(s)
        print(name, 'step', k, 'newton', ['%.2e' % v for v in s.residual_history],
              'krylov', s.engine.stats['krylov_iters'][-len(s.residual_history):])
        for key in ('x', 'sigma', 'F', 'b', 'p'):
            gk = 's%d_%s' % (k, key)
            if key not in f or gk not in g:
                continue
            if key == 'x':
                e = H.rel_err(f[key] - x0, g[gk] - x0)
            else:
                e = H.rel_err(f[key], g[gk])
            print('   %-6s rel err %.3e' % (key, e))
            assert e < 1e-6, (name, k, key, e)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize('name', CASES)
def test_mixed_form_matches_oracle(name):
    _run_against_fixture(name, {})


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['bar_j2', 'terzaghi', 'indent_nh'])
def test_mixed_form_with_fp64_assembly_matches_oracle(name):
    s = _run_against_fixture(name, dict(mixed_assembly='fp64'))
    assert s.engine.vals64 is not None


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['bar_j2', 'indent_nh'])
def test_mixed_form_fp64_fallback_matches_oracle(name):
    """FP32 solves capped at one Krylov iteration leave relres > 0.1, so every
    correction is solved again in FP64 (FP64 tangent, Jacobian and hierarchy)."""
    def cap(eng):
        eng.mixed_inner_maxiter = 1
    s = _run_against_fixture(name, dict(ir_steps=1), cap)
    print(name, 'fallbacks', s.engine.stats.get('fp64_fallbacks'))
    assert s.engine.stats.get('fp64_fallbacks', 0) > 0

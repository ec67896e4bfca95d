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


@pytest.mark.gpu
def test_chord_step_and_skipped_trial_tangent():
    """Elastic skeletons in the mixed form solve a load step's last correction on the
    previous Jacobian (a chord step) and skip the particle tangent in the line-search
    trial of the correction solved to pass the Newton test.  Forcing every trial to skip
    it makes the Newton iteration compute the tangent again wherever a Jacobian is
    assembled, which must give the same bits; both runs meet the oracle bar of the
    indentation fixture (line search on)."""
    import torch
    import geowarp_amd.solvers as gs
    g = np.load(os.path.join(GOLD, 'indent_nh.npz'))
    total = int(g['total_steps'][0])
    out = {}
    for forced in (False, True):
        cfg = H.MINI['indent_nh']()
        cfg['iteration_dict'] = dict(cfg['iteration_dict'], precision='mixed', line_search=True)
        s = H.make(gs, cfg)
        assert s.final_chord
        calls = {'tangent': 0, 'plain': 0}
        stage = s.engine.stage

        def counted(args, want_tangent=True, _stage=stage, _calls=calls):
            _calls['tangent' if want_tangent else 'plain'] += 1
            return _stage(args, want_tangent=want_tangent)
        s.engine.stage = counted
        if forced:
            ew = s._ew_forcing

            def always_skip(*a, _ew=ew, _s=s):
                eta = _ew(*a)
                _s._aims_to_pass = True
                return eta
            s._ew_forcing = always_skip
        x0 = s.x0.cpu().numpy()
        for k in range(int(g['n_steps'][0])):
            s.advance_one_step(k, total)
            assert s.newton_converged
            f = H.fields(s)
            for key in ('x', 'sigma', 'p'):
                ref = g['s%d_%s' % (k, key)]
                e = H.rel_err(f[key] - x0, ref - x0) if key == 'x' else H.rel_err(f[key], ref)
                assert e < 1e-6, (forced, k, key, e)
        torch.cuda.synchronize()
        out[forced] = (s.x_particles.clone(), s.particle_Cauchy_stress_array.clone(), s.particle_pressure_array.clone(),
                       dict(calls), s.engine.stats.get('chord_steps', 0))
        print('forced skip' if forced else 'default', 'stage calls', calls, 'chord steps', out[forced][4])
    (xa, sa, pa, ca, na), (xb, sb, pb, cb, nb) = out[False], out[True]
    assert na > 0 and nb == na
    assert cb['plain'] > 0 and cb['plain'] >= ca['plain'] and cb['tangent'] >= ca['tangent']
    assert torch.equal(xa, xb) and torch.equal(sa, sb) and torch.equal(pa, pb)


@pytest.mark.gpu
def test_mixed_form_fields_at_the_benchmark_tolerance():
    """The benchmark's indentation (Neo-Hookean, indenter from the surface, physical penalty,
    line search) on a 16 x 16 x 8-cell block, load steps 0-7 to 1e-8 relative residual with
    the default options of each form (the mixed form's Eisenstat-Walker forcing, chord last
    correction and kept coarse operators; the FP64 form's fixed 1e-10 inner tolerance): the
    mixed form's displacement, stress and pore pressure stay within 1e-6 of the FP64 form's."""
    import torch
    import geowarp_amd.solvers as gs
    from geowarp_amd import problems as P
    out = {}
    for form in ('fp64', 'mixed'):
        cfg = P.indent3d(n_cells=(16, 16, 8), dx=1.0 / 64, total_steps=40, depth=0.05, footing=0.05,
                         material='Neo-Hookean', precision=form, tol=1e-8, n_iter=12, dp_apex='return',
                         shape_factor=0.1, footing_start='surface', penalty_form='physical')
        s = P.build_solver(gs, cfg, device='cuda:0')
        cfg['particle_flags'](s.x_particles, s.particle_traction_flag_array, s.particle_pressure_boundary_flag_array)
        s.line_search = True
        for k in range(8):
            s.advance_one_step(k, cfg['total_steps'])
            assert s.newton_converged, (form, k)
        torch.cuda.synchronize()
        out[form] = (s.x_particles - s.x0, s.particle_Cauchy_stress_array.clone(), s.particle_pressure_array.clone(),
                     s.engine.stats.get('chord_steps', 0))
    (u0, s0, p0, _), (u1, s1, p1, chords) = out['fp64'], out['mixed']
    rel = [float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in ((u1, u0), (s1, s0), (p1, p0))]
    print('mixed vs FP64 after 8 load steps: displacement %.2e stress %.2e pressure %.2e, chord steps %d'
          % (rel[0], rel[1], rel[2], chords))
    assert chords > 0
    assert max(rel) < 1e-6

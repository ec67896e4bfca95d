"""GPU parity: the HIP solvers against the oracle's golden fixtures.

Fixtures (tests/golden/*.npz) hold the oracle's state after each load step of a
reduced workload and its coloured-FD Jacobian at the first Newton iterate.
Bars: FP64 fields within 1e-8 relative (positions, stress, F, b, pressure);
Jacobian within 1e-5 relative (the oracle's Jacobian is a central FD with
h = 1e-6 dx, whose truncation error dominates).
"""
import os

import numpy as np
import pytest

import harness as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CASES = ['bar_nh', 'bar_hencky', 'bar_j2', 'bar_dp', 'bar_j2_platen', 'terzaghi', 'indent_nh', 'cantilever',
         'indent_dp', 'indent_dp_cone', 'c1_bar_nh', 'terzaghi_log', 'bar_norsand']
FIELD_TOL = 1e-8


def _gold(name):
    return np.load(os.path.join(GOLD, name + '.npz'))


@pytest.mark.gpu
@pytest.mark.parametrize('name', CASES)
def test_first_jacobian_matches_fd(name):
    import geowarp_amd.solvers as gs
    from geowarp_amd.solvers._engine import bsr_to_scipy_dofmajor
    import scipy.sparse as sp
    g = _gold(name)
    cfg = H.MINI[name]()
    s = H.make(gs, cfg)
    total = cfg['total_steps']
    if cfg.get('dts') is not None:
        s.dt = cfg['dts'][0]
    s.prepare_step(0, total)
    u_new = s._inc if hasattr(s, '_inc') else s._new
    args = s._args(0, total)
    eng = s.engine
    eng.stage(args, want_tangent=True)
    eng.residual(s.x_particles, s.GIMP_lp)
    eng.jacobian(s.x_particles, s.GIMP_lp)
    r = s.rhs.cpu().numpy()
    K = bsr_to_scipy_dofmajor(eng.row_ptr, eng.col_idx, eng.vals, eng.n_nodes, eng.nb)
    n = K.shape[0]
    Kref = sp.coo_matrix((g['jac_val'], (g['jac_row'], g['jac_col'])), shape=(n, n)).tocsr()
    rref = g['jac_res']
    print(name, 'residual rel err', H.rel_err(r, rref))
    assert H.rel_err(r, rref) < 1e-12
    D = (K - Kref).toarray()
    err = np.linalg.norm(D) / np.linalg.norm(Kref.toarray())
    print(name, 'jacobian rel err (Frobenius)', err)
    assert err < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize('name', CASES)
def test_load_steps_match_oracle(name):
    import geowarp_amd.solvers as gs
    g = _gold(name)
    cfg = H.MINI[name]()
    s = H.make(gs, cfg)
    total = int(g['total_steps'][0])
    for k in range(int(g['n_steps'][0])):
        if cfg.get('dts') is not None:
            s.dt = cfg['dts'][k]
        s.advance_one_step(k, total)
        f = H.fields(s)
        print(name, 'step', k, 'newton', ['%.2e' % v for v in s.residual_history],
              'krylov', s.engine.stats['krylov_iters'][-len(s.residual_history):])
        for key in ('x', 'sigma', 'F', 'b', 'p', 'alpha', 'force'):
            gk = 's%d_%s' % (k, key)
            if gk not in g or key not in f:
                continue
            ref = g[gk]
            if key == 'alpha' and np.abs(ref).max() == 0.0:
                assert np.abs(f[key]).max() == 0.0
                continue
            if key == 'x':
                e = H.rel_err(f[key] - g['s0_x'] * 0 - _x0(s), ref - _x0(s))  # displacement, not position
            else:
                e = H.rel_err(f[key], ref)
            print('   %-6s rel err %.3e' % (key, e))
            assert e < FIELD_TOL, (name, k, key, e)


def _x0(s):
    return s.x0.cpu().numpy()

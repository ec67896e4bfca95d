"""The C++/OpenMP restatement of the 3D u-p Neo-Hookean assembly (oracle/cpu_port.cpp,
the benchmark's CPU baseline) against the NumPy oracle: the same residual, and an
exact Jacobian (forward-mode dual numbers over the particle unknowns) equal to the
oracle's coloured central differences."""
import numpy as np
import pytest
import scipy.sparse as sp

import harness as H


@pytest.fixture(scope='module')
def port():
    import oracle.cpu_port as cp
    cp.build()
    return cp


def _state(cfg_name='indent_nh'):
    import oracle.solvers as osv
    cfg = H.MINI[cfg_name]()
    s = H.make(osv, cfg)
    st, f, u0, steps = s.prepare_step(0, cfg['total_steps'])
    return cfg, s, st, f, u0, steps


def test_residual_equals_numpy_oracle(port):
    cfg, s, st, f, u0, _ = _state()
    A = port.Assembler(s, st, 0, cfg['total_steps'], n_threads=2)
    rng = np.random.default_rng(1)
    for u in (u0, u0 + 1e-4 * rng.standard_normal(u0.size) * (~s.boundary)):
        r = A.assemble(u)
        r_ref = f(u)
        assert np.linalg.norm(r - r_ref) <= 1e-12 * np.linalg.norm(r_ref)


def test_jacobian_equals_coloured_fd_of_oracle(port):
    from oracle.solvers import fd_jacobian
    cfg, s, st, f, u0, steps = _state()
    A = port.Assembler(s, st, 0, cfg['total_steps'], n_threads=2)
    A.assemble(u0)
    K = A.matrix()
    g = np.load(H.__file__.replace('harness.py', 'golden/indent_nh.npz'))
    Kref = sp.coo_matrix((g['jac_val'], (g['jac_row'], g['jac_col'])), shape=K.shape).tocsr()
    assert np.linalg.norm((K - Kref).toarray()) <= 1e-6 * np.linalg.norm(Kref.toarray())
    # away from the first iterate (every particle moved): live central differences
    rng = np.random.default_rng(2)
    u = u0 + 2e-4 * rng.standard_normal(u0.size) * (~s.boundary)
    A.assemble(u)
    K = A.matrix()
    Kfd = fd_jacobian(lambda v: f(v), u, s.n_grid, 3, 4, s.n_nodes, s.boundary, s.act, steps)
    assert np.linalg.norm((K - Kfd).toarray()) <= 1e-6 * np.linalg.norm(Kfd.toarray())


def test_thread_count_does_not_change_the_result(port):
    cfg, s, st, f, u0, _ = _state()
    out = []
    for nt in (1, 3):
        A = port.Assembler(s, st, 0, cfg['total_steps'], n_threads=nt)
        r = A.assemble(u0)
        out.append((r, A.Kwin.copy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_rejects_other_materials(port):
    cfg, s, st, f, u0, _ = _state('indent_dp')
    with pytest.raises(ValueError):
        port.Assembler(s, st, 0, cfg['total_steps'])

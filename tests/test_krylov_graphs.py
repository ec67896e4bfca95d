"""GPU: the Arnoldi step graphs of the GMRES solvers (gm_krylov_graphs) give the same
bits as the eager launches.

A Newton system of the 3D u-p indentation (Neo-Hookean) in a deformed state is solved with the
multigrid-preconditioned FGMRES (gm_gmres_mg) and the block-Jacobi GMRES (gm_gmres),
once with the graphs off and three times with them on (the first solve captures every
step, the later ones replay them), with eager work and synchronisations between the
replays.  Every replayed solution and iteration count equals the eager one."""
import pytest

import harness as H
from geowarp_amd import problems as P


def _deformed_nh(precision):
    """3D u-p indentation (Neo-Hookean skeleton) after one load step, with the
    stage, residual and Jacobian of the next step's first Newton iterate."""
    import geowarp_amd.solvers as gs
    cfg = P.indent3d(n_cells=(8, 8, 6), dx=0.125, total_steps=10, depth=0.02, footing=0.5, material='Neo-Hookean',
                     precision=precision, tol=1e-10, n_iter=6, shape_factor=0.1, footing_start='surface')
    s = H.make(gs, cfg)
    total = cfg['total_steps']
    s.line_search = True
    s.advance_one_step(0, total)
    s.prepare_step(1, total)
    args = s._args(1, total)
    s.engine.stage(args, want_tangent=True)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
@pytest.mark.parametrize('method', ['mg', 'gmres'])
def test_gmres_graph_replays_equal_eager_solves(precision, method):
    import torch
    from geowarp_amd import _lib as L
    lib = L.lib()
    s = _deformed_nh(precision)
    eng = s.engine
    eng.residual(s.x_particles, s.GIMP_lp)
    eng.jacobian(s.x_particles, s.GIMP_lp)
    fp32 = precision == 'mixed'
    b = eng.rhs.float() if fp32 else eng.rhs.clone()
    tol = 1e-4 if fp32 else 1e-9  # an FP32 solve stops far above FP64 round-off
    prev = lib.gm_krylov_graphs(-1)
    try:
        lib.gm_krylov_graphs(0)
        x_e = torch.zeros_like(b)
        its_e, rel_e = eng.krylov(b, x_e, fp32, tol, 500, method=method, restart=40)
        torch.cuda.synchronize()
        lib.gm_krylov_graphs(1)
        for rep in range(3):
            x_g = torch.full_like(b, float('nan'))
            x_g.zero_()
            its_g, rel_g = eng.krylov(b, x_g, fp32, tol, 500, method=method, restart=40)
            torch.cuda.synchronize()
            print(precision, method, 'replay', rep, 'iterations', its_g, 'eager', its_e, 'relres', rel_g, rel_e)
            assert its_g == its_e and rel_g == rel_e
            assert torch.equal(x_g, x_e)
            # eager work of the program between replays
            junk = torch.randn(1 << 20, device=b.device)
            float(junk.abs().sum())
            torch.cuda.synchronize()
        # the multigrid solve converges; block-Jacobi GMRES only makes progress on this
        # penalty system within the iteration cap (the property tested is the equality)
        assert its_e > 2 and rel_e <= (tol if method == 'mg' else 1e-2)
    finally:
        lib.gm_krylov_graphs(prev)

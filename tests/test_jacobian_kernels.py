"""GPU: the 3D u-p node-Jacobian kernels agree bit for bit.

k_node_jacobian_up3 (lanes per column and column dof of a bin, the default)
and k_node_jacobian (one wavefront per block row, GM_JACOBIAN_ROW_WAVE=1) add
the same particle contributions to every block entry in the same order, so the
assembled matrices must be identical, in FP64 and in the FP32 mixed form, on a
state with stretched particle domains (3-node supports) and Drucker-Prager
plasticity."""
import os

import numpy as np
import pytest

import harness as H
from geowarp_amd import problems as P


def _assemble(s, total, row_wave):
    if row_wave:
        os.environ['GM_JACOBIAN_ROW_WAVE'] = '1'
    else:
        os.environ.pop('GM_JACOBIAN_ROW_WAVE', None)
    try:
        eng = s.engine
        eng.jacobian(s.x_particles, s.GIMP_lp)
        return eng.vals[:eng.nnzb * eng.nb * eng.nb].clone()
    finally:
        os.environ.pop('GM_JACOBIAN_ROW_WAVE', None)


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_node_jacobian_kernels_agree_bitwise(precision):
    import torch
    import geowarp_amd.solvers as gs
    cfg = P.indent3d(n_cells=(8, 8, 6), dx=0.125, total_steps=10, depth=0.02, footing=0.5, material='Drucker-Prager',
                     precision=precision, tol=1e-10, n_iter=6, dp_apex='return', shape_factor=0.1,
                     footing_start='surface')
    cfg['material_dict']['dilation_angle'] = 30.0
    s = H.make(gs, cfg)
    total = cfg['total_steps']
    s.line_search = True
    s.advance_one_step(0, total)  # deformed state: particle domains straddle cell faces
    s.prepare_step(1, total)
    args = s._args(1, total)
    s.engine.penalty = None
    s.engine.stage(args, want_tangent=True)
    a = _assemble(s, total, row_wave=False)
    b = _assemble(s, total, row_wave=True)
    torch.cuda.synchronize()
    nnz = int((a != 0).sum())
    print(precision, 'nnzb', s.engine.nnzb, 'nonzeros', nnz, 'max |a-b|', float((a - b).abs().max()))
    assert nnz > 1000
    assert torch.equal(a, b)

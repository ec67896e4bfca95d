"""GPU: the 3D u-p node-Jacobian kernels assemble the same matrix.

k_node_jacobian_up3 (lanes per column and column dof of a bin) and
k_node_jacobian (one wavefront per block row, GM_JACOBIAN_ROW_WAVE=1) add the
same particle contributions to every block entry in the same order, so their
matrices are identical bit for bit.  The binned form (per-bin products on the
matrix cores, then a fixed-order gather over bins; the default) sums in another
order: in FP64 it agrees with the row form to round-off, and its FP32 form (the
mixed form's) to FP32 round-off.  The state has stretched particle domains
(3-node supports) and Drucker-Prager plasticity.
"""
import os

import numpy as np
import pytest

import harness as H
from geowarp_amd import problems as P


def _assemble(s, kernel, row_wave=False):
    eng = s.engine
    old = eng.jacobian_kernel
    eng.jacobian_kernel = kernel
    if row_wave:
        os.environ['GM_JACOBIAN_ROW_WAVE'] = '1'
    try:
        eng.jacobian(s.x_particles, s.GIMP_lp)
        return eng.vals[:eng.nnzb * eng.nb * eng.nb].clone()
    finally:
        eng.jacobian_kernel = old
        os.environ.pop('GM_JACOBIAN_ROW_WAVE', None)


def _deformed(precision):
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
    return s


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_node_jacobian_kernels_agree(precision):
    import torch
    s = _deformed(precision)
    a = _assemble(s, 'rows')
    b = _assemble(s, 'rows', row_wave=True)
    c = _assemble(s, 'bins')
    torch.cuda.synchronize()
    nnz = int((a != 0).sum())
    a64, c64 = a.double(), c.double()
    rel = float((c64 - a64).norm() / a64.norm())
    rel_max = float((c64 - a64).abs().max() / a64.abs().max())
    print(precision, 'nnzb', s.engine.nnzb, 'nonzeros', nnz, 'rows vs row-wave max |a-b|', float((a - b).abs().max()),
          'bins vs rows: Frobenius', rel, 'max', rel_max, 'nonzero patterns equal', bool(((a != 0) == (c != 0)).all()))
    assert nnz > 1000
    assert torch.equal(a, b)
    tol = 1e-13 if precision == 'fp64' else 1e-6
    assert rel < tol and rel_max < tol


@pytest.mark.gpu
def test_binned_residual_equals_node_row_residual():
    """The per-bin residual (gm_node_residual_binned, the 3D u-p default) sums the same
    particle terms as the node-row gather in another order: equal to FP64 round-off."""
    import torch
    s = _deformed('fp64')
    eng = s.engine
    out = {}
    for kernel in ('rows', 'bins'):
        eng.jacobian_kernel = kernel
        eng.residual(s.x_particles, s.GIMP_lp)
        out[kernel] = eng.rhs.clone()
    eng.jacobian_kernel = 'bins'
    torch.cuda.synchronize()
    a, b = out['rows'], out['bins']
    rel = float((a - b).norm() / a.norm())
    print('residual bins vs rows', rel, 'norm', float(a.norm()))
    assert float(a.norm()) > 0 and rel < 1e-13

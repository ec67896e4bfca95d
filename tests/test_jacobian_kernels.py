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


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_binned_jacobian_variants_write_the_same_bits(precision):
    """Every per-bin variant of the binned assembly (particles per group, workgroups
    per CU, one workgroup per bin or a persistent sweep with prefetch) and both
    gather variants (per-lane reads, chunks staged in LDS in two round sizes) give the same matrix bit
    for bit (groups of 4 particles split every bin in two or more); the workspace and the matrix are filled with NaN before each pair, so
    an entry a variant fails to write shows.  The tuner returns variants in range
    and leaves the same matrix."""
    import torch
    from geowarp_amd import _lib as L
    from geowarp_amd.solvers._engine import _stream
    s = _deformed(precision)
    eng = s.engine
    lib = L.lib()
    t32 = precision == 'mixed' and eng._tangent_fp32()
    v32 = precision == 'mixed'
    g, b, pat = L.ctypes.byref(eng.grid), L.ctypes.byref(eng.bins), L.ctypes.byref(eng.pattern)
    need = int(lib.gm_node_jacobian_work_bytes(g, int(t32)))
    w = eng._bins_workspace(need)
    n = eng.nnzb * eng.nb * eng.nb
    vals = torch.empty(n, dtype=torch.float32 if v32 else torch.float64, device=eng.vals.device)
    counts = eng.cell_start[1:] - eng.cell_start[:-1]
    print(precision, 'bins', int((counts > 0).sum()), 'max particles per bin', int(counts.max()))
    ref = None
    for vb in range(4):
        for vg in range(3):
            w.fill_(255)
            vals.fill_(float('nan'))
            L.check(lib.gm_jacobian_bins_variant(g, b, L.ptr(eng.D), int(t32), vb, L.ptr(w), need, _stream()),
                    'gm_jacobian_bins_variant')
            L.check(lib.gm_jacobian_gather_variant(g, b, pat, int(t32), vg, L.ptr(w), L.ptr(eng.freem),
                                                   L.ptr(vals), int(v32), _stream()), 'gm_jacobian_gather_variant')
            out = vals.clone()
            assert not bool(torch.isnan(out).any()), (vb, vg)
            if ref is None:
                ref = out
            else:
                assert torch.equal(out, ref), (vb, vg)
    with pytest.raises(L.GMError):
        L.check(lib.gm_jacobian_bins_variant(g, b, L.ptr(eng.D), int(t32), 4, L.ptr(w), need, _stream()), 'variant')
    bb, bg = eng.autotune_jacobian()
    torch.cuda.synchronize()
    print(precision, eng.stats['jacobian_tune'])
    assert 0 <= bb < 4 and 0 <= bg < 3
    got = eng.vals64 if (v32 and not t32) else eng.vals
    assert torch.equal(got[:n], ref.to(got.dtype))


@pytest.mark.gpu
def test_binned_p2g_matches_node_p2g():
    """The binned 3D P2G (gm_p2g_binned: per bin, then a fixed-order gather per node) sets
    the same activation and pressure-boundary flags as the node-centric gm_p2g, and its
    consistent-mass right-hand side sums the same terms in another order (FP64 round-off),
    on the deformed state with flagged surface particles."""
    import torch
    s = _deformed('fp64')
    eng = s.engine
    pbc = s.particle_pressure_boundary_flag_array.to(torch.uint8)
    assert int(pbc.sum()) > 0
    bflag0 = eng.bflag.clone()
    bflag0.view(-1, 4)[:, 3] = 0  # the pressure dofs: set again by the P2G's boundary rule
    out = {}
    for binned in (False, True):
        eng.binned_p2g = binned
        eng.bflag.copy_(bflag0)
        mrhs = eng.p2g(s.x_particles, s.GIMP_lp, 1e-7, pbc, s.dz, s.p_mass, s.particle_pressure_array, mass=True)
        torch.cuda.synchronize()
        out[binned] = (eng.act.clone(), eng.bflag.clone(), mrhs.clone(), eng.freem.clone())
    eng.binned_p2g = True
    (a0, f0, r0, m0), (a1, f1, r1, m1) = out[False], out[True]
    rel = float((r1 - r0).norm() / r0.norm())
    print('active dofs', int(a0.sum()), 'pressure-boundary dofs set', int((f0 != bflag0).sum()), 'rhs rel', rel)
    assert int(a0.sum()) > 0 and int((f0 != bflag0).sum()) > 0
    assert torch.equal(a0, a1) and torch.equal(f0, f1) and torch.equal(m0, m1)
    assert float(r0.norm()) > 0 and rel < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize('dim', [2, 3])
def test_pattern_from_support_sets_equals_particle_pattern(dim):
    """gm_bsr_pattern's row masks from each bin's set of support shapes (the default) equal
    those formed particle by particle, bit for bit, on a deformed state (stretched domains:
    2- and 3-node supports)."""
    import torch
    import geowarp_amd.solvers as gs
    from geowarp_amd import _lib as L
    if dim == 3:
        s = _deformed('fp64')
    else:
        cfg = H.MINI['bar_dp']()
        s = H.make(gs, cfg)
        s.advance_one_step(0, cfg['total_steps'])
        s.prepare_step(1, cfg['total_steps'])
    eng = s.engine
    lib = L.lib()
    out = {}
    try:
        for by_sets in (0, 1):
            L.check(lib.gm_pattern_kernel(by_sets), 'gm_pattern_kernel')
            eng.begin_step(s.x_particles, s.GIMP_lp)
            torch.cuda.synchronize()
            n = eng.n_nodes
            out[by_sets] = (eng.row_mask.clone(), eng.row_ptr[:n + 1].clone(), eng.col_idx[:eng.nnzb].clone())
    finally:
        lib.gm_pattern_kernel(1)
    ext = eng.ext[:s.x_particles.shape[0]]
    print(dim, 'nnzb', eng.nnzb, 'particles with a 3-node axis', int((ext != 0).sum()), 'of', ext.numel())
    assert int((ext != 0).sum()) > 0
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)

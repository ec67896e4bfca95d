"""GPU: the consistent-mass projection's conjugate gradients (gm_pcg) against BiCGStab
(gm_bicgstab) and against SciPy's direct solve of the same matrix, on the 3D u-p
indentation after one load step: cold and warm starts, diagonal and lumped (row-sum)
preconditioners.  Every solution satisfies ||M x - b|| <= 1e-13 ||b|| in the solvers'
own stopping rule and agrees with the direct solve to 1e-11.  CG stops on the device, so
the interval at which the host reads its state leaves the iterate's bits unchanged."""
import numpy as np
import pytest

import harness as H
from geowarp_amd import problems as P


@pytest.mark.gpu
def test_pcg_mass_projection_matches_bicgstab_and_direct_solve():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    import geowarp_amd.solvers as gs
    from geowarp_amd import _lib as L
    from geowarp_amd.solvers._engine import _stream
    cfg = P.indent3d(n_cells=(8, 8, 6), dx=0.125, total_steps=10, depth=0.02, footing=0.5, material='Neo-Hookean',
                     precision='fp64', tol=1e-10, n_iter=8, shape_factor=0.1, footing_start='surface')
    s = H.make(gs, cfg)
    total = cfg['total_steps']
    s.advance_one_step(0, total)
    p_prev = s._new.view(-1, s.engine.nb)[:, -1].clone()
    # the next load step's projection system (P2G of the moved particles)
    eng = s.engine
    s.reset()
    s._sync_grid()
    s.set_boundary_dofs(s.dofStruct.boundary_flag_array, s.n_grid_x, s.n_grid_y, s.n_nodes, s.dx, s.dy, s.end_x,
                        s.end_y)
    eng.begin_step(s.x_particles, s.GIMP_lp)
    eng.set_boundary_from_dofmajor(s.dofStruct.boundary_flag_array)
    pbc = s.particle_pressure_boundary_flag_array.to(torch.uint8)
    mrhs = eng.p2g(s.x_particles, s.GIMP_lp, 1e-7, pbc, s.dz, s.p_mass, s.particle_pressure_array, mass=True)
    lib = L.lib()
    L.check(lib.gm_fix_empty_diag(eng.n_nodes, L.ptr(eng.row_ptr), L.ptr(eng.col_idx), L.ptr(eng.mvals), L.ptr(mrhs),
                                  _stream()), 'gm_fix_empty_diag')
    n = eng.n_nodes
    nnz = int(eng.row_ptr[n].item())
    M = sp.csr_matrix((eng.mvals[:nnz].cpu().numpy(), eng.col_idx[:nnz].cpu().numpy(), eng.row_ptr.cpu().numpy()),
                      shape=(n, n))
    b = mrhs.cpu().numpy()
    assert abs(M - M.T).max() <= 1e-14 * abs(M).max()  # symmetric
    x_ref = spla.spsolve(M.tocsc(), b)
    ws = eng.workspace(1, False)
    out = {}
    for name, opts, x0 in (('pcg', 0, None), ('pcg lumped', 1, None), ('pcg warm', 2, p_prev),
                           ('pcg lumped warm', 3, p_prev), ('bicgstab', None, None)):
        x = torch.zeros(n, dtype=torch.float64, device=mrhs.device) if x0 is None else x0.clone()
        it, rr = L.c_int(0), L.c_double(0.0)
        head = (1, 0, n, L.ptr(eng.row_ptr), L.ptr(eng.col_idx), L.ptr(eng.mvals), L.ptr(mrhs), L.ptr(x),
                L.ctypes.byref(ws), 1e-13, 5000, 4)
        tail = (L.ctypes.byref(it), L.ctypes.byref(rr), _stream())
        if opts is None:
            L.check(lib.gm_bicgstab(*head, *tail), name)
        else:
            L.check(lib.gm_pcg(*head, opts, *tail), name)
        xv = x.cpu().numpy()
        true_rel = np.linalg.norm(M @ xv - b) / np.linalg.norm(b)
        err = np.linalg.norm(xv - x_ref) / np.linalg.norm(x_ref)
        out[name] = (it.value, rr.value, true_rel, err)
        print(name, 'iterations', it.value, 'relres', rr.value, 'true relres %.2e' % true_rel,
              'vs direct %.2e' % err)
        assert rr.value <= 1e-13 and true_rel <= 2e-13 and err <= 1e-11, name
    assert out['pcg'][0] < 2 * out['bicgstab'][0] + 8  # one SpMV per CG step, two per BiCGStab step
    # CG stops on the device at the first iterate that meets the tolerance: how often the
    # host reads its state changes neither that iterate nor the iteration count
    res = []
    for every in (1, 16, 64):
        x = p_prev.clone()
        it, rr = L.c_int(0), L.c_double(0.0)
        head = (1, 0, n, L.ptr(eng.row_ptr), L.ptr(eng.col_idx), L.ptr(eng.mvals), L.ptr(mrhs), L.ptr(x),
                L.ctypes.byref(ws), 1e-13, 5000, every)
        L.check(lib.gm_pcg(*head, 2, L.ctypes.byref(it), L.ctypes.byref(rr), _stream()), 'pcg every %d' % every)
        res.append((x.clone(), it.value, rr.value))
        print('check every', every, 'iterations', it.value, 'relres', rr.value)
    assert res[0][1] == out['pcg warm'][0]
    for xr, its, rel in res[1:]:
        assert its == res[0][1] and rel == res[0][2] and torch.equal(xr, res[0][0])

"""Workload generators (geowarp_amd.problems), on the CPU.

* The batched specimens' platen hook sets the boundary flags from the current
  particle tops alone: flags left by an earlier, more compressed state do not
  survive (a load path recomputed from its initial state must see the same
  boundary conditions as the first time).
* The platen rows are the node rows nearest to each specimen's top surface.
"""
import numpy as np

from geowarp_amd import problems as P


def _flags(cfg, xp, lp):
    g = cfg['background_grid_dict']
    nx1, ny1 = g['n_grid_x'] + 1, g['n_grid_y'] + 1
    n_nodes = nx1 * ny1
    boundary = np.zeros(2 * n_nodes, dtype=bool)
    cfg['set_boundary_dofs'](boundary, g['n_grid_x'], n_nodes)
    return boundary, n_nodes, nx1


def _platen(cfg, boundary, xp, lp, n_nodes, nx1, step):
    g = cfg['background_grid_dict']
    pres = np.zeros(2 * n_nodes)
    cfg['load_dict']['prescribed_increment'](xp, lp, boundary, pres, g['n_grid_x'], n_nodes, g['dx'], step,
                                             cfg['total_steps'])
    return pres


def test_platen_hook_is_stateless():
    cfg = P.specimens2d([3, 5], [30.0, 35.0], [5.0, 10.0], total_steps=4, strain=0.06)
    x0 = cfg['x_particles']
    dx = cfg['background_grid_dict']['dx']
    lp = np.full_like(x0, 0.25 * dx)
    boundary, n_nodes, nx1 = _flags(cfg, x0, lp)
    _platen(cfg, boundary, x0, lp, n_nodes, nx1, 0)
    first = boundary.copy()
    # a state compressed by two cells: the platen rows move down
    xc = x0.copy()
    xc[:, 1] -= 2.0 * dx * (xc[:, 1] - xc[:, 1].min()) / (xc[:, 1].max() - xc[:, 1].min())
    _platen(cfg, boundary, xc, lp, n_nodes, nx1, 3)
    assert boundary.sum() > first.sum()
    # back at the initial state: exactly the flags of the first call
    _platen(cfg, boundary, x0, lp, n_nodes, nx1, 0)
    assert np.array_equal(boundary, first)


def test_platen_rows_are_nearest_to_the_top():
    widths = [3]
    cfg = P.specimens2d(widths, [30.0], [5.0], total_steps=4, strain=0.06)
    x0 = cfg['x_particles']
    dx = cfg['background_grid_dict']['dx']
    lp = np.full_like(x0, 0.25 * dx)
    for shift, row_expected in ((0.0, 7), (0.004, 7), (0.006, 6), (0.014, 6), (0.016, 5)):
        boundary, n_nodes, nx1 = _flags(cfg, x0, lp)
        xs = x0.copy()
        xs[:, 1] -= shift
        top = float((xs[:, 1] + lp[:, 1]).max())
        assert abs(top / dx - round(top / dx)) > 1e-6 or shift == 0.0
        pres = _platen(cfg, boundary, xs, lp, n_nodes, nx1, 0)
        by = boundary[n_nodes:].reshape(-1, nx1)
        py = pres[n_nodes:].reshape(-1, nx1)
        col = 4 + 1  # an inner node column of the specimen (x0 = pad 1 + gap 3)
        rows = np.nonzero(by[2:, col])[0] + 2  # above the fixed base rows 0..1
        assert rows.min() == row_expected, (shift, rows)
        assert np.all(py[rows, col] < 0.0)
        assert not by[2:row_expected, col].any()


def test_cantilever_boundary_and_load_flags_follow_the_driver():
    """cantilever_beam_cpGIMP.py:98-133 written out on the node and particle grid:
    x fixed for node columns 0-1; y fixed only at those columns' node row
    int((end_y - d0/2 + dx/2) / dx) = row 38 (end_y 10, d0 1, dx 0.25); the
    loaded particles are the last particle column within dx/PPD of the
    mid-height y = 9.5 (two particles at PPD 6)."""
    import geowarp_amd.solvers._base as B
    cfg = P.cantilever()
    g = cfg['background_grid_dict']
    nx1 = g['n_grid_x'] + 1
    n_nodes = nx1 * (g['n_grid_y'] + 1)
    boundary = np.zeros(2 * n_nodes, dtype=bool)
    cfg['set_boundary_dofs'](boundary, g['n_grid_x'], n_nodes)
    bx = boundary[:n_nodes].reshape(-1, nx1)
    by = boundary[n_nodes:].reshape(-1, nx1)
    assert bx[:, :2].all() and not bx[:, 2:].any()
    assert by[38, :2].all() and by.sum() == 2
    x = B.init_particles_rectangle(2, (g['start_x'], g['start_y']), (g['end_x'], g['end_y']), g['dx'],
                                   (g['n_grid_x'], g['n_grid_y']), 6)
    assert x.shape[0] == cfg['particles_dict']['n_particles'] == 5760
    flag = np.zeros(x.shape[0], dtype=bool)
    cfg['external_flags'](x, flag)
    assert flag.sum() == 2
    assert np.allclose(np.sort(x[flag, 1]), [9.5 - 0.25 / 12, 9.5 + 0.25 / 12])
    assert np.allclose(x[flag, 0], 10.25 - 0.25 / 12)


def test_load_loop_writes_one_vtk_file_per_step(tmp_path):
    """run_load_steps drives advance_one_step and writes the initial state plus
    one file per step (the drivers' post-processing loop)."""
    from geowarp_amd import vtk

    class Fake:
        def __init__(self):
            self.x_particles = np.zeros((3, 2))
            self.particle_Cauchy_stress_array = np.zeros((3, 3, 3))
            self.calls = []
            self.newton_iterations = 2

        def advance_one_step(self, step, total):
            self.calls.append((step, total))
            self.x_particles = self.x_particles + 1.0

    s = Fake()
    iters = P.run_load_steps(s, dict(total_steps=4), n_steps=3, vtk_pattern=str(tmp_path / 'p_%d.vtk'))
    assert s.calls == [(0, 4), (1, 4), (2, 4)] and iters == [2, 2, 2]
    pts, f = vtk.read_points(str(tmp_path / 'p_3.vtk'))
    assert np.all(pts[:, :2] == 3.0) and set(f) == {'stress_xx', 'stress_yy'}
    assert (tmp_path / 'p_0.vtk').exists()

"""Legacy-VTK particle writer (stand-in for the drivers' meshio calls,
terzaghi_cpGIMP.py:176-189): header layout, big-endian binary blocks, round
trip in both encodings, and the driver-style Mesh(...).write call."""
import numpy as np
import pytest

from geowarp_amd import vtk


@pytest.mark.parametrize('binary', [True, False])
def test_round_trip_2d_points_with_scalars_and_vectors(tmp_path, binary):
    rng = np.random.default_rng(0)
    x = rng.uniform(size=(37, 2))
    sig = rng.normal(size=(37, 3, 3))
    p = rng.normal(size=37)
    path = tmp_path / ('p_%d.vtk' % binary)
    vtk.Mesh(points=x, cells=[], point_data={'stress_xx': sig[:, 0, 0], 'stress_yy': sig[:, 1, 1],
                                              'pressure': p, 'u': x * 2, 'F': sig.reshape(37, 9)}).write(
        str(path), binary=binary)
    pts, f = vtk.read_points(str(path))
    assert np.array_equal(pts[:, :2], x) and np.all(pts[:, 2] == 0.0)
    assert np.array_equal(f['stress_xx'], sig[:, 0, 0])
    assert np.array_equal(f['stress_yy'], sig[:, 1, 1])
    assert np.array_equal(f['pressure'], p)
    assert np.array_equal(f['u'][:, :2], 2 * x)
    assert np.array_equal(f['F'], sig.reshape(37, 9))


def test_binary_layout_is_big_endian_legacy_vtk(tmp_path):
    x = np.array([[1.5, 2.0, 3.25], [4.0, 5.0, 6.0]])
    path = tmp_path / 'b.vtk'
    vtk.write_points(str(path), x, {'s': np.array([7.0, 8.0])})
    raw = path.read_bytes()
    head = raw.split(b'\n')[:5]
    assert head[0] == b'# vtk DataFile Version 4.2'
    assert head[2] == b'BINARY'
    assert head[3] == b'DATASET UNSTRUCTURED_GRID'
    assert head[4] == b'POINTS 2 double'
    start = sum(len(h) + 1 for h in head)
    block = raw[start:start + 48]
    assert np.array_equal(np.frombuffer(block, dtype='>f8').reshape(2, 3), x)
    assert b'CELLS 0 0\nCELL_TYPES 0\nPOINT_DATA 2\nSCALARS s double 1\nLOOKUP_TABLE default\n' in raw


def test_rejects_cells_and_mismatched_data(tmp_path):
    with pytest.raises(ValueError):
        vtk.Mesh(points=np.zeros((2, 3)), cells=[('vertex', [[0]])])
    with pytest.raises(ValueError):
        vtk.write_points(str(tmp_path / 'x.vtk'), np.zeros((3, 2)), {'a': np.zeros(2)})

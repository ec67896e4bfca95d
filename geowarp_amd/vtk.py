"""Legacy-VTK particle output, a stand-in for the reference drivers' meshio calls.

The drivers write one file per load step with

    meshio.Mesh(points=x, cells=[], point_data={'stress_xx': ..., 'pressure': ...}).write(path)

(problems/consolidation/terzaghi_cpGIMP.py:176-189,
problems/bar_compaction/bar_compaction_cpGIMP.py:124-137).  meshio is not a
dependency here; ``Mesh`` takes the same keyword arguments and ``write`` emits a
legacy VTK (version 4.2) unstructured grid of vertex-free points: POINTS (2D
coordinates padded with z = 0, as VTK requires three components) and one
SCALARS array per point-data entry (vectors of 2 or 3 components as VECTORS,
padded likewise).  Binary files are big-endian, as the legacy format demands.
Torch tensors and numpy arrays are both accepted.
"""
import numpy as np

__all__ = ['Mesh', 'write_points', 'read_points']


def _np(a):
    if hasattr(a, 'detach'):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _pad3(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    if a.shape[1] == 3:
        return a
    out = np.zeros((a.shape[0], 3))
    out[:, :a.shape[1]] = a
    return out


def _block(f, arr, binary):
    if binary:
        f.write(np.ascontiguousarray(arr, dtype='>f8').tobytes())
        f.write(b'\n')
    else:
        flat = np.asarray(arr, dtype=np.float64).reshape(arr.shape[0], -1)
        for row in flat:
            f.write((' '.join(repr(float(v)) for v in row) + '\n').encode())


def write_points(path, points, point_data=None, binary=True, title='geowarp_amd particles'):
    """Write particles as a legacy-VTK unstructured grid without cells."""
    pts = _pad3(_np(points))
    n = pts.shape[0]
    with open(path, 'wb') as f:
        f.write(b'# vtk DataFile Version 4.2\n')
        f.write(title.encode()[:255] + b'\n')
        f.write(b'BINARY\n' if binary else b'ASCII\n')
        f.write(b'DATASET UNSTRUCTURED_GRID\n')
        f.write(('POINTS %d double\n' % n).encode())
        _block(f, pts, binary)
        f.write(b'CELLS 0 0\nCELL_TYPES 0\n')
        if point_data:
            f.write(('POINT_DATA %d\n' % n).encode())
            for name, val in point_data.items():
                v = _np(val).astype(np.float64)
                if v.shape[0] != n:
                    raise ValueError('point data %r has %d entries for %d points' % (name, v.shape[0], n))
                key = str(name).replace(' ', '_')
                if v.ndim == 1 or (v.ndim == 2 and v.shape[1] == 1):
                    f.write(('SCALARS %s double 1\nLOOKUP_TABLE default\n' % key).encode())
                    _block(f, v.reshape(n, 1), binary)
                elif v.ndim == 2 and v.shape[1] in (2, 3):
                    f.write(('VECTORS %s double\n' % key).encode())
                    _block(f, _pad3(v), binary)
                else:
                    f.write(('FIELD FieldData 1\n%s %d %d double\n' % (key, v.reshape(n, -1).shape[1], n)).encode())
                    _block(f, v.reshape(n, -1), binary)


class Mesh:
    """meshio.Mesh(points=..., cells=[], point_data={...}) for particle output;
    ``write(path)`` writes legacy VTK (binary unless binary=False)."""

    def __init__(self, points, cells=None, point_data=None, **unused):
        if cells:
            raise ValueError('only cell-free particle meshes are written (the drivers pass cells=[])')
        self.points = points
        self.cells = []
        self.point_data = dict(point_data or {})

    def write(self, path, binary=True):
        write_points(path, self.points, self.point_data, binary=binary)


def read_points(path):
    """Read back a file written by write_points: (points (n, 3), {name: array})."""
    with open(path, 'rb') as f:
        data = f.read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b'\n', pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    def values(count, binary):
        nonlocal pos
        if binary:
            arr = np.frombuffer(data, dtype='>f8', count=count, offset=pos).astype(np.float64)
            pos += 8 * count + 1
            return arr
        out = []
        while len(out) < count:
            out.extend(float(t) for t in line().split())
        return np.array(out)

    assert line().startswith('# vtk DataFile')
    line()
    binary = line().strip() == 'BINARY'
    assert line().strip() == 'DATASET UNSTRUCTURED_GRID'
    n = int(line().split()[1])
    pts = values(3 * n, binary).reshape(n, 3)
    fields = {}
    while pos < len(data):
        toks = line().split()
        if not toks or toks[0] in ('CELLS', 'CELL_TYPES', 'POINT_DATA'):
            continue
        if toks[0] == 'SCALARS':
            line()  # LOOKUP_TABLE
            fields[toks[1]] = values(n, binary)
        elif toks[0] == 'VECTORS':
            fields[toks[1]] = values(3 * n, binary).reshape(n, 3)
        elif toks[0] == 'FIELD':
            name, ncomp = line().split()[:2]
            fields[name] = values(int(ncomp) * n, binary).reshape(n, int(ncomp))
    return pts, fields

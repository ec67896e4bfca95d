"""Test infrastructure (the oracle): build and call the C++17/OpenMP restatement
(oracle/cpu_port.cpp) of one Newton iteration's assembly of the 3D u-p
Neo-Hookean indentation residual (mpm_utils.py:1020-1294) with its exact
Jacobian, from the NumPy oracle's per-step state.

Used by the tests (against the NumPy oracle's residual and coloured-FD Jacobian)
and by the benchmark's CPU baseline; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'cpu_port.cpp')
LIB = os.path.join(HERE, 'libcpu_port.so')
# x86-64-v3 (AVX2, FMA): every host this runs on has it, and the code vectorises the dual numbers
FLAGS = ['-O3', '-march=x86-64-v3', '-std=c++17', '-fopenmp', '-fPIC', '-shared']


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ('mu', 'lam', 'p_vol', 'p_rho', 'phi0', 'mobility', 'dt', 'gravity_z',
                                               'pen_par', 'h_cur', 'fminx', 'fmaxx', 'fminy', 'fmaxy')]


def build():
    tmp = LIB + '.tmp'
    r = subprocess.run(['g++'] + FLAGS + ['-o', tmp, SRC], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('g++ failed:\n' + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.op_assemble_nh3d.restype = ctypes.c_int
        L.op_assemble_nh3d.argtypes = [ctypes.c_int, P, P, P, P, P, P, P, P, P, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(Params), P, P, P, P, ctypes.c_int]
        L.op_params_size.restype = ctypes.c_int
        if L.op_params_size() != ctypes.sizeof(Params):
            raise RuntimeError('cpu_port: parameter layout mismatch')
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def params_of(s, step, total):
    """Parameters of an oracle quasi_static_coupled_solver_3d load step (mirrors
    oracle/mpm.py residual_coupled_3d)."""
    fminx, fmaxx, fminy, fmaxy, h0, incr, pf = s.footing
    q = Params()
    q.mu, q.lam = s.mat.mu, s.mat.lam
    q.p_vol, q.p_rho, q.phi0 = s.p_vol, s.p_rho, s.phi_initial
    q.mobility, q.dt = s.mobility_constant, s.dt
    q.gravity_z = -s.gravity_mag * (step + 1.0) / total
    q.pen_par = s.mat.E * pf * (s.dx / s.PPD) * (s.dy / s.PPD)
    q.h_cur = h0 + (step + 1.0) * incr
    q.fminx, q.fmaxx, q.fminy, q.fmaxy = fminx, fmaxx, fminy, fmaxy
    return q


class Assembler:
    """Per-step arrays of an oracle 3D coupled solver, after prepare_step(step, total)
    returned the stencil `st`; assemble(u) gives (rhs, K) at the nodal unknowns u
    (dof-major), K = -dR/du between free dofs and identity elsewhere, as the
    oracle's fd_jacobian defines it."""

    def __init__(self, s, st, step, total, n_threads=None):
        if s.mat.name != 'Neo-Hookean':
            raise ValueError('cpu_port covers the Neo-Hookean 3D u-p form')
        self.s, self.step, self.total = s, step, total
        self.n = st.n
        self.idx = np.ascontiguousarray(st.idx, dtype=np.int64)
        self.W = np.ascontiguousarray(st.W)
        self.G = np.ascontiguousarray(st.G)
        self.Wa = np.ascontiguousarray(st.Wa)
        self.support = np.ascontiguousarray(st.support, dtype=np.uint8)
        self.x = np.ascontiguousarray(s.x_particles)
        self.F_old = np.ascontiguousarray(s.F_old)
        self.old = np.ascontiguousarray(s.old_solution)
        self.nx1, self.ny1, self.nz1 = s.n_grid[0] + 1, s.n_grid[1] + 1, s.n_grid[2] + 1
        self.nn = self.nx1 * self.ny1 * self.nz1
        self.free = (~s.boundary) & s.act
        self.freem = np.ascontiguousarray(self.free, dtype=np.uint8)
        self.q = params_of(s, step, total)
        self.n_threads = int(n_threads or len(os.sched_getaffinity(0)))
        self.Kwin = np.zeros(self.nn * 125 * 16)
        self.used = np.zeros(self.nn * 125, dtype=np.uint8)
        self._coo = None

    def assemble(self, u, want_jac=True):
        """C++ assembly; returns rhs (and the window blocks kept in self.Kwin / used)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        rhs = np.zeros(4 * self.nn)
        if want_jac:
            self.Kwin[:] = 0.0
            self.used[:] = 0
        rc = lib().op_assemble_nh3d(self.n, _p(self.idx), _p(self.W), _p(self.G), _p(self.Wa), _p(self.support),
                                    _p(self.x), _p(self.F_old), _p(self.old), _p(u), self.nx1, self.ny1, self.nz1,
                                    ctypes.byref(self.q), _p(self.freem), _p(rhs),
                                    _p(self.Kwin) if want_jac else None, _p(self.used) if want_jac else None,
                                    self.n_threads)
        if rc != 0:
            raise RuntimeError('cpu_port: a particle reaches a node outside the grid')
        return rhs

    def matrix(self):
        """Window blocks of the last assemble() as the oracle's dof-major CSR matrix."""
        nn = self.nn
        I, slot = np.nonzero(self.used.reshape(nn, 125))
        dx, dy, dz = slot % 5 - 2, (slot // 5) % 5 - 2, slot // 25 - 2
        J = I + dx + dy * self.nx1 + dz * self.nx1 * self.ny1
        blocks = self.Kwin.reshape(nn, 125, 4, 4)[I, slot]
        r = (I[:, None, None] + nn * np.arange(4)[None, :, None]).repeat(4, axis=2)
        c = (J[:, None, None] + nn * np.arange(4)[None, None, :]).repeat(4, axis=1)
        r, c, v = r.ravel(), c.ravel(), blocks.ravel()
        keep = self.free[r] & self.free[c]
        fix = np.nonzero(~self.free)[0]
        r = np.concatenate([r[keep], fix])
        c = np.concatenate([c[keep], fix])
        v = np.concatenate([v[keep], np.ones(fix.size)])
        return sp.coo_matrix((v, (r, c)), shape=(4 * nn, 4 * nn)).tocsr()

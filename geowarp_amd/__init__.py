"""geowarp_amd: MI355X-native implicit cpGIMP MPM load-step Newton loop.

Drop-in for the solver classes and kernel entry points of GeoWarp
(solvers/quasi_static_*_solver_*.py, mpm/gimp.py, mpm/mpm_utils.py,
materials/plasticity.py, solvers/sparse_differentiation.py).  Numerical work
runs in libgeowarp_amd.so (hand-written HIP for gfx950) behind a C ABI.
"""
from ._lib import GMError, LIB_PATH  # noqa: F401

__version__ = '0.1.0'

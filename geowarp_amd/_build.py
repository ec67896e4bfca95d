"""Build recipe for libgeowarp_amd.so (hipcc, gfx950).

The shared library is built in-tree next to this file so that it travels with
the source tree.  ``build()`` always recompiles.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libgeowarp_amd.so')
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith('.hip'))  # every HIP source of csrc/
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result']


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + FLAGS + ['-I', CSRC, '-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))

"""Write the oracle golden fixtures used by the GPU parity tests.

Run from the repository root:  python tests/golden/gen_fixtures.py [names...]
Each fixture holds, for a reduced workload, the oracle's state after every
load step (positions, Cauchy stress, F, b_e, pressure, hardening variable) and
the oracle's coloured-FD Jacobian (-dR/du, reference masking) and residual at
the first Newton iterate of the first load step.  Workloads with a 'dts' list
set the solver's time step before each load step; KEEP limits a long fixture
to some steps and fields.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, '..', '..')))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, '..')))

import oracle.solvers as osv  # noqa: E402
import harness as H  # noqa: E402

STEPS = {'bar_nh': 3, 'bar_hencky': 2, 'bar_j2': 3, 'bar_dp': 3, 'bar_j2_platen': 4, 'terzaghi': 2, 'indent_nh': 2,
         'indent_dp': 2, 'indent_dp_cone': 2, 'cantilever': 3, 'c1_bar_nh': 20, 'terzaghi_log': 8, 'bar_norsand': 3}
# long fixtures keep a subset: (steps, fields)
KEEP = {'c1_bar_nh': ((0, 4, 9, 14, 19), ('x', 'sigma'))}


def gen(name):
    t0 = time.time()
    cfg = H.MINI[name]()
    total = cfg['total_steps']
    dts = cfg.get('dts')
    # Jacobian at the first iterate of step 0 (separate instance: prepare_step mutates state)
    sj = H.make(osv, cfg)
    if dts is not None:
        sj.dt = dts[0]
    K, r, u0 = sj.first_jacobian(0, total)
    K = K.tocoo()
    s = H.make(osv, cfg)
    out = dict(jac_row=K.row.astype(np.int32), jac_col=K.col.astype(np.int32), jac_val=K.data, jac_res=r,
               n_steps=np.array([STEPS[name]]), total_steps=np.array([total]))
    keep_steps, keep_fields = KEEP.get(name, (None, None))
    for k in range(STEPS[name]):
        if dts is not None:
            s.dt = dts[k]
        s.advance_one_step(k, total)
        out['s%d_newton' % k] = np.array(s.residual_history)
        if keep_steps is not None and k not in keep_steps:
            continue
        f = H.fields(s)
        for key, v in f.items():
            if keep_fields is None or key in keep_fields:
                out['s%d_%s' % (k, key)] = v
    path = os.path.join(HERE, name + '.npz')
    np.savez_compressed(path, **out)
    print('%-14s %6.1fs  %s  (%d KB)' % (name, time.time() - t0, path, os.path.getsize(path) // 1024))


if __name__ == '__main__':
    names = sys.argv[1:] or list(STEPS)
    for n in names:
        gen(n)

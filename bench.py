"""Benchmark: implicit cpGIMP MPM load steps on one MI355X.

Workload (default): 3D u-p indentation, soil block 1 x 1 x 0.5 m on 64 x 64 x 32
cells (dx = 1/64), 2x2x2 particles per cell = 1,048,576 particles, 4 dofs per
node, E 20 MPa, nu 0.3, k/mu_f 1e-6, no gravity, rigid 0.2 x 0.2 m corner
indenter pushed 0.05 m in 40 load steps, starting at the top particle row, with
the textbook penalty sign.  Material: Neo-Hookean (--material Drucker-Prager
selects phi 30, c 5 kPa, psi 0: on this stress-free block the indenter drives
about a fifth of the particles past the tension capacity and the Newton
iterations stall, so that run reports its failed load step and no value).
A "step" is one load step converged to 1e-8 relative Newton residual (P2G +
consistent-mass projection + Newton: particle stage, residual, Jacobian,
multigrid-preconditioned flexible GMRES, backtracking update + G2P; a load step
whose Newton iteration fails is halved up to --substeps times).  --warmup load
steps run untimed, then --steps load steps are timed (defaults 5 and 20: load
steps 5-24).

The timed form is the mixed-precision solver (FP32 Jacobian and Krylov, FP16
smoother copies, FP64 residual); the plain FP64 form runs the same load steps
from the same state (as many as the --budget allows) and both are reported,
with the mixed/FP64 time ratio over the compared steps and the field errors at
the last of them.  value = Newton iterations per second of the mixed form over
the timed load steps; every timed step must converge, or value is null.  The
mixed form solves the last correction of a load step (the one applied after the
residual has passed the Newton test) on the previous iteration's Jacobian
(chord_iters counts them; they stay Newton iterations of the count, each
evaluating the residual, solving and updating as the others do).

roofline: the dominant kernel of the timed form, the per-bin phase of the
Jacobian assembly (k_jac_bin: the 27 x 27 block products of every particle bin
on the matrix cores), timed live with HIP events on its launch stream over
repeated launches on the last tangent of the timed region; achieved =
algorithmic flops per launch (2 x 60 per particle and block pair x 729 pairs)
over that time, against the dense FP32 matrix peak; traffic is the counted HBM
bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
(profiles/), or null.  roofline_smoother: the level-0 FP16 multigrid smoother
sweep (k_level), and roofline_krylov_spmv: the FP32 block-CSR SpMV of the
Krylov operator, likewise against the HBM peak with their algorithmic bytes.
cpu_baseline: one Newton iteration of the same workload's first load step on its
16 x 16 x 8-cell sub-block, on the host: the C++/OpenMP FP64 restatement of the
assembly (oracle/cpu_port.cpp: residual + exact Jacobian, every core this
process may use) and a SciPy SuperLU solve, scaled linearly in particle count.

--extras adds: c2_j2_bar, the 2D J2 bar of 256 x 1024 cells (FP64 and mixed);
c3_consolidation, the 2D consolidation column against Terzaghi's series;
adjoint, the batched 64-specimen inverse-analysis gradient with a central
finite-difference check of one component.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'Newton steps/s + adjoint grads/s, mixed vs FP64 at 1e-8 rel. residual, 1 GPU'
HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFS = 157.3  # MI355X dense FP32 matrix
MFMA_FP64_PEAK_TFS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20, help='timed load steps')
    ap.add_argument('--warmup', type=int, default=5, help='untimed load steps before the timed ones')
    ap.add_argument('--cells', type=str, default='64,64,32')
    ap.add_argument('--material', type=str, default='Neo-Hookean')
    ap.add_argument('--dp-apex', type=str, default='return',
                    help="Drucker-Prager apex rule: 'return' (apex return) or 'reference' (no return for P_trial >= 0)")
    ap.add_argument('--shape-factor', type=float, default=0.1,
                    help='Drucker-Prager hyperbolic shape factor (0: the cone with an apex)')
    ap.add_argument('--footing-start', type=str, default='surface',
                    help="'surface': the indenter starts at the top particle row (every step pushes one "
                         "increment); 'reference': end_z - dz/PPD as the reference driver")
    ap.add_argument('--penalty-form', type=str, default='physical',
                    help="indenter penalty: 'physical' (pushes penetrating soil down) or 'reference' (the "
                         "reference's sign, a negative stiffness under the indenter)")
    ap.add_argument('--substeps', type=int, default=4, help='levels of load-step halving on Newton failure')
    ap.add_argument('--gravity', type=float, default=0.0,
                    help='gravity magnitude (m/s^2; the reference driver and the default: 0, a stress-free block)')
    ap.add_argument('--forcing-max', type=float, default=None,
                    help='cap of the mixed form\'s Eisenstat-Walker inner tolerance (the solver\'s default: '
                         '1e-3 for elastic materials, 1e-4 for plastic ones)')
    ap.add_argument('--mg-lag', type=int, default=None,
                    help='0: rebuild the coarse multigrid operators for every Newton correction; 1: keep those of '
                         'a load step\'s first correction; 2: keep them across load steps too (default: the '
                         'solver\'s, 2 for the mixed form and 1 for FP64)')
    ap.add_argument('--budget', type=float, default=540.0,
                    help='wall-clock seconds for the whole run; the optional blocks are skipped past it')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fp64', action='store_true')
    ap.add_argument('--verbose', action='store_true')
    ap.add_argument('--extras', action='store_true',
                    help='also run the 2D J2 bar, the consolidation column and the batched adjoint blocks')
    ap.add_argument('--adjoint-steps', type=int, default=30,
                    help='load steps of the batched-specimen adjoint measurement (with --extras; 0 = skip)')
    ap.add_argument('--krylov', type=str, default='mg', help="'mg' (multigrid GMRES) or 'gmres' (block Jacobi)")
    ap.add_argument('--c2-steps', type=int, default=2, help='load steps of the 2D J2 bar (with --extras; 0 = skip)')
    ap.add_argument('--c3-steps', type=int, default=101,
                    help='load steps of the 2D consolidation column (with --extras; 0 = skip)')
    return ap.parse_args()


def build_solver(cells, precision, args, tol=1e-8):
    import geowarp_amd.solvers as gs
    from geowarp_amd import problems as P
    dx = 1.0 / 64
    cfg = P.indent3d(n_cells=cells, dx=dx, total_steps=40, depth=0.05, footing=0.2 * cells[0] / 64.0,
                     material=args.material, precision=precision, tol=tol, n_iter=12, krylov=args.krylov,
                     dp_apex=args.dp_apex, shape_factor=args.shape_factor, footing_start=args.footing_start,
                     penalty_form=args.penalty_form, gravity=args.gravity)
    cfg['iteration_dict']['substeps'] = args.substeps
    if args.mg_lag is not None:
        cfg['iteration_dict']['mg_lag'] = int(args.mg_lag)
    if args.forcing_max is not None:
        cfg['iteration_dict']['forcing_max'] = args.forcing_max
    s = P.build_solver(gs, cfg, device='cuda:0')
    cfg['particle_flags'](s.x_particles, s.particle_traction_flag_array, s.particle_pressure_boundary_flag_array)
    s.line_search = True
    s.linear_maxiter = 1500
    return s, cfg


def run_steps(s, cfg, first, count, log=None):
    """Advance load steps first .. first+count-1; stop at the first step whose
    Newton iteration (with its sub-steps) does not converge.  Returns per-step
    records."""
    import torch
    recs = []
    s.engine.stats['krylov_iters'].clear()
    for k in range(first, first + count):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n_kry = len(s.engine.stats['krylov_iters'])
        n_chord = s.engine.stats.get('chord_steps', 0)
        s.advance_one_step(k, cfg['total_steps'])
        torch.cuda.synchronize()
        h = s.residual_history
        rec = dict(step=k, seconds=round(time.perf_counter() - t0, 4), newton_iters=s.step_newton_iterations,
                   substeps=s.substep_count, krylov_iters=int(sum(s.engine.stats['krylov_iters'][n_kry:])),
                   chord_iters=s.engine.stats.get('chord_steps', 0) - n_chord,
                   converged=bool(s.newton_converged),
                   relative_residual=(h[-1] / h[0]) if h and h[0] > 0 else None,
                   indentation_force=float(s.indentation_force.item()))
        recs.append(rec)
        if log:
            print(log, json.dumps(rec), file=sys.stderr, flush=True)
        if not rec['converged']:
            break
    return recs


def summarize(recs, count):
    ok = len(recs) == count and all(r['converged'] for r in recs)
    sec = sum(r['seconds'] for r in recs)
    it = sum(r['newton_iters'] for r in recs)
    out = dict(seconds=round(sec, 4), load_steps=len(recs), all_converged=ok, newton_iters=it,
               krylov_iters=sum(r['krylov_iters'] for r in recs), substeps=sum(r['substeps'] for r in recs),
               chord_iters=sum(r['chord_iters'] for r in recs),
               newton_per_s=round(it / sec, 4) if sec > 0 else None,
               ms_per_load_step=round(1e3 * sec / max(len(recs), 1), 3))
    if not ok and recs:
        bad = recs[-1]
        out['failed_step'] = dict(step=bad['step'], relative_residual=bad['relative_residual'])
    return out


def spmv_roofline(s, reps=50):
    import torch
    from geowarp_amd import _lib as L
    eng = s.engine
    fp32 = eng.precision == 'mixed'
    sz = 4 if fp32 else 8
    n = eng.n_nodes * eng.nb
    dt = torch.float32 if fp32 else torch.float64
    x = torch.randn(n, dtype=dt, device=eng.dev)
    y = torch.zeros(n, dtype=dt, device=eng.dev)
    stream = L.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = L.lib()

    def launch():
        L.check(lib.gm_bsr_spmv(eng.nb, int(fp32), int(fp32), eng.n_nodes, L.ptr(eng.row_ptr), L.ptr(eng.col_idx),
                                L.ptr(eng.vals), L.ptr(x), L.ptr(y), stream), 'gm_bsr_spmv')

    for _ in range(5):
        launch()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    avg_ms = e0.elapsed_time(e1) / reps
    nb = eng.nb
    bytes_ = eng.nnzb * (nb * nb * sz + 4) + 4 * (eng.n_nodes + 1) + 2 * n * sz
    achieved = bytes_ / (avg_ms * 1e-3) / 1e9
    # reference stream for counter calibration: y += a x over 32 Mi doubles
    # (algorithmic traffic 512 MiB read + 256 MiB written) in the k_axpy_d kernel
    cx = torch.ones(1 << 25, dtype=torch.float64, device=eng.dev)
    cy = torch.zeros(1 << 25, dtype=torch.float64, device=eng.dev)
    L.check(lib.gm_axpy(cx.numel(), 1.0, L.ptr(cx), 0, L.ptr(cy), stream), 'gm_axpy')
    torch.cuda.synchronize()
    lanes = 16 if nb >= 4 else (8 if nb == 3 else 4)
    return dict(kernel='k_spmv<%d,%d,%s>' % (nb, lanes, 'float' if fp32 else 'double'), avg_us=avg_ms * 1e3,
                bytes_per_launch=bytes_, nnzb=eng.nnzb, achieved_GBs=achieved)


def smoother_roofline(s, reps=50):
    """Time the level-0 multigrid smoother sweep (k_level, Jacobi mode: out = x + omega
    Dinv (b - A x), from the FP16 copy in the mixed form) on the hierarchy of the last
    Jacobian, with HIP events on the launch stream."""
    import ctypes
    import torch
    from geowarp_amd import _lib as L
    eng = s.engine
    fp32 = eng.precision == 'mixed'
    mg = eng.multigrid(fp32)[0]
    lv = mg.levels[0]
    nb = mg.nb
    n = eng.n_nodes * nb
    dt = torch.float32 if fp32 else torch.float64
    sz = 4 if fp32 else 8
    x = torch.randn(n, dtype=dt, device=eng.dev)
    b = torch.randn(n, dtype=dt, device=eng.dev)
    out = torch.zeros(n, dtype=dt, device=eng.dev)
    stream = L.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = L.lib()
    addr = ctypes.addressof(lv)

    def launch():
        L.check(lib.gm_mg_level_op(nb, int(fp32), addr, 0, 2, L.ptr(x), L.ptr(b), mg.omega, L.ptr(out), stream),
                'gm_mg_level_op')

    for _ in range(5):
        launch()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    avg_ms = e0.elapsed_time(e1) / reps
    lo = bool(lv.vals_lo)
    nnzb = eng.nnzb
    mat = nnzb * (nb * nb * (2 if lo else sz) + 4) + 4 * (eng.n_nodes + 1) + (eng.n_nodes * nb * nb * 4 if lo else 0)
    bytes_ = mat + eng.n_nodes * nb * nb * sz + 3 * n * sz  # + dinv, x, b, out
    return dict(kernel='k_level<%d,16,Jacobi sweep,%s>' % (nb, 'FP16 copy' if lo else ('float' if fp32 else 'double')),
                avg_us=avg_ms * 1e3, bytes_per_launch=bytes_, achieved_GBs=bytes_ / (avg_ms * 1e-3) / 1e9)


def jacobian_roofline(s, reps=20):
    """Time the per-bin phase of the binned Jacobian assembly (k_jac_bin through
    gm_jacobian_bins: every bin's 27 x 27 block products on the matrix cores) on the
    last tangent of the timed region, with HIP events on the launch stream.
    Algorithmic flops per launch: 2 x 60 per (particle, block pair) -- 4 rows x 15
    non-zero (column dof, shape component) terms -- x 729 block pairs per particle."""
    import torch
    from geowarp_amd import _lib as L
    eng = s.engine
    fp32 = eng._tangent_fp32()
    lib = L.lib()
    g, b = L.ctypes.byref(eng.grid), L.ctypes.byref(eng.bins)
    need = int(lib.gm_node_jacobian_work_bytes(g, int(fp32)))
    w = eng._bins_workspace(need)
    stream = L.c_void_p(torch.cuda.current_stream().cuda_stream)

    def launch():
        L.check(lib.gm_jacobian_bins(g, b, L.ptr(eng.D), int(fp32), L.ptr(w), need, stream), 'gm_jacobian_bins')

    mixed = eng.precision == 'mixed'
    tgt, v32 = (eng.vals64, False) if (mixed and not fp32) else (eng.vals, mixed)

    def assemble():
        eng._assemble(fp32, tgt, v32)

    def avg(fn):
        for _ in range(3):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    avg_ms = avg(launch)
    # the whole assembly (per-bin products + gather into the matrix), same stream
    asm_ms = avg(assemble)
    tune = eng.stats.get('jacobian_tune') or {}
    n_p = int(s.n_particles)
    flops = 2.0 * 60 * 729 * n_p
    nonempty = int((eng.cell_start[1:] != eng.cell_start[:-1]).sum().item())
    sz = 4 if fp32 else 8
    bytes_ = n_p * (17 * 15 * sz + 48) + nonempty * 27 * 27 * 16 * sz
    return dict(kernel='k_jac_bin<%s> (gm_jacobian_bins, variant %s)' % ('float' if fp32 else 'double',
                                                                        tune.get('bins_variant', 0)),
                avg_us=avg_ms * 1e3, assembly_ms=asm_ms, tune=tune,
                flops_per_launch=flops, achieved_TFs=flops / (avg_ms * 1e-3) / 1e12, bytes_per_launch=bytes_,
                achieved_GBs=bytes_ / (avg_ms * 1e-3) / 1e9, nonempty_bins=nonempty)


def adjoint_bench(steps, total=30, every=5, n_specimens=64, krylov='mg'):
    """Gradient of a least-squares reaction-force misfit with respect to every
    specimen's friction angle and cohesion (2 x 64 components) through the
    hand-written adjoint: synthetic force curves from the FP64 forward at the
    true parameters (untimed), then one timed forward + adjoint sweep at a trial
    parameter set."""
    import torch
    import geowarp_amd.solvers as gs
    from geowarp_amd import problems as P
    from geowarp_amd.adjoint import LoadPathAdjoint, LoadPathFunction
    W, phi, coh = P.specimen_counts(n_specimens, seed=0)
    # Newton to 1e-12 relative residual: the finite-difference check below resolves loss changes of
    # ~1e-7 relative, which needs forces converged well beyond that
    cfg = P.specimens2d(W, phi, coh, total_steps=total, strain=0.03, tol=1e-12, n_iter=14)
    cfg['iteration_dict']['krylov'] = krylov
    s = P.build_solver(gs, cfg, device='cuda:0')
    cfg['set_boundary_dofs'](s.dofStruct.boundary_flag_array, s.n_grid_x, s.n_nodes)
    dev = s.device
    g = torch.as_tensor(cfg['groups'], device=dev)
    adj = LoadPathAdjoint(s, total, n_steps=steps, groups=g, checkpoint_every=every)
    with torch.no_grad():
        f_obs = LoadPathFunction.apply(torch.tensor(phi, device=dev), torch.tensor(coh, device=dev), adj)
    phi_t = torch.tensor(phi + 2.0, device=dev, requires_grad=True)
    coh_t = torch.tensor(coh + 3.0, device=dev, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f = LoadPathFunction.apply(phi_t, coh_t, adj)
    loss = ((f - f_obs) ** 2).sum()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    unconverged = list(adj.stats['unconverged_steps'])
    kits = adj.stats['krylov_iters']
    # central finite difference (h = 1e-6 deg) of one component: the largest specimen's friction angle
    i = int(np.argmax(W))
    h = 1e-6

    def loss_at(dphi):
        p = phi + 2.0
        p[i] += dphi
        with torch.no_grad():
            ff = LoadPathFunction.apply(torch.tensor(p, device=dev), torch.tensor(coh + 3.0, device=dev), adj)
        return float(((ff - f_obs) ** 2).sum())

    fd = (loss_at(h) - loss_at(-h)) / (2 * h)
    ad = float(phi_t.grad[i])
    return dict(workload='%d 2D Drucker-Prager specimens (seed 0, %d particles, %d-%d cells wide), load steps '
                         '0-%d of %d, checkpoint every %d, loss = sum of squared reaction-force misfits'
                         % (n_specimens, s.n_particles, int(W.min()), int(W.max()), steps - 1, total, every),
                seconds_forward=round(t1 - t0, 3), seconds_adjoint=round(t2 - t1, 3),
                grads_per_s=round(1.0 / (t2 - t0), 4), gradient_components=2 * n_specimens,
                forward_steps=adj.stats['forward_steps'], replayed_steps=adj.stats['replayed_steps'],
                adjoint_steps=adj.stats['adjoint_steps'],
                transposed_solve_iters=dict(min=min(kits), max=max(kits), mean=round(sum(kits) / len(kits), 1)),
                unconverged_forward_steps=unconverged,
                grad_norm=dict(phi=float(phi_t.grad.norm()), coh=float(coh_t.grad.norm())),
                fd_check=dict(component='dloss/dphi of specimen %d (%d cells wide)' % (i, int(W[i])), h=h,
                              adjoint=ad, central_fd=fd, rel_err=abs(ad - fd) / max(abs(fd), 1e-300)))


def c2_bench(steps):
    """2D plane-strain J2 bar (256 x 1024 cells of 1.953125 mm, 1,048,576 particles,
    E 10 MPa, sigma_y 20 kPa, H 100 kPa, top platen 10 % of the height in 50
    steps): load steps 0..steps-1 in the FP64 and the mixed form from the same
    initial state; times, Newton iterations and field errors of the mixed form."""
    import torch
    import geowarp_amd.solvers as gs
    from geowarp_amd import problems as P
    out, solvers = {}, {}
    for form in ('fp64', 'mixed'):
        cfg = P.bar_j2(precision=form)
        s = P.build_solver(gs, cfg, device='cuda:0')
        cfg['set_boundary_dofs'](s.dofStruct.boundary_flag_array, s.n_grid_x, s.n_nodes)
        s.line_search = True
        from geowarp_amd._lib import GMError
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters = 0
        err = None
        try:
            for k in range(steps):
                s.advance_one_step(k, cfg['total_steps'])
                iters += s.newton_iterations
        except GMError as e:  # a diverged form is reported, not fatal to the benchmark
            err = str(e)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        hist = s.residual_history
        conv = err is None and bool(hist and hist[0] > 0 and hist[-1] / hist[0] < s.tol)
        out[form] = dict(seconds=round(dt, 4), newton_iters=iters, newton_per_s=round(iters / dt, 4),
                         converged=conv, last_residuals=['%.2e' % v for v in hist])
        if err is not None:
            out[form]['error'] = err
        solvers[form] = s
    if 'error' in out['fp64'] or 'error' in out['mixed']:
        out['workload'] = '2D J2 bar 256x1024 cells, load steps 0-%d' % (steps - 1)
        return out
    a, b = solvers['mixed'], solvers['fp64']
    du = float(torch.linalg.norm((a.x_particles - a.x0) - (b.x_particles - b.x0)) / torch.linalg.norm(b.x_particles - b.x0))
    ds = float(torch.linalg.norm(a.particle_Cauchy_stress_array - b.particle_Cauchy_stress_array)
               / torch.linalg.norm(b.particle_Cauchy_stress_array))
    out['mixed_vs_fp64_speedup'] = round(out['fp64']['seconds'] / out['mixed']['seconds'], 3)
    out['mixed_vs_fp64_field_rel_err'] = dict(displacement=du, stress=ds)
    out['workload'] = ('2D J2 bar 256x1024 cells, %d particles, platen 10%% of the height in 50 steps, load steps '
                       '0-%d to 1e-8 relative residual' % (b.n_particles, steps - 1))
    return out


def _terzaghi_series(z, H, Tv, n_terms=400):
    """Terzaghi's excess pore pressure u/u0 at depth z below a drained top, impermeable
    base at depth H: sum_m 2/M sin(M z/H) exp(-M^2 Tv), M = pi (2m+1)/2."""
    M = np.pi * (2 * np.arange(n_terms) + 1) / 2
    return (2 / M[None, :] * np.sin(M[None, :] * z[:, None] / H) * np.exp(-M[None, :] ** 2 * Tv)).sum(1)


def c3_bench(steps):
    """2D u-p consolidation column (Terzaghi), 32 x 320 cells of 1/32 m, 40,960
    particles, Neo-Hookean skeleton E 10 MPa, nu 0.3, mobility 1e-4, drained top,
    impermeable base: a 10 kPa surface load in load step 0 (dt 1e-3 s, undrained),
    then `steps - 1` consolidation steps at log-spaced times up to t = 150 s
    (time factor T_v ~ 2), Newton to 1e-8 relative residual or a 1e-11 absolute floor
    (twice the residual's round-off level on this column).  FP64 and mixed forms
    from the same initial state: times, Newton iterations, field errors of the mixed
    form; and the FP64 pore pressure against Terzaghi's series solution scaled by the
    undrained pressure of load step 0, at T_v ~ 0.05, 0.2, 0.5 and 1."""
    import torch
    import geowarp_amd.solvers as gs
    from geowarp_amd import problems as P
    t_end, t1 = 150.0, 0.05
    tk = np.concatenate([[1e-3], 1e-3 + np.logspace(np.log10(t1), np.log10(t_end), steps - 1)])
    dts = np.diff(np.concatenate([[0.0], tk]))
    H, E, nu, mob = 10.0, 1.0e4, 0.3, 1e-4
    cv = mob * E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))  # consolidation coefficient, m^2/s
    Tv = cv * (tk - tk[0]) / H ** 2
    targets = [0.05, 0.2, 0.5, 1.0]
    check_steps = {int(np.argmin(np.abs(Tv - t))): t for t in targets if Tv[-1] >= t}
    out, solvers, p0, series = {}, {}, None, []
    # stopping: 1e-8 relative, or below an absolute floor of 1e-11 -- twice the round-off floor
    # of this column's residual (about 5e-12), which the relative test cannot pass when the
    # step's first residual is itself small (about 4e-4 in the later consolidation steps)
    abs_floor = 1e-11
    for form in ('fp64', 'mixed'):
        cfg = P.terzaghi(precision=form, total_steps=steps, tol=1e-8)
        cfg['iteration_dict']['abs_floor'] = abs_floor
        s = P.build_solver(gs, cfg, device='cuda:0')
        cfg['particle_flags'](s.x_particles, s.particle_traction_flag_array, s.particle_pressure_boundary_flag_array)
        z = H - s.x0[:, 1].cpu().numpy()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters, unconverged = 0, 0
        for k in range(steps):
            s.dt = float(dts[k])
            s.advance_one_step(k, steps)
            iters += s.newton_iterations
            unconverged += int(not s.newton_converged)
            if form == 'fp64':
                if k == 0:
                    p0 = float(s.particle_pressure_array.mean())
                if k in check_steps:
                    pn = s.particle_pressure_array.cpu().numpy()
                    ua = _terzaghi_series(z, H, Tv[k]) * p0
                    series.append(dict(step=k, time_factor=round(float(Tv[k]), 4),
                                       rel_l2=float(np.linalg.norm(pn - ua) / np.linalg.norm(ua)),
                                       max_abs_over_p0=float(np.abs(pn - ua).max() / p0),
                                       base_kPa=dict(mpm=float(pn[np.argmax(z)]), series=float(ua[np.argmax(z)]))))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[form] = dict(seconds=round(dt, 3), newton_iters=iters, unconverged_steps=unconverged)
        solvers[form] = s
    a, b = solvers['mixed'], solvers['fp64']
    out['mixed_vs_fp64_speedup'] = round(out['fp64']['seconds'] / out['mixed']['seconds'], 3)
    out['mixed_vs_fp64_field_rel_err'] = field_errors(fields(a), fields(b))
    out['fp64_vs_terzaghi'] = dict(undrained_pressure_kPa=p0, consolidation_coefficient=cv, checks=series)
    out['workload'] = ('2D u-p consolidation column 32x320 cells, %d particles, %d load steps (log-spaced times to '
                       't = %g s) to 1e-8 relative residual or %g absolute' % (b.n_particles, steps, t_end, abs_floor))
    return out


def measured_traffic(cells, material, kernel='k_spmv'):
    """HBM bytes per launch of a kernel (k_spmv: the FP32 SpMV; k_level: the FP16
    smoother sweep; k_jac_bin: the per-bin Jacobian products) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/<kernel>_c4_indentation*),
    scaled by the calibration kernel; None for other workloads."""
    path = os.path.join(ROOT, 'profiles', '%s_c4_indentation_traffic.json' % kernel)
    if tuple(cells) != (64, 64, 32) or material != 'Neo-Hookean' or not os.path.isfile(path):
        return None
    with open(path) as f:
        return json.load(f)['traffic_bytes_per_launch']


def fields(s):
    """Comparable end-of-step fields of a solver (device copies)."""
    return dict(u=(s.x_particles - s.x0).clone(), sigma=s.particle_Cauchy_stress_array.clone(),
                p=s.particle_pressure_array.clone())


def field_errors(fa, fb):
    """Relative L2 differences of fields fa against fb (dicts from fields())."""
    import torch

    def rel(x, y):
        ny = torch.linalg.norm(y)
        return float(torch.linalg.norm(x - y) / ny) if ny > 0 else float(torch.linalg.norm(x))
    return dict(displacement=rel(fa['u'], fb['u']), stress=rel(fa['sigma'], fb['sigma']),
                pore_pressure=rel(fa['p'], fb['p']))


def cpu_baseline(args):
    """One Newton iteration of the timed workload's first load step on its 16 x 16 x 8-cell
    corner sub-block (same material, indenter and penalty), on the host: the C++/OpenMP
    restatement of the assembly (residual + exact Jacobian, oracle/cpu_port.cpp) on every
    core this process may use, then a sparse direct solve (SciPy SuperLU, the reference's
    host solver family), repeated until 10 s of work (at most 30 times), the median taken;
    scaled linearly in particle count."""
    if args.material != 'Neo-Hookean':
        return None
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import harness as H
    import oracle.solvers as osv
    import oracle.cpu_port as cp
    from scipy.sparse.linalg import spsolve
    from geowarp_amd import problems as P
    cells = (16, 16, 8)
    cfg = P.indent3d(n_cells=cells, dx=1.0 / 64, total_steps=40, depth=0.05, footing=0.2 * cells[0] / 64.0,
                     material=args.material, tol=1e-8, footing_start=args.footing_start,
                     penalty_form=args.penalty_form)
    s = H.make(osv, cfg)
    st, f, u0, _ = s.prepare_step(0, cfg['total_steps'])
    A = cp.Assembler(s, st, 0, cfg['total_steps'])
    ts, ta = [], []
    while len(ts) < 3 or (sum(ts) < 10.0 and len(ts) < 30):
        t0 = time.perf_counter()
        r = A.assemble(u0)
        K = A.matrix()
        t1 = time.perf_counter()
        spsolve(K.tocsc(), r)
        ts.append(time.perf_counter() - t0)
        ta.append(t1 - t0)
    t = sorted(ts)[len(ts) // 2]
    n_full = 64 * 64 * 32 * 8
    t_full = t * n_full / s.n_particles
    model = ''
    try:
        with open('/proc/cpuinfo') as fh:
            model = next((ln.split(':', 1)[1].strip() for ln in fh if ln.startswith('model name')), '')
    except OSError:
        pass
    return dict(value=round(1.0 / t_full, 6), unit='Newton steps/s', cores=A.n_threads, kind='port',
                cpu=model,
                sample='C++/OpenMP FP64 restatement (oracle/cpu_port.cpp, %d threads) of one Newton iteration: '
                       'residual + exact Jacobian (%.3f s) + SciPy SuperLU solve (1 thread), median of %d '
                       '(%.1f s of work) = %.3f s on the 16x16x8-cell sub-block (%d particles, %d dofs) of the '
                       'same workload, scaled linearly to %d particles (the direct solve grows faster than '
                       'linearly, so the scaled figure flatters the CPU)'
                       % (A.n_threads, sorted(ta)[len(ta) // 2], len(ts), sum(ts), t, s.n_particles, 4 * A.nn,
                          n_full))


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the path is single-device (multi-GPU sharding is a design note, '
              'not built); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    t_start = time.perf_counter()
    import torch
    cells = tuple(int(c) for c in args.cells.split(','))
    torch.cuda.set_device(0)
    log = 'bench' if args.verbose else None

    # warm-up load steps 0 .. W-1 in the mixed form; both forms start the timed steps from its state
    smx, cfg = build_solver(cells, 'mixed', args)
    warm = run_steps(smx, cfg, 0, args.warmup, log=log and 'warmup')
    out = {'metric': METRIC, 'unit': 'Newton steps/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f32+f64 (FP32 Jacobian/Krylov/FP16 smoother copies, FP64 residual and refinement)',
           'data': 'synthetic'}
    shape = ('shape factor %g' % args.shape_factor) if args.material == 'Drucker-Prager' else ''
    out['config'] = {'workload': '3D u-p indentation %dx%dx%d cells, %d particles, %s%s, indenter 0.2 x 0.2 m pushed '
                                 '0.05 m in 40 load steps (start: %s), load steps %d-%d timed to 1e-8 relative '
                                 'residual' % (cells + (smx.n_particles, args.material,
                                                        (' phi 30 c 5 psi 0 apex %s, %s' % (args.dp_apex, shape))
                                                        if args.material == 'Drucker-Prager' else '',
                                                        args.footing_start, args.warmup,
                                                        args.warmup + args.steps - 1)),
                     'nodes': smx.n_nodes, 'dofs': smx.n_matrix_size, 'substep_levels': args.substeps}
    if args.gravity:
        out['config']['gravity'] = args.gravity
    out['warmup_steps'] = summarize(warm, args.warmup)
    if not out['warmup_steps']['all_converged']:
        out['value'] = None
        out['error'] = 'a warm-up load step did not converge (see warmup_steps.failed_step)'
        print(json.dumps(out))
        return
    snap = smx.state_dict()

    # timed: mixed form (the value), keeping end-of-step fields for the comparison
    from geowarp_amd.solvers._engine import PhaseTimer
    mixed_fields = {}
    smx.engine.timers = PhaseTimer()
    recs = []
    smx.engine.stats['krylov_iters'].clear()
    for k in range(args.warmup, args.warmup + args.steps):
        r = run_steps(smx, cfg, k, 1, log=log and 'mixed')
        recs += r
        mixed_fields[k] = fields(smx)
        if not r[-1]['converged']:
            break
    phases = smx.engine.timers.summary()
    smx.engine.timers = None
    res_mx = summarize(recs, args.steps)
    res_mx['phases_ms'] = {k: round(v, 2) for k, v in phases.items()}
    res_mx['per_step'] = recs
    out['mixed'] = res_mx
    out['config']['nnzb'] = smx.engine.nnzb
    out['value'] = res_mx['newton_per_s'] if res_mx['all_converged'] else None
    out['ms_per_step'] = res_mx['ms_per_load_step']
    if not res_mx['all_converged']:
        out['error'] = 'mixed form: load step %s did not converge' % res_mx.get('failed_step')

    # timed: FP64 form from the same state, as many of the same steps as the budget allows
    if not args.no_fp64:
        s64, _ = build_solver(cells, 'fp64', args)
        # one untimed FP64 load step first, so that the kernel tuners (Jacobian variants, SpMV
        # workgroup size) run outside the timed region as the mixed form's did in its warm-up
        s64.load_state_dict(snap)
        run_steps(s64, cfg, args.warmup, 1)
        s64.load_state_dict(snap)
        s64.engine.timers = PhaseTimer()
        recs = []
        for k in range(args.warmup, args.warmup + args.steps):
            elapsed = time.perf_counter() - t_start
            if recs:
                per = sum(x['seconds'] for x in recs) / len(recs)
                if elapsed + 1.5 * per > 0.8 * args.budget:
                    break
            r = run_steps(s64, cfg, k, 1, log=log and 'fp64')
            recs += r
            if not r[-1]['converged']:
                break
        phases = s64.engine.timers.summary()
        s64.engine.timers = None
        res64 = summarize(recs, len(recs))
        res64['phases_ms'] = {k: round(v, 2) for k, v in phases.items()}
        res64['per_step'] = recs
        res64['jacobian_tune'] = s64.engine.stats.get('jacobian_tune')
        out['fp64'] = res64
        n = len(recs)
        if n and res64['all_converged']:
            mx_sec = sum(x['seconds'] for x in res_mx['per_step'][:n])
            out['mixed_vs_fp64_speedup'] = round(res64['seconds'] / mx_sec, 3) if mx_sec > 0 else None
            out['mixed_vs_fp64_steps_compared'] = [args.warmup, args.warmup + n - 1]
            last = args.warmup + n - 1
            if last in mixed_fields:
                out['mixed_vs_fp64_field_rel_err'] = field_errors(mixed_fields[last], fields(s64))
        del s64
        torch.cuda.empty_cache()

    # roofline of the dominant kernel of the timed form (the per-bin Jacobian products), of the
    # level-0 smoother sweep and of the Krylov SpMV (last Jacobian of the timed region)
    jr = jacobian_roofline(smx)
    smo = smoother_roofline(smx)
    roof = spmv_roofline(smx)
    peak = MFMA_FP32_PEAK_TFS if smx.engine._tangent_fp32() else MFMA_FP64_PEAK_TFS
    out['roofline'] = {'bound': 'mfma', 'achieved': round(jr['achieved_TFs'], 2), 'peak': peak, 'unit': 'TFLOP/s',
                       'frac': round(jr['achieved_TFs'] / peak, 4),
                       'traffic': measured_traffic(cells, args.material, 'k_jac_bin'),
                       'kernel': jr['kernel'], 'avg_us': round(jr['avg_us'], 2),
                       'algorithmic_flops': jr['flops_per_launch'],
                       'bytes_per_launch': jr['bytes_per_launch'],
                       'hbm_frac_of_those_bytes': round(jr['achieved_GBs'] / HBM_PEAK_GBS, 4)}
    out['jacobian_assembly_ms'] = round(jr['assembly_ms'], 3) if jr['assembly_ms'] else None
    out['jacobian_tune'] = jr['tune']
    out['roofline_smoother'] = {'bound': 'hbm', 'achieved': round(smo['achieved_GBs'], 1), 'peak': HBM_PEAK_GBS,
                                'unit': 'GB/s', 'frac': round(smo['achieved_GBs'] / HBM_PEAK_GBS, 4),
                                'traffic': measured_traffic(cells, args.material, 'k_level'),
                                'kernel': smo['kernel'], 'avg_us': round(smo['avg_us'], 2),
                                'algorithmic_bytes': smo['bytes_per_launch']}
    out['roofline_krylov_spmv'] = {'bound': 'hbm', 'achieved': round(roof['achieved_GBs'], 1), 'peak': HBM_PEAK_GBS,
                                   'unit': 'GB/s', 'frac': round(roof['achieved_GBs'] / HBM_PEAK_GBS, 4),
                                   'traffic': measured_traffic(cells, args.material),
                                   'kernel': roof['kernel'], 'avg_us': round(roof['avg_us'], 2),
                                   'algorithmic_bytes': roof['bytes_per_launch']}
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(args)
    if args.extras:
        if args.c2_steps > 0 and time.perf_counter() - t_start < args.budget:
            out['c2_j2_bar'] = c2_bench(args.c2_steps)
        if args.c3_steps > 0 and time.perf_counter() - t_start < args.budget:
            out['c3_consolidation'] = c3_bench(args.c3_steps)
        if args.adjoint_steps > 0 and time.perf_counter() - t_start < args.budget:
            out['adjoint'] = adjoint_bench(args.adjoint_steps, krylov=args.krylov)
    out['wall_seconds'] = round(time.perf_counter() - t_start, 1)
    print(json.dumps(out))


if __name__ == '__main__':
    main()

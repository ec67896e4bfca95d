"""Workload set-ups: the benchmark configurations and reduced versions of them.

Each builder returns a dict with the solver class name, the constructor dicts,
the Dirichlet callback and a particle-flag set-up function.  Callbacks and flag
functions only use slicing, so they act in place on numpy arrays (oracle) and
torch tensors (device solvers) alike.  They re-express the reference drivers'
Warp kernels without Warp:
  bar compaction      problems/bar_compaction/bar_compaction_cpGIMP.py:93-107
  Terzaghi column     problems/consolidation/terzaghi_cpGIMP.py:104-149
  3D indentation      problems/3D_indentation/3d_indentation_coupled_cpGIMP.py:317-358

Configurations (the sizes in brackets are the full benchmark ones):
  bar_nh      2D Neo-Hookean bar compaction, 8 x 80 cells, dx 0.0125, 2x2 PPC
              (2560 particles), E 1 MPa, nu 0.3, rho 0.08 t/m^3, gravity ramped
              over 20 steps, roller walls, fixed base.
  bar_j2      2D J2 bar, 0.5 m x 2 m on 256 x 1024 cells of 1.953125 mm, 2x2 PPC
              (1,048,576 particles), E 10 MPa, nu 0.3, sigma_y 20 kPa, H 100 kPa,
              top-platen displacement 10% of the height in 50 steps (prescribed
              y increments on the node rows at and above the current top).
  terzaghi    2D u-p column 1 m x 10 m on 32 x 320 cells, 2x2 PPC (40,960
              particles), E 10 MPa, nu 0.3, k/mu_f 1e-4 m^2/(kPa s), phi0 0.5,
              10 kPa surface traction, drained top, undrained base.
  indent3d    3D u-p indentation, soil block 1 x 1 x 0.5 m on 64 x 64 x 32 cells
              (dx 1/64), 2x2x2 PPC (1,048,576 particles), Drucker-Prager phi 30,
              c 5 kPa, psi 0, E 20 MPa, nu 0.3, k/mu_f 1e-6, rigid 0.2 x 0.2 m
              indenter at the corner, 0.05 m in 40 steps, penalty factor 1000.
              Background grid padded by 1 cell below x/y and 3 cells above.
              The particle-domain box (lp clipping, gimp.py:489-505) ends at the
              initial soil surface, so heaving particle domains do not reach
              into the empty nodes above it (clip_top=False keeps the grid top,
              as the reference driver does).
"""
import numpy as np


def _grid2d(nx_cells, ny_cells, dx, pad_lo=1, pad_hi=2):
    nx = nx_cells + pad_lo + pad_hi
    ny = ny_cells + pad_lo + pad_hi
    sx = pad_lo * dx
    sy = pad_lo * dx
    ex = sx + nx_cells * dx
    ey = sy + ny_cells * dx
    return dict(n_grid_x=nx, n_grid_y=ny, max_x=nx * dx, max_y=ny * dx, dx=dx, dy=dx, start_x=sx, end_x=ex,
                start_y=sy, end_y=ey, domain_min_x=sx, domain_max_x=ex, domain_min_y=sy, domain_max_y=ey)


def _view2d(arr, k, n_nodes, nx1):
    return arr[k * n_nodes:(k + 1) * n_nodes].reshape(n_nodes // nx1, nx1)


def make_bar_bc(ix_lo, ix_hi, iy_lo):
    """Roller walls at node_ix <= ix_lo and >= ix_hi, fixed base at node_iy <= iy_lo
    (bar_compaction_cpGIMP.py:93-107 for a wider bar)."""

    def set_boundary_dofs(boundary, n_grid_x, n_nodes):
        nx1 = n_grid_x + 1
        bx = _view2d(boundary, 0, n_nodes, nx1)
        by = _view2d(boundary, 1, n_nodes, nx1)
        bx[:, :ix_lo + 1] = True
        bx[:, ix_hi:] = True
        by[:iy_lo + 1, :] = True

    return set_boundary_dofs


def bar_nh(n_cells_x=8, n_cells_y=80, dx=0.0125, total_steps=20, tol=1e-10):
    g = _grid2d(n_cells_x, n_cells_y, dx)
    PPD = 2
    npart = n_cells_x * n_cells_y * PPD * PPD
    return dict(
        solver='quasi_static_solver_2d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=npart, PPD=PPD, p_vol=(dx / PPD) ** 2, p_rho=0.08),
        iteration_dict=dict(n_iter=10, tol=tol, tol_mode='relative', solver_name='warp'),
        material_dict=dict(youngs_modulus=1000.0, poisson_ratio=0.3, material_name='Neo-Hookean'),
        load_dict=dict(gravity_mag=10.0, traction_value_x=0.0, traction_value_y=0.0, point_load_value_x=0.0,
                       point_load_value_y=0.0),
        set_boundary_dofs=make_bar_bc(1, 1 + n_cells_x, 1),
        total_steps=total_steps,
    )


def make_platen(total_disp):
    """Displacement-controlled rigid platen on top of a 2D bar: every load step,
    the node rows at and above the current top of the particle domains
    (max y + lp_y) are fixed in y with the increment total_disp / total_steps;
    x stays free (frictionless platen).  Driver hook for quasi_static_solver_2d's
    load_dict['prescribed_increment']."""

    def hook(x, lp, boundary, pres, n_grid_x, n_nodes, dx, current_step, total_steps):
        nx1 = n_grid_x + 1
        top = float((x[:, 1] + lp[:, 1]).max())
        row = int(np.ceil(top / dx - 1e-8))
        by = _view2d(boundary, 1, n_nodes, nx1)
        py = _view2d(pres, 1, n_nodes, nx1)
        by[row:, :] = True
        py[row:, :] = total_disp / float(total_steps)

    return hook


def bar_j2(n_cells_x=256, n_cells_y=1024, dx=1.953125e-3, total_steps=50, tol=1e-8, precision='fp64',
           platen_strain=0.1):
    """J2 bar compressed by a top platen by platen_strain of its height."""
    g = _grid2d(n_cells_x, n_cells_y, dx)
    PPD = 2
    npart = n_cells_x * n_cells_y * PPD * PPD
    return dict(
        solver='quasi_static_solver_2d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=npart, PPD=PPD, p_vol=(dx / PPD) ** 2, p_rho=2.0),
        iteration_dict=dict(n_iter=12, tol=tol, tol_mode='relative', precision=precision),
        material_dict=dict(youngs_modulus=1.0e4, poisson_ratio=0.3, material_name='J2',
                           kappa=np.sqrt(2.0 / 3.0) * 20.0, H=100.0),
        load_dict=dict(gravity_mag=10.0, traction_value_x=0.0, traction_value_y=0.0, point_load_value_x=0.0,
                       point_load_value_y=0.0,
                       prescribed_increment=make_platen(-platen_strain * n_cells_y * dx) if platen_strain else None),
        set_boundary_dofs=make_bar_bc(1, 1 + n_cells_x, 1),
        total_steps=total_steps,
    )


def norsand_box(n_cells_x=4, n_cells_y=8, dx=0.0125, total_steps=10, strain=0.02, tol=1e-10, precision='fp64',
                norsand=None, prestrain=-0.002):
    """2D plane-strain NorSand specimen in a rigid box (x fixed on the side walls,
    fixed base) compressed by a top platen by `strain` of its height
    (oedometric loading), no gravity.  The particles start at the isotropic
    NorSand pressure p0, in equilibrium with the walls, at the elastic
    volumetric strain `prestrain` (b_old = exp(2 prestrain / 3) I and ev0 =
    prestrain, so that the return's tension cut at zero volumetric strain is
    |prestrain| away); the material is the one of the reference's NorSand
    driver (triaxial_NorSand.py:8-30) unless `norsand` overrides keys."""
    g = _grid2d(n_cells_x, n_cells_y, dx)
    PPD = 2
    npart = n_cells_x * n_cells_y * PPD * PPD
    return dict(
        solver='quasi_static_solver_2d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=npart, PPD=PPD, p_vol=(dx / PPD) ** 2, p_rho=2.0),
        iteration_dict=dict(n_iter=12, tol=tol, tol_mode='relative', precision=precision),
        material_dict=dict(youngs_modulus=1.0e5, poisson_ratio=0.3, material_name='NorSand',
                           norsand=dict(dict(ev0=prestrain), **(norsand or {}))),
        load_dict=dict(gravity_mag=0.0, traction_value_x=0.0, traction_value_y=0.0, point_load_value_x=0.0,
                       point_load_value_y=0.0, prescribed_increment=make_platen(-strain * n_cells_y * dx)),
        set_boundary_dofs=make_bar_bc(1, 1 + n_cells_x, 1),
        initial_volumetric_strain=prestrain,
        total_steps=total_steps,
    )


def make_cantilever_bc(end_y, d0, dx):
    """cantilever_beam_cpGIMP.py:98-113: x fixed on node columns ix <= 1; y fixed
    on those columns at the node row int((end_y - d0/2 + dx/2)/dx) only."""

    def set_boundary_dofs(boundary, n_grid_x, n_nodes):
        nx1 = n_grid_x + 1
        bx = _view2d(boundary, 0, n_nodes, nx1)
        by = _view2d(boundary, 1, n_nodes, nx1)
        bx[:, :2] = True
        by[int((end_y - d0 / 2.0 + 0.5 * dx) / dx), :2] = True

    return set_boundary_dofs


def cantilever(n_grid=60, max_len=15.0, l0=10.0, d0=1.0, PPD=6, E=12000.0, nu=0.2, tip_load=-50.0, total_steps=50,
               tol=1e-8, precision='fp64'):
    """Plane-strain cantilever (cantilever_beam_cpGIMP.py:11-167): beam l0 x d0
    clamped on the left (x on node columns 0-1, y on the mid-height node row of
    those columns), Hencky elasticity, a point load on the particles of the free
    end at mid-height (set_external_force_flag, :116-133) ramped over the load
    steps."""
    dx = max_len / n_grid
    sx = dx
    g = dict(n_grid_x=n_grid, n_grid_y=n_grid, max_x=max_len, max_y=max_len, dx=dx, dy=dx, start_x=sx,
             end_x=sx + l0, start_y=l0 - d0, end_y=l0, domain_min_x=sx, domain_max_x=max_len - dx, domain_min_y=dx,
             domain_max_y=max_len - dx)
    npart = int(round(l0 / dx)) * int(round(d0 / dx)) * PPD * PPD

    def external_flags(x, ext_flag):
        tip = x[:, 0] > sx + l0 - dx / PPD
        mid = (x[:, 1] > l0 - d0 / 2.0 - dx / PPD) & (x[:, 1] < l0 - d0 / 2.0 + dx / PPD)
        ext_flag[tip & mid] = True

    return dict(
        solver='quasi_static_solver_2d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=npart, PPD=PPD, p_vol=(dx / PPD) ** 2, p_rho=1.0),
        iteration_dict=dict(n_iter=10, tol=tol, tol_mode='relative', precision=precision),
        material_dict=dict(youngs_modulus=E, poisson_ratio=nu, material_name='Hencky elasticity'),
        load_dict=dict(gravity_mag=0.0, traction_value_x=0.0, traction_value_y=0.0, point_load_value_x=0.0,
                       point_load_value_y=tip_load),
        set_boundary_dofs=make_cantilever_bc(l0, d0, dx),
        external_flags=external_flags,
        total_steps=total_steps,
    )


def run_load_steps(solver, cfg, n_steps=None, first=0, vtk_pattern=None, fields=('stress_xx', 'stress_yy'),
                   on_step=None):
    """The drivers' load loop (e.g. bar_compaction_cpGIMP.py:128-137,
    terzaghi_cpGIMP.py:180-189): advance_one_step(step, total) for each step and,
    with vtk_pattern (e.g. 'out/particles_%d.vtk'), one legacy-VTK particle file
    per step (and one for the initial state) holding the named point data
    (stress_xx, stress_yy, pressure).  on_step(step, solver) runs after each
    step.  Returns the Newton iteration counts."""
    from . import vtk

    total = cfg['total_steps']
    n = total - first if n_steps is None else n_steps

    def dump(k):
        if vtk_pattern is None:
            return
        sig = solver.particle_Cauchy_stress_array
        avail = {'stress_xx': lambda: sig[:, 0, 0], 'stress_yy': lambda: sig[:, 1, 1],
                 'stress_zz': lambda: sig[:, 2, 2],
                 'pressure': lambda: solver.particle_pressure_array}
        vtk.Mesh(points=solver.x_particles, cells=[],
                 point_data={f: avail[f]() for f in fields if f in avail}).write(vtk_pattern % k)

    if first == 0:
        dump(0)
    iters = []
    for step in range(first, first + n):
        solver.advance_one_step(step, total)
        iters.append(solver.newton_iterations)
        dump(step + 1)
        if on_step is not None:
            on_step(step, solver)
    return iters


def make_column_bc():
    """Terzaghi column: roller walls (x fixed at node_ix <= 1 or >= 2 for a one-cell
    column -> at the two wall lines), fixed base in y (terzaghi_cpGIMP.py:104-119)."""

    def set_boundary_dofs(boundary, n_grid_x, n_nodes):
        nx1 = n_grid_x + 1
        bx = _view2d(boundary, 0, n_nodes, nx1)
        by = _view2d(boundary, 1, n_nodes, nx1)
        bx[:, :2] = True
        by[:1, :] = True
        return None

    return set_boundary_dofs


def make_wall_column_bc(ix_lo, ix_hi):
    def set_boundary_dofs(boundary, n_grid_x, n_nodes):
        nx1 = n_grid_x + 1
        bx = _view2d(boundary, 0, n_nodes, nx1)
        by = _view2d(boundary, 1, n_nodes, nx1)
        bx[:, :ix_lo + 1] = True
        bx[:, ix_hi:] = True
        by[:1, :] = True

    return set_boundary_dofs


def terzaghi(n_cells_x=32, n_cells_y=320, width=1.0, total_steps=101, dt=None, precision='fp64', tol=1e-10):
    dx = width / n_cells_x
    pad_lo, pad_hi = 1, 2
    nx = n_cells_x + pad_lo + pad_hi
    ny = n_cells_y + 2
    g = dict(n_grid_x=nx, n_grid_y=ny, max_x=nx * dx, max_y=ny * dx, dx=dx, dy=dx, start_x=dx,
             end_x=dx + n_cells_x * dx, start_y=0.0, end_y=n_cells_y * dx, domain_min_x=dx,
             domain_max_x=dx + n_cells_x * dx, domain_min_y=0.0, domain_max_y=ny * dx)
    PPD = 2
    npart = n_cells_x * n_cells_y * PPD * PPD
    end_y = n_cells_y * dx

    def flags(x, traction_flag, pbc_flag):
        top = x[:, 1] > end_y - dx / PPD  # terzaghi_cpGIMP.py:127-149
        traction_flag[top] = True
        pbc_flag[top] = True

    return dict(
        solver='quasi_static_coupled_solver_2d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=npart, PPD=PPD, p_vol=(dx / PPD) ** 2, p_rho=2.0),
        iteration_dict=dict(n_iter=10, tol=tol, tol_mode='relative', dt=dt if dt is not None else 1.0,
                            precision=precision),
        material_dict=dict(youngs_modulus=1.0e4, poisson_ratio=0.3, material_name='Neo-Hookean'),
        porous_media_dict=dict(phi_initial=0.5, mobility_constant=1e-4),
        load_dict=dict(gravity_mag=0.0, traction_value_x=0.0, traction_value_y=-10.0, point_load_value_x=0.0,
                       point_load_value_y=0.0),
        set_boundary_dofs=make_wall_column_bc(1, 1 + n_cells_x),
        particle_flags=flags,
        total_steps=total_steps,
    )


def make_indent_bc():
    """3d_indentation_coupled_cpGIMP.py:317-342: x fixed at node_ix <= 1 or beyond
    end_x, y likewise, z fixed at the base."""

    def set_boundary_dofs(boundary, n_grid_x, n_grid_y, n_nodes, dx, dy, end_x, end_y):
        nx1, ny1 = n_grid_x + 1, n_grid_y + 1
        nz1 = n_nodes // (nx1 * ny1)
        bx = boundary[0:n_nodes].reshape(nz1, ny1, nx1)
        by = boundary[n_nodes:2 * n_nodes].reshape(nz1, ny1, nx1)
        bz = boundary[2 * n_nodes:3 * n_nodes].reshape(nz1, ny1, nx1)
        ix = np.arange(nx1)
        iy = np.arange(ny1)
        fx = (ix <= 1) | ((ix + 0.5) * dx > end_x)
        fy = (iy <= 1) | ((iy + 0.5) * dy > end_y)
        for i in np.nonzero(fx)[0]:
            bx[:, :, int(i)] = True
        for j in np.nonzero(fy)[0]:
            by[:, int(j), :] = True
        bz[0, :, :] = True

    return set_boundary_dofs


def indent3d(n_cells=(64, 64, 32), dx=1.0 / 64, total_steps=40, depth=0.05, footing=0.2, material='Drucker-Prager',
             E=2.0e4, precision='fp64', tol=1e-8, n_iter=10, mobility=1e-6, dt=1.0, gravity=0.0,
             dp_apex='reference', shape_factor=0.0, clip_top=True, krylov=None, footing_start='reference',
             penalty_form='reference'):
    ncx, ncy, ncz = n_cells
    nx, ny, nz = ncx + 4, ncy + 4, ncz + 3
    sx, sy, sz = dx, dx, 0.0
    ex, ey, ez = sx + ncx * dx, sy + ncy * dx, ncz * dx
    g = dict(n_grid_x=nx, n_grid_y=ny, n_grid_z=nz, max_x=nx * dx, max_y=ny * dx, max_z=nz * dx, dx=dx, dy=dx,
             dz=dx, start_x=sx, end_x=ex, start_y=sy, end_y=ey, start_z=sz, end_z=ez, domain_min_x=sx,
             domain_max_x=ex, domain_min_y=sy, domain_max_y=ey, domain_min_z=sz,
             domain_max_z=ez if clip_top else nz * dx)
    PPD = 2
    npart = ncx * ncy * ncz * PPD ** 3

    # indenter face before load step 0: the driver's end_z - dz/PPD (3d_indentation_coupled_cpGIMP.py:84),
    # which lies a quarter cell below the top particle row at 2 particles per cell, so that load
    # step 0 pushes that quarter cell plus one increment; 'surface' starts at the top particle row
    # so that every load step pushes exactly one increment (depth in total_steps steps)
    h0 = {'reference': ez - dx / PPD, 'surface': ez - dx / (2 * PPD)}[footing_start]

    def flags(x, traction_flag, pbc_flag):
        top = x[:, 2] > ez - dx / PPD  # 3d_indentation_coupled_cpGIMP.py:348-358
        pbc_flag[top] = True

    return dict(
        solver='quasi_static_coupled_solver_3d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=npart, PPD=PPD, p_vol=(dx / PPD) ** 3, p_rho=2.0),
        iteration_dict=dict(n_iter=n_iter, tol=tol, tol_mode='relative', dt=dt, precision=precision,
                            **({'krylov': krylov} if krylov else {})),
        material_dict=dict(youngs_modulus=E, poisson_ratio=0.3, material_name=material, friction_angle=30.0,
                           dilation_angle=0.0, cohesion=5.0, dp_apex=dp_apex, shape_factor=shape_factor),
        porous_media_dict=dict(phi_initial=0.5, mobility_constant=mobility),
        load_dict=dict(gravity_mag=gravity, traction_value_x=0.0, traction_value_y=0.0, traction_value_z=0.0,
                       point_load_value_x=0.0, point_load_value_y=0.0, point_load_value_z=0.0,
                       footing_region_min_x=sx, footing_region_max_x=sx + footing, footing_region_min_y=sy,
                       footing_region_max_y=sy + footing, footing_initial_height=h0,
                       footing_incr_disp=-depth / total_steps, penalty_factor=1000.0, penalty_form=penalty_form),
        set_boundary_dofs=make_indent_bc(),
        particle_flags=flags,
        total_steps=total_steps,
    )


def build_solver(module, cfg, **kw):
    """Instantiate cfg['solver'] from `module` (geowarp_amd.solvers or oracle.solvers)."""
    cls = getattr(module, cfg['solver'])
    if 'x_particles' in cfg:
        kw.setdefault('x_particles', cfg['x_particles'])
    if cfg['solver'] == 'quasi_static_solver_2d':
        s = cls(cfg['background_grid_dict'], cfg['particles_dict'], cfg['iteration_dict'], cfg['material_dict'],
                cfg['load_dict'], **kw)
    else:
        s = cls(cfg['background_grid_dict'], cfg['particles_dict'], cfg['iteration_dict'], cfg['material_dict'],
                cfg['porous_media_dict'], cfg['load_dict'], cfg['set_boundary_dofs'], **kw)
    ev0 = cfg.get('initial_volumetric_strain')
    if ev0:
        # isotropic elastic pre-strain: b_old = exp(2 ev0 / 3) b_old, F_old unchanged
        b = s.b_old if hasattr(s, 'b_old') else s.left_Cauchy_Green_old
        b *= np.exp(2.0 * ev0 / 3.0)
    return s


def _like(ref, arr):
    """numpy bool/float array as the array type of `ref` (numpy or torch)."""
    if isinstance(ref, np.ndarray):
        return arr
    import torch
    return torch.as_tensor(arr, device=ref.device)


def _to_np(a):
    return a if isinstance(a, np.ndarray) else a.detach().cpu().numpy()


def specimen_counts(n_specimens=64, seed=0, min_count=1e3, max_count=5e4, phi_range=(25.0, 40.0),
                    coh_range=(0.0, 20.0)):
    """Specimen sizes and true parameters of the batched inverse analysis:
    numpy default_rng(seed), drawn in the order counts, phi, c; count
    N = exp(U(ln min, ln max)) gives W = round(sqrt(N / 8)) cells wide, 2W tall
    (8 W^2 particles at 2x2 per cell)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(min_count), np.log(max_count), n_specimens))
    phi = rng.uniform(phi_range[0], phi_range[1], n_specimens)
    coh = rng.uniform(coh_range[0], coh_range[1], n_specimens)
    W = np.maximum(1, np.round(np.sqrt(counts / 8.0)).astype(int))
    return W, phi, coh


def specimens2d(widths, phi, coh, dx=0.01, total_steps=30, strain=0.03, E=1.0e4, nu=0.1, gap=3, tol=1e-10,
                n_iter=12, precision='fp64'):
    """Independent 2D plane-strain Drucker-Prager specimens (W_i x 2W_i cells,
    2x2 particles per cell) side by side on one background grid, `gap` empty
    cells apart.  Each specimen sits in a box with frictionless walls (x fixed
    on its side node lines), a fixed base and a top platen that moves down by
    strain * height over total_steps load steps (prescribed y increments on the
    node rows at and above the row nearest to its current top, so that the
    platen keeps a support weight of at least half a cell on the top
    particles).  Per-particle friction angle and
    cohesion come from the specimen (solver.friction_angle_array /
    cohesion_array); `groups` maps particles to specimens."""
    widths = [int(w) for w in widths]
    PPD = 2
    pad_lo, pad_hi = 1, 3
    x0 = []
    cx = pad_lo + gap
    for w in widths:
        x0.append(cx)
        cx += w + gap
    nx = cx + pad_lo
    hmax = 2 * max(widths)
    ny = hmax + pad_lo + pad_hi
    g = dict(n_grid_x=nx, n_grid_y=ny, max_x=nx * dx, max_y=ny * dx, dx=dx, dy=dx, start_x=0.0, end_x=nx * dx,
             start_y=pad_lo * dx, end_y=(pad_lo + hmax) * dx, domain_min_x=0.0, domain_max_x=nx * dx,
             domain_min_y=pad_lo * dx, domain_max_y=ny * dx)
    xs, grp = [], []
    for i, w in enumerate(widths):
        sx, sy = x0[i] * dx, pad_lo * dx
        ii, jj, a, b = np.meshgrid(np.arange(w), np.arange(2 * w), np.arange(PPD), np.arange(PPD), indexing='ij')
        px = sx + (ii + (0.5 + a) / PPD) * dx
        py = sy + (jj + (0.5 + b) / PPD) * dx
        xs.append(np.stack([px.reshape(-1), py.reshape(-1)], 1))
        grp.append(np.full(px.size, i))
    x = np.concatenate(xs)
    groups = np.concatenate(grp)
    nx1 = nx + 1
    # column ownership: node columns x0 .. x0+w belong to specimen i
    owner = np.full(nx1, -1)
    for i, w in enumerate(widths):
        owner[x0[i]:x0[i] + w + 1] = i
    heights = np.array([2 * w * dx for w in widths])
    dy_step = -strain * heights / float(total_steps)

    def set_boundary_dofs(boundary, n_grid_x, n_nodes):
        bx = _view2d(boundary, 0, n_nodes, nx1)
        by = _view2d(boundary, 1, n_nodes, nx1)
        inner = np.zeros(nx1, dtype=bool)
        for i, w in enumerate(widths):
            inner[x0[i] + 1:x0[i] + w] = True
        bx[:, _like(bx, ~inner)] = True
        by[:pad_lo + 1, :] = True

    def platen(xp, lp, boundary, pres, n_grid_x, n_nodes, dx_, current_step, total_steps_):
        top = _to_np(xp[:, 1] + lp[:, 1])
        tops = np.full(len(widths), -np.inf)
        np.maximum.at(tops, groups, top)
        rows = np.floor(tops / dx_ + 0.5).astype(int)  # node row nearest to the top surface
        iy = np.arange(ny + 1)[:, None]
        col_row = np.where(owner >= 0, rows[np.maximum(owner, 0)], ny + 2)
        m = (iy >= col_row[None, :]) & (owner[None, :] >= 0)
        val = np.where(owner >= 0, dy_step[np.maximum(owner, 0)] * total_steps / float(total_steps_), 0.0)
        by = _view2d(boundary, 1, n_nodes, nx1)
        py = _view2d(pres, 1, n_nodes, nx1)
        by[pad_lo + 1:, :] = False  # the platen rows of earlier steps (the base rows stay fixed)
        by[_like(by, m)] = True
        py[_like(py, m)] = _like(py, np.broadcast_to(val[None, :], m.shape)[m])

    n = x.shape[0]
    return dict(
        solver='quasi_static_solver_2d',
        background_grid_dict=g,
        particles_dict=dict(n_particles=n, PPD=PPD, p_vol=(dx / PPD) ** 2, p_rho=2.0),
        iteration_dict=dict(n_iter=n_iter, tol=tol, tol_mode='relative', precision=precision),
        material_dict=dict(youngs_modulus=E, poisson_ratio=nu, material_name='Drucker-Prager',
                           friction_angle=float(np.mean(phi)), dilation_angle=0.0, cohesion=float(np.mean(coh))),
        load_dict=dict(gravity_mag=0.0, traction_value_x=0.0, traction_value_y=0.0, point_load_value_x=0.0,
                       point_load_value_y=0.0, prescribed_increment=platen),
        set_boundary_dofs=set_boundary_dofs,
        total_steps=total_steps,
        x_particles=x,
        groups=groups,
        phi=np.asarray(phi, dtype=np.float64),
        coh=np.asarray(coh, dtype=np.float64),
    )

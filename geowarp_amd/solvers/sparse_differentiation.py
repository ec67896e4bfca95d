"""Sparse-Jacobian helpers with the reference's names (GeoWarp
solvers/sparse_differentiation.py).

The reference builds its Jacobian column group by column group: 25 (2D) or
125 (3D) seed vectors each select nodes five apart on every axis, a reverse
sweep of the tape gives e^T dR/du, and ``from_jacobian_to_vector_*`` scatters
that row combination back into COO triplets.  The solvers here assemble the
Jacobian analytically in block-CSR instead; these functions keep the
reference's data formats for drivers that still use them:

* ``pick_grid_nodes_*`` / ``precompute_seed_vectors_*``: the same node sets
  and seed vectors (as torch tensors on ``device``);
* ``from_jacobian_to_vector_parallel_*``: the same COO scatter, vectorised;
* ``export_coo``: the assembled Jacobian of a solver in the reference's
  dof-major COO convention (K = -dR/du on free dofs, identity on fixed or
  inactive ones).
"""
import numpy as np
import torch


def _default_device(device):
    if device is not None:
        return torch.device(device)
    return torch.device('cuda') if torch.cuda.is_available() else torch.device('cpu')


def pick_grid_nodes_2d(n_grid_x, n_grid_y, first_grid_node):
    """sparse_differentiation.py:8-29: nodes (i0 + 5a, j0 + 5b) on the grid,
    ordered row by row."""
    j0, i0 = divmod(int(first_grid_node), n_grid_x + 1)
    jj, ii = np.meshgrid(np.arange(j0, n_grid_y + 1, 5), np.arange(i0, n_grid_x + 1, 5), indexing='ij')
    return (jj * (n_grid_x + 1) + ii).reshape(-1)


def pick_grid_nodes_3d(n_grid_x, n_grid_y, n_grid_z, first_grid_node):
    """sparse_differentiation.py:31-56: nodes (i0 + 5a, j0 + 5b, k0 + 5c),
    ordered layer, row, column."""
    sxy = (n_grid_x + 1) * (n_grid_y + 1)
    k0, rem = divmod(int(first_grid_node), sxy)
    j0, i0 = divmod(rem, n_grid_x + 1)
    kk, jj, ii = np.meshgrid(np.arange(k0, n_grid_z + 1, 5), np.arange(j0, n_grid_y + 1, 5),
                             np.arange(i0, n_grid_x + 1, 5), indexing='ij')
    return (kk * sxy + jj * (n_grid_x + 1) + ii).reshape(-1)


def _seeds(offsets, picker, nb, n_nodes, n_matrix_size, max_selector_length, device):
    sel = [[] for _ in range(nb)]
    es = [[] for _ in range(nb)]
    for first in offsets:
        nodes = picker(first)
        for k in range(nb):
            s = nodes + k * n_nodes
            padded = np.zeros(max_selector_length, dtype=np.int32)  # ndarray.resize pads with zeros
            m = min(len(s), max_selector_length)
            padded[:m] = s[:m]
            sel[k].append(torch.from_numpy(padded).to(device))
            e = torch.zeros(n_matrix_size, dtype=torch.float64)
            e[torch.from_numpy(s)] = 1.0
            es[k].append(e.to(device))
    return sel, es


def precompute_seed_vectors_2d(n_grid_x, n_grid_y, n_nodes, n_matrix_size, max_selector_length, device=None):
    """sparse_differentiation.py:59-93: returns (selector_x, selector_y, e_x, e_y),
    25 entries each, first node c + r*(n_grid_x+1) for c in 0..4 (outer), r in 0..4."""
    offs = [c + r * (n_grid_x + 1) for c in range(5) for r in range(5)]
    sel, es = _seeds(offs, lambda f: pick_grid_nodes_2d(n_grid_x, n_grid_y, f), 2, n_nodes, n_matrix_size,
                     max_selector_length, _default_device(device))
    return sel[0], sel[1], es[0], es[1]


def precompute_seed_vectors_coupled_2d(n_grid_x, n_grid_y, n_nodes, n_matrix_size, max_selector_length,
                                       device=None):
    """sparse_differentiation.py:96-144: (selector_x, selector_y, selector_p, e_x, e_y, e_p)."""
    offs = [c + r * (n_grid_x + 1) for c in range(5) for r in range(5)]
    sel, es = _seeds(offs, lambda f: pick_grid_nodes_2d(n_grid_x, n_grid_y, f), 3, n_nodes, n_matrix_size,
                     max_selector_length, _default_device(device))
    return sel[0], sel[1], sel[2], es[0], es[1], es[2]


def precompute_seed_vectors_coupled_3d(n_grid_x, n_grid_y, n_grid_z, n_nodes, n_matrix_size, max_selector_length,
                                       device=None):
    """sparse_differentiation.py:147-210: 125 entries each, first node
    c + r*(nx+1) + l*(nx+1)(ny+1), c outer, l inner."""
    sxy = (n_grid_x + 1) * (n_grid_y + 1)
    offs = [c + r * (n_grid_x + 1) + l * sxy for c in range(5) for r in range(5) for l in range(5)]
    sel, es = _seeds(offs, lambda f: pick_grid_nodes_3d(n_grid_x, n_grid_y, n_grid_z, f), 4, n_nodes,
                     n_matrix_size, max_selector_length, _default_device(device))
    return tuple(sel) + tuple(es)


# ---------------------------------------------------------------------- COO scatter
def _scatter(jac, rows, cols, vals, grid, n_nodes, n_matrix_size, selector, boundary, activate, nb, check_nan):
    """Shared body of from_jacobian_to_vector_parallel_*: for every selected
    row dof r (entry 0 of the selector, and every later non-zero entry) and
    every node within +-2 on each axis, write (r, c, -jac[c]) for each free
    dof c of that node at slot k*W*n_matrix_size + r*W + offset (W = 5^dim)."""
    dev = jac.device
    dim = len(grid)
    W = 5 ** dim
    sel = selector.long()
    ok = sel > 0
    ok[0] = ok[0] | (sel[0] == 0)
    r = sel[ok]
    free = (~boundary.bool()) & activate.bool()
    r = r[free[r]]
    if r.numel() == 0:
        return
    node = r % n_nodes
    stride = [1, grid[0] + 1, (grid[0] + 1) * (grid[1] + 1)]
    coord = []
    rem = node
    for a in range(dim - 1, -1, -1):
        coord.insert(0, rem // stride[a])
        rem = rem % stride[a]
    d = torch.arange(5, device=dev) - 2
    if dim == 2:
        di = d.view(5, 1).expand(5, 5)   # i varies fastest in the slot offset i + 5j
        dj = d.view(1, 5).expand(5, 5)
        offs = [di.t().reshape(-1), dj.t().reshape(-1)]
    else:
        di = d.view(5, 1, 1).expand(5, 5, 5)
        dj = d.view(1, 5, 1).expand(5, 5, 5)
        dk = d.view(1, 1, 5).expand(5, 5, 5)
        offs = [t.permute(2, 1, 0).reshape(-1) for t in (di, dj, dk)]  # slot i + 5j + 25k
    adj = [coord[a].view(-1, 1) + offs[a].view(1, -1) for a in range(dim)]
    inside = torch.ones_like(adj[0], dtype=torch.bool)
    for a in range(dim):
        inside &= (adj[a] >= 0) & (adj[a] <= grid[a])
    adj_node = sum(adj[a] * stride[a] for a in range(dim))
    slot = r.view(-1, 1) * W + torch.arange(W, device=dev).view(1, -1)
    rr = r.view(-1, 1).expand_as(slot)
    for k in range(nb):
        c = (adj_node + k * n_nodes).clamp(0, free.numel() - 1)
        m = inside & free[c]
        if check_nan:
            m &= ~torch.isnan(jac[c])
        s = slot[m] + k * W * n_matrix_size
        rows[s] = rr[m].to(rows.dtype)
        cols[s] = c[m].to(cols.dtype)
        vals[s] = -jac[c[m]].to(vals.dtype)


def from_jacobian_to_vector_parallel_2d(jacobian_wp, rows, cols, vals, n_grid_x, n_grid_y, n_nodes, n_matrix_size,
                                        selector_wp, boundary_flag_array, activate_flag_array):
    """sparse_differentiation.py:213-262."""
    _scatter(jacobian_wp, rows, cols, vals, (n_grid_x, n_grid_y), n_nodes, n_matrix_size, selector_wp,
             boundary_flag_array, activate_flag_array, 2, False)


def from_jacobian_to_vector_parallel_coupled_2d(jacobian_wp, rows, cols, vals, n_grid_x, n_grid_y, n_nodes,
                                                n_matrix_size, selector_wp, boundary_flag_array, activate_flag_array):
    """sparse_differentiation.py:265-322."""
    _scatter(jacobian_wp, rows, cols, vals, (n_grid_x, n_grid_y), n_nodes, n_matrix_size, selector_wp,
             boundary_flag_array, activate_flag_array, 3, False)


def from_jacobian_to_vector_parallel_coupled_3d(jacobian_wp, rows, cols, vals, n_grid_x, n_grid_y, n_grid_z, n_nodes,
                                                n_matrix_size, selector_wp, boundary_flag_array, activate_flag_array):
    """sparse_differentiation.py:325-404 (NaN entries of the sweep are skipped)."""
    _scatter(jacobian_wp, rows, cols, vals, (n_grid_x, n_grid_y, n_grid_z), n_nodes, n_matrix_size, selector_wp,
             boundary_flag_array, activate_flag_array, 4, True)


# ---------------------------------------------------------------------- export
def export_coo(solver):
    """The solver's current assembled Jacobian as dof-major COO (rows, cols,
    vals) torch tensors, zeros dropped: the matrix the reference hands to its
    sparse solver after the scatter and set_diagnal_component_* steps."""
    eng = solver.engine
    nb, n = eng.nb, eng.n_nodes
    rp = eng.row_ptr.long()
    nnzb = int(rp[-1].item())
    brow = torch.repeat_interleave(torch.arange(n, device=rp.device), rp[1:] - rp[:-1])
    bcol = eng.col_idx[:nnzb].long()
    v = eng.vals[:nnzb * nb * nb].view(nnzb, nb, nb).to(torch.float64)
    a = torch.arange(nb, device=rp.device)
    rows = (brow.view(-1, 1, 1) + n * a.view(1, nb, 1)).expand(nnzb, nb, nb).reshape(-1)
    cols = (bcol.view(-1, 1, 1) + n * a.view(1, 1, nb)).expand(nnzb, nb, nb).reshape(-1)
    vals = v.reshape(-1)
    keep = vals != 0
    return rows[keep], cols[keep], vals[keep]

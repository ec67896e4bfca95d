// Plain-old-data argument blocks shared by the C ABI (include/geowarp_amd.h)
// and the HIP kernels.  Pointers are device pointers; every array is
// row-major.  Nodal vectors are node-interleaved: value k of node n sits at
// n*nb + k (nb = dofs per node).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gm_grid {
  int dim;       /* 2 or 3 */
  int nx, ny, nz; /* background cells per axis (nz = 0 in 2D) */
  int n_nodes;   /* (nx+1)(ny+1)(nz+1) */
  int nb;        /* dofs per node: dim (single phase) or dim+1 (u-p) */
  double h[3];   /* spacing used by the shape functions (2D: dx on both axes) */
  double inv_h[3];
} gm_grid;

/* Per-step particle binning: particles sorted by the node index of their
   stencil base (lo_x, lo_y[, lo_z]). */
typedef struct gm_bins {
  int n_particles;
  int* cell_start;   /* n_nodes + 1 */
  int* cell_count;   /* n_nodes + 1 (scratch) */
  int* sorted;       /* n_particles: particle ids ordered by base node, then id */
  uint8_t* ext;      /* n_particles, particle order: bit a set => 3 nodes along axis a, else 2 */
  int* scan_tmp;     /* >= 1025 ints of scratch */
  /* bin-sorted copies (index k = position in `sorted`), written by gm_bin_particles */
  uint8_t* ext_s;    /* n_particles */
  double* x_s;       /* n_particles x dim */
  double* lp_s;      /* n_particles x dim */
} gm_bins;

/* Per-step block-CSR pattern over nodes (5^dim window around each node). */
typedef struct gm_pattern {
  int* row_ptr;       /* n_nodes + 1 */
  int* col_idx;       /* nnzb */
  uint64_t* row_mask; /* n_nodes * 2: window-offset bitmask of each row */
  int nnzb;
} gm_pattern;

/* NorSand parameters (triaxial_NorSand.py:243-262): kappa, tilde_lambda, M, N,
   beta, critical specific volume vc0, reference pressure p0 (negative in
   compression), initial specific volume v0, hardening modulus h, shear modulus
   mu, initial volumetric strain ev0, and the tolerance of the tension / yield
   tests. */
typedef struct gm_norsand_params {
  double kappa, tilde_lambda, M, N, beta, vc0, p0, v0, h, mu, ev0, tol;
} gm_norsand_params;

/* Arguments of the per-particle stress/tangent stage. */
typedef struct gm_stage_args {
  int n_particles;
  int coupled;       /* 0: single phase, 1: u-p */
  int want_tangent;  /* write the particle tangent D */
  int tangent_fp32;  /* store D as float instead of double */
  const double* x;   /* n_p x dim */
  const double* lp;  /* n_p x dim */
  const double* F_old; double* F_new;   /* n_p x 9 */
  const double* b_old; double* b_new;   /* n_p x 9 */
  double* sigma;                        /* n_p x 9 */
  const double* alpha_old; double* alpha_new; /* n_p (J2 hardening), may be null */
  const uint8_t* ext_flag;   /* n_p point-load flags, may be null */
  const uint8_t* trac_flag;  /* n_p traction flags, may be null */
  const double* u_new;  /* single phase: increment; u-p: new solution (node-interleaved) */
  const double* u_old;  /* u-p: old solution (node-interleaved) */
  const int* sorted;    /* bin order (gm_bins.sorted) or null; Y and D rows follow it */
  double* Y;            /* n_p x nY particle "stress vector", bin-sorted rows */
  void* D;              /* n_p x nY x nZ tangent (double or float), bin-sorted rows */
  double* pen_force;    /* n_p per-particle indentation force (3D), may be null */
  /* material */
  int mat_kind; int dp_apex;
  double lam, mu, E, kappa, H, phi, psi, coh, shape;
  /* loading / porous media */
  double p_vol, p_rho, phi0, mobility, dt, PPD;
  double gravity[3];     /* already ramped by (step+1)/total */
  double point_load[3];  /* already ramped */
  double traction[3];    /* per-particle line load already multiplied by h/PPD */
  double footing[4];     /* min_x, max_x, min_y, max_y */
  double footing_height; /* current indenter height */
  double pen_par;        /* penalty stiffness E*pf*(dx/PPD)(dy/PPD) */
  int penalty;           /* 3D indenter active */
  const double* p_phi;   /* per-particle Drucker-Prager friction angle, may be null (then phi) */
  const double* p_coh;   /* per-particle cohesion, may be null (then coh) */
  gm_norsand_params ns;  /* NorSand (mat_kind 4): alpha_old/new carry the image pressure pi */
} gm_stage_args;

/* Krylov workspace (device pointers).  Vectors hold n_rows*nb entries of the
   solve precision; dinv holds n_rows*nb*nb; part >= 2*max(n_rows/16+1, 1024)
   doubles; sc >= 16 doubles. */
typedef struct gm_krylov_ws {
  void* r; void* rhat; void* p; void* v; void* s; void* t; void* y; void* z;  /* n_rows*nb each, TV */
  void* dinv;     /* n_rows*nb*nb, TM */
  double* part;   /* >= 2*max(spmv blocks, 1024) */
  double* sc;     /* >= 16 */
} gm_krylov_ws;


/* Arguments of the 2D plane-strain local map used by the adjoint
   (gm_local_eval_2d / gm_local_vjp_2d).  Particle arrays are in particle order;
   nodal vectors (u, cot) are node-interleaved (value c of node n at 2n+c).
   Optional pointers may be null (zero cotangent / uniform parameter). */
typedef struct gm_local_args {
  int n_particles;
  int nx1;                 /* n_grid_x + 1 */
  int mat_kind, dp_apex;   /* 0 Hencky, 1 J2, 2 Neo-Hookean, 3 Drucker-Prager */
  double h;                /* dx (used on both axes) */
  double shape;
  double p_vol, p_rho;
  double gravity[2], point_load[2];  /* ramped by (step+1)/total */
  double lp0, dmin[2], dmax[2];
  double E, nu, kappa, H, phi, psi, coh;
  const double* p_phi;     /* per-particle friction angle (optional) */
  const double* p_coh;     /* per-particle cohesion (optional) */
  const double* x;         /* n x 2 */
  const double* lp;        /* n x 2 */
  const double* F_old;     /* n x 9 */
  const double* b_old;     /* n x 9 */
  const double* alpha_old; /* n (optional) */
  const uint8_t* ext_flag; /* n (optional) */
  const double* u;         /* converged nodal increment, 2 per node */
  const double* cot;       /* nodal cotangent of the residual contributions (optional) */
  const double* cot_scale; /* n per-particle factor on that cotangent (optional) */
  const double* xbar;      /* n x 2 cotangent of x_new (optional) */
  const double* lpbar;     /* n x 2 (optional) */
  const double* Fbar;      /* n x 5: xx, xy, yx, yy, zz (optional) */
  const double* bbar;      /* n x 4: xx, xy (symmetric pair), yy, zz (optional) */
  const double* abar;      /* n (optional) */
} gm_local_args;

/* One level of the geometric multigrid preconditioner on the background grid
   (level 0 = the Jacobian's own grid and pattern; level l+1 has every second
   node of level l per axis, ceil(n/2) cells).  Patterns are block-CSR with a
   5^dim window bitmask per row (row_mask, 2 words per row).  vals/dinv and the
   vectors x, b, r hold the solve precision (double or float). */
typedef struct gm_mg_level {
  int nx, ny, nz;       /* cells per axis (nz = 0 in 2D) */
  int n_nodes;
  int nnzb;
  const int* row_ptr;
  const int* col_idx;
  const uint64_t* row_mask;
  void* vals;           /* nnzb * nb * nb: level 0 in the solve precision, coarser levels FP64 */
  void* dinv;           /* n_nodes * nb * nb */
  uint8_t* freem;       /* n_nodes * nb: 1 = free dof */
  void* x; void* b; void* r;  /* n_nodes * nb work vectors (level 0 x/b unused) */
  const void* dense_inv;      /* coarsest level only, may be null: (n_nodes*nb)^2 FP64 inverse, row-major */
  int gamma;                  /* coarse-grid corrections from this level per cycle (<= 1: V, 2: W) */
  /* optional, FP32 hierarchies only: an FP16 copy of vals, each (row, column dof)
     sub-row of a block row scaled by its largest magnitude,
     vals[k][r][c] ~ vals_lo[k][r][c] * row_scale[(row * nb + r) * nb + c], filled by
     gm_mg_setup; when set, the V-cycle's smoothing and residual products on this level read it */
  void* vals_lo;              /* nnzb * nb * nb halves, or null */
  float* row_scale;           /* n_nodes * nb * nb */
} gm_mg_level;

#ifdef __cplusplus
}
#endif

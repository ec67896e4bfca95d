/* geowarp_amd.h -- C ABI of libgeowarp_amd.so (MI355X / gfx950).
 *
 * Drop-in replacement for the numerical core of GeoWarp's implicit cpGIMP MPM
 * load-step Newton loop.  The reference implements each phase as a Warp
 * kernel (@wp.kernel, compiled at run time) launched from the solver classes;
 * each entry point below cites the reference interface it replaces.
 *
 * Conventions
 *   - every pointer is a device pointer unless stated otherwise; arrays are
 *     row-major; particle tensors are (n_particles, ...) FP64;
 *   - nodal vectors are node-interleaved: value k of node n at n*nb + k
 *     (the Python layer converts to and from the reference's dof-major order,
 *     dof = node + k*n_nodes);
 *   - every call is ordered on `stream` (a hipStream_t) and returns 0 on
 *     success or a negative code; gm_last_error() then describes the failure;
 *   - the library allocates no device memory: buffers and workspaces are
 *     passed in (the Python layer allocates them with torch);
 *   - calls that return a host value (gm_bsr_pattern's block count, the
 *     Krylov solvers' iteration counts and residuals) synchronise `stream`;
 *     the GMRES solvers also read each Arnoldi step's Hessenberg column on
 *     the host (Givens rotations and the stopping test run there), and CG
 *     reads its state every `check_every` iterations.
 *   The argument structs gm_grid, gm_bins, gm_pattern, gm_stage_args and
 *   gm_krylov_ws are defined in geowarp_amd/csrc/gm_types.h.
 */
#ifndef GEOWARP_AMD_H
#define GEOWARP_AMD_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../geowarp_amd/csrc/gm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Library version and the message of the last failed call (thread-local). */
int gm_version(void);
const char* gm_last_error(void);

/* Exclusive prefix sum of n ints (out has n+1 entries; out[n] = total).
   tmp >= 1025 ints.  Helper of the binning and pattern passes (no reference
   counterpart; the reference sorts COO triplets in wps.bsr_set_from_triplets,
   qs2d:202). */
int gm_exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t stream);

/* Sort particle ids by the node index of their cpGIMP stencil base
   (gimp.py:33-49 floors), record each particle's per-axis extent and write
   bin-sorted copies of x, lp and the extents (gm_bins.x_s, lp_s, ext_s) that
   the node-gather kernels read.
   pbase: n_particles ints of scratch; err: one int, set to the number of
   particles whose stencil leaves the grid (the reference does not check).
   Replaces the per-particle index arithmetic repeated in every kernel of
   mpm/mpm_utils.py (e.g. :103-113, :1088-1100). */
int gm_bin_particles(const gm_grid* g, const double* x, const double* lp, gm_bins* bins, int* pbase, int* err,
                     hipStream_t stream);

/* Per-load-step node block-CSR pattern: block (I,J) exists when a particle's
   support holds both nodes.  row_nnzb: n_nodes+1 ints of scratch; returns -4
   with pat->nnzb set when max_nnzb is too small.  Replaces the fixed-slot COO
   layout of sparse_differentiation.py:213-401 (5^d slots per row and dof
   type) and wps.bsr_set_from_triplets (qs2d:202, qc3d:336). */
int gm_bsr_pattern(const gm_grid* g, const gm_bins* bins, gm_pattern* pat, int* row_nnzb, int max_nnzb,
                   hipStream_t stream);

/* How gm_bsr_pattern forms its row masks (no counterpart in the reference): 1 (the
   default) from each bin's set of particle support shapes, held in bins->cell_count
   (the binning's scratch); 0 particle by particle.  Both give the same pattern. */
int gm_pattern_kernel(int by_sets);

/* Per-particle stage of one residual evaluation: gather the nodal increment
   (and pressures), update F, b_e, Cauchy stress, hardening variable; write the
   particle stress vector Y and (want_tangent) its analytic tangent D, both in
   bin-sorted row order when a->sorted is given.
   Replaces the particle part of assemble_residual_2d_hencky / _J2 /
   assemble_residual_coupled_2d_neohookean / _3d_neohookean
   (mpm_utils.py:416-1294) and the coloured reverse sweeps that differentiate
   them (qs2d:171-190, qc2d:197-224, qc3d:285-321); materials from
   plasticity.py:8-37 and triaxial_DruckerPrager.py:308-412. */
int gm_particle_stage(const gm_grid* g, const gm_stage_args* a, hipStream_t stream);

/* Node-gather residual r_I = sum_p L(geom_I) Y_p, zero at fixed/inactive dofs
   (freem = !boundary && active, node-interleaved).  Replaces the atomic
   scatter of the assemble_residual_* kernels (e.g. mpm_utils.py:575-596,
   1252-1294). */
int gm_node_residual(const gm_grid* g, const gm_bins* bins, const double* Y, const uint8_t* freem, double* rhs,
                     int coupled, hipStream_t stream);

/* The same residual for the 3D u-p form (mpm_utils.py:1020-1294) summed per particle
   bin (the 27 nodes of each bin, its particles in bin order) into `work` (at least
   gm_residual_work_bytes(g): n_nodes x 108 doubles), then per node over the <= 27 bins
   holding it in a fixed order; masked like gm_node_residual. */
size_t gm_residual_work_bytes(const gm_grid* g);
int gm_node_residual_binned(const gm_grid* g, const gm_bins* bins, const double* Y, const uint8_t* freem,
                            double* rhs, void* work, size_t work_bytes, hipStream_t stream);

/* Node-gather Jacobian into the pattern: K = -dR/du between free & active
   dofs, identity on the others (reference sign and masking:
   sparse_differentiation.py:253-262, mpm_utils.py:1299-1341).  D and vals may
   be FP32 (tangent_fp32 / vals_fp32) for the mixed-precision form. */
int gm_node_jacobian(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, const void* D, int tangent_fp32,
                     const uint8_t* freem, void* vals, int vals_fp32, int coupled, hipStream_t stream);

/* The same matrix for the 3D u-p form (the reference's qc3d:285-321 coloured
   reverse sweeps), assembled per particle bin on the matrix cores: each bin's
   27 x 27 node blocks go to the workspace `work` (at least
   gm_node_jacobian_work_bytes(g, tangent_fp32) bytes: n_nodes x 11664 values
   of the tangent's precision), then every block of the pattern sums the bins
   that hold both of its nodes in a fixed order.  Arithmetic in the tangent's
   precision; -1 for another form or a smaller workspace. */
size_t gm_node_jacobian_work_bytes(const gm_grid* g, int tangent_fp32);
int gm_node_jacobian_binned(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, const void* D,
                            int tangent_fp32, const uint8_t* freem, void* vals, int vals_fp32, void* work,
                            size_t work_bytes, hipStream_t stream);
/* Its two phases, callable separately (gm_node_jacobian_binned is the pair): the
   per-bin products into `work`, and the per-block gather from `work` into vals. */
int gm_jacobian_bins(const gm_grid* g, const gm_bins* bins, const void* D, int tangent_fp32, void* work,
                     size_t work_bytes, hipStream_t stream);
int gm_jacobian_gather(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, int tangent_fp32,
                       const void* work, const uint8_t* freem, void* vals, int vals_fp32, hipStream_t stream);
/* The same phases with an explicit kernel variant (the two above use the one the
   tuner below chose for the tangent's precision, else variant 0).  Per-bin variants
   0-3 (particles per group, workgroups per CU, grid): FP32 (8, 2, one per bin),
   (8, 2, persistent), (4, 4, persistent), (4, 3, persistent); FP64 (4, 2, one per
   bin), (4, 2, persistent), (8, 1, one per bin), (8, 1, persistent).  Gather
   variants 0 (per-lane reads of the workspace), 1 and 2 (bin chunks staged in LDS,
   27 or 9 at a time in FP32, 9 or 3 in FP64).
   Every variant writes the same bits; -1 for a variant out of range. */
int gm_jacobian_bins_variant(const gm_grid* g, const gm_bins* bins, const void* D, int tangent_fp32, int variant,
                             void* work, size_t work_bytes, hipStream_t stream);
int gm_jacobian_gather_variant(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, int tangent_fp32,
                               int variant, const void* work, const uint8_t* freem, void* vals, int vals_fp32,
                               hipStream_t stream);
/* Tile and launch-shape tuner of the binned assembly: times every per-bin and
   gather variant on this tangent (`reps` launches each after a warm-up, HIP events
   on the stream; ms[0..3] per-bin, ms[4..6] gather, milliseconds per launch), keeps
   the fastest of each for later assemblies in this precision in this process and
   returns them.  vals and work end holding the assembly.  No reference counterpart:
   the reference's Jacobian (qc3d:285-321) has no launch parameters to choose. */
int gm_jacobian_autotune(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, const void* D,
                         int tangent_fp32, const uint8_t* freem, void* vals, int vals_fp32, void* work,
                         size_t work_bytes, int reps, int* best_bins, int* best_gather, float* ms,
                         hipStream_t stream);

/* Load-step sensitivity of the 3D u-p Neo-Hookean form to Young's modulus: per
   particle (bin-sorted row kq, particle sorted[kq]) the derivative of its stress
   vector Y with respect to E at fixed nodal unknowns, dP/dE = -V A^T sigma / E,
   dT_z/dE = -pen_force / E, dS/dE = -S / E, zero elsewhere (the stress and the
   penalty are proportional to E, the PPP coefficient to 1/E).  Gathered like the
   residual it is dR/dE of the adjoint dL/dE = dL/dE|_u + lambda^T dR/dE.  The
   reference differentiates only the explicit term (qc3d:323-325); no counterpart
   for the rest. */
int gm_stage_de_3d(int n, const int* sorted, const double* F_old, const double* F_new, const double* sigma,
                const double* pen_force, const double* Y, double p_vol, double E, double* dY, hipStream_t stream);

/* 3D consistent-mass matrix M_IJ = sum_p m w_I w_J (mpm_utils.py:327-383) in the
   pattern, assembled per particle bin into `work` (at least gm_mass_work_bytes(g):
   n_nodes x 729 doubles) and gathered per row in a fixed order; gm_p2g then
   leaves mvals null and fills only the activation flags and the right-hand side. */
size_t gm_mass_work_bytes(const gm_grid* g);
int gm_mass_binned(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, double p_mass, double* mvals,
                   void* work, size_t work_bytes, hipStream_t stream);

/* 3D u-p P2G without the mass matrix (P2G_coupled_3d, mpm_utils.py:260-386: activation
   where w_I > threshold, the pressure-boundary dof of every node at or above a flagged
   particle among the 27 of its stencil, and mrhs_I = sum_p w_I m_p p_p sum_J w_J), the
   same flags and right-hand side as gm_p2g with mvals null, summed per particle bin into
   `work` (at least gm_p2g_work_bytes(g): n_nodes x 27 doubles and bytes) and gathered per
   node in a fixed order.  act and bflag are only set, as by gm_p2g; mrhs may be null
   (flags only), else p_part (particle pressures by particle id) is required. */
size_t gm_p2g_work_bytes(const gm_grid* g);
int gm_p2g_binned(const gm_grid* g, const gm_bins* bins, double threshold, const uint8_t* pbc_flag,
                  double node_h_last, double p_mass, const double* p_part, uint8_t* act, uint8_t* bflag,
                  double* mrhs, void* work, size_t work_bytes, hipStream_t stream);

/* Preconditioner copy of the 3D u-p Newton matrix: adds coef * w_I * w_J to
   the zz entry of block (I, J) for every
   particle the indenter penalises (z >= height inside the footing box
   footing = {min_x, max_x, min_y, max_y}; predicate of mpm_utils.py:1242),
   over the node rows of the box box_lo[a] .. box_lo[a] + box_cnt[a] - 1 and
   free z-dofs only.  coef = 2 pen_par reverses the sign of the reference's
   penalty term in K = -dR/du (mpm_utils.py:1240-1246, which makes K indefinite
   under the indenter).  vals is a copy of the gm_node_jacobian output for
   the multigrid hierarchy; the Krylov operator keeps K (gm_gmres_mg op_vals).
   No reference counterpart (the reference solves with a direct or Jacobi-
   BiCGStab solver, qc3d:335-363). */
int gm_penalty_pc(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, const uint8_t* freem, void* vals,
                  int vals_fp32, const int* box_lo, const int* box_cnt, const double* footing, double height,
                  double coef, hipStream_t stream);

/* P2G: activation flags (threshold < 0: every support node, P2G_2d
   mpm_utils.py:83-136; else w > threshold, P2G_coupled_2d/3d :209-223,
   :341-347), pressure-Dirichlet flags above flagged particles (:225-235,
   :349-352; node_h_last = dy in 2D, dz in 3D), and optionally the consistent
   mass matrix in the pattern (b = 1) and its rhs (:237-257, :354-383). */
int gm_p2g(const gm_grid* g, const gm_bins* bins, const gm_pattern* pat, double threshold, const uint8_t* pbc_flag,
           double node_h_last, double p_mass, const double* p_part, uint8_t* act, uint8_t* bflag, double* mvals,
           double* mrhs, hipStream_t stream);

/* freem[i] = !bflag[i] && act[i] over n dofs. */
int gm_free_mask(int n, const uint8_t* bflag, const uint8_t* act, uint8_t* freem, hipStream_t stream);

/* P2G_set_nodal_solution / _3d (mpm_utils.py:387-414). */
int gm_set_nodal_pressure(int n_nodes, int nb, const uint8_t* bflag, const uint8_t* act, int require_active,
                          const double* xp, double* old_solution, double* new_solution, hipStream_t stream);

/* G2P_2d / G2P_coupled_2d / G2P_coupled_3d (mpm_utils.py:1351-1590): move
   particles, commit F, b_e and the hardening variable, interpolate pressure. */
int gm_g2p(const gm_grid* g, int n, double* x, const double* lp, const double* u_new, const double* u_old,
           double* F_old, const double* F_new, double* b_old, const double* b_new, double* alpha_old,
           const double* alpha_new, double* p_part, int coupled, hipStream_t stream);

/* update_GIMP_lp_2d / _3d (gimp.py:428-507).  dmin, dmax: host arrays of 3. */
int gm_update_lp(int n, int dim, const double* F_new, double* lp, double lp0, const double* x, const double* dmin,
                 const double* dmax, hipStream_t stream);

/* Block-Jacobi BiCGStab on a node block-CSR matrix (nb = 1..4), FP64
   (fp32 = 0) or FP32 (fp32 = 1) matrix and vectors, FP64 reductions; x0 = 0;
   stops when ||r|| <= tol ||b|| (checked every check_every iterations) or
   after maxiter.  iters, relres: host outputs.  Replaces wps.bsr_set_from_triplets
   + wp.optim.linear.preconditioner('diag') + bicgstab (qs2d:201-204,
   qc3d:335-338, 410-412) and the host SciPy / pyamg / pypardiso options
   (qs2d:205-220, qc3d:339-361). */
int gm_bicgstab(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                const void* b, void* x, gm_krylov_ws* ws, double tol, int maxiter, int check_every, int* iters,
                double* relres, hipStream_t stream);
/* Preconditioned conjugate gradients for symmetric positive definite node block-CSR
   systems (the consistent-mass projection), same arguments, workspace and stopping rule
   as gm_bicgstab: one SpMV and two reductions per iteration, every scalar on the device.
   options bit 0: the inverse row sums as the preconditioner
   (nb = 1 only; for a consistent-mass matrix the lumped nodal masses) instead of the
   block-Jacobi inverse; bit 1: warm start from the x passed in (else x0 = 0); the test
   stays ||r|| <= tol ||b||.  Stops early, with relres above tol, when p.Ap is not
   positive (the matrix is not positive definite).  Replaces the reference's
   Jacobi-BiCGStab of the mass projection (qc3d:410-412, qc2d:288-293). */
int gm_pcg(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals, const void* b,
           void* x, gm_krylov_ws* ws, double tol, int maxiter, int check_every, int options, int* iters,
           double* relres, hipStream_t stream);


/* Restarted right-preconditioned GMRES(m) with block-Jacobi and classical Gram-Schmidt
   Arnoldi (reorthogonalised when a pass removes more than half of the vector's norm) on
   the same matrices (the robust alternative to BiCGStab, no breakdown).  V:
   (m+1)*n_rows*nb entries of the solve precision, z and w: n_rows*nb each,
   dinv: n_rows*nb*nb, part: >= max(64*(m+1), 2048) doubles, dh: >= 3*(m+2)
   doubles.  Replaces the same reference solves as gm_bicgstab. */
int gm_gmres(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals, const void* b,
             void* x, void* V, void* z, void* w, void* dinv, double* part, double* dh, int m, double tol, int maxiter,
             int* iters, double* relres, hipStream_t stream);

/* y = A x on the block-CSR matrix (FP64 accumulation; FP32 or FP64 operands). */
int gm_bsr_spmv(int nb, int mat_fp32, int vec_fp32, int n_rows, const int* row_ptr, const int* col_idx,
                const void* vals, const void* x, void* y, hipStream_t stream);

/* out = b - A x in FP64 with A possibly FP32 (iterative-refinement residual);
   ax: n_rows*nb doubles of scratch. */
int gm_residual_refine(int nb, int mat_fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                       const double* b, const double* x, double* ax, double* out, hipStream_t stream);

/* Deterministic reductions into a device scalar: squared 2-norm and sum
   (the reference's np.linalg.norm(rhs.numpy()), qs2d:227-231, and the
   indentation-force atomics, mpm_utils.py:1247-1248).  part >= 2048 doubles. */
int gm_norm2(long n, const double* v, double* part, double* out, hipStream_t stream);
int gm_sum(long n, const double* v, double* part, double* out, hipStream_t stream);

/* gm_norm2 without a workspace argument: no counterpart in the reference; for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops).  The partial sums live in a static 16 KiB device array
   of the library (one per GPU, kept until the process ends); the result is
   bitwise that of gm_norm2.  0 <= n <= 2^40, else -1 and gm_last_error().
   Calls that overlap on one GPU (unordered streams) race on that array. */
int gm_norm2_noalloc(long n, const double* v, double* out, hipStream_t stream);

/* y += a*x (x FP64 or FP32): from_increment_to_solution (mpm_utils.py:1343-1348). */
int gm_axpy(long n, double a, const void* x, int x_fp32, double* y, hipStream_t stream);

/* FP64 -> FP32 copy. */
int gm_d2f(long n, const double* a, float* b, hipStream_t stream);

/* Identity rows for empty diagonals of a scalar (nb = 1) matrix and zero rhs
   there (nodes that no particle weights; the reference leaves them empty in
   the consistent-mass system, qc2d:288, qc3d:410). */
int gm_fix_empty_diag(int n_rows, const int* row_ptr, const int* col_idx, double* vals, double* rhs,
                      hipStream_t stream);

/* Per-particle cpGIMP weights and gradients in the reference's output order
   (get_GIMP_shape_function_and_gradient_2d/3d, gimp.py:25-112, 197-315):
   out[p] = (w0 per axis, w1 per axis, w2 per axis, g0 per axis, g1, g2), 6*dim
   doubles; avg (may be null) likewise for the cell-average variants
   (gimp.py:115-193, 318-425); base (may be null) = first stencil node per axis.
   h: host array of dim spacings (2D: dx twice, as gimp.py:26-31). */
int gm_gimp_weights(int n, int dim, const double* x, const double* lp, const double* h, double* out, double* avg,
                    int* base, hipStream_t stream);

/* Standalone principal-strain return mapping, one particle per lane:
   mat_kind 1 = J2 (plasticity.py:8-37, with linear hardening H),
   3 = Drucker-Prager (triaxial_DruckerPrager.py:308-412; dp_apex 0 = the
   reference's no-return for P_trial >= 0, 1 = apex return).  eps_tr, eps: n x 3;
   dE (may be null): n x 9 consistent tangent d eps / d eps_tr; alpha_old/new
   (may be null): J2 equivalent plastic strain. */
int gm_return_map(int n, int mat_kind, int dp_apex, double lam, double mu, double kappa, double H, double phi,
                  double psi, double coh, double shape, const double* eps_tr, double* eps, double* dE,
                  const double* alpha_old, double* alpha_new, hipStream_t stream);

/* NorSand principal-strain return mapping, one particle per lane
   (return_mapping_NorSand, triaxial_NorSand.py:243-400): eps_tr n x 3 trial
   principal strains, eps_total_trace n total volumetric strains (specific
   volume update), pi_old n image pressures; writes eps (n x 3), pi_new (n) and,
   if res_last is not null, the 4 local residuals of the last Newton iterate
   (n x 4, the reference's saved_local_residual[14]). */
int gm_return_map_norsand(int n, const gm_norsand_params* prm, const double* eps_tr, const double* eps_total_trace,
                          const double* pi_old, double* eps, double* pi_new, double* res_last, hipStream_t stream);

/* ---------------------------------------------------------------- adjoint (2D plane strain)
   No single reference function corresponds: the reference differentiates its
   load step with the Warp tape (qs2d:158-192 reverse sweeps; qc3d:270-327 for
   the E gradient).  These entry points are the pieces of a hand-written
   adjoint of quasi_static_solver_2d's converged load step. */

/* Per-particle local map (residual contributions assemble_residual_2d_*,
   mpm_utils.py:416-780; G2P_2d :1351-1419; update_GIMP_lp_2d gimp.py:428-462)
   evaluated in double: acc[p] = sum_I cot_I . r_I^p (may be null); state (may be
   null): 18 doubles per particle = x_new(2), lp_new(2), F_new(xx,xy,yx,yy,zz),
   b_new(xx,xy,yy,zz), alpha_new, sigma(xx,xy,yy,zz). */
int gm_local_eval_2d(const gm_local_args* args, double* acc, double* state, hipStream_t stream);

/* Vector-Jacobian products of the local map: g[p*25 + d] = d/d(input d) of
   sum_I cot_I . r_I^p + xbar.x_new + lpbar.lp_new + Fbar:F_new + bbar:b_new +
   abar alpha_new, for the 25 inputs 0-1 x, 2-3 lp, 4-8 F_old(xx,xy,yx,yy,zz),
   9-12 b_old(xx,xy=yx,yy,zz), 13 alpha_old, 14-17 H(xx,xy,yx,yy), 18-24
   E, nu, kappa, H, phi, psi, cohesion.  Forward-mode dual numbers, one
   direction per input. */
int gm_local_vjp_2d(const gm_local_args* args, double* g, hipStream_t stream);

/* K^T in K's block pattern (structurally symmetric, ascending columns per row),
   FP64 (fp32 = 0) or FP32 values; used for the transposed (adjoint) solve. */
int gm_bsr_transpose(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                     void* vals_t, hipStream_t stream);

/* ---------------------------------------------------------------- geometric multigrid
   Multilevel preconditioner for the Newton systems.  The reference's multilevel
   option is pyamg smoothed aggregation + BiCGStab on the host
   (qs2d:205-213, qc2d:288-293); here: geometric coarsening of the background
   grid (every second node per axis), trilinear prolongation restricted to free
   dofs, Galerkin coarse operators P^T A P built on the device, damped
   block-Jacobi smoothing. */

/* Build levels 1..n_levels-1 from level 0 (whose pattern, values and free mask
   are the Jacobian's) and every level's block-diagonal inverse.  Coarse
   patterns (row_ptr, col_idx, row_mask: full 5^dim windows clipped to the
   grid) and all buffers are the caller's.  Level 0 holds the solve precision;
   the Galerkin products, the coarse operators (levels 1.. vals) and a coarsest
   dense inverse are FP64 in both precisions (the smoothers of an FP32 hierarchy
   run in FP32, or from the FP16 copies when vals_lo is set).  ap_work: level-0
   n_nodes * 4^dim * nb * nb doubles. */
int gm_mg_setup(int dim, int nb, int fp32, int n_levels, gm_mg_level* levels, void* ap_work, hipStream_t stream);

/* x = M^-1 b for one V-cycle (nu_pre / nu_post damped block-Jacobi sweeps with
   factor omega, coarse_sweeps sweeps on the coarsest level). */
int gm_mg_vcycle(int dim, int nb, int fp32, int n_levels, const gm_mg_level* levels, int nu_pre, int nu_post,
                 int coarse_sweeps, double omega, const void* b, void* x, hipStream_t stream);

/* One fused smoother-side operation on a single level, as the V-cycle runs it: mode 1
   out = b - A x, mode 2 (damped block-Jacobi sweep) out = x + omega Dinv (b - A x), from the
   level's FP16 copy when it has one.  level_index 0 means the level matrix holds the solve
   precision, > 0 FP64.  No reference counterpart (the reference has no multigrid). */
int gm_mg_level_op(int nb, int fp32, const gm_mg_level* level, int level_index, int mode, const void* x,
                   const void* b, double omega, void* out, hipStream_t stream);

/* Flexible GMRES(m) right-preconditioned by the V-cycle.  The operator is
   op_vals (in level 0's pattern and precision) when it is not null, else level
   0's own matrix (op_vals lets the hierarchy be built from a modified copy of
   the system, see gm_penalty_pc); other arguments as gm_gmres (V: (m+1) * n
   vectors, Z: m * n preconditioned directions, z, w: n; part >= 64(m+1) and
   >= 2048 doubles; dh >= 3(m+2) doubles). */
int gm_gmres_mg(int dim, int nb, int fp32, int n_levels, const gm_mg_level* levels, const void* op_vals,
                int nu_pre, int nu_post, int coarse_sweeps, double omega, const void* b, void* x, void* V, void* Z,
                void* z, void* w, double* part, double* dh, int m, double tol, int maxiter, int* iters,
                double* relres, hipStream_t stream);
/* Arnoldi step graphs of the GMRES solvers (gm_gmres, gm_gmres_mg, gm_gmres_up): each
   step's launches (preconditioner, operator, Gram-Schmidt passes, normalisation, copy of
   the Hessenberg column to pinned host memory) are captured once per step index as a HIP
   graph and replayed, keyed on every pointer and size they use.  on = 1 / 0 turns them on
   or off for later solves (default on; GM_KRYLOV_GRAPHS=0 in the environment starts off),
   -1 only reads the setting; returns the previous setting.  Both give the same bits.  No
   reference counterpart (the reference's solve is Warp's bicgstab, qc3d:335-338). */
int gm_krylov_graphs(int on);


/* u-p systems (nb = dim + 1): block lower-triangular preconditioner
   z_u = MG(A_uu)^-1 v_u, z_p = S^-1 (v_p - K_pu z_u), S the diagonal Schur
   approximation C - D blockdiag(A)^-1 G.  gm_up_setup extracts the displacement
   block into ulevels[0] (which shares the system's row_ptr/col_idx/row_mask and
   owns vals/freem with nb-1 dofs per node), builds the hierarchy and writes
   sinv (n_rows); gm_gmres_up runs flexible GMRES(m) with it (vu, zu: n_rows *
   (nb-1) work vectors; other arguments as gm_gmres_mg). */
int gm_up_setup(int dim, int nb, int fp32, int n_rows, int nnzb, const int* row_ptr, const int* col_idx,
                const void* vals, const uint8_t* freem, int n_levels, gm_mg_level* ulevels, void* ap_work,
                void* sinv, hipStream_t stream);
int gm_gmres_up(int dim, int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                const void* sinv, int n_levels, const gm_mg_level* ulevels, int nu_pre, int nu_post,
                int coarse_sweeps, double omega, void* vu, void* zu, const void* b, void* x, void* V, void* Z,
                void* z, void* w, double* part, double* dh, int m, double tol, int maxiter, int* iters,
                double* relres, hipStream_t stream);

/* Row equilibration of the Newton system: K <- diag(1/|K_ii|) K in place and
   rhs_out = diag(1/|K_ii|) rhs_in (the solution is unchanged; displacement
   and pressure rows then weigh alike in the Krylov residual).  No reference
   counterpart: the reference hands K to a direct or Jacobi-BiCGStab solver. */
int gm_row_scale(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, void* vals,
                 const double* rhs_in, double* rhs_out, hipStream_t stream);

/* SpMV autotuner (no reference counterpart): times gm_bsr_spmv's kernel on the
   given matrix for workgroups of 64, 128, 256 and 512 threads (reps launches
   each, on the stream), keeps the fastest for later SpMVs with this block size
   and precision (process-wide) and returns it in *best with the four average
   times (ms[4], milliseconds).  The choice never changes a result bit. */
int gm_spmv_autotune(int nb, int fp32, int n_rows, const int* row_ptr, const int* col_idx, const void* vals,
                     const void* x, void* y, int reps, int* best, float* ms, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* GEOWARP_AMD_H */

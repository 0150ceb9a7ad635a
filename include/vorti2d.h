/*
 * vorti2d.h -- C ABI of libvorti2d, the MI355X (gfx950) metric / assembly /
 * Krylov hot path of the vorticity-streamfunction Newton solver.
 *
 * Every function returns 0 on success and -1 on failure; v2d_last_error()
 * then holds a message.  Pointers are plain host pointers unless the name ends
 * in _dev.  Arrays follow the reference conventions: mesh arrays (imax, jmax)
 * column-major (i fastest), field arrays of length N = imax*jmax in node order
 * k = imax*j + i, Newton system field-blocked (psi rows/cols [0,N), omega
 * [N,2N)), COO indices 0-based int32, values float64.
 */
#ifndef VORTI2D_H
#define VORTI2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failure on this thread ("" if none). */
const char *v2d_last_error(void);
/* Library ABI version (major*100 + minor). */
int v2d_version(void);

/* ------------------------------------------------------------------------
 * f2py-equivalent core entry points (host pointers).  Together they replace
 * the compiled module vorti2d._core.vorti2d_core (vorti2d/__init__.py:13-15).
 * ---------------------------------------------------------------------- */

/* Replaces node_neighbors (src/fortran/vorti2d_core.f90:33-61).  1-based i, j;
 * out = k, e, w, n, s, ne, nw, se, sw (1-based pointers, seam wrapped). */
int v2d_node_neighbors(int i, int j, int imax, int out[9]);

/* Replaces compute_metrics (vorti2d_core.f90:66-189; called at
 * vorti2d/solver.py:72-74).  xg, yg: (imax, jmax) column-major.  out: 10*N
 * doubles, planes jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys,
 * yphys.  Requires imax >= 4, jmax >= 4. */
int v2d_compute_metrics(int imax, int jmax, double dksi, double deta, const double *xg,
                        const double *yg, double *out);

/* Device-pointer, stream-ordered form of v2d_compute_metrics.  stream is a
 * hipStream_t (NULL = default stream). */
int v2d_compute_metrics_dev(int imax, int jmax, double dksi, double deta, const double *d_xg,
                            const double *d_yg, double *d_out, void *stream);

/* Replaces assemble_coo (vorti2d_core.f90:219-383; called at
 * vorti2d/solver.py:121-127).  metrics: 10 pointers in compute_metrics order.
 * Rows [r0, r1) of the field-blocked system in the reference emission order.
 * Unlike the reference, the capacity maxnnz is checked (error, nothing
 * written) and *nnz is the number of entries written. */
int v2d_assemble_coo(int imax, int jmax, double re, double invdtau, double inv2dt, double urot,
                     double dksi, double deta, const double *const *metrics, const double *psi,
                     const double *ome, const double *omeold, const double *omeoldold, int r0,
                     int r1, int maxnnz, int ff_bc, double ca, double sa, int *coo_i, int *coo_j,
                     double *coo_v, int *nnz, double *bvec);

/* Replaces compute_metrics_local (vorti2d_core.f90:392-495; called at
 * vorti2d/dist_solver.py:92-93).  il0, il1 are 1-based owned columns. */
int v2d_compute_metrics_local(int gxm, int jmax, int il0, int il1, double dksi, double deta,
                              const double *xgl, const double *ygl, double *out);

/* Replaces assemble_coo_local (vorti2d_core.f90:510-661; called at
 * vorti2d/dist_solver.py:310-317).  lg: 2*gxm*jmax int32 local-to-global map. */
int v2d_assemble_coo_local(int gxm, int jmax, int il0, int il1, double re, double invdtau,
                           double inv2dt, double urot, double dksi, double deta,
                           const double *const *metrics, const double *psi, const double *ome,
                           const double *omeold, const double *omeoldold, int r0, int nloc,
                           const int *lg, int maxnnz, int ff_bc, double ca, double sa,
                           int *coo_i, int *coo_j, double *coo_v, int *nnz, double *bvec);

/* Replaces compute_force_coeffs (vorti2d/forces.py:82-194).  Arrays of length
 * N.  ref_length <= 0 selects the wall-estimated diameter.  out: cd, cl, cm,
 * cdp, cdf, clp, clf, cmp, cmf, ref_length. */
int v2d_force_coeffs(int imax, int jmax, double re, double dksi, double deta,
                     const double *beta, const double *gama, const double *detadx,
                     const double *detady, const double *xphys, const double *yphys,
                     const double *ome, double ref_length, double x0, double y0, double *out);

/* The force coefficients of v2d_force_coeffs (same arguments, meaning, errors
 * and results: the same kernel) without any device allocation.  No
 * counterpart in the reference; it is for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  The wall
 * rows the integral reads are staged in a fixed device array of the library,
 * 13 * V2D_NOALLOC_MAX_IMAX + 16 doubles (1.7 MB) on every GPU it runs on,
 * held until the process ends; imax > V2D_NOALLOC_MAX_IMAX is rejected.
 * Calls that overlap on one GPU race on that array. */
#define V2D_NOALLOC_MAX_IMAX 16385
int v2d_force_coeffs_noalloc(int imax, int jmax, double re, double dksi, double deta,
                             const double *beta, const double *gama, const double *detadx,
                             const double *detady, const double *xphys, const double *yphys,
                             const double *ome, double ref_length, double x0, double y0,
                             double *out);

/* ------------------------------------------------------------------------
 * Device-resident solver context.  Replaces the per-iteration host work of
 * Solver._assemble_and_solve_once (vorti2d/solver.py:117-134): the assembly
 * (core.assemble_coo), the COO->CSR pattern of MumpsSolver.assemble
 * (vorti2d/petsc_solver.py:76-95) and the MUMPS solve (:97-115).  Mesh,
 * metrics, state, Jacobian and factorisation stay in HBM; only scalars cross
 * the boundary per Newton step.  One HIP stream per context; calls on one
 * context are not re-entrant.
 * ---------------------------------------------------------------------- */
typedef struct v2d_ctx v2d_ctx;

enum { V2D_FP64 = 0, V2D_MIXED = 1 };

/* Upload the mesh, compute the metrics on the device, build the fixed
 * Jacobian pattern for (ff_bc, ca, sa) and allocate the solver.  precision:
 * V2D_FP64 (FP64 Jacobian, exact FP64 block LU: the reference's exact Newton
 * step) or V2D_MIXED (FP32 Jacobian and block-LU preconditioner, FP32 GMRES,
 * FP64 residual and state). */
int v2d_ctx_create(v2d_ctx **out, int imax, int jmax, const double *xg, const double *yg,
                   int ff_bc, double ca, double sa, int precision);
/* A new context in the given precision on src's mesh and far-field setting,
 * with src's state (psi, omega and the two older omega planes), after src's
 * queued work.  Mesh, state and the Jacobian pattern are copied on the device
 * (no host pattern build); nothing is shared, so either context may be
 * destroyed first.  No counterpart in the reference: for an FP64 adjoint
 * context beside a mixed-precision solve. */
int v2d_ctx_create_like(v2d_ctx **out, v2d_ctx *src, int precision);
int v2d_ctx_destroy(v2d_ctx *ctx);
/* Bytes of device memory held by the context. */
int v2d_ctx_device_bytes(v2d_ctx *ctx, int64_t *bytes);
/* D2H copy of the 10 metric planes (10*N doubles). */
int v2d_ctx_metrics(v2d_ctx *ctx, double *out);
/* State upload / download; any pointer may be NULL to skip that plane. */
int v2d_ctx_set_state(v2d_ctx *ctx, const double *psi, const double *ome, const double *omeold,
                      const double *omeoldold);
int v2d_ctx_get_state(v2d_ctx *ctx, double *psi, double *ome, double *omeold, double *omeoldold);
/* Set the initial state of vorti2d/solver.py:103-114 on the device: zeros,
 * far-field psi = ca*y - sa*x. */
int v2d_ctx_init_state(v2d_ctx *ctx);
/* Linear-solver knobs for V2D_MIXED: refactor the FP32 block LU every
 * refactor_every Newton steps (0 = only when GMRES needs more than
 * gmres_max_iter iterations), inner relative tolerance, iteration cap. */
int v2d_ctx_set_solver(v2d_ctx *ctx, int refactor_every, double gmres_rtol, int gmres_max_iter);
/* Inexact-Newton forcing for V2D_MIXED (no counterpart in the reference, whose
 * Krylov solves use a fixed ksp_rtol): the GMRES tolerance of Newton step k is
 * max(gmres_rtol, eta_k) with eta_k = min(eta_max, 0.9 (||R_k|| /
 * ||R_k-1||)^2) and the Eisenstat-Walker safeguard; with r_target > 0 (an
 * absolute ||R||_2 the caller is content with), eta_k >= 0.5 r_target /
 * ||R_k|| as well (up to eta_max), so that steps close to the target do not
 * solve more accurately than the target needs.  eta_max <= 0 restores the
 * fixed tolerance.  The history restarts with every new state. */
int v2d_ctx_set_forcing(v2d_ctx *ctx, double eta_max, double r_target);
/* Kernel of the block-LU Gauss-Jordan update: VALU on 64- or 128-wide tiles,
 * or the matrix cores (MFMA 16x16x4, FP32 or FP64) on 64-wide tiles (the
 * default).  A tuning knob with no counterpart in the reference: results do not
 * depend on it beyond floating-point rounding order.  Drops a stored
 * factorisation. */
enum { V2D_UPDATE_VALU64 = 0, V2D_UPDATE_VALU128 = 1, V2D_UPDATE_MFMA64 = 2 };
int v2d_ctx_set_update_kernel(v2d_ctx *ctx, int kind);
/* Kernel of the block-LU solve sweeps: two launches per level step
 * (V2D_SWEEP_LEVELS), or one persistent launch for the whole solve
 * (V2D_SWEEP_PERSISTENT: one workgroup per compute unit, each owning a slice
 * of every level and forming each level's whole right-hand side itself,
 * synchronised through per-workgroup progress words).  Default: persistent
 * for FP32 solves when its two chains fill the compute units, launch pairs
 * otherwise and for FP64.  The persistent launch needs all its workgroups
 * resident at once; if another persistent launch holds the GPU its bounded
 * waits expire and the solve is redone with per-level launches, which the
 * context then keeps.  V2D_SWEEP_FUSED runs every FP32 level step
 * whose right-hand side couples one neighbouring level as one launch (each
 * workgroup forms the right-hand side itself) and the other steps as launch
 * pairs; FP64 solves use launch pairs.  A tuning knob with no counterpart in
 * the reference: all compute the same right-hand sides, and the same sums in
 * the same order except for the rows of a level past 16 per compute unit,
 * which the fused step sums in sixteenths. */
enum { V2D_SWEEP_LEVELS = 0, V2D_SWEEP_PERSISTENT = 1, V2D_SWEEP_FUSED = 2 };
int v2d_ctx_set_sweep_kernel(v2d_ctx *ctx, int kind);
/* Bound of each wait of the persistent sweep, in polls (default 2^21, a few
 * tenths of a second).  When a wait expires the solve is redone with launch
 * pairs, which the context then keeps; 0 makes every wait expire at once,
 * without polling: a way to exercise that path.  No counterpart in the
 * reference. */
int v2d_ctx_set_sweep_spin_limit(v2d_ctx *ctx, int polls);
/* The sweep kernel the context's solves use now (it changes from
 * V2D_SWEEP_PERSISTENT to V2D_SWEEP_LEVELS after a persistent solve whose
 * waits expired) and the number of persistent solves redone with launch pairs
 * so far.  No counterpart in the reference. */
int v2d_ctx_get_sweep_kernel(v2d_ctx *ctx, int *kind, int64_t *fallbacks);
/* Linear solver of the Newton systems (the exact FP64 step, the FP32
 * preconditioner of the mixed path, the adjoint's transposed solves).
 * V2D_SOLVER_MULTIFRONTAL: nested-dissection multifrontal LU (the structure of
 * the reference's MUMPS direct solve, linsolve="mumps":
 * vorti2d/petsc_solver.py:37-62, vorti2d/dist_solver.py:133-137) -- the grid's
 * node graph is dissected by grid lines into a tree of dense fronts, factored
 * height by height on the device.  V2D_SOLVER_BLOCK_LU: the block-tridiagonal
 * LU over radial levels with dense per-level inverses.  Both are exact LUs
 * without pivoting; they differ in rounding.  Selecting a solver drops a
 * stored factorisation; each solver's storage is allocated at its first
 * factorisation. */
enum { V2D_SOLVER_BLOCK_LU = 0, V2D_SOLVER_MULTIFRONTAL = 1 };
int v2d_ctx_set_linear_solver(v2d_ctx *ctx, int kind);
int v2d_ctx_get_linear_solver(v2d_ctx *ctx, int *kind);
/* Schedule of the multifrontal factorisation's Gauss-Jordan steps (no
 * counterpart in the reference; a tuning choice of this library): paired != 0
 * (the default) runs them in pairs, a step that updates only the tiles which
 * form the next pivot panels followed by a step that applies both steps'
 * panels to every other tile in one rank-128 product (one pass over each
 * front per two steps); 0 runs every step over the whole front.  The two
 * schedules give the same factorisation up to rounding.  Takes effect at the
 * next factorisation. */
int v2d_ctx_set_step_pairing(v2d_ctx *ctx, int paired);
/* Multifrontal fronts of a single 64-pivot Gauss-Jordan step (no counterpart
 * in the reference; a tuning choice of this library): fused != 0 (the
 * default) assembles each such front, adds the Jacobian entries it owns and
 * sweeps it in one workgroup of one launch per tree height; 0 takes them
 * through the level assembly, entry scatter and step launches like the larger
 * fronts.  Same factorisation up to rounding.  Takes effect at the next
 * factorisation. */
int v2d_ctx_set_front_fusion(v2d_ctx *ctx, int fused);
/* Sizes of the multifrontal analysis of a context's grid (built on first
 * call): out[0..7] = fronts, tree heights, fronts swept by the matrix-core
 * Gauss-Jordan steps, front storage (elements), front-vector elements,
 * Gauss-Jordan panel elements, Jacobian entries, distinct front positions they
 * fill; *flops = floating-point operations of one factorisation.  No
 * counterpart in the reference (MUMPS reports its own analysis statistics). */
int v2d_ctx_mf_info(v2d_ctx *ctx, int64_t *out, double *flops);

/* Factorise the last assembled Jacobian and solve one system with the
 * multifrontal LU, timing every launch with HIP events on the context's
 * stream.  out[4 g .. 4 g + 3] for the kernel groups g = 0..10 (front assembly,
 * entry scatter, small-front sweep, first Gauss-Jordan step (pivot inverse,
 * pivot rows and panels), full Gauss-Jordan steps, light steps, heavy steps,
 * one-step fronts assembled and swept in one launch, forward gather, forward
 * front products, backward front products): milliseconds, launches,
 * algorithmic flops, algorithmic bytes; out[44], out[45]: the factorisation's
 * and the solve's event spans (ms).  The
 * factorisation is kept.  No counterpart in the reference. */
int v2d_ctx_profile_mf(v2d_ctx *ctx, double *out);

/* Copy of the factored fronts (front storage elements, FP64; each front's
 * nf x ld block row-major at its offset fo, holding [G W; X C]), for
 * inspection.  No counterpart in the reference. */
int v2d_ctx_mf_fronts(v2d_ctx *ctx, double *F);

/* The multifrontal analysis alone, on the host (no device call), for
 * inspection: ff_flags[i] are the far-field outflow bits of node (i, jmax-1)
 * (1: omega row couples j-1, j-2; 2: psi row does; NULL: none).  No
 * counterpart in the reference. */
typedef struct v2d_mf_analysis v2d_mf_analysis;
int v2d_mf_analyse(int imax, int jmax, const uint8_t *ff_flags, v2d_mf_analysis **out);
/* out[0..15]: fronts, heights, big fronts, front storage elements, front
 * vector elements, pivot-vector elements, front-unknown entries, child-map
 * entries, child-list entries, distinct assembly positions, Jacobian entries,
 * rows-table entries, 0, 0, 0, 0; *flops as in v2d_ctx_mf_info. */
int v2d_mf_sizes(const v2d_mf_analysis *a, int64_t *out, double *flops);
/* fronts: 14 int64 per front (fo, vo, zo, uo, mo, ns, nsp, nb, nf, ld, nch,
 * co, big, height); fu: the field-blocked unknown of every front row (-1 on
 * pivot padding); ch: child lists; map: per front and child, the child's
 * coupling row of every parent front row (-1: none); aoff: the front-storage
 * offset of every assembly position, aseg (positions + 1) its range in avid,
 * avid: slot-layout value indices (the value plane times N plus the node).
 * Any pointer may be NULL. */
int v2d_mf_export(const v2d_mf_analysis *a, int64_t *fronts, int *fu, int *ch, int *map, int64_t *aoff,
                  int *aseg, int *avid);
int v2d_mf_destroy(v2d_mf_analysis *a);
/* Mixed precision, refactor_every = 0: when GMRES on a stored FP32
 * factorisation gives up.  V2D_REFACTOR_FIXED (the default): once the
 * geometric projection of its convergence exceeds gmres_max_iter iterations,
 * a function of the iterates alone, so the Newton trajectory does not depend
 * on GPU load.  V2D_REFACTOR_MEASURED: once it exceeds (last factorisation
 * time / mean GMRES iteration time) + the iterations a fresh factorisation
 * needed, both measured on the context's stream (cheaper where refactoring is
 * cheap, but load-dependent).  No counterpart in the reference. */
enum { V2D_REFACTOR_FIXED = 0, V2D_REFACTOR_MEASURED = 1 };
int v2d_ctx_set_refactor_policy(v2d_ctx *ctx, int policy);
/* Mixed precision: how GMRES launches its device work.  V2D_GRAPH_APPLY (the
 * default): the FP32 preconditioner apply of every iteration is one replay of
 * a HIP graph captured over fixed device buffers (multifrontal LU, or the block
 * LU's per-level sweeps), instead of its launches one by one.
 * V2D_GRAPH_ITERATION: the whole iteration -- apply, SpMV, both Gram-Schmidt
 * rounds, the norm, the scaling of the new basis vector and the copies of the
 * Hessenberg column to pinned host memory -- is one replay of a graph captured
 * per basis index on first use (the block LU's persistent sweep falls back to
 * eager launches).  V2D_GRAPH_OFF: every kernel launched eagerly.  Same
 * kernels in the same order in every mode, so the same results bit for bit;
 * graphs take the launches off the host thread.  Changing the mode drops the
 * captured graphs.  No counterpart in the reference. */
enum { V2D_GRAPH_OFF = 0, V2D_GRAPH_APPLY = 1, V2D_GRAPH_ITERATION = 2 };
int v2d_ctx_set_graph(v2d_ctx *ctx, int mode);
/* Factorise the last assembled Jacobian once with every Gauss-Jordan update
 * launch bracketed by HIP events on the context's stream (no counterpart in the
 * reference; measurement only).  out[5] = {update launches, summed update
 * kernel milliseconds, algorithmic flops of those launches, algorithmic bytes
 * of those launches, wall milliseconds of the whole factorisation}. */
int v2d_ctx_profile_factor(v2d_ctx *ctx, double *out);

/* Device assembly of the Jacobian values and the FP64 residual-form RHS for
 * the current state (the scalars of assemble_coo). */
int v2d_ctx_assemble(v2d_ctx *ctx, double re, double invdtau, double inv2dt, double urot);
/* ||rhs||_2 in FP64 of the last assembly. */
int v2d_ctx_rhs_norm(v2d_ctx *ctx, double *out);
/* Number of Jacobian entries (equals the reference nnz). */
int v2d_ctx_nnz(v2d_ctx *ctx, int64_t *nnz);
/* Export of the last assembly in the reference COO order (all 2N rows) and
 * the RHS; values widened to FP64. */
int v2d_ctx_export_coo(v2d_ctx *ctx, int *coo_i, int *coo_j, double *coo_v, double *bvec);
/* Canonical CSR of the last assembly (columns ascending, 2N+1 row pointers). */
int v2d_ctx_export_csr(v2d_ctx *ctx, int64_t *indptr, int *indices, double *data);
/* y = A x with the last assembly (host vectors of 2N doubles). */
int v2d_ctx_spmv(v2d_ctx *ctx, const double *x, double *y);
/* Factor the last assembly with the block LU of the context precision. */
int v2d_ctx_factor(v2d_ctx *ctx);
/* x = M b with the current factorisation (host vectors of 2N doubles). */
int v2d_ctx_solve(v2d_ctx *ctx, const double *b, double *x);

/* One inner iteration of _assemble_and_solve_once (vorti2d/solver.py:117-134)
 * on the device: assemble, solve A d = b, psi += d[:N], ome += d[N:].
 * res = ||d||_2 / (2N) (the reference convergence measure), rnorm = ||b||_2 of
 * this assembly, lin_iters = GMRES iterations (0 for the direct FP64 step). */
int v2d_ctx_newton_step(v2d_ctx *ctx, double re, double invdtau, double inv2dt, double urot,
                        double *res, double *rnorm, int *lin_iters);
/* BDF2 history shift omeoldold <- omeold <- ome (vorti2d/solver.py:186-188). */
int v2d_ctx_bdf2_shift(v2d_ctx *ctx);
/* Force coefficients of the current state (vorti2d/solver.py:272-283). */
int v2d_ctx_forces(v2d_ctx *ctx, double re, double ref_length, double x0, double y0,
                   double *out);
/* Accumulated device time (ms) of assembly, factorisation, linear solve and
 * whole Newton steps since creation or the last reset; counts[0] = steps,
 * counts[1] = factorisations, counts[2] = GMRES iterations. */
int v2d_ctx_timers(v2d_ctx *ctx, double *ms4, int64_t *counts3, int reset);
/* Average duration (ms) of the last `launches` assembly kernels, measured with
 * HIP events on the context stream around a loop of that many launches. */
int v2d_ctx_time_assembly(v2d_ctx *ctx, double re, double invdtau, double inv2dt, double urot,
                          int launches, double *ms_per_launch);
/* Same for the stencil SpMV y = A x on the last assembly (x = ones). */
int v2d_ctx_time_spmv(v2d_ctx *ctx, int launches, double *ms_per_launch);

/* ---------------------------------------------------------------------
 * Discrete adjoint of the steady force coefficients.  None of these has a
 * counterpart in the reference, which computes no derivatives; they serve the
 * gradient of J = wd * Cd + wl * Cl (Cd, Cl as in vorti2d/forces.py:82-194,
 * body axes, given reference length) with respect to the angle of attack,
 * the Reynolds number and every mesh coordinate:
 *     dJ/dtheta = dJ/dtheta|explicit + lambda^T db/dtheta,  A^T lambda = dJ/dU,
 * with b the residual-form RHS of the assembly and A = -db/dU its Newton
 * matrix.
 * --------------------------------------------------------------------- */
/* x = A^-T b with the current factorisation (host vectors of 2N doubles). */
int v2d_ctx_solve_transpose(v2d_ctx *ctx, const double *b, double *x);
/* g_m (10 planes of N, overwritten) = lambda^T db/d(metric planes) at the
 * current state; gscal[0] = lambda . db/dRe, gscal[1] = lambda . db/dalpha
 * (radians).  lam: host vector of 2N (psi rows then omega rows). */
int v2d_ctx_residual_vjp(v2d_ctx *ctx, double re, double inv2dt, double urot, const double *lam,
                         double *g_m, double *gscal);
/* Metric backward: g_x, g_y (N each, i fastest) = g_m . d(compute_metrics)/d(xg, yg)
 * for the 10 output planes of v2d_compute_metrics (host arrays). */
int v2d_metrics_vjp(int imax, int jmax, double dksi, double deta, const double *xg,
                    const double *yg, const double *g_m, double *g_x, double *g_y);
/* Force backward at the current state: g_ome (N), g_m (10 planes) =
 * dJ/d(omega, metric planes); out3 = {Cd, Cl, explicit dJ/dRe}. */
int v2d_ctx_forces_vjp(v2d_ctx *ctx, double re, double ref_length, double wd, double wl,
                       double *g_ome, double *g_m, double *out3);
/* The whole adjoint at the stored (converged, steady) state of an FP64
 * context: refactorises the Newton matrix, solves A^T lambda = dJ/dU and
 * returns g_x, g_y = dJ/d(xg, yg) (N each), gscal = {dJ/dalpha per degree,
 * dJ/dRe}, coef = {Cd, Cl}. */
int v2d_ctx_adjoint_forces(v2d_ctx *ctx, double re, double urot, double ref_length, double wd,
                           double wl, double *g_x, double *g_y, double *gscal, double *coef);

/* ---------------------------------------------------------------------
 * Lockstep batch of contexts of uneven mesh sizes (no counterpart in the
 * reference, whose steady loop vorti2d/solver.py:136-145 runs one mesh).  A
 * batch has up to V2D_BATCH_MAX slots of one precision; the caller owns the
 * contexts and keeps them alive while they sit in a slot.  One batched Newton
 * step of the active filled slots runs one assembly launch, one residual-norm
 * launch pair, per case the factorisation (when due) and the preconditioner
 * applies on the case's own stream, mixed precision the GMRES iterations in
 * lockstep with one SpMV, two multi-dot / reduce / multi-axpy rounds, one norm
 * round and one normalisation launch per iteration for all iterating cases,
 * and one update launch: every case's iterates equal those of
 * v2d_ctx_newton_step on its own context bit for bit.  Mixed-precision cases
 * need the fixed refactor policy; a case with the persistent sweep kernel is
 * switched to launch pairs (same results).
 * --------------------------------------------------------------------- */
#define V2D_BATCH_MAX 16
typedef struct v2d_batch v2d_batch;
int v2d_batch_create(v2d_batch **out, int precision, int slots);
int v2d_batch_destroy(v2d_batch *b);
/* Put ctx into a slot (NULL empties it). */
int v2d_batch_set_case(v2d_batch *b, int slot, v2d_ctx *ctx);
/* One Newton step of every filled slot with active[slot] != 0 (active NULL:
 * all).  re, invdtau, inv2dt, urot: per-slot scalars of v2d_ctx_newton_step
 * (the last three may be NULL for zeros); res, rnorm, lin_iters: per-slot
 * outputs as there. */
int v2d_batch_newton_step(v2d_batch *b, const int *active, const double *re, const double *invdtau,
                          const double *inv2dt, const double *urot, double *res, double *rnorm,
                          int *lin_iters);
/* Workgroup sizes (threads) of the batched pointwise launches: assembly, SpMV,
 * vector updates.  Results do not depend on them.  autotune times 128, 256,
 * 512 and 1024 threads on the filled slots (re: the assembly's Reynolds
 * number; the slots' Jacobians are reassembled at their current states),
 * keeps the fastest of each and remembers them for batches of the same size
 * class (total nodes rounded up to a power of two). */
int v2d_batch_get_blocks(v2d_batch *b, int *threads3);
int v2d_batch_set_blocks(v2d_batch *b, const int *threads3);
int v2d_batch_autotune(v2d_batch *b, double re, int *threads3);

/* Discrete adjoint of an unsteady BDF2 run (vorti2d/solver.py:169-190: one
 * converged level per physical time step, history shift after each).  No
 * counterpart in the reference.  For J = sum_n (wd_n Cd_n + wl_n Cl_n) over
 * the levels n of the run, the reverse sweep visits the levels from the last
 * to the first; at level n,
 *     A_n^T lambda_n = dJ_n/dU^n + C_{n+1}^T lambda_{n+1} + C_{n+2}^T lambda_{n+2},
 * with A_n = -db_n/dU^n and C_{n+1}, C_{n+2} the dependence of the later
 * levels' BDF2 terms on omega^n (4 inv2dt and -inv2dt on the interior omega
 * rows), and dJ/dtheta = dJ/dtheta|explicit + sum_n lambda_n^T db_n/dtheta.
 * begin clears the sweep (FP64 contexts only).  step: the context's state
 * must hold level n's converged psi, omega and its history omega^(n-1),
 * omega^(n-2) (v2d_ctx_set_state); it refactorises that level's Newton
 * matrix, solves for lambda_n with the two later levels' lambdas it keeps, and
 * accumulates the gradients; wd, wl: this level's objective weights (0, 0:
 * none); coef (may be NULL) = {Cd_n, Cl_n}.  end: g_x, g_y = dJ/d(xg, yg) (N
 * each, i fastest), gscal = {dJ/dalpha per degree, dJ/dRe}. */
int v2d_ctx_adjoint_bdf2_begin(v2d_ctx *ctx);
int v2d_ctx_adjoint_bdf2_step(v2d_ctx *ctx, double re, double inv2dt, double urot, double ref_length,
                              double wd, double wl, double *coef);
int v2d_ctx_adjoint_bdf2_end(v2d_ctx *ctx, double *g_x, double *g_y, double *gscal);

#ifdef __cplusplus
}
#endif

#endif /* VORTI2D_H */

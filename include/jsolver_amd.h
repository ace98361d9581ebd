/*
 * jsolver_amd C ABI: geometric multigrid for  div(D grad phi) - lambda phi = rhs
 * on ghost-padded (m+3)-per-axis grids, as hand-written HIP kernels for
 * MI355X (gfx950).
 *
 * The reference (davef96/jsolver) is a pure Python/JAX package with no FFI;
 * each entry point below replaces one logical operation of its solver path,
 * cited as file:line into the reference.  The Python package jsolver_amd
 * binds these through ctypes (jsolver_amd/_lib.py) behind the reference's own
 * function names (solve_2d, solve_3d, do_BCs, grid_1D).
 *
 * Conventions
 *   - Every function returns 0 on success, otherwise a nonzero status; the
 *     message is available from jsv_last_error() (thread-local).
 *   - Array pointers are device pointers to C-contiguous row-major arrays in
 *     the reference's layout: 2D (my+3, mx+3), 3D (mz+3, my+3, mx+3).
 *     Inputs are never modified; the caller owns them.
 *   - A plan owns its per-level device buffers (allocated once, at creation)
 *     and is bound to no stream: every call takes the stream it runs on.  A
 *     plan is not thread-safe; calls on one plan must not overlap.
 *   - "stream" is a hipStream_t passed as void* (NULL = default stream).
 */
#ifndef JSOLVER_AMD_H
#define JSOLVER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jsv_plan jsv_plan;

/* precision of the multigrid cycles */
#define JSV_FP64 0   /* every cycle in FP64 (the reference's arithmetic) */
#define JSV_MIXED 1  /* FP32 cycles inside FP64 iterative refinement */

/* stop rules */
#define JSV_STOP_DISTANCE 0     /* max|phi_old - phi| <= epsilon (solver_2d.py:465-477) */
#define JSV_STOP_REL_RESIDUAL 1 /* ||rhs - L phi||_2 / ||rhs||_2 <= rtol on the unknowns */

typedef struct jsv_config {
    int fixed;              /* 1: exactly max_iters cycles (solver_2d.py:461-463) */
    double epsilon;         /* distance threshold, default 1e-12 */
    int gamma;              /* 1 V-cycle, 2 W-cycle */
    int fcycle;             /* F-cycle (solver_2d.py:609-611) */
    int skip_pre;           /* skip level-0 pre-smoothing after the first cycle */
    int nu1, nu2;           /* pre/post smoothing sweeps */
    int64_t max_iters;      /* max outer cycles (int32 max = unbounded) */
    int stop;               /* JSV_STOP_* (dynamic mode only) */
    double rtol;            /* relative-residual threshold for JSV_STOP_REL_RESIDUAL */
    int inner_cycles;       /* JSV_MIXED: FP32 cycles per FP64 refinement step;
                               0 = predicted from the measured contraction (8 first,
                               then 1..64: to the tolerance, or in fixed mode to five
                               decades below the current residual) */
    int check_every;        /* JSV_FP64 with JSV_STOP_REL_RESIDUAL: cycles between
                               residual checks (1 = every cycle, the exact count);
                               0 = spaced by the predicted cycles-to-tolerance */
    double *history;        /* optional host array: per-cycle distance or relative residual */
    int64_t history_len;    /* capacity of history */
} jsv_config;

typedef struct jsv_result {
    int64_t count;          /* cycles performed (FP32 cycles in JSV_MIXED) */
    int64_t outer_steps;    /* FP64 refinement steps (JSV_MIXED), else = count */
    double distance;        /* last max|phi_old - phi| (JSV_STOP_DISTANCE) */
    double rel_residual;    /* last ||rhs - L phi|| / ||rhs|| when computed, else -1 */
    int nonfinite;          /* 1 if a NaN/Inf was seen in the stop quantity */
} jsv_result;

int jsv_abi_version(void);
const char *jsv_last_error(void);

/* Plan creation: level layout and grid spacing.
 * Replaces Solver_2D.__init__/compute_shapes/compute_idel (solver_2d.py:247-324)
 * and Solver_3D's (solver_3d.py:248-336).  dim = 2 ignores mz/idel_z.
 * Errors (nonzero status): resolution not a power of 2, coarsest level > 5. */
int jsv_plan_create(int dim, int mx, int my, int mz, double idel_x, double idel_y,
                    double idel_z, int precision, jsv_plan **out);
/* Batched 2D plan: `batch` independent problems on the same grid, stored as
 * consecutive (my+3, mx+3) arrays (array shape (batch, my+3, mx+3)) and solved
 * together, each with its own coefficients, right-hand side and stop rule.
 * No reference counterpart (the reference solves one problem per call);
 * 3D requires batch = 1. */
int jsv_plan_create_batched(int dim, int mx, int my, int mz, int batch, double idel_x,
                            double idel_y, double idel_z, int precision, jsv_plan **out);
int jsv_plan_destroy(jsv_plan *plan);
int jsv_plan_levels(const jsv_plan *plan, int *levels);
/* shape of level l in array-axis order (nz, ny, nx); nz = 1 in 2D */
int jsv_plan_level_shape(const jsv_plan *plan, int level, int shape[3]);
/* bytes of device memory the plan holds */
int jsv_plan_device_bytes(const jsv_plan *plan, int64_t *bytes);
/* Launch-schedule knobs; results never depend on them.  No reference
 * counterpart.
 *   JSV_TUNE_SLAB_BYTES   working set of one z-slab of the 3D two-launch
 *                         smoothing schedule (default 0 = whole level per launch)
 *   JSV_TUNE_TWO_PASS_MIN level size (elements) from which the restriction
 *                         stores first-half residuals and gathers them in a
 *                         second pass (default 2^20; 2^13 for 3D FP32 levels;
 *                         the tuned value applies to both precisions)
 *   JSV_TUNE_FUSED_MIN_M  3D levels with mz >= this run each RBGS sweep as one
 *                         fused out-of-place z-marching kernel (default 256;
 *                         0 = never)
 *   JSV_TUNE_ZCHUNK       planes per workgroup of that kernel (0 = by level:
 *                         one wave of workgroups for the pair kernel, 16 in
 *                         FP32 / 32 in FP64 for the point kernel)
 *   JSV_TUNE_SWEEP_ROWS   rows per workgroup tile of that kernel: 4, 6, 8, 16
 *                         (0 = 16 in FP32, 4 in FP64)
 *   JSV_TUNE_SWEEP_THREADS threads per workgroup of that kernel: 256 or 1024
 *                         (0 = 1024).  Pair kernel, FP32 two-pair sweeps: 0
 *                         (default) the plain sweep and the one storing the
 *                         first-half residuals in 10-row tiles of 768
 *                         threads, the one with the prolongation in 14-row
 *                         tiles of 1024; 768: all in 768; 1024: all but the
 *                         residual-storing one in 1024
 *   JSV_TUNE_GRAPHS       1 (default): solves replay each cycle as a captured
 *                         hipGraph; 0: launch every kernel
 *   JSV_TUNE_COARSE_M     levels whose largest resolution is <= this (and
 *                         every coarser level) run each coarse sub-cycle as
 *                         one single-workgroup launch per problem
 *                         (default 3D: 8 for FP64, 16 for FP32; 2D: 64;
 *                         0 = off; the tuned value applies to both)
 *   JSV_TUNE_COARSE_LDS   1 (default): that launch keeps its coarsest levels
 *                         (up to 150 KB) in LDS; 0: all levels in HBM
 *   JSV_TUNE_SWEEP_VARIANT kernel of the fused sweep: 2 (default, 0) one
 *                         thread per red/black column pair with the
 *                         coefficients loaded one plane ahead (levels with
 *                         mx <= 512, rows 2, 4, 5, 6 or 14; others use 1);
 *                         1 one thread per ring point
 *   JSV_TUNE_SWEEP_PAIRS  column pairs per thread of the pair kernel: 0
 *                         (default) 2 for FP32 sweeps of levels with mx >= 256,
 *                         else 1; 1; 2
 *   JSV_TUNE_FUSE_PROLONG in 3D the first post-smoothing pair sweep of a
 *                         level applies the prolongation from the coarser
 *                         level itself: 2 (default) FP32 and FP64 levels, 1
 *                         FP32 levels only, 0 never (separate pass)
 *   JSV_TUNE_COARSE_SOLO  ops of that launch on LDS levels touching at most
 *                         this many points run in one wave without workgroup
 *                         barriers (default 64, at most 64; 0 = never)
 *   JSV_TUNE_COARSE_THREADS workgroup size of that launch: 256, 512 or 1024;
 *                         0 (default): 512 for 2D FP64 programs, 1024 for
 *                         the others
 *   JSV_TUNE_FUSE_HALVES  levels smoothed with one launch per RBGS half (2D
 *                         without D_xy; 3D below JSV_TUNE_FUSED_MIN_M): 1
 *                         (default) the restriction into such a level also
 *                         runs the first half of its first sweep (its phi is
 *                         zero then), and the first half of the
 *                         post-smoothing applies the prolongation on the fly;
 *                         0: separate launches
 *   JSV_TUNE_DEVICE_STOP  1 (default): an FP64 dynamic solve of one problem
 *                         that tests every cycle (JSV_STOP_DISTANCE, or
 *                         JSV_STOP_REL_RESIDUAL with check_every = 1) runs the
 *                         test on the device inside each cycle and reads it
 *                         back once per block of cycles; 0: the host reads
 *                         every test.  Count and phi are the same either way;
 *                         a relative residual summed inside the first sweep
 *                         of a cycle may differ from the host's sum in the
 *                         last bits
 *   JSV_TUNE_COARSE_REG   2 (default): in 2D without D_xy, the coarse
 *                         program's one-wave runs on levels of at most 64
 *                         points keep those levels in registers (neighbours
 *                         by lane shuffles), and so do runs from a level of
 *                         at most 256 points (a 2 x 2 block per lane) down
 *                         through at most three levels below it; 1: the
 *                         former only; 0: in LDS
 *   JSV_TUNE_ZIGZAG       2 (default): every pair sweep except the one with
 *                         the residual sum marches each z-chunk in the
 *                         opposite direction of that chunk's previous pair
 *                         sweep (the Infinity Cache still holds the planes it
 *                         touched last), and neighbouring chunks march in
 *                         opposite directions, so both reach their shared
 *                         boundary planes at the same time; 1: the whole
 *                         level marches opposite to its previous pair sweep,
 *                         for sweeps without the prolongation or the residual
 *                         sum (with or without the stored first-half
 *                         residuals); 0: always upward
 *   JSV_TUNE_FUSE_RESID   1 (default): the last pre-smoothing pair sweep of a
 *                         3D level with the two-pass restriction also stores
 *                         the residuals that restriction reads, at the red
 *                         points of its tiles' inner rows and planes (that
 *                         sweep marches upward); a small pass adds the tile
 *                         edge rows and planes.  0: one residual pass over
 *                         the level after the sweep.  Same bits either way */
#define JSV_TUNE_SLAB_BYTES 0
#define JSV_TUNE_TWO_PASS_MIN 1
#define JSV_TUNE_FUSED_MIN_M 2
#define JSV_TUNE_ZCHUNK 3
#define JSV_TUNE_SWEEP_ROWS 4
#define JSV_TUNE_SWEEP_THREADS 5
#define JSV_TUNE_GRAPHS 6
#define JSV_TUNE_COARSE_M 7
#define JSV_TUNE_SWEEP_VARIANT 8
#define JSV_TUNE_FUSE_PROLONG 9
#define JSV_TUNE_COARSE_LDS 10
#define JSV_TUNE_SWEEP_PAIRS 11
#define JSV_TUNE_COARSE_SOLO 12
#define JSV_TUNE_COARSE_THREADS 13
#define JSV_TUNE_FUSE_HALVES 14
#define JSV_TUNE_DEVICE_STOP 15
#define JSV_TUNE_COARSE_REG 16
#define JSV_TUNE_ZIGZAG 17
#define JSV_TUNE_FUSE_RESID 18
int jsv_plan_set_tuning(jsv_plan *plan, int key, int64_t value);

/* Per-level coefficients.  Replaces compute_lambda / compute_diffusion_tensor /
 * compute_discretization (solver_2d.py:326-421, solver_3d.py:338-429).
 * lam / dxx / dyy / dzz / dxy: device arrays of the fine shape, or NULL to use
 * the scalar (lam_scalar, d_scalar[0..2] = D_xx, D_yy, D_zz).  D is either all
 * scalar or all arrays (solver_2d.py:444-447).  dxy (2D only) requires arrays. */
int jsv_plan_set_operator(jsv_plan *plan, const double *lam, double lam_scalar,
                          const double *dxx, const double *dyy, const double *dzz,
                          const double d_scalar[3], const double *dxy, void *stream);

/* Whole solve.  Replaces Solver_2D.solve / Solver_3D.solve (solver_2d.py:423-479,
 * solver_3d.py:431-486).  phi_out may equal phi_in.  Synchronises the stream
 * in dynamic mode (it reads the stop quantity back) and at the end. */
int jsv_solve(jsv_plan *plan, const double *phi_in, const double *rhs, double *phi_out,
              const jsv_config *cfg, jsv_result *res, void *stream);
/* Batched solve: counts (host int64[batch], may be NULL) receives each
 * problem's cycle count, stopvals (host double[batch], may be NULL) its last
 * stop quantity; res summarises (count = max over the batch).  A problem that
 * has met its stop rule is frozen while the others continue, so every
 * problem's result equals its unbatched solve. */
int jsv_solve_batched(jsv_plan *plan, const double *phi_in, const double *rhs, double *phi_out,
                      const jsv_config *cfg, jsv_result *res, int64_t *counts, double *stopvals,
                      void *stream);

/* ---- single operations on the plan's level buffers (parity tests, adjoint) --
 * prec selects the hierarchy: JSV_FP64 (double buffers) or JSV_MIXED (float). */
#define JSV_BUF_PHI 0
#define JSV_BUF_RHS 1
/* copy a caller device array (double for JSV_FP64, float for JSV_MIXED) into /
 * out of a level buffer */
int jsv_level_upload(jsv_plan *plan, int prec, int level, int which, const void *src, void *stream);
int jsv_level_download(jsv_plan *plan, int prec, int level, int which, void *dst, void *stream);
/* sweeps x red_black_gauss_seidel on a level (solver_2d.py:670-692, solver_3d.py:679-706) */
int jsv_smooth(jsv_plan *plan, int prec, int level, int sweeps, void *stream);
/* rhs[level+1] = restrict(rhs - L phi) on level (defect, solver_2d.py:616-647,
 * solver_3d.py:622-652) */
int jsv_restrict(jsv_plan *plan, int prec, int level, void *stream);
/* phi[level] += prolongate(phi[level+1]) (solver_2d.py:612, 649-668;
 * solver_3d.py:618, 654-677); nu2 selects the skipped interpolations */
int jsv_prolong_add(jsv_plan *plan, int prec, int level, int nu2, void *stream);
/* one multigrid cycle on the level-0 buffers (multigrid_cycle, solver_2d.py:582-614),
 * including the leading do_BCs(phi) */
int jsv_cycle(jsv_plan *plan, int prec, const jsv_config *cfg, int skip_pre, void *stream);

/* ---- reductions and residuals (no live counterpart in the reference; basis:
 * the dead compute_residual/get_l2_residual, solver_2d.py:496-545, with the
 * correct z term instead of solver_3d.py:517's y coefficients) ---- */
/* out_r (may be NULL): rhs - L phi on the unknowns, zero elsewhere (FP64);
 * phi's boundary nodes and ghosts are taken as the zero Dirichlet values.
 * norm2_out (host, may be NULL): ||rhs - L phi||_2^2 over the unknowns, one
 * value per batched problem. */
int jsv_residual(jsv_plan *plan, const double *phi, const double *rhs, double *out_r,
                 double *norm2_out, void *stream);
/* max |a - b| over n elements (distance, solver_2d.py:481-494); host result */
int jsv_max_abs_diff(const double *a, const double *b, int64_t n, double *out, void *stream);

/* jsv_max_abs_diff that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc).  out[s] (host, nseg doubles) receives
 * max |a - b| over segment s of n elements (elements s*n .. s*n + n - 1);
 * nseg = 1 gives jsv_max_abs_diff's result.  The result words live in a
 * fixed device array of JSV_NOALLOC_MAX_SEGMENTS words per GPU, held until
 * the process ends; more segments are rejected with an error.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other host
 * threads, other streams) race.  Synchronises the stream. */
#define JSV_NOALLOC_MAX_SEGMENTS 1024
int jsv_max_abs_diff_noalloc(const double *a, const double *b, int64_t n, int64_t nseg, double *out,
                             void *stream);

/* Streaming-bandwidth reference, no counterpart in the reference: dst[i] =
 * sum over a < k of src[a*n + i] (k = 1, 2 or 6 source arrays of n floats,
 * consecutive in src; 16-byte aligned; n a multiple of 4) with 16-byte
 * non-temporal loads and stores.  k = 6 reads and writes in the RBGS sweep's
 * 6:1 ratio.  blocks: workgroups of 256 threads (0 = one per 1024 floats).
 * Bytes moved per launch: 4 n (k + 1). */
int jsv_stream_bench(const float *src, int k, int64_t n, float *dst, int blocks, void *stream);

/* ---- boundary handler (do_BCs, boundary_handler_2d.py:3-29 /
 * boundary_handler_3d.py:3-35), in place on a caller array.  vals_* may be
 * NULL (zero).  3D: like the reference, the z faces take vals_yb/vals_ye. ---- */
int jsv_do_bcs_2d(double *arr, int ny, int nx, int rim, int direction,
                  const double *vals_xb, const double *vals_xe,
                  const double *vals_yb, const double *vals_ye, void *stream);
int jsv_do_bcs_3d(double *arr, int nz, int ny, int nx, int rim, int direction,
                  const double *vals_xb, const double *vals_xe,
                  const double *vals_yb, const double *vals_ye, void *stream);

/* Non-zero Dirichlet values: the values phi_b holds on the level-0 boundary
 * nodes (index 1 and m+1 per axis, as do_BCs(arr, vals_...) sets them,
 * boundary_handler_2d.py:3-29) moved to the right-hand side of the unknowns
 * next to them: rhs_out = rhs - sum over boundary-node neighbours q of
 * A(p, q) phi_b[q] (the operator set on the plan; D_xy corners included).
 * A zero-boundary solve of rhs_out is the solve with those boundary values;
 * the reference's solver has no such option (it calls do_BCs with zeros,
 * README.md:90).  Arrays of the plan's fine shape (batched 2D: the batch). */
int jsv_dirichlet_lift(jsv_plan *plan, const double *phi_b, const double *rhs, double *rhs_out,
                       void *stream);

/* ---- reverse pass of fixed-mode solves (FP64; 2D with or without D_xy, 3D) ----
 * Replaces the reference's jax.vjp over a fixed solve
 * (examples/benchmark.py:1184-1226; tests/test_solver_2d.py:200-207).
 * jsv_forward_checkpoints runs cfg->max_iters cycles from phi0 and stores the
 * iterate entering every cycle and the final one: ckpt holds
 * (max_iters + 1) x n doubles (n = fine array size).  jsv_adjoint takes the
 * cotangent g_out of the final phi and returns the cotangents of phi0, rhs,
 * lambda and the D components (any output may be NULL); each cycle's interior
 * states are recomputed from ckpt, not stored.  g_lam needs an array lambda
 * and g_dxx.. array D on the plan (scalar parts: jsv_adjoint_full). */
int jsv_forward_checkpoints(jsv_plan *plan, const double *phi0, const double *rhs,
                            const jsv_config *cfg, double *ckpt, void *stream);
int jsv_adjoint(jsv_plan *plan, const double *phi0, const double *rhs, const double *ckpt,
                const jsv_config *cfg, const double *g_out, double *g_phi0, double *g_rhs,
                double *g_lam, double *g_dxx, double *g_dyy, double *g_dzz, void *stream);
/* jsv_adjoint with every operator input of the reference's jax.vjp
 * (tests/test_solver_2d.py:200-207 differentiates scalar lambda and D;
 * examples/benchmark.py:1223-1229 also D_xy):
 *   g_dxy     cotangent of the 2D D_xy array (plans whose operator has D_xy;
 *             solver_2d.py:411-416, 686), or NULL;
 *   g_scalar  device double[4], or NULL: the cotangents of the operator's
 *             scalar parts, [0] lambda, [1..3] D_xx, D_yy, D_zz (solver_2d.py:
 *             417-421); entries of parts set from arrays are 0.
 * Array outputs of scalar parts are an error. */
int jsv_adjoint_full(jsv_plan *plan, const double *phi0, const double *rhs, const double *ckpt,
                     const jsv_config *cfg, const double *g_out, double *g_phi0, double *g_rhs,
                     double *g_lam, double *g_dxx, double *g_dyy, double *g_dzz, double *g_dxy,
                     double *g_scalar, void *stream);
/* device bytes held by the plan's reverse-pass buffers (tape included) */
int jsv_adjoint_device_bytes(const jsv_plan *plan, int64_t *bytes);

/* ---- tangent (forward-mode) pass of solves (FP64) ----
 * Replaces the reference's jax.jvp over a solve (examples/benchmark.py:1146-1182;
 * tests/test_solver_2d.py:173-176, check_grads(modes=["fwd"])).  Runs
 * cfg->max_iters cycles (with cfg's cycle shape; cfg->fixed and the stop rule
 * are ignored: a dynamic solve's tangent is that of its cycle count) on the
 * primal (phi0, rhs) with plan's operator, and alongside them the tangent
 * cycles driven by (phi0_dot, rhs_dot) and by the tangent operator held by
 * tangent_plan: a second plan of the same shape whose operator was set from
 * the tangents (D_dot, lambda_dot[, D_xy_dot]) with jsv_plan_set_operator (the
 * coefficient setup is linear in D and lambda).  phi0_dot / rhs_dot may be
 * NULL (zero); phi_out may be NULL.  Both plans' level buffers are
 * overwritten.  Unbatched plans; D_xy supported. */
int jsv_tangent(jsv_plan *plan, jsv_plan *tangent_plan, const double *phi0, const double *rhs,
                const double *phi0_dot, const double *rhs_dot, const jsv_config *cfg,
                double *phi_out, double *phi_dot_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* JSOLVER_AMD_H */

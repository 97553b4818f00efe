/* gpumatting.h -- C ABI of libgpumatting.so (MI355X / gfx950, HIP).
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns a
 * gm_status (0 = success); gm_last_error() gives the message of the last
 * failure on the calling thread.  Device pointers are HIP device memory;
 * "stream" is a hipStream_t (NULL = the null stream).  Nothing inside an
 * iteration synchronises with the host except where stated.
 *
 * Each entry point names the reference interface (rocketman768/gpumatting,
 * file:line) whose role it takes.  The C++ drop-in headers BandedMatrix.h,
 * Vector.h, Solve.h and Matting.h wrap these calls under the reference's own
 * function names and parameter orders.
 */
#ifndef GPUMATTING_H
#define GPUMATTING_H

#include <stddef.h>
#include "BandedMatrix.h"
#include "CompressedMatrix.h"
#include "CoordMatrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without pulling in the HIP headers */
typedef struct ihipStream_t* gm_stream_t;

typedef enum gm_status_e {
   GM_OK = 0,
   GM_ERR_INVALID = 1,       /* bad argument (null pointer, size, pitch, pad) */
   GM_ERR_OOM = 2,           /* device or pinned allocation failed */
   GM_ERR_HIP = 3,           /* a HIP runtime call failed */
   GM_ERR_BREAKDOWN = 4,     /* solver breakdown: p'Ap <= 0 (matrix not SPD) */
   GM_ERR_NONFINITE = 5,     /* NaN/Inf met in the iteration */
   GM_ERR_NOT_CONVERGED = 6, /* tolerance not reached within max_iter */
   GM_ERR_UNSUPPORTED = 7    /* valid request this build does not implement */
} gm_status;

const char* gm_last_error(void);
const char* gm_version(void);
/* Number of HIP devices the runtime sees (0 when none); no device context is created. */
int gm_device_count(void);

/* ---------------------------------------------------------------------- */
/* Matrices and vectors (BandedMatrix.cu, Vector.cu)                        */

/* bmCopyToDevice, BandedMatrix.cu:23-41: allocate dA (pitched, zero padded)
 * and copy hA's bands and offsets.  dA->apitch is in elements. */
gm_status gm_bm_copy_to_device(BandedMatrix* dA, const BandedMatrix* hA);
/* bmDeviceFree, BandedMatrix.cu:43-47 */
gm_status gm_bm_device_free(BandedMatrix* dA);
/* Allocate a zeroed device matrix with the given host band offsets (no
 * reference counterpart: the reference only ever copied a host matrix). */
gm_status gm_bm_device_alloc(BandedMatrix* dA, int rows, int nbands, const int* host_bands);
/* Copy a device matrix back into a host matrix with the same rows/nbands
 * (hA->a holds nbands*hA->apitch floats). */
gm_status gm_bm_copy_to_host(BandedMatrix* hA, const BandedMatrix* dA);

/* vecDeviceMalloc, Vector.cu:23-30: zeroed, *dx points past the left pad. */
gm_status gm_vec_device_malloc(float** dx, int length, int left_pad, int right_pad);
/* vecCopyToDevice, Vector.cu:37-41 */
gm_status gm_vec_copy_to_device(float** dx, const float* hx, int length, int left_pad, int right_pad);
/* vecDeviceFree, Vector.cu:43-46 */
gm_status gm_vec_device_free(float* dx, int left_pad);
/* The device-to-host copy the reference's driver makes with cudaMemcpy
 * (main.cu:257, 261, 282): length floats from dx (past any pad) to hx; synchronous. */
gm_status gm_vec_copy_to_host(float* hx, const float* dx, int length);

/* ---------------------------------------------------------------------- */
/* Kernels as host launchers (BandedMatrix.cu, Vector.cu)                   */

/* bmAx_k, BandedMatrix.cu:107-141,194-198: b = A x.  x readable at
 * i + bands[k] (zero padded, BandedMatrix.cu:145-146). */
gm_status gm_bm_ax(float* b, const BandedMatrix* A, const float* x, gm_stream_t stream);
/* bmAxpy_k, BandedMatrix.cu:152-189,203-207: b = A x + y (add) or A x - y. */
gm_status gm_bm_axpy(float* b, const BandedMatrix* A, const float* x, const float* y, int add,
                     gm_stream_t stream);
/* jacobi_k, BandedMatrix.cu:61-98 (diagonal taken from the band with offset 0,
 * for any band count): xx = omega*(x + (b - A x)/a_ii) + (1-omega)*x. */
gm_status gm_bm_jacobi_sweep(float* xx, const float* x, const BandedMatrix* A, const float* b,
                             float omega, gm_stream_t stream);

/* vecAdd_k / vecSub_k, Vector.cu:142-150 */
gm_status gm_vec_add(float* r, const float* x, const float* y, int len, gm_stream_t stream);
gm_status gm_vec_sub(float* r, const float* x, const float* y, int len, gm_stream_t stream);
/* vecScaleConst_k, Vector.cu:175 */
gm_status gm_vec_scale_const(float* r, const float* x, float s, int len, gm_stream_t stream);
/* vecScale_k, Vector.cu:180 (s is a device scalar) */
gm_status gm_vec_scale(float* r, const float* x, const float* s, int len, gm_stream_t stream);
/* innerProd_k, Vector.cu:195-265: *result (device) = <x,y>; deterministic
 * two-stage sum (float products, double accumulation), no atomics on data. */
gm_status gm_inner_prod(float* result, const float* x, const float* y, int len, gm_stream_t stream);
/* gm_inner_prod without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): with it, bmAx, the
 * vector ops and this call run the reference's CG loop (main.cu:481-518) with
 * no allocation at all.  Its partial sums and counter live in a static 33 KB
 * device array that every GPU holds until the process ends; any len >= 0
 * fits (the grid is capped at 4096 workgroups).  Same kernel and grid as
 * gm_inner_prod, so the same result bit for bit.  Calls that overlap on one
 * GPU (several streams or threads at once) share that array and race. */
gm_status gm_inner_prod_noalloc(float* result, const float* x, const float* y, int len, gm_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Solvers                                                                 */

/* gradSolve, main.cu:406-456: steepest descent, fixed iteration count,
 * alpha (device, padded by pad) updated in place. */
gm_status gm_grad_solve(float* alpha, const BandedMatrix* L, const float* b, int iterations, int pad,
                        gm_stream_t stream);
/* cgSolve, main.cu:458-526: CG with r = L alpha - b, restart every
 * restart_interval iterations, fixed iteration count (stops early only on an
 * exactly zero residual or a breakdown). */
gm_status gm_cg_solve(float* alpha, const BandedMatrix* L, const float* b, int pad, int iterations,
                      int restart_interval, gm_stream_t stream);
/* jacobi, main.cu:528-556: damped Jacobi sweeps, result in x. */
gm_status gm_jacobi(float* x, const BandedMatrix* A, const float* b, float omega, int pad,
                    int iterations, gm_stream_t stream);
/* gaussSeidel_host, Solve.cpp:3-33 / Solve.h:17: host pointers in and out
 * (x padded by the matrix's largest |offset|); the sweep runs on the GPU in
 * wavefront order and gives the lexicographic Gauss-Seidel result. */
gm_status gm_gauss_seidel_host(float* x, const BandedMatrix* A, const float* b, size_t iterations);
/* jacobi_host, Solve.cpp:35-85 / Solve.h:31: host pointers; result always in x. */
gm_status gm_jacobi_host(float* x, const BandedMatrix* A, const float* b, size_t iterations, int xpad,
                         float omega);

typedef enum gm_method_e {
   GM_METHOD_CG = 0,    /* conjugate gradient */
   GM_METHOD_PCG = 1,   /* Jacobi-preconditioned CG */
   GM_METHOD_GD = 2     /* steepest descent, exact line search */
} gm_method;

typedef enum gm_precision_e {
   GM_PREC_FP32 = 0,    /* FP32 diagonals, FP32 vectors */
   GM_PREC_FP16 = 1,    /* Jacobi-scaled FP16 diagonals, FP32 accumulate, FP64 residual replacement */
   GM_PREC_BF16 = 2     /* as FP16 with BF16 diagonals */
} gm_precision;

typedef struct gm_solve_opts_s {
   int method;          /* gm_method */
   int precision;       /* gm_precision (16-bit forms: PCG on the scaled system) */
   int max_iter;        /* iteration cap (> 0) */
   double tol;          /* stop when ||b - A x|| / ||b|| <= tol; <= 0: run max_iter */
   int restart;         /* > 0: recompute r = A x - b every `restart` iterations */
   int refine_every;    /* FP64 residual replacement period: > 0 that many iterations; 0 = default
                           (50 for the 16-bit forms and refine_mode 2, none for plain FP32);
                           < 0 = none */
   int check_every;     /* iterations enqueued between host status polls (0 -> 16) */
   int verify;          /* 1: confirm convergence with the FP64 true residual */
   int refine_mode;     /* 0: FP32 x, residual replacement; 1: iterative refinement with an FP64
                           master x (outer r = b - A x in FP64, inner solves of A d = r to 1e-4);
                           2: FP64 master x inside the iteration: at every residual replacement
                           (refine_every, default 50) the FP32 x is added into the master and
                           restarts from zero, and r = b - A x is recomputed from the master in
                           FP64, keeping the search direction while the recursive residual has
                           not drifted.  Modes 1 and 2 remove the FP32-x floor (~u*|A||x|/|b|)
                           of the true residual; true_relres is then that of the FP64 master,
                           and alpha is returned rounded to FP32.  Mode 2: CG/PCG only */
} gm_solve_opts;

typedef struct gm_solve_info_s {
   int iterations;      /* iterations performed */
   int status;          /* gm_status of the solve */
   double relres;       /* last residual norm / ||b|| the iteration saw */
   double true_relres;  /* FP64-accumulated ||b - A x|| / ||b|| at exit (-1 if not computed) */
   int replacements;    /* FP64 residual replacements performed */
} gm_solve_info;

void gm_solve_opts_default(gm_solve_opts* o);

/* Solver entry: (L + lambda D) alpha = lambda D b as assembled by
 * gm_levin_laplacian, or any SPD banded system.  alpha: device, padded by
 * `pad` (>= max |band offset|), initial guess in, solution out. */
gm_status gm_solve(float* alpha, const BandedMatrix* L, const float* b, int pad, const gm_solve_opts* opts,
                   gm_solve_info* info, gm_stream_t stream);

/* Project autotuner (no reference counterpart): times the fused CG step on
 * this matrix for every (tile geometry, rows per thread, workgroups per CU)
 * candidate with HIP events and records the fastest for solves of the same
 * shape and precision (FP32: also the plain product gm_bm_ax).  Clobbers the calling stream's solver workspace (call before a
 * solve, not during). */
typedef struct gm_tune_result_s {
   int rows_per_thread;
   int tile_rows;       /* 0: 1-D row tiles; > 0: 2-D pixel tiles of that many image rows */
   int wg_per_cu;
   double step_us;      /* mean duration of one fused CG step with that shape */
   int spmv_rows_per_thread;  /* FP32 only: the plain product y = A x, tuned the same way */
   int spmv_tile_rows;
   int spmv_wg_per_cu;
   double spmv_us;
   int update_wg_per_cu;   /* grid of the CG update kernel */
   double update_us;
   int spmv_direct;        /* 1: the plain product's LDS-free form (x read through the caches);
                              2: its long-run form (4 rows per tile, each lane rows_per_thread / 4
                              runs of 4 pixels 256 pixels apart) */
   int step_walk;          /* 16-bit symmetric step: 1 = column-walk form (workgroups walk down
                              columns of tiles and keep the rows shared with the tile above in LDS);
                              2 = column walk with the whole iteration in one launch (the update of
                              x and r fused into the next step; 4-row tiles).  FP32 Jacobi-PCG with
                              W % 4 == 0: 2 = the same one-launch iteration on the symmetric scaled
                              FP32 operator S L S (12 upper bands, FP64 residual replacement through
                              L every 50 iterations), 0 = the Jacobi step + update on L's 25 bands */
   double fused_us;        /* that one-launch iteration's time (0: not timed): the mean of the
                              two launches a solve alternates, */
   double fused_update_us; /* ... one that updates x (its own update and the one deferred by
                              the launch before: 64 B per row with 16-bit bands) */
   double fused_defer_us;  /* ... and one that defers its update of x (52 B per row) */
} gm_tune_result;
gm_status gm_autotune(const BandedMatrix* L, int precision, int reps, gm_tune_result* out, gm_stream_t stream);
/* Text table of every candidate timed by the last gm_autotune on this thread. */
const char* gm_autotune_report(void);
/* Pin a shape without timing: (rows_per_thread, tile_rows) one of (2,0) (4,0)
 * (4,4) (4,8) (4,16) for FP32 (2-D shapes need W % rows_per_thread == 0); for
 * 16-bit, (4,4) (4,8) (4,16) when W % 4 == 0 (the symmetric 12-band form),
 * else (4,0) (8,0). */
gm_status gm_tune_set(int W, int N, int precision, int rows_per_thread, int tile_rows, int wg_per_cu);
/* As gm_tune_set, also choosing the symmetric 16-bit step's column-walk form
 * (column_walk = 1; 16-bit with W % 4 == 0 only), or that walk with one
 * launch per CG iteration (column_walk = 2; tile_rows 4 and W % 4 == 0 only;
 * for FP32 the fused iteration on the symmetric scaled FP32 operator). */
gm_status gm_tune_set_step(int W, int N, int precision, int rows_per_thread, int tile_rows, int wg_per_cu,
                           int column_walk);

/* Pin the plain product's (gm_bm_ax) form for FP32 Levin matrices of this
 * shape without timing: the staged forms (rows_per_thread, tile_rows) as for
 * gm_tune_set's FP32 shapes, direct = 1 with (4,4) (4,8) (2,8) (2,16) for
 * the LDS-free form (x 4*rows_per_thread-byte aligned, else a staged form
 * runs), or direct = 2 with (8,4) (12,4) (16,4) for the long-run form
 * (W % 4 == 0; out, x and z 16-byte aligned, else a staged form runs).  No
 * reference counterpart. */
gm_status gm_spmv_tune_set(int W, int N, int rows_per_thread, int tile_rows, int direct, int wg_per_cu);

/* FP64-accumulated ||b - A x|| / ||b|| (synchronises; writes *relres). */
gm_status gm_true_relres(const BandedMatrix* A, const float* x, const float* b, double* relres,
                         gm_stream_t stream);

/* ---------------------------------------------------------------------- */
/* Levin matting Laplacian (Matting.cu)                                      */

typedef enum gm_assembly_mode_e {
   GM_ASSEMBLY_REFERENCE = 0, /* bit-identical to Matting.cu:39-219's float/double rounding */
   GM_ASSEMBLY_ACCURATE = 1   /* exact double covariance sums (same definition) */
} gm_assembly_mode;

/* GPU assembly, gather form (replaces hostLevinLaplacian Matting.cu:39-219 and
 * the unfinished levinLaplacian kernel Matting.cu:240-382).
 * L: device matrix with 25 bands in the layout bands[5(dy+2)+(dx+2)] = dy*W+dx
 * (main.cu:161-190) and rows = W*H; every entry is written (no pre-zeroing).
 * im: device, nch (3 or 4) interleaved floats per pixel, pitch in pixels.
 * Data term: per-pixel weights lam[i] and targets tgt[i] (device, W*H, may be
 * NULL = none): L_ii += lam_i, b_i = lam_i*tgt_i.  dinv (nullable) receives
 * 1/L_ii (0 where L_ii == 0). */
gm_status gm_levin_laplacian(BandedMatrix* L, float* b, float* dinv, const float* im, int nch,
                             size_t im_pitch, const float* lam, const float* tgt, float eps, int mode,
                             int W, int H, gm_stream_t stream);
/* Same with the reference's u8 scribble rule (Matting.cu:205-216): 255 ->
 * L_ii += gamma, b_i = gamma; 0 -> L_ii += gamma, b_i = 0; else b_i = 0.
 * scribs: device, indexed with im_pitch like the reference (:205). */
gm_status gm_levin_laplacian_u8(BandedMatrix* L, float* b, float* dinv, const float* im, int nch,
                                size_t im_pitch, const unsigned char* scribs, float eps, float gamma,
                                int mode, int W, int H, gm_stream_t stream);
/* hostLevinLaplacian, Matting.cu:39-46 signature: host L (pre-allocated,
 * 25 bands, apitch >= rows), host b, host float4 RGBA image with pitch in
 * float4 elements, host u8 scribbles, eps (the reference calls it lambda),
 * gamma = 1.  Runs the reference-rounding GPU assembly and copies back
 * (synchronous). */
gm_status gm_host_levin_laplacian(BandedMatrix* L, float* b, float eps, const float* im4,
                                  const unsigned char* scribs, int W, int H, size_t im_pitch);

/* Adjoint of the assembly for (L(im) + diag(lam)) alpha = lam * tgt (no
 * reference counterpart: the differentiation path).  mu = (L + diag(lam))^-1
 * dLoss/dalpha (an adjoint solve with gm_solve), alpha the forward solution.
 * Writes dLoss/dim (W*H*3 floats, RGB interleaved, dense), and if non-NULL
 * dLoss/dlam_i = mu_i (tgt_i - alpha_i) and dLoss/dtgt_i = mu_i lam_i.
 * The image derivative is that of GM_ASSEMBLY_ACCURATE's Laplacian, in FP64. */
gm_status gm_levin_backward(float* grad_im, float* grad_lam, float* grad_tgt, const float* im, int nch,
                            size_t im_pitch, const float* mu, const float* alpha, const float* lam, const float* tgt,
                            float eps, int W, int H, gm_stream_t stream);

/* 5x5-window band offsets (main.cu:161-190) and the vector pad
 * ((2W+2)/32)*32+32 (main.cu:228-229). */
void gm_levin_bands(int W, int* bands25);
int gm_levin_pad(int W);

/* ---------------------------------------------------------------------- */
/* Batched solve: many images of uneven size, several right-hand sides each  */
/* (no reference counterpart).                                              */

typedef struct gm_batch_s gm_batch;

/* Allocate a batch: nimages images (widths[i] x heights[i]) with nrhs <= 8
 * constraint sets each; precision GM_PREC_FP32, GM_PREC_FP16 or GM_PREC_BF16
 * (16-bit off-diagonal bands, FP32 diagonals and vectors, FP64 residual
 * replacement).  Each image has 1 to 2^25 - 1 pixels (GM_ERR_INVALID
 * otherwise, before any device call). */
gm_status gm_batch_create(gm_batch** batch, int nimages, const int* widths, const int* heights, int nrhs,
                          int precision);
gm_status gm_batch_destroy(gm_batch* batch);
/* Assemble image i's Laplacian (accurate rounding, no data term) from a device
 * image (nch 3|4 floats per pixel, pitch in pixels). */
gm_status gm_batch_assemble(gm_batch* batch, int image, const float* im, int nch, size_t im_pitch, float eps,
                            gm_stream_t stream);
/* Set system (image, rhs): weights lam and targets tgt (device, W*H, NULL =
 * none), initial guess x0 (device or NULL = 0). */
gm_status gm_batch_set_rhs(gm_batch* batch, int image, int rhs, const float* lam, const float* tgt, const float* x0,
                           gm_stream_t stream);
/* Jacobi-PCG on every system at once: one launch per CG iteration for all
 * images and right-hand sides (the update of iteration j and the step of
 * j + 1 fused), hipGraph-captured chunks of check_every (rounded up to even)
 * iterations; per-system convergence. */
gm_status gm_batch_solve(gm_batch* batch, const gm_solve_opts* opts, gm_stream_t stream);
/* Device pointer of system (image, rhs)'s solution (valid until destroy). */
gm_status gm_batch_alpha(gm_batch* batch, int image, int rhs, float** alpha);
/* Per-system result of the last gm_batch_solve. */
gm_status gm_batch_info(gm_batch* batch, int image, int rhs, gm_solve_info* info);
/* Workgroups per CU of the batched launches, 1..64, at most 16384 in all
 * (default 32: more than fit at once, so workgroups whose ranges lie in
 * converged images end early and the hardware balances the rest, and the
 * images that converge last are split into short ranges; 4 per CU, one
 * resident wave, is 10 % slower on uneven batches). */
gm_status gm_batch_set_grid(gm_batch* batch, int per_cu);
/* Tile width of the fused iteration: 60 (60 x 4 tiles), 124 (124 x 2, the
 * default) or 252 (252 x 1; batches of up to 4 right-hand sides). */
gm_status gm_batch_set_tiles(gm_batch* batch, int tile_width);
/* Time the fused iteration at every tile width and 4, 8, 16 and 32
 * workgroups per CU on this batch's own matrices and keep the fastest.  With
 * opts NULL each candidate runs reps full-load launches (*best_us: time per
 * launch); with solve options, reps whole solves with them (*best_us: time
 * per solve), which also weighs a solve's tail, where only the slowest images
 * still run.  Assemble and set the right-hand sides first; the initial
 * guesses are kept.  Synchronous. */
gm_status gm_batch_autotune(gm_batch* batch, const gm_solve_opts* opts, int reps, double* best_us, gm_stream_t stream);
/* The tile width and workgroups per CU in use, and the tuned time per launch
 * (0 before gm_batch_autotune). */
gm_status gm_batch_get_config(gm_batch* batch, int* tile_width, int* per_cu, double* tuned_us);
int gm_batch_tiles(gm_batch* batch);

/* Release cached solver workspaces (device memory held between solves). */
gm_status gm_release_workspaces(void);

/* ---------------------------------------------------------------------- */
/* Sparse matrices and superpixel multilevel operators                     */
/* (CompressedMatrix.cu, CoordMatrix.cu, gridoperators.h, SLIC.h)          */
/*                                                                         */
/* The reference's sparse layer is unfinished and never called (element    */
/* counts passed as byte counts, a row bound off by one, y[0] added to     */
/* every row, a float atomicAdd product, host-to-host copies); these entry */
/* points give the operations it names their evident meaning.  Every       */
/* product is deterministic (fixed summation order, FP64 accumulation of   */
/* FP32 products).                                                         */

/* csmInit, CompressedMatrix.cu:20-31: rows x cols, room for `length`
 * entries, p zeroed, nnz = 0 (device arrays). */
gm_status gm_csm_init(CompressedMatrix* m, int rows, int cols, size_t length);
/* Free what gm_csm_init (or a builder below) allocated; zeroes *m. */
gm_status gm_csm_free(CompressedMatrix* m);
/* cooInit, CoordMatrix.cu:20-29 */
gm_status gm_coo_init(CoordMatrix* m, int rows, int cols, size_t length);
/* cooRelease, CoordMatrix.cu:31-40 */
gm_status gm_coo_release(CoordMatrix* m);
/* cooAppend, CoordMatrix.cu:42-48: append n host entries (i, j, k) after the
 * nnz already held; indices are checked against rows / cols and the total
 * against length (nothing is appended on error).  Synchronous. */
gm_status gm_coo_append(CoordMatrix* m, const int* i, const int* j, const float* k, size_t n);
/* CSR form of a COO matrix (no reference counterpart): entries sorted by
 * (row, column), repeats kept in their COO order.  Allocates *out (free with
 * gm_csm_free).  Synchronous. */
gm_status gm_coo_to_csm(CompressedMatrix* out, const CoordMatrix* in, gm_stream_t stream);
/* csmAx, CompressedMatrix.cu:118-165: b = A x (b: rows, x: cols floats). */
gm_status gm_csm_ax(float* b, const CompressedMatrix* A, const float* x, gm_stream_t stream);
/* csmAxpy, CompressedMatrix.cu:45-85: b = A x + y (add) or A x - y. */
gm_status gm_csm_axpy(float* b, const CompressedMatrix* A, const float* x, const float* y, int add,
                      gm_stream_t stream);
/* cooAxpy, CoordMatrix.cu:50-87: b += A x + y (the reference accumulates onto
 * b).  Sorts a copy of the entries on every call: convert once with
 * gm_coo_to_csm for repeated products.  Synchronous. */
gm_status gm_coo_axpy(float* b, const CoordMatrix* A, const float* x, const float* y, gm_stream_t stream);

/* First level of downsampleOps, gridoperators.h:18-57: the restriction P
 * (nlabels x npix, P[labels[u], u] = 1) of a superpixel labelling, as CSR
 * (each row's pixels ascending).  labels: npix device ints in [0, nlabels)
 * (GM_ERR_INVALID otherwise).  Allocates *P.  Synchronous.  P x sums a pixel
 * vector over each superpixel (gm_csm_ax); gm_sp_prolong is P^T. */
gm_status gm_sp_restriction(CompressedMatrix* P, const int* labels, int npix, int nlabels, gm_stream_t stream);
/* e[u] = xc[labels[u]] for u < npix: the prolongation P^T xc (no reference
 * counterpart: the reference stops at building P); NaN where a label is
 * outside [0, nlabels). */
gm_status gm_sp_prolong(float* e, const float* xc, const int* labels, int npix, int nlabels, gm_stream_t stream);
/* The Galerkin coarse operator Ac = P A P^T (nlabels x nlabels CSR, repeats
 * merged: one entry per (row, column), the fine entries (pixel, band) summed
 * in FP64 in a fixed order) of a device banded matrix (any band count;
 * entries stored as 0 are skipped) and a labelling as above -- the operator
 * the reference's multilevel direction coarsens with.  Allocates *Ac.
 * Synchronous. */
gm_status gm_sp_galerkin(CompressedMatrix* Ac, const BandedMatrix* A, const int* labels, int nlabels,
                         gm_stream_t stream);
/* Later levels of downsampleOps, gridoperators.h:59-138 (there unfinished):
 * merge superpixels pairwise, greedily by ascending squared distance of their
 * mean colours, while that distance is at most `threshold` (the reference:
 * 0.0111); only superpixels that share a 4-neighbour pixel edge merge (the
 * reference's "need to check ... spatially adjacent"), and each merges at most
 * once per level.  Each level maps label c to map[c] (merged pairs take new
 * labels in merge order, the rest follow in label order) and the means are
 * recomputed from the pixels (its "need to update means").  im: host RGB
 * (nch floats per pixel, pitch pixels per row); labels: host, updated in
 * place to the last level's labels; *nlabels in/out.  Host computation. */
gm_status gm_sp_merge_levels(int* labels, int* nlabels, const float* im, int nch, size_t pitch, int W, int H,
                             int nlevels, float threshold);
/* SLIC superpixels (slicSegmentation, SLIC.h:43; algorithm of Achanta et al.,
 * "SLIC Superpixels Compared to State-of-the-Art Superpixel Methods", TPAMI
 * 2012): CIELAB k-means over (l, a, b, x, y) with grid step S = round(sqrt(W H
 * / K)), seeds moved to the lowest gradient of their 3x3 neighbourhood,
 * `iters` assignment / update rounds (distance (dc/m)^2 + (ds/S)^2 over the
 * seeds whose 2S x 2S window holds the pixel), then connected components:
 * every 4-connected piece of a label becomes a superpixel, and pieces smaller
 * than a quarter of S^2 join the superpixel of the pixel left of (or else
 * above) their first pixel in raster order.  Final labels are consecutive
 * from 0 in raster order of first appearance.  im: device RGB in [0, 1] (nch
 * floats per pixel, pitch pixels per row), labels: W*H device ints.
 * Synchronous. */
gm_status gm_slic(int* labels, int* nlabels, const float* im, int nch, size_t pitch, int W, int H, int K,
                  float compactness, int iters, gm_stream_t stream);
/* The stages of gm_slic, for callers that keep the colour planes or the raw
 * clustering (no reference counterpart: slicSegmentation, SLIC.h:43, runs
 * them as one call).
 * gm_slic_grid: the grid gm_slic uses for K superpixels: step S and nsx x nsy
 * seeds (host computation).
 * gm_slic_lab: CIELAB planes (L, a, b; 3*W*H device floats) of the RGB image,
 * evaluated in FP64 and rounded to multiples of 2^-16.
 * gm_slic_raw: seeds and `iters` assignment / update rounds from Lab planes:
 * labels are seed indices before connectivity; seeds (optional; 5 planes of
 * maxseeds >= nsx*nsy floats: l, a, b, x, y) receives the final seeds.
 * Synchronous.
 * gm_slic_connect: connectivity in place (pieces below minsize pixels join
 * their left / upper neighbour; consecutive numbering).  Synchronous. */
gm_status gm_slic_grid(int W, int H, int K, int* S, int* nsx, int* nsy);
gm_status gm_slic_lab(float* lab, const float* im, int nch, size_t pitch, int W, int H, gm_stream_t stream);
gm_status gm_slic_raw(int* labels, float* seeds, int maxseeds, const float* lab, int W, int H, int K,
                      float compactness, int iters, gm_stream_t stream);
gm_status gm_slic_connect(int* labels, int* nlabels, int W, int H, int minsize, gm_stream_t stream);

/* ---- host file I/O of the matting driver (no device work) ---------------
 * Netpbm: binary or ASCII, header tokens separated by whitespace with '#'
 * comments, 8-bit samples (maxval 1..255).  Returned buffers are malloc'd:
 * release them with gm_free. */
void gm_free(void* p);
/* pgmread, ppm.h:32-129: "P5" / "P2" grey image, w*h samples row-major.
 * maxval may be NULL. */
gm_status gm_pgm_read(const char* path, unsigned char** data, int* w, int* h, int* maxval);
/* ppmread, ppm.h:178-278: "P6" / "P3" RGB image, 3*w*h samples row-major.
 * ppmread_float4 (ppm.h:309-337) normalises each sample by maxval. */
gm_status gm_ppm_read(const char* path, unsigned char** rgb, int* w, int* h, int* maxval);
/* pgmwrite, ppm.h:345-418: "P5" (binary = 1) or "P2"; comment line
 * "# <comment> " when comment is not NULL; maxval 255. */
gm_status gm_pgm_write(const char* path, const unsigned char* data, int w, int h, const char* comment, int binary);
/* pgmwrite_float's quantisation, ppm.h:430-451: v < 0 -> 0, v > 1 -> 255,
 * else (unsigned char)(255.f * v); NaN -> 0. */
void gm_quantize_u8(const float* v, int n, unsigned char* out);
/* pgmwrite_float, ppm.h:430-451: quantise as gm_quantize_u8, then pgmwrite. */
gm_status gm_pgm_write_float(const char* path, const float* data, int w, int h, const char* comment, int binary);
/* computeError, main.cu:558-577: mean over n pixels of (clamp(alpha) - gt)^2
 * with alpha > 1 scored as 1 and alpha < 0 as 0; each term in FP32, the sum
 * in FP64.  Host arrays. */
gm_status gm_clamped_mse(const float* alpha, const float* gt, int n, double* mse);
/* dump, main.cu:334-355 ("A.bin"): a HOST matrix's bands one after another,
 * `rows` FP32 values each (pitch padding dropped), no header. */
gm_status gm_bm_dump(const BandedMatrix* A, const char* path);
/* Reader of that format (main.cu:334-355; matlab/laplacianFromBmDump.m reads
 * it in MATLAB): fills a HOST matrix whose nbands, rows, apitch and a are set;
 * the file must hold exactly nbands * rows floats; padding is zeroed. */
gm_status gm_bm_load(BandedMatrix* A, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* GPUMATTING_H */

/*
 * sipg_hip.h -- C-ABI of the MI355X-native matrix-free SIPG/SEM Poisson
 * solver (libsipg_hip.so, gfx950).
 *
 * Plain C: opaque handles, plain pointers and sizes.  Device pointers are
 * ordinary HIP device allocations (sipg_malloc or hipMalloc).  Every
 * function returns an int status (SIPG_OK = 0) and never exits the process;
 * sipg_strerror() turns a status into text.
 *
 * Vector layouts.  The reference ("mode_vector" order) stores the dof
 * (i1,i2,e) of element e = dimx*y + x at ((i1*(p+1)+i2)*noe + e), noe =
 * dimx*dimy, i1 the x-node, i2 the y-node (mode_vector.hpp:108-120,
 * mesh_info.hpp:231-235).  Device vectors of this library are element-major:
 * the same dof lives at (e*(p+1)^2 + i1*(p+1)+i2), so each element's tile is
 * contiguous and the operator's halo reads are whole tiles.  sipg_upload /
 * sipg_download take and return host data in the reference order;
 * sipg_to_device_layout / sipg_to_reference_layout convert device arrays
 * (e.g. a reference mode_vector's device array).  Vector kernels (dot, axpy,
 * ...) do not depend on the layout.
 *
 * Penalty: `pen` is the reference kernels' pen (sipg_sem_2d_gpu_kernels.hpp:
 * 243-244, 308-309); the physical SIPG sigma of sigma/h*int[u][v] is 2*pen.
 */
#ifndef SIPG_HIP_H
#define SIPG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    SIPG_OK = 0,
    SIPG_ERR_INVALID = 1,      /* bad argument (sizes, null pointers, degree) */
    SIPG_ERR_ALIAS = 2,        /* input and output of an apply overlap        */
    SIPG_ERR_UNSUPPORTED = 3,  /* degree / dtype combination not built         */
    SIPG_ERR_HIP = 4,          /* a HIP runtime call failed                    */
    SIPG_ERR_NOMEM = 5,        /* device allocation failed                     */
    SIPG_NOT_CONVERGED = 6,    /* iteration cap reached before the tolerance   */
    SIPG_ERR_NONFINITE = 7,    /* NaN/Inf met in a residual or dot product     */
    SIPG_ERR_BREAKDOWN = 8,    /* CG breakdown: p.Ap <= 0 (A not SPD)          */
    SIPG_ERR_BUDGET = 9,       /* request exceeds the configured memory budget */
    SIPG_ERR_CANCELLED = 10,   /* request cancelled or timed out               */
    SIPG_ERR_NOT_FOUND = 11    /* unknown request / handle id                  */
};

/* ---- enums ------------------------------------------------------------- */
enum { SIPG_F64 = 0, SIPG_F32 = 1 };                 /* vector element type */
enum { SIPG_TOL_ABS = 0, SIPG_TOL_REL = 1 };          /* CG stopping test    */

typedef struct sipg_op sipg_op;   /* operator: mesh, degree, penalty, tables */

/* ---- library ----------------------------------------------------------- */
const char *sipg_strerror(int status);
/* Highest polynomial degree an operator accepts: 63, the reference's
 * 64-point LGL tables (device_const.hpp:10-18). */
int sipg_max_degree(void);
/* Highest degree of the compiled fused operator kernels (16).  Above it the
 * operator runs on run-time tables (two launches per apply), CG and the
 * mixed CG work as below it, and the preconditioned solvers, batches and
 * the server return SIPG_ERR_UNSUPPORTED. */
int sipg_max_fused_degree(void);
/* Device count / selection helpers (thin HIP wrappers, no torch). */
int sipg_device_count(int *count);
int sipg_set_device(int device);
int sipg_device_synchronize(void);

/* Host-only basis tables for degree p (no device work): LGL nodes and
 * weights (p+1 each, LGL_quadrature_table.hpp:84-146), D[mode*(p+1)+node]
 * = phi'_mode(z_node) (sem_function.hpp:86-110) and the 1-D stiffness K1
 * (laplacian_operator.hpp:31-43), (p+1)^2 each.  Any pointer may be NULL.
 * Valid for 1 <= p <= 63. */
int sipg_basis_tables(int p, double *nodes, double *weights, double *D, double *K1);

/* ---- device memory (library allocator; plain HIP device memory) -------- */
int sipg_malloc(void **ptr, size_t bytes);
int sipg_free(void *ptr);
int sipg_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sipg_memcpy_d2h(void *dst, const void *src, size_t bytes);
int sipg_memcpy_d2d(void *dst, const void *src, size_t bytes);
int sipg_memset_zero(void *dst, size_t bytes);

/* ---- operator ---------------------------------------------------------- */
/* Replaces the reference's process-global __constant__ tables
 * (device_const.hpp:10-18, 70-174) and the launch set-up of the
 * sipg_sem_2d constructor (sipg_sem_2d.hpp:96-113): the op owns its degree
 * tables in device memory, so ops of different degrees coexist.
 * stream: a hipStream_t, or NULL for the default stream. */
int sipg_op_create(sipg_op **op, int dimx, int dimy, int p, double pen, int dtype, void *stream);
int sipg_op_destroy(sipg_op *op);
/* n = dimx*dimy*(p+1)^2 */
int sipg_op_size(const sipg_op *op, int64_t *n);
int sipg_op_set_stream(sipg_op *op, void *stream);
/* Host copies of the op's tables (lengths p+1, p+1, (p+1)^2, (p+1)^2):
 * LGL nodes/weights (LGL_quadrature_table.hpp:84-146), D[mode*(p+1)+node]
 * (device_const.hpp:115-127), K1 (laplacian_operator.hpp:31-43). */
int sipg_op_tables(const sipg_op *op, double *nodes, double *weights, double *D, double *K1);

/* Contraction kernel of the op's applies.  No counterpart in the
 * reference: SIPG_KERNEL_VALU runs the element contractions on the vector
 * ALUs with an element's lines spread over a workgroup, SIPG_KERNEL_WAVE
 * on the vector ALUs with whole elements per wave (p = 8), SIPG_KERNEL_MFMA
 * on the matrix cores (FP32, p = 15); a kernel that does not exist for the
 * op's dtype and degree falls back to SIPG_KERNEL_VALU.  SIPG_KERNEL_AUTO
 * (the default) picks the fastest.  A mixed-precision solve passes the
 * setting of the FP64 op to its FP32 inner cycles.  sipg_op_kernel reports
 * the kernel the op's applies in `dtype` use (the op's own dtype, or FP32
 * for the inner cycles of an FP64 op). */
enum { SIPG_KERNEL_AUTO = 0, SIPG_KERNEL_VALU = 1, SIPG_KERNEL_MFMA = 2, SIPG_KERNEL_WAVE = 3 };
int sipg_op_set_kernel(sipg_op *op, int kernel);
int sipg_op_kernel(const sipg_op *op, int dtype, int *kernel);

/* y = A x.  Replaces sipg_sem_2d<T>::_mvm (sipg_sem_2d.hpp:167-201): the
 * three launches volume + flux_term6a + flux_term6b become one fused
 * kernel.  x, y: device arrays of n elements of the op's dtype; they must
 * not overlap (SIPG_ERR_ALIAS).  Asynchronous on the op's stream. */
int sipg_apply(sipg_op *op, const void *x, void *y);

/* ---- CG vector kernels (op's dtype, length n, op's stream) ------------- */
/* *result = a.b, accumulated in double, deterministic order.  Replaces
 * cublas_dot (cublas_wrapper.hpp:62-84).  Synchronous (result on host). */
int sipg_dot(sipg_op *op, const void *a, const void *b, double *result);
/* y += alpha x */
int sipg_axpy(sipg_op *op, double alpha, const void *x, void *y);
/* p = r + beta p   (beta_kernel, conjugate_gradient_kernels.hpp:26-38) */
int sipg_xpby(sipg_op *op, const void *r, double beta, void *p);
/* x += alpha p; r -= alpha Ap   (alpha_kernel, conjugate_gradient_kernels.hpp:5-23) */
int sipg_alpha_update(sipg_op *op, double alpha, void *x, const void *p, void *r, const void *Ap);

/* Operator-free forms of the same vector kernels, for callers that hold
 * only raw vectors (e.g. a CG over a user-defined abs_mvm).  n elements of
 * dtype; stream as in sipg_op_create. */
int sipg_vec_dot(int dtype, int64_t n, const void *a, const void *b, double *result, void *stream);
int sipg_vec_axpy(int dtype, int64_t n, double alpha, const void *x, void *y, void *stream);
int sipg_vec_xpby(int dtype, int64_t n, const void *r, double beta, void *p, void *stream);
/* y = x on the device (copy(), mode_vector.hpp:314-359, into an existing
 * buffer of the same length; overlapping buffers rejected) */
int sipg_vec_copy(int dtype, int64_t n, const void *x, void *y, void *stream);
int sipg_vec_alpha_update(int dtype, int64_t n, double alpha, void *x, const void *p, void *r,
                          const void *Ap, void *stream);
/* sipg_vec_dot without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops
 * that must not call hipMalloc, harnesses that forbid it).  Same arguments,
 * result and summation order as sipg_vec_dot, for any n >= 0.  Its scratch
 * (the block sums and the sum, 1025 doubles) is a fixed device array of
 * the library, one per device, held until the process ends; calls that
 * overlap in time on one device race on it. */
int sipg_vec_dot_noalloc(int dtype, int64_t n, const void *a, const void *b, double *result, void *stream);

/* ---- CG solve ---------------------------------------------------------- */
/* Unpreconditioned CG on A x = b with device-resident scalars.  Replaces
 * conjugate_gradient (conjugate_gradient_gpu.hpp:18-109).  x: initial
 * guess in, solution out.  tol_mode SIPG_TOL_ABS stops on sqrt(r.r) <= tol
 * (the reference's test, :63,69), SIPG_TOL_REL on sqrt(r.r) <= tol*||b||.
 * maxit <= 0 means n (the reference's cap, :65).  On return *iters is the
 * iteration count and *res_norm the final sqrt(r.r) of the recursive
 * residual.  Returns SIPG_OK, SIPG_NOT_CONVERGED, SIPG_ERR_BREAKDOWN or
 * SIPG_ERR_NONFINITE (x then holds the last iterate). */
int sipg_cg(sipg_op *op, void *x, const void *b, double tol, int tol_mode, int maxit,
            int *iters, double *res_norm);

/* Extended CG.  Same iteration as sipg_cg plus, optionally, residual
 * replacement (van der Vorst & Ye 2000): whenever the recursive sqrt(r.r)
 * has fallen by `replace_factor` since the last true residual, r is
 * recomputed as b - A x between the x/r and the p update of an iteration
 * (the increments since the last one are kept apart and folded into x
 * there); with `true_residual_stop` the iteration finishes only on a true
 * residual below the tolerance (once the recursive one meets it, the true
 * residual is inserted within a few iterations; if that one fails, b - A x
 * of the current iterate is also evaluated every 32 iterations without
 * replacing r, and the first that meets the tolerance ends the solve --
 * the workspace holds one more vector for it).  This has no counterpart
 * in the reference, whose CG stops on the recursive residual only; it is
 * what makes a relative residual near the FP64 floor reachable. */
typedef struct {
    double tol;              /* stopping tolerance                              */
    int tol_mode;            /* SIPG_TOL_ABS or SIPG_TOL_REL                    */
    int maxit;               /* <= 0: n                                         */
    double replace_factor;   /* in (0,1): recompute r at each such drop; 0: off */
    int true_residual_stop;  /* 1: confirm convergence on b - A x               */
    int restart;             /* 0: keep the search direction at each true
                                residual (reliable update); 1: restart from
                                p = r at each (iterative refinement, a fresh
                                Krylov space per cycle); 2: keep it while each
                                true residual improves on the previous one by
                                2x or more, then restart at every later one  */
} sipg_cg_opts;

typedef struct {
    int status;              /* same codes as the return value                  */
    int iters;               /* CG iterations                                   */
    int replacements;        /* true-residual recomputations                    */
    int pad;
    double res_norm;         /* final sqrt(r.r) of the iterated residual        */
    double true_res_norm;    /* ||b - A x|| at the last recomputation, or -1    */
    double b_norm;           /* ||b||                                           */
} sipg_cg_info;

int sipg_cg_ex(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info);

/* Mixed-precision solve (no counterpart in the reference): the CG
 * iteration runs in FP32 (operator apply, vector updates, FP64-accumulated
 * dots) on an FP32 twin of `op`; each time the FP32 residual has fallen by
 * `replace_factor` (default 0.1 when 0 is passed) the FP32 iterate is
 * folded into the FP64 solution and the true residual b - A x is
 * recomputed in FP64 (defect correction).  With opts->restart = 1 each
 * cycle restarts the FP32 CG from that residual (iterative refinement);
 * with 0 the search direction is kept (a reliable update, Sleijpen & van der
 * Vorst 1996; Clark et al. 2010), which is cheaper while eps32 * kappa(A)
 * is small but loses convergence once it is not; 2 switches from the first
 * to the second when the true residual stagnates.  The final convergence
 * test is on the FP64 true residual.  op must be SIPG_F64; x (in/out) and b
 * are FP64 device vectors. */
int sipg_cg_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info);

/* Block-Jacobi preconditioned CG: replaces preconditioned_conjugate_gradient
 * (conjugate_gradient_gpu.hpp:116-213) with block_diagonal_preconditioner_2d
 * (laplacian_operator.hpp:95-186).  M is the exact diagonal block of A of an
 * interior element, (K U W + W U K) with K the 1-D stiffness plus the
 * element's own face terms, applied to every element and inverted by fast
 * diagonalisation; the reference instead inverts a dense block built with
 * pen truncated to an integer and keeps part of it (mode_matrix.hpp:495-516).
 * opts and info as for sipg_cg_ex: res_norm and the stopping test are on
 * r.r (not r.z), with the same reliable updates and true-residual stop.
 * x is in/out in the op's dtype.  SIPG_ERR_INVALID when the interior block
 * is not positive definite (penalty too small). */
int sipg_pcg(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info);

/* Mixed-precision block-Jacobi PCG (no counterpart in the reference): as
 * sipg_cg_mixed, with the FP32 inner cycles preconditioned as in sipg_pcg
 * (FP32 tables of the op's FP32 twin).  op must be SIPG_F64. */
int sipg_pcg_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info);

/* PCG with a two-level preconditioner: the block-Jacobi blocks of sipg_pcg
 * plus an exact coarse solve on the continuous bilinear functions of the
 * element vertices (zero on the boundary),
 *   M2^-1 r = M_BJ^-1 r + P0^T A0^-1 P0 r,
 * A0 = P0 A P0^T (the Q1 stiffness: on continuous functions every face term
 * of the SIPG form vanishes), inverted by discrete sine transforms.  The
 * method is preconditioned_conjugate_gradient (conjugate_gradient_gpu.hpp:
 * 116-213); the coarse level has no counterpart in the reference, whose
 * preconditioner is block-diagonal only.  The iteration count no longer
 * grows with the mesh.  Options and results as sipg_pcg.  With
 * replace_factor > 0 the x increments of the first 4 iterations (the
 * coarse-level ones, large against the final solution) are accumulated in
 * compensated sums (a low-order part beside each entry, in one more
 * workspace vector) and then folded into x, so that b - A x keeps following
 * the recursive residual and a replacement every 100x drop is enough. */
int sipg_pcg2(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info);
/* z = M2^-1 r (two-level, as sipg_pcg2), asynchronous on the op's stream. */
int sipg_prec2_apply(sipg_op *op, const void *r, void *z);
/* sipg_pcg_mixed with the two-level preconditioner: FP32 two-level PCG
 * cycles (the coarse solve in FP64) inside FP64 iterative refinement on the
 * true residual.  Options and results as sipg_pcg_mixed. */
int sipg_pcg2_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info);

/* z = M^-1 r with the preconditioner of sipg_pcg (device vectors of the
 * op's dtype, r and z distinct); replaces _prec_mvm (abs_mvm.hpp:15). */
int sipg_prec_apply(sipg_op *op, const void *r, void *z);

/* ||b - A x||_2 / ||b||_2 in FP64 (op must be SIPG_F64). */
int sipg_true_relres(sipg_op *op, const void *x, const void *b, double *relres);

/* ---- batches of independent solves ------------------------------------ */
/* A batch holds any number of independent SIPG problems (instance i: a
 * dimx[i] x dimy[i] mesh of degree p[i] with penalty pen[i]) and solves them
 * together, FP64.  No counterpart in the reference, whose degree tables are
 * process globals (device_const.hpp:10-18) so that one process handles one
 * degree at a time; each instance follows the single-problem semantics of
 * sipg_cg_ex.  Vectors of a batch are one device array of sipg_batch_size
 * elements: instance i occupies [offset_i, offset_i + n_i), element-major
 * (see the layout note above).  stream: as in sipg_op_create, except that a
 * null stream makes the batch create its own (CG iterations are replayed
 * from a captured hipGraph, which the null stream cannot hold). */
typedef struct sipg_batch sipg_batch;
int sipg_batch_create(sipg_batch **out, int ninst, const int *dimx, const int *dimy, const int *p,
                      const double *pen, void *stream);
int sipg_batch_destroy(sipg_batch *b);
int sipg_batch_size(const sipg_batch *b, int64_t *total_n);
int sipg_batch_offset(const sipg_batch *b, int inst, int64_t *offset, int64_t *n);
/* Chooses the operator kernel's block shape for every degree of the batch
 * by timing the candidates (16x4, 8x4, 8x2 elements) on the batch's own
 * instances, and records the choice in a process-wide cache keyed by the
 * device architecture, the degree and the size class of that degree's
 * work (the binary order of its element count).  On first use without a
 * call, a batch whose every degree is in the cache takes those shapes
 * without timing; otherwise it is timed as here.  sipg_batch_tuning
 * reports the choice for degree p and its time per apply of all instances
 * of that degree (as timed when the cache entry was made; -1 before
 * tuning). */
int sipg_batch_tune(sipg_batch *b);
int sipg_batch_tuning(const sipg_batch *b, int p, int *bx, int *by, double *us_per_apply);
/* Fixes the block shape (bx x by elements, one of the candidates) for the
 * instances of degree p, in place of tuning (for callers that tune offline,
 * and for tests of every shape). */
int sipg_batch_set_shape(sipg_batch *b, int p, int bx, int by);
/* use_graphs 1 (the default): the solvers capture 8 CG iterations in a
 * hipGraph once and replay it; 0: they launch the same kernels in the same
 * order directly on the batch stream each time (identical results; for
 * callers or tools that cannot use graphs).  No counterpart in the
 * reference. */
int sipg_batch_set_graphs(sipg_batch *b, int use_graphs);
/* y = A_i x_i for every instance (x, y: batch vectors, not overlapping). */
int sipg_batch_apply(sipg_batch *b, const void *x, void *y);
/* Fourier forcing of instance `inst` into its part of the batch vector rhs
 * (as sipg_rhs_fourier; square meshes). */
int sipg_batch_rhs_fourier(sipg_batch *b, int inst, int nmodes, const double *a, const int *m,
                           const int *n, void *rhs);
/* Host data of one instance in the reference order <-> its part of a batch
 * vector (synchronous). */
int sipg_batch_upload(sipg_batch *b, int inst, const double *host, void *dev);
int sipg_batch_download(sipg_batch *b, int inst, const void *dev, double *host);
/* CG on every instance: x (in/out) and rhs are batch vectors.  Uses tol,
 * tol_mode, maxit (<= 0: n_i per instance), replace_factor and
 * true_residual_stop of opts with the meaning of sipg_cg_ex (reliable
 * updates keeping the search direction; restart is not used).  info, if not
 * null, receives one record per instance.  Returns SIPG_OK when every
 * instance converged, else the first failing instance's status. */
int sipg_batch_cg(sipg_batch *b, void *x, const void *rhs, const sipg_cg_opts *opts, sipg_cg_info *info);

/* sipg_batch_cg with the block-Jacobi preconditioner of sipg_pcg for every
 * instance (preconditioned_conjugate_gradient, conjugate_gradient_gpu.hpp:
 * 116-213, batched).  Every instance must use the standard penalty
 * pen = 5 (p+1)^2, whose preconditioner tables are compiled in; otherwise
 * SIPG_ERR_UNSUPPORTED.  Options and results as sipg_batch_cg. */
int sipg_batch_pcg(sipg_batch *b, void *x, const void *rhs, const sipg_cg_opts *opts, sipg_cg_info *info);

/* ---- serving: continuous batching of a request stream ------------------- */
/* A server solves a stream of independent FP64 problems (square meshes).
 * No counterpart in the reference.  Submitted requests wait in arrival
 * order and are admitted as soon as a slot (up to 4096 on the device) and
 * device memory within hbm_budget_bytes are free: a request of n unknowns
 * needs 48 n bytes (solution, right-hand side and four CG vectors).  Each
 * call of sipg_server_progress advances every admitted request by a few CG
 * iterations (one replay of a captured hipGraph), confirms finished ones
 * on their true residual, retires them and admits new ones; it blocks
 * until that work is done.  A request's result does not depend on the other
 * requests in flight.  Requests whose recursion meets the tolerance are
 * checked on b - A x (as sipg_cg_ex with true_residual_stop). */
typedef struct sipg_server sipg_server;
int sipg_server_create(sipg_server **out, double hbm_budget_bytes, void *stream);
int sipg_server_destroy(sipg_server *s);
/* Block shape of the operator kernel for degree p (a sipg_batch_tune
 * result; 8x4 by default), while no request of degree p is running. */
int sipg_server_set_shape(sipg_server *s, int p, int bx, int by);
/* Solver of every request: SIPG_METHOD_CG (the default) or SIPG_METHOD_PCG,
 * the block-Jacobi preconditioned CG of sipg_batch_pcg
 * (preconditioned_conjugate_gradient, conjugate_gradient_gpu.hpp:116-213).
 * Only while the server holds no request (else SIPG_ERR_INVALID).  Under
 * PCG a request of n unknowns needs 56 n bytes (one more vector), submit
 * refuses a penalty other than 5 (p+1)^2 with SIPG_ERR_UNSUPPORTED, and so
 * does set_shape a block shape the preconditioner kernel lacks. */
enum { SIPG_METHOD_CG = 0, SIPG_METHOD_PCG = 1 };
int sipg_server_set_method(sipg_server *s, int method);
/* Queue a request: dim x dim mesh of degree p, penalty pen, Fourier forcing
 * (as sipg_rhs_fourier), tolerance (tol, tol_mode), maxit <= 0: n.
 * SIPG_ERR_BUDGET if it can never fit the budget. */
int sipg_server_submit(sipg_server *s, int dim, int p, double pen, int nmodes, const double *a,
                       const int *m, const int *n, double tol, int tol_mode, int maxit, int64_t *id);
/* One round (see above); *n_done = completed requests not yet taken. */
int sipg_server_progress(sipg_server *s, int *n_done);
/* Oldest completed request: its id and record, and (if x_host is not null,
 * with room for n_cap values) its solution in the reference order; releases
 * its memory.  SIPG_ERR_NOT_FOUND if none has completed. */
int sipg_server_take(sipg_server *s, int64_t *id, sipg_cg_info *info, double *x_host, int64_t n_cap);
int sipg_server_stats(const sipg_server *s, int *queued, int *running, int *done_held, double *bytes_used);
/* When a request was submitted, admitted (-1: never, it was cancelled in
 * the queue) and completed, in seconds on a monotonic clock that starts at
 * sipg_server_create: admitted - submitted is the queueing delay,
 * completed - admitted the service time. */
typedef struct {
    double submitted;
    double admitted;
    double completed;
} sipg_request_times;
/* Id and number of unknowns of the oldest completed request, without
 * taking it (so a caller can size the solution buffer, or take it without
 * one); SIPG_ERR_NOT_FOUND if none has completed. */
int sipg_server_peek(const sipg_server *s, int64_t *id, int64_t *n);
/* sipg_server_take that also returns the request's times (times may be
 * null). */
int sipg_server_take_ex(sipg_server *s, int64_t *id, sipg_cg_info *info, double *x_host, int64_t n_cap,
                        sipg_request_times *times);
/* Cancels request id.  A queued request completes at once, a running one
 * at the next sipg_server_progress with its iterate so far; either is then
 * taken like any other request, with status SIPG_ERR_CANCELLED (a request
 * cancelled in the queue returns x = 0).  SIPG_ERR_NOT_FOUND for an unknown
 * or already completed id. */
int sipg_server_cancel(sipg_server *s, int64_t id);
/* Requests not completed within `seconds` of their submission are
 * cancelled as by sipg_server_cancel, checked at each
 * sipg_server_progress; 0 (the default) means no limit. */
int sipg_server_set_timeout(sipg_server *s, double seconds);

/* ---- right-hand side and error norms (device) -------------------------- */
/* Seeded Fourier forcing u = sum_k a_k sin(m_k pi x) sin(n_k pi y) drawn
 * from std::mt19937_64(seed): K = kmin + r%(kmax-kmin+1), then per mode
 * a = -1 + 2*((r>>11)*2^-53), m = 1 + r%mmax, n = 1 + r%mmax, one raw
 * 64-bit draw r each, in that order.  Arrays must hold kmax entries. */
int sipg_forcing_draw(uint64_t seed, int kmin, int kmax, int mmax,
                      int *nmodes, double *a, int *m, int *n);
/* b = load vector of the Fourier problem on the op's square mesh, with
 * homogeneous Dirichlet data (sipg_sem_2d.hpp:257-262 with g = 0), in the
 * op's dtype, computed on the device. */
int sipg_rhs_fourier(sipg_op *op, int nmodes, const double *a, const int *m, const int *n, void *b);
/* Error norms of the device vector x (op's dtype and layout) against the
 * Fourier solution u = sum_k a_k sin(m_k pi x) sin(n_k pi y), computed on
 * the device: out[0] = max nodal error, out[1] = L2 error, out[2] = H1 error
 * (L2 plus both derivative errors), by GLL quadrature on each element as
 * sipg_sem_2d::err_norms (sipg_sem_2d.hpp:311-370; members max_err, L2_err,
 * H1_err, sipg_sem_2d.hpp:73-81).  Square meshes. */
int sipg_err_norms_fourier(sipg_op *op, const void *x, int nmodes, const double *a, const int *m, const int *n,
                           double *out);

/* ---- general problem set-up: any f, Dirichlet data g, any exact solution */
/* A function of the point (x, y) with a caller context, evaluated on the
 * host (the reference's FLOAT_TYPE (*)(FLOAT_TYPE, FLOAT_TYPE) arguments,
 * sipg_sem_2d.hpp:83-92, with a context pointer added). */
typedef double (*sipg_fn2)(double x, double y, void *ctx);
/* Physical coordinates of every GLL node of the op's square mesh of the
 * unit square, element-major like device vectors (x[k], y[k] for k = e
 * (p+1)^2 + i1 (p+1) + i2): element corners as build_square_mesh.hpp:
 * 250-288 places them (element e = row * dim + col, row 0 at the top), plus
 * (z_i + 1) / (2 dim) per direction (sipg_sem_2d.hpp:250-255).  x, y: host
 * arrays of the op's length. */
int sipg_quad_points(sipg_op *op, double *x, double *y);
/* b = load vector of -lap u = f with Dirichlet data g, as
 * sipg_sem_2d::compute_rhs (sipg_sem_2d.hpp:232-305): (w_i1 h/2)(w_i2 h/2)
 * f at every node, plus on the four sides of the square pen w_i g and the
 * consistency terms -/+ w_i phi_j'(side) g.  f, g: host arrays of the op's
 * length holding f and g at the points of sipg_quad_points (g is read on
 * the boundary nodes of boundary elements only; g NULL: g = 0).  b: device
 * vector of the op (its dtype).  Assembled on the host in double in the
 * reference's order of terms.  Square meshes. */
int sipg_rhs_nodal(sipg_op *op, const double *f, const double *g, void *b);
/* As sipg_rhs_nodal with f and g as functions evaluated at the node points
 * (compute_rhs(f, u_ex), sipg_sem_2d.hpp:232-235).  g NULL: g = 0. */
int sipg_rhs(sipg_op *op, sipg_fn2 f, sipg_fn2 g, void *ctx, void *b);
/* Error norms of the device vector x against an exact solution given by
 * nodal samples u, ux = du/dx, uy = du/dy at the points of sipg_quad_points
 * (host arrays of the op's length): out = {max nodal error, L2 error, H1
 * error} as sipg_err_norms_fourier (sipg_sem_2d.hpp:311-370), reduced on
 * the device.  Square meshes. */
int sipg_err_norms_nodal(sipg_op *op, const void *x, const double *u, const double *ux, const double *uy,
                         double *out);
/* As sipg_err_norms_nodal with the exact solution and its derivatives as
 * functions (err_norms(u, f, u_ex, dx_u_ex, dy_u_ex), sipg_sem_2d.hpp:
 * 311-316). */
int sipg_err_norms(sipg_op *op, const void *x, sipg_fn2 u, sipg_fn2 ux, sipg_fn2 uy, void *ctx, double *out);

/* The solution as text, as sipg_sem_2d::print_result (sipg_sem_2d.hpp:
 * 375-403): per element (row-major from the top-left), per node (x index
 * outer, y inner) one line "x<TAB>y<TAB>u", each field printf "%10g" (the
 * ostream's setw(10) at its default precision), and an empty line after
 * each element.  path NULL or "": standard error (where the reference
 * prints).  x: device vector of the op.  Square meshes. */
int sipg_write_result(sipg_op *op, const void *x, const char *path);

/* ---- measurement -------------------------------------------------------- */
/* Runs `reps` back-to-back applies y = A x on the op's stream between two
 * HIP events and returns the mean launch-to-launch time in milliseconds
 * (the reference's event-timed driver: mvm_performance_test.cu:44-65). */
int sipg_time_apply(sipg_op *op, const void *x, void *y, int reps, double *ms_per_apply);

/* ---- layout helpers ---------------------------------------------------- */
/* Host FP64 data in the reference order <-> device vector of the op's dtype
 * in the element-major order (synchronous; converts FP64 <-> FP32). */
int sipg_upload(sipg_op *op, const double *host, void *dev);
int sipg_download(sipg_op *op, const void *dev, double *host);
/* Device array in the reference order <-> element-major device order, in
 * the op's dtype (asynchronous on the op's stream; arrays must not overlap).
 * Replace copy() of mode_vector.hpp:314-359 when data is already on the
 * device in the reference's layout. */
int sipg_to_device_layout(sipg_op *op, const void *ref, void *dev);
int sipg_to_reference_layout(sipg_op *op, const void *dev, void *ref);

#ifdef __cplusplus
}
#endif
#endif

// sipg_internal.hpp -- shared definitions of libsipg_hip (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstddef>
#include <vector>

#include "../../include/sipg_hip.h"

#define SIPG_HIP_TRY(expr)                                  \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) {                             \
            sipg_internal::set_last_hip_error(_e);          \
            return _e == hipErrorOutOfMemory ? SIPG_ERR_NOMEM : SIPG_ERR_HIP; \
        }                                                   \
    } while (0)

namespace sipg_internal {

// Highest degree with a compiled fused kernel.
constexpr int kMaxDegree = 16;
// Highest degree of the run-time-table operator above it (64 LGL points,
// the reference's table size).
constexpr int kMaxGenericDegree = 63;

void set_last_hip_error(hipError_t e);

// ---- host basis numerics (basis.cpp) --------------------------------------
// GLL nodes/weights on [-1,1], p+1 points (LGL_quadrature_table.hpp:84-146).
void gll(int p, double *nodes, double *weights);
// D[mode*(p+1)+node] = phi'_mode(z_node) (sem_function.hpp:86-110).
void dphi(int p, const double *nodes, double *D);
// K1[a*(p+1)+b] = sum_k w_k D_a(k) D_b(k) (laplacian_operator.hpp:31-43).
void stiffness_1d(int p, const double *w, const double *D, double *K1);

// Device table layout used by every kernel, in the op's dtype:
//   [0, N1)              w       LGL weights
//   [N1, N1+N1^2)        K1      1-D stiffness, row-major
//   [.., +N1)            D0      D0[i] = D(i, node 0)  = phi'_i(-1)
//   [.., +N1)            DP      DP[i] = D(i, node p)  = phi'_i(+1)
//   [.., +N1)            z       LGL nodes
inline int table_len(int p) { return (p + 1) * (p + 1) + 4 * (p + 1); }
void build_tables(int p, std::vector<double> &tab);

// ---- CG scalar state, resident on the device ------------------------------
// The iteration launched k-th (counted by the host) reads slot k%2 and its
// last kernel writes slot (k+1)%2, so no kernel reads a word that a block of
// the same launch writes.  An iteration that finds its slot finished is a
// no-op that only carries the slot forward; host-side launch parity and the
// slots therefore always agree.
struct cg_slot {
    double rr;      // r.r entering the iteration
    double alpha;   // of the iteration that produced this slot
    double beta;
    double pAp;
    double rz;      // preconditioned CG: r.z entering the iteration (CG: r.r)
    int it;         // iterations completed
    int done;       // 0 running, 1 finished (tolerance met or iteration cap)
    int checked;    // 1 when rr is a true residual b - A x
    int flags;      // kSlot* bits
};
// cg_slot::flags.  kSlotReplaced / kSlotRestart are set on the slot an
// iteration reads, between its x/r update and its p update, when r was
// just replaced by the true residual (and p is to restart from r);
// kSlotWantCheck is set by the p update when the recursive residual meets
// the tolerance but a true residual is required to finish.
constexpr int kSlotReplaced = 1;
constexpr int kSlotRestart = 2;
constexpr int kSlotWantCheck = 4;
// Batched and served solves: failed final true-residual checks that gain
// less than 10% on the best failed one (counted over the whole solve) before
// the solve is declared stuck at the floor of FP64 CG and ends, not
// converged.  Near the floor the true residual of successive checks
// scatters around a fixed level, so a lucky low one does not reset the
// count.
constexpr int kStallChecks = 2;
// The same for PCG.  Its residual 2-norm swings several-fold from one
// iteration to the next while its envelope falls, and a final check lands
// a few iterations after the recursion met the tolerance, at a random point
// of that swing: two checks in a row can miss a 10% gain well above the
// FP64 floor.
constexpr int kStallChecksPcg = 4;
// Once the recursion has met the tolerance, a final check is requested when
// it has fallen 4x below the last true residual, and also after this many
// iterations since the last true residual was requested: near the FP64
// floor that 4x drop can take thousands of iterations while the true
// residual creeps down by more than 10% per check, and the stall count
// above only advances with checks.  it / 40 is about a quarter of a decade
// of the solve's own average rate (ten decades in `it` iterations), so a
// solve that is still converging gains more than 10% between checks.
inline int force_check_interval(int it) { return it / 40 > 64 ? it / 40 : 64; }
// The stall count ends a solve only once its best failed check is also
// this many iterations old: a quarter of the solve so far.  A solve that is
// converging sets a new best (10% below the last) well within that window,
// even when single checks land on the peaks of its residual's swings; one
// at the FP64 floor stops about a quarter of its iterations after arriving.
inline int stall_window(int it) { return it / 4 > 128 ? it / 4 : 128; }
// The floor rules act only once the last true residual (r.r = rr_ref) is
// below kFloorWatchRel of |b| in norm.  The FP64 floor of this operator
// lies far below that for any mesh that fits in memory (about 8e-11 at
// 256 x 256 elements of degree 8, 7e-10 at 1024 x 1024, growing about as
// the mesh side to the power 1.5); above it the residual 2-norm of PCG
// swings up to ~100x over a hundred iterations early in a solve (it
// minimises the energy norm of the error, not the residual), which the
// rules would read as a stall.
constexpr double kFloorWatchRel = 1e-6;
inline bool floor_in_reach(double rr_ref, double bb) {
    return rr_ref <= kFloorWatchRel * kFloorWatchRel * bb;
}
struct cg_state {
    cg_slot s[2];
    double bb;       // b.b
    double stop;     // stopping threshold on sqrt(r.r) of the recursion
    double stop_true;  // threshold for a true residual (batched solver: the
                       // recursion's is tightened after a failed check)
    double rr0;      // initial r.r
    double rr_ref;   // r.r of the last true residual
    int maxit;
    int err;         // 0, 2 breakdown (p.Ap <= 0), 3 non-finite -- terminal
    int nrep;        // true-residual recomputations so far
    int true_stop;   // 1: finishing requires a true residual below the tolerance
};

// ---- batched solves (kernels_batch.hip, batch.cpp) ------------------------
// One instance of a batch: its vectors start at `off` in the concatenated
// arrays (element-major within the instance).
struct batch_inst {
    int64_t off;
    int dimx, dimy;
    double pen;
};
// Per-instance solver parameters.
struct batch_ctl {
    double tol;
    int tol_mode, maxit;
};
// A BX x BY block of an instance for the operator kernel.
struct batch_item {
    int inst, x0, y0, pad;
};
// A range [begin, end) of the concatenated vectors, inside one instance.
struct batch_chunk {
    int inst, pad;
    int64_t begin, end;
};
// Candidate block shapes of the batched operator (the autotuner picks one
// per degree): 16x4, 8x4, 8x2 elements.
constexpr int kBatchShapes = 3;
extern const int kBatchShapeBX[kBatchShapes];
extern const int kBatchShapeBY[kBatchShapes];
bool batch_shape_supported(int p, int shape);

// Reduction partial buffers hold at most this many block sums; consumers
// add up at most this many themselves (more are first folded into one by a
// separate reduction).  The wave-owned apply has one block sum per 7 x 4
// elements at p = 8: 2368 at 256 x 256.
constexpr int kMaxPartials = 4096;
constexpr int kVecThreads = 256;

// ---- launchers (kernels_apply.hip, kernels_mfma.hip) ----------------------
// The contraction kernel an apply of (dtype, p) runs for a requested
// SIPG_KERNEL_* setting: SIPG_KERNEL_VALU or SIPG_KERNEL_MFMA.
int resolve_kernel(int dtype, int p, int kernel);
// Number of thread blocks the fused apply uses for a mesh (= partial count).
int apply_num_blocks(int dtype, int p, int dimx, int dimy, int kernel = SIPG_KERNEL_AUTO);
// y = A x; with partials != nullptr also writes per-block sums of x.y and
// returns early when st->s[slot].done is set.
int launch_apply(int dtype, int p, const void *x, void *y, const void *tab, int dimx, int dimy,
                 double pen, double *partials, const cg_state *st, int slot, hipStream_t s,
                 int kernel = SIPG_KERNEL_AUTO);
// the vector-ALU apply with wave-owned elements (kernels_apply_wave.hip)
bool wave_apply_available(int dtype, int p);
int wave_apply_num_blocks(int dtype, int p, int dimx, int dimy);
int launch_apply_wave(int dtype, int p, const void *x, void *y, int dimx, int dimy, double pen, double *partials,
                      const cg_state *st, int slot, hipStream_t s);
// the matrix-core apply (FP32, p = 15)
bool mfma_apply_available(int dtype, int p);
int mfma_apply_num_blocks(int dimx, int dimy);
int launch_apply_mfma(int dtype, int p, const void *x, void *y, int dimx, int dimy, double pen, double *partials,
                      const cg_state *st, int slot, hipStream_t s);
// the run-time-table apply for kMaxDegree < p <= kMaxGenericDegree
// (kernels_generic.hip).  Its face buffer (4 (p+1) values per element)
// lives in the op's table allocation after the tables: such ops allocate
// generic_table_len elements of their dtype for d_tab, of which the first
// generic_face_offset hold the tables.
int generic_apply_num_blocks(int dimx, int dimy);
size_t generic_face_offset(int p, size_t elem);
size_t generic_table_len(int dtype, int p, int dimx, int dimy);
int launch_apply_generic(int dtype, int p, const void *x, void *y, const void *tab, int dimx, int dimy, double pen,
                         double *partials, const cg_state *st, int slot, hipStream_t s);

// ---- launchers (kernels_vec.hip) ------------------------------------------
int vec_num_blocks(int64_t n);
int launch_dot_partials(int dtype, const void *a, const void *b, int64_t n, double *partials,
                        int nblk, hipStream_t s);
int launch_reduce_partials(const double *partials, int np, double *out, hipStream_t s);
int launch_axpy(int dtype, double alpha, const void *x, void *y, int64_t n, hipStream_t s);
int launch_copy(int dtype, const void *x, void *y, int64_t n, hipStream_t s);
int launch_xpby(int dtype, const void *r, double beta, void *p, int64_t n, hipStream_t s);
int launch_alpha_update(int dtype, double alpha, void *x, const void *p, void *r, const void *Ap,
                        int64_t n, hipStream_t s);
// CG pieces with device-resident scalars
int launch_cg_init_finalize(const double *part_rr, const double *part_bb, int np, cg_state *st,
                            double tol, int tol_mode, int maxit, int true_stop, hipStream_t s);
int launch_cg_update_xr(int dtype, void *x, const void *p, void *r, const void *Ap, int64_t n,
                        const double *part_pAp, int np_pAp, double *part_rr, int nblk,
                        cg_state *st, int slot, hipStream_t s, void *xlo = nullptr);
int launch_cg_update_p(int dtype, const void *r, void *p, int64_t n, const double *part_rr,
                       int np_rr, int nblk, cg_state *st, int slot, hipStream_t s);
int launch_residual(int dtype, const void *b, const void *Ax, void *r, int64_t n, hipStream_t s);
// out = (out dtype)(b - Ax) with b, Ax in in_dtype; optionally p = out;
// block sums of (b-Ax).(b-Ax) and, when pbb != nullptr, of b.b.
int launch_true_residual(int in_dtype, int out_dtype, const void *b, const void *Ax, void *out,
                         void *p_copy, int64_t n, double *prr, double *pbb, int nblk, hipStream_t s);
// x += xl; xl = 0   (x in x_dtype, xl in xl_dtype; f64+=f64, f64+=f32, f32+=f32)
int launch_accumulate(int x_dtype, int xl_dtype, void *x, void *xl, int64_t n, hipStream_t s, void *xlo = nullptr);
// x += xi with the rounding error added to xlo (both x's dtype), xi = 0
int launch_comp_fold(int dtype, void *x, void *xlo, void *xi, int64_t n, hipStream_t s);
// out = x + (xl + xlo), and the peek: with prr the block sums of a true residual's
// r.r that did not replace r, finish slot `slot` on it if it meets st->stop
int launch_sum(int dtype, const void *x, const void *xl, void *out, int64_t n, hipStream_t s, const void *xlo = nullptr);
int launch_peek_finalize(const double *prr, int np, cg_state *st, int slot, hipStream_t s);
// Marks the slot an iteration reads as replaced (r is now b - A x and the
// rr block sums are its own), optionally with a restart p = r; a no-op
// for a finished slot.
int launch_mark_replaced(cg_state *st, int slot, int restart, hipStream_t s);

// ---- launchers (kernels_pcg.hip) ------------------------------------------
// z = M^-1 r (block-Jacobi: the interior element block's exact inverse), with
// block sums of r.r and r.z.  mode 1: first r -= alpha Ap (alpha = slot.rz /
// p.Ap from the apply's sums; no-op once the slot is finished); mode 2: also
// p = z; mode 0: neither.  pcg_num_blocks = number of block sums written.
// (pen selects the kernel: the standard 5 (p+1)^2 has its coefficients
// compiled in, any other reads them from prec)
int pcg_num_blocks(int dtype, int p, double pen, int mode, int64_t noe);
// E4 non-null: also the two-level restriction's corner sums of r (as
// launch_coarse_solve's first step), from the same read of r, where
// pcg_rz_fuses_restrict (the standard penalty's kernels, where the LDS has
// room); ignored elsewhere
bool pcg_rz_fuses_restrict(int dtype, int p, double pen);
int launch_pcg_rz(int dtype, int p, double pen, int mode, void *r, const void *Ap, void *z, void *pcopy,
                  int64_t noe, const void *prec, const double *part_pAp, int np_pAp, double *part_rr,
                  double *part_rz, cg_state *st, int slot, hipStream_t s, double *E4 = nullptr);
// p = z + beta p (and x += alpha p unless r was replaced this iteration);
// writes the next slot
int launch_pcg_update_p(int dtype, void *x, void *p, const void *z, int64_t n, const double *part_pAp,
                        int np_pAp, const double *part_rr, const double *part_rz, int np, int nblk,
                        cg_state *st, int slot, hipStream_t s, void *xlo = nullptr);
// x += alpha p, before a true residual replaces r in the same iteration
int launch_pcg_x(int dtype, void *x, const void *p, int64_t n, const double *part_pAp, int np_pAp, int nblk,
                 const cg_state *st, int slot, hipStream_t s, void *xlo = nullptr);
int launch_pcg_init_finalize(const double *part_rr, const double *part_rz, int np, const double *part_bb,
                             int npb, cg_state *st, double tol, int tol_mode, int maxit, int true_stop,
                             hipStream_t s, const double *part_c = nullptr, int npc = 0);
// ---- coarse level of the two-level preconditioner (kernels_coarse.hip) ----
// Device pieces, in one allocation per op: DST matrices Qx ((dimx-1)^2),
// Qy ((dimy-1)^2), 1/Lambda ((dimy-1) x (dimx-1)), the GLL positions
// xi = (z+1)/2 (p+1), corner sums E4 (4 noe), S, T1, T2 ((dimy-1) x
// (dimx-1)), the zero-bordered vertex grid Cpad ((dimy+1) x (dimx+1)) and
// the S.C block partials.
struct coarse_tables {
    double *Qx = nullptr, *Qy = nullptr, *Linv = nullptr, *xi = nullptr, *E4 = nullptr, *S = nullptr,
           *T1 = nullptr, *T2 = nullptr, *Cpad = nullptr, *part = nullptr;
};
int coarse_num_partials(int dimx, int dimy);
// C = A0^-1 P0 r into Cpad, with the S.C partials (nothing when a side has
// one element: no interior vertex); have_e4: the corner sums are already in
// c.E4 (launch_pcg_rz wrote them); st/slot: inside a solve, the product
// kernels return at once when that slot is done
int launch_coarse_solve(int dtype, int p, const void *r, int dimx, int dimy, const coarse_tables &c, hipStream_t s,
                        bool have_e4 = false, const cg_state *st = nullptr, int slot = 0);
// v += P0^T C
int launch_coarse_add(int dtype, int p, void *v, int dimx, int64_t n, const coarse_tables &c, int nblk,
                      hipStream_t s);
// the PCG p update with the coarse level (r.z gains S.C, p gains P0^T C)
int launch_pcg2_update_p(int dtype, int p, void *x, void *pv, const void *z, int64_t n, const double *part_pAp,
                         int np_pAp, const double *part_rr, const double *part_rz, int np, int npc, int dimx,
                         const coarse_tables &c, int nblk, cg_state *st, int slot, hipStream_t s, void *xlo = nullptr);
// host tables of the coarse solve: Q (n x n, n = dim - 1), and the 1-D
// eigenvalues of K and of the GLL-quadrature Q1 mass
void coarse_dst(int dim, std::vector<double> &Q, std::vector<double> &lamK, std::vector<double> &lamM, int p);

// ---- block-Jacobi tables (precond.cpp) ------------------------------------
// [S (row-major, S[i][a]), Linv[a][c] = 1/(lambda_a + lambda_c)] of the
// interior element block K U W + W U K, K = K1 + own face terms (see
// kernels_pcg.hip); SIPG_ERR_INVALID when that block is not positive
// definite (penalty too small).
int build_block_jacobi(int p, double pen, std::vector<double> &tab);
int launch_convert(int from_dtype, int to_dtype, const void *src, void *dst, int64_t n,
                   hipStream_t s);
// batched launchers (FP64); counts live in device memory so a captured
// graph stays valid while the lists shrink; max_* size the grids.  With
// item_part (the fused p.Ap variant) blocks of finished instances are
// skipped, which needs states/slot.
int launch_batch_apply(int p, int shape, const double *X, double *Y, const batch_inst *inst,
                       const batch_item *items, const int *count, int max_items, double *item_part,
                       const cg_state *states, int slot, hipStream_t s);
int launch_batch_alpha(const int *list, const int *count, int max_inst, const int2 *item_range,
                       const double *item_part, cg_state *states, int slot, hipStream_t s);
int launch_batch_update_xr(double *X, const double *Pv, double *Rv, const double *APv,
                           const batch_chunk *chunks, const int *count, int max_chunks,
                           double *chunk_part, const cg_state *states, int slot, hipStream_t s);
int launch_batch_beta(const int *list, const int *count, int max_inst, const int2 *chunk_range,
                      const double *chunk_part, const int2 *repl_range, const double *repl_part,
                      cg_state *states, int slot, hipStream_t s);
int launch_batch_update_p(const double *Rv, double *Pv, const batch_chunk *chunks, const int *count,
                          int max_chunks, const cg_state *states, int slot, hipStream_t s);
int launch_batch_residual(const double *B, const double *AX, double *Rv, double *Pv,
                          const batch_chunk *chunks, const int *count, int max_chunks, double *part_rr,
                          double *part_bb, hipStream_t s);
// (item_range / item_rz: the preconditioner's r.z item sums for PCG, or
// nullptr for CG, where r.z = r.r)
int launch_batch_init(const int *list, const int *count, int max_inst, const int2 *chunk_range,
                      const double *part_rr, const double *part_bb, cg_state *states, const batch_ctl *ctl,
                      const int2 *item_range, const double *item_rz, hipStream_t s);
// ---- batched PCG (kernels_batch_pcg.hip) ----------------------------------
// z = M^-1 r over the listed blocks of degree p (shape as the operator's;
// Pv non-null: also p = z), with per-item r.r and r.z; finished instances
// skipped when states is non-null.  Requires pen = 5 (p+1)^2.
bool batch_prec_supported(int p, double pen, int shape);
int launch_batch_prec(int p, int shape, const double *R, double *Zq, double *Pv, const batch_inst *inst,
                      const batch_item *items, const int *count, int max_items, double *item_rr, double *item_rz,
                      const cg_state *states, int slot, hipStream_t s);
// beta = r.z(new) / r.z(old) per instance from the item sums; next slot
int launch_batch_beta_pcg(const int *list, const int *count, int max_inst, const int2 *item_range,
                          const double *item_rr, const double *item_rz, cg_state *states, int slot, hipStream_t s);
// x += z; z = 0 over the listed chunks (group update)
int launch_batch_fold(double *X, double *Z, const batch_chunk *chunks, const int *count, int max_chunks,
                      hipStream_t s);
// sets kSlotReplaced on the listed instances (after their r = b - A x)
int launch_batch_mark(const int *list, const int *count, int max_inst, cg_state *states, int slot,
                      hipStream_t s);
// reference order <-> element-major device order (to_device = 1: ref -> dev)
int launch_layout(int dtype, int to_device, const void *src, void *dst, int64_t noe, int nq,
                  hipStream_t s);
// block partials of the error norms (4 x nb: max, L2, dx, dy); D = phi_i'(z_k)
int launch_err_norms(int dtype, int p, int dim, const double *tab_f64, const double *D_dev, int nmodes,
                     const double *a_dev, const int *m_dev, const int *n_dev, const void *u, double *part, int nb,
                     hipStream_t s);
// as launch_err_norms against nodal samples of u_ex, d/dx u_ex, d/dy u_ex
// (device arrays, element-major, one value per node)
int launch_err_norms_nodal(int dtype, int p, int dim, const double *tab_f64, const double *D_dev,
                           const double *ue, const double *uxe, const double *uye, const void *u, double *part,
                           int nb, hipStream_t s);
int launch_rhs_fourier(int dtype, int p, int dim, const void *tab_f64, int nmodes,
                       const double *a_dev, const int *m_dev, const int *n_dev, void *b,
                       hipStream_t s);

}  // namespace sipg_internal

// ---- the opaque operator --------------------------------------------------
struct sipg_op {
    int dimx = 0, dimy = 0, p = 0, dtype = SIPG_F64;
    double pen = 0;
    int64_t n = 0;
    int kernel = SIPG_KERNEL_AUTO;    // contraction kernel setting (sipg_op_set_kernel)
    hipStream_t stream = nullptr;
    std::vector<double> h_tab;        // host copy (double) of the tables
    void *d_tab = nullptr;            // tables in the op's dtype
    void *d_tab64 = nullptr;          // tables in double (for RHS and FP64 helpers)
    double *d_part = nullptr;         // 4 x kMaxPartials reduction partials
    double *d_scalar = nullptr;       // small device scalars
    sipg_internal::cg_state *d_state = nullptr;
    sipg_internal::cg_state *h_state = nullptr;   // pinned mirror
    // cached CG workspace, allocated on first use
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // FP32 twin for the mixed-precision solver (owned)
    sipg_op *twin = nullptr;
    // block-Jacobi tables in the op's dtype, built on first use
    void *d_prec = nullptr;
    void *d_coarse = nullptr;         // the coarse level's tables and work arrays (one allocation)
    sipg_internal::coarse_tables coarse;   // pointers into d_coarse
};

namespace sipg_internal {
// ---- shared host helpers (abi.cpp) ----------------------------------------
inline size_t dtype_size(int dtype) { return dtype == SIPG_F64 ? 8 : 4; }
inline bool overlaps(const void *a, const void *b, size_t bytes) {
    const char *pa = (const char *)a, *pb = (const char *)b;
    return pa < pb + bytes && pb < pa + bytes;
}
// grows op->ws to at least `bytes` (contents not kept)
int ensure_ws(sipg_op *op, size_t bytes);
// Partial-buffer layout inside op->d_part
struct parts {
    double *pAp, *rr, *bb, *misc, *rz;
    int napply;
};
parts part_layout(sipg_op *op);
// Error norms: the block count of the norm kernels, and the host reduction
// of their 4 x nb partials (left in the r.r region of op->d_part) into
// out = {max, L2, H1}
int norms_num_blocks(sipg_op *op);
int norms_finish(sipg_op *op, int nb, double *out);
// Element corners of the square mesh as the mesh builder places them (x
// summed along each row from 0, y stepped down from 1 once per row,
// build_square_mesh.hpp:250-288), element e = row * dim + col.
void mesh_corners(int dim, std::vector<double> &xe, std::vector<double> &ye);
}  // namespace sipg_internal

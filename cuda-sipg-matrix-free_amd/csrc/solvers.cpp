// solvers.cpp -- device-resident conjugate gradient drivers.
//
// Reference: conjugate_gradient (conjugate_gradient_gpu.hpp:18-109) runs the
// loop on the host: two blocking cuBLAS dots per iteration, alpha and beta
// computed on the host, five launches.  Here the scalars live in a
// cg_state on the device (kernels_vec.hip), an iteration is three
// launches (fused apply + p.Ap, x/r update + r.r, p update), and the host
// only enqueues chunks of iterations and polls a copy of the state one chunk
// behind, so the GPU queue never drains while the host decides.
//
// One engine serves three entry points:
//   sipg_cg        the reference's semantics: recursive residual only;
//   sipg_cg_ex     + residual replacement / true-residual finish (FP64 or
//                  FP32 op);
//   sipg_cg_mixed  FP32 iteration with reliable updates into an FP64
//                  solution and an FP64 true residual.
#include "sipg_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

using namespace sipg_internal;

namespace {

struct Engine {
    // inner iteration
    sipg_op *opI = nullptr;          // op of the iteration precision
    void *xI = nullptr;              // iterate (x itself, or the FP32 correction x^)
    void *r = nullptr, *p = nullptr, *Ap = nullptr;
    // true residual (outer precision)
    sipg_op *opO = nullptr;          // op used for b - A x (FP64 op in mixed mode)
    void *xO = nullptr;              // solution vector in outer precision
    const void *b = nullptr;         // right-hand side in outer precision
    void *AxO = nullptr;             // scratch for A x in outer precision
    void *xpk = nullptr;             // scratch for xO + xI (peeks; null: no peeks)
    void *xL = nullptr;              // rounding errors of the xI updates (compensated
                                     // sum; null: plain updates)
    // The first kCompIters iterations update xI with compensation (their
    // increments carry the coarse solution, a large part of x); then xI is
    // folded into xO with its rounding error kept in xL, and the later,
    // much smaller increments accumulate plainly in a fresh xI.  (At
    // 256 x 256, p = 8, 4, 8 and 16 such iterations give the same iteration
    // counts and true residuals; each costs two more vector streams.)
    static constexpr int kCompIters = 4;
    bool mixed = false;              // fold xI into xO at each true residual
    int restart = 0;                 // p = r at a true residual: 0 never, 1 always,
                                     // 2 once the true residual stagnates
    double replace_factor = 0.0;
    bool true_stop = false;
    int maxit = 0;
    // preconditioned CG (block-Jacobi): z = M^-1 r, and the device tables
    bool pcg = false;
    void *z = nullptr;
    const void *prec = nullptr;
    // two-level: the coarse level of kernels_coarse.hip on top of the blocks
    const coarse_tables *coarse = nullptr;
    int ncp = 0;                     // its S.C partials
    // where the block-Jacobi kernel leaves the restriction's corner sums
    // (null: the coarse solve restricts r itself)
    double *e4() const {
        return coarse && pcg_rz_fuses_restrict(opI->dtype, opI->p, opI->pen) ? coarse->E4 : nullptr;
    }

    hipStream_t s = nullptr;
    cg_state *st = nullptr;
    cg_state *hst = nullptr;         // pinned, 2 entries
    parts P{};
    int nb = 0;
    bool fold = false;
    int64_t launched = 0;

    int64_t noe() const { return (int64_t)opI->dimx * opI->dimy; }

    int init(double tol, int tol_mode) {
        int rc;
        // r = b - A x ; p = r ; r.r ; b.b    (conjugate_gradient_gpu.hpp:40-60)
        if ((rc = launch_apply(opO->dtype, opO->p, xO, AxO, opO->d_tab, opO->dimx, opO->dimy,
                               opO->pen, nullptr, nullptr, 0, s, opO->kernel)))
            return rc;
        if (mixed) SIPG_HIP_TRY(hipMemsetAsync(xI, 0, (size_t)opI->n * dtype_size(opI->dtype), s));
        if (xL) SIPG_HIP_TRY(hipMemsetAsync(xL, 0, (size_t)opI->n * dtype_size(opI->dtype), s));
        if (pcg) {
            // r = b - A x ; z = M^-1 r ; p = z ; r.r, r.z, b.b
            // (conjugate_gradient_gpu.hpp:136-160)
            if ((rc = launch_true_residual(opO->dtype, opI->dtype, b, AxO, r, nullptr, opI->n, P.rr, P.bb, nb, s)))
                return rc;
            if ((rc = launch_pcg_rz(opI->dtype, opI->p, opI->pen, 2, r, nullptr, z, p, noe(), prec, nullptr, 0, P.rr, P.rz,
                                    nullptr, 0, s, e4())))
                return rc;
            if (coarse) {      // p = z_BJ + P0^T A0^-1 P0 r; r.z gains S.C
                if ((rc = launch_coarse_solve(opI->dtype, opI->p, r, opI->dimx, opI->dimy, *coarse, s,
                                              e4() != nullptr)) ||
                    (rc = launch_coarse_add(opI->dtype, opI->p, p, opI->dimx, opI->n, *coarse, nb, s)))
                    return rc;
            }
            return launch_pcg_init_finalize(P.rr, P.rz, npc(2), P.bb, nb, st, tol, tol_mode, maxit,
                                            true_stop ? 1 : 0, s, coarse ? coarse->part : nullptr, ncp);
        }
        if ((rc = launch_true_residual(opO->dtype, opI->dtype, b, AxO, r, p, opI->n, P.rr, P.bb, nb, s)))
            return rc;
        return launch_cg_init_finalize(P.rr, P.bb, nb, st, tol, tol_mode, maxit, true_stop ? 1 : 0, s);
    }

    // One CG iteration.  With `replace`, r is recomputed as b - A x between
    // the x/r update and the p update (after folding the accumulated
    // increments into x), so the p update forms beta and p from the true
    // residual: the reliable update of van der Vorst & Ye.  `restart`
    // makes that p update restart from p = r.
    int npc(int mode) const { return pcg_num_blocks(opI->dtype, opI->p, opI->pen, mode, noe()); }

    // One PCG iteration (conjugate_gradient_gpu.hpp:161-205): the apply with
    // p.Ap, then r -= alpha Ap fused with z = M^-1 r and r.r, r.z, then
    // x += alpha p and p = z + beta p.  With `replace`, x is updated first
    // and r is recomputed as b - A x (then z and its sums) before the p
    // update, as in the CG iteration.
    int iteration_pcg(bool replace, bool restart_p) {
        const int slot = (int)(launched & 1);
        int rc;
        if (xL && launched == kCompIters && (rc = launch_comp_fold(opO->dtype, xO, xL, xI, opO->n, s))) return rc;
        void *xLu = xL && launched < kCompIters ? xL : nullptr;     // compensate this iteration's update
        if ((rc = launch_apply(opI->dtype, opI->p, p, Ap, opI->d_tab, opI->dimx, opI->dimy, opI->pen,
                               P.pAp, st, slot, s, opI->kernel)))
            return rc;
        const double *src = P.pAp;
        int np = P.napply;
        if (fold) {
            if ((rc = launch_reduce_partials(P.pAp, P.napply, P.misc, s))) return rc;
            src = P.misc;
            np = 1;
        }
        if ((rc = launch_pcg_rz(opI->dtype, opI->p, opI->pen, 1, r, Ap, z, nullptr, noe(), prec, src, np, P.rr, P.rz, st,
                                slot, s, replace ? nullptr : e4())))
            return rc;
        if (coarse && !replace &&
            (rc = launch_coarse_solve(opI->dtype, opI->p, r, opI->dimx, opI->dimy, *coarse, s, e4() != nullptr, st,
                                      slot)))
            return rc;
        int nrz = npc(1);
        if (replace) {
            if ((rc = launch_pcg_x(opI->dtype, xI, p, opI->n, src, np, nb, st, slot, s, xLu))) return rc;
            if (mixed && (rc = launch_accumulate(opO->dtype, opI->dtype, xO, xI, opI->n, s, xL))) return rc;
            if ((rc = launch_apply(opO->dtype, opO->p, xO, AxO, opO->d_tab, opO->dimx, opO->dimy,
                                   opO->pen, nullptr, nullptr, 0, s, opO->kernel)))
                return rc;
            if ((rc = launch_true_residual(opO->dtype, opI->dtype, b, AxO, r, nullptr, opI->n, P.rr,
                                           nullptr, nb, s)))
                return rc;
            if ((rc = launch_pcg_rz(opI->dtype, opI->p, opI->pen, 0, r, nullptr, z, nullptr, noe(), prec, nullptr, 0, P.rr,
                                    P.rz, nullptr, 0, s, e4())))
                return rc;
            if (coarse && (rc = launch_coarse_solve(opI->dtype, opI->p, r, opI->dimx, opI->dimy, *coarse, s,
                                                    e4() != nullptr, st, slot)))
                return rc;
            nrz = npc(0);
            if ((rc = launch_mark_replaced(st, slot, restart_p ? 1 : 0, s))) return rc;
        }
        if (coarse)
            rc = launch_pcg2_update_p(opI->dtype, opI->p, xI, p, z, opI->n, src, np, P.rr, P.rz, nrz, ncp, opI->dimx,
                                      *coarse, nb, st, slot, s, xLu);
        else
            rc = launch_pcg_update_p(opI->dtype, xI, p, z, opI->n, src, np, P.rr, P.rz, nrz, nb, st, slot, s, xLu);
        if (rc) return rc;
        ++launched;
        return SIPG_OK;
    }

    // A peek between two iterations: the true residual of xO + xI (the
    // iterate a fold would give) without replacing r; when it meets the
    // tolerance the slot the next iteration reads is finished on it, and
    // the final fold produces that same iterate.  Ap is free between
    // iterations and takes A (xO + xI).
    int peek() {
        const int slot = (int)(launched & 1);
        int rc;
        if ((rc = launch_sum(opO->dtype, xO, xI, xpk, opO->n, s, xL))) return rc;
        if ((rc = launch_apply(opO->dtype, opO->p, xpk, AxO, opO->d_tab, opO->dimx, opO->dimy, opO->pen, nullptr,
                               nullptr, 0, s, opO->kernel)))
            return rc;
        if ((rc = launch_true_residual(opO->dtype, opO->dtype, b, AxO, nullptr, nullptr, opO->n, P.misc, nullptr, nb,
                                       s)))
            return rc;
        return launch_peek_finalize(P.misc, nb, st, slot, s);
    }

    int iteration(bool replace, bool restart_p) {
        if (pcg) return iteration_pcg(replace, restart_p);
        const int slot = (int)(launched & 1);
        int rc;
        if ((rc = launch_apply(opI->dtype, opI->p, p, Ap, opI->d_tab, opI->dimx, opI->dimy, opI->pen,
                               P.pAp, st, slot, s, opI->kernel)))
            return rc;
        const double *src = P.pAp;
        int np = P.napply;
        if (fold) {
            if ((rc = launch_reduce_partials(P.pAp, P.napply, P.misc, s))) return rc;
            src = P.misc;
            np = 1;
        }
        if ((rc = launch_cg_update_xr(opI->dtype, xI, p, r, Ap, opI->n, src, np, P.rr, nb, st, slot, s, xL)))
            return rc;
        if (replace) {
            if (mixed && (rc = launch_accumulate(opO->dtype, opI->dtype, xO, xI, opI->n, s, xL))) return rc;
            if ((rc = launch_apply(opO->dtype, opO->p, xO, AxO, opO->d_tab, opO->dimx, opO->dimy,
                                   opO->pen, nullptr, nullptr, 0, s, opO->kernel)))
                return rc;
            if ((rc = launch_true_residual(opO->dtype, opI->dtype, b, AxO, r, nullptr, opI->n, P.rr,
                                           nullptr, nb, s)))
                return rc;
            if ((rc = launch_mark_replaced(st, slot, restart_p ? 1 : 0, s))) return rc;
        }
        if ((rc = launch_cg_update_p(opI->dtype, r, p, opI->n, P.rr, nb, nb, st, slot, s))) return rc;
        ++launched;
        return SIPG_OK;
    }

    int run(sipg_cg_info *info) {
        hipEvent_t ev[2];
        SIPG_HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        SIPG_HIP_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        int rc = SIPG_OK;
        // state after init
        SIPG_HIP_TRY(hipMemcpyAsync(&hst[0], st, sizeof(cg_state), hipMemcpyDeviceToHost, s));
        SIPG_HIP_TRY(hipStreamSynchronize(s));
        cg_state cur = hst[0];
        int64_t cur_launched = 0;          // launches reflected in `cur`
        int nrep_enq = 0;
        int chunk = 4;
        int inflight = 0, buf = 0;
        int64_t pending_launched[2] = {0, 0};
        // adaptive restart: the direction is kept while each true residual
        // is well below the previous one (the recursion and the true
        // residual agree); once a replacement interval gains less than 2x in
        // the true residual norm, later replacements restart from p = r
        bool restarting = restart == 1;
        int seen_rep = 0;
        double last_ref = cur.rr_ref;
        // the FP64 floor: the true residual no longer follows the
        // recursion.  A true residual is requested only once the recursion
        // has fallen 16x (a final check) or 1/replace_factor^2 (a periodic
        // replacement) in r.r below the last one; while the solve converges
        // the true residual falls with it.  kStallChecks (PCG:
        // kStallChecksPcg) consecutive ones that gain less than 4x in r.r on
        // the previous one, the first of them stall_window iterations ago,
        // end the solve there, not converged.  (Not in the mixed-precision
        // solver, whose refinement cycles each restart the FP32 Krylov
        // space from the true residual.)
        const bool floor_watch = opI == opO;
        bool check_pending = false;
        double ref_at_rep = 0.0;           // the true r.r before the pending one
        int lag_count = 0, lag_it = 0, rep_it = 0;
        double rec_min = -1.0;             // lowest recursive r.r since the last true residual
        double rec_mark = -1.0;            // rec_min at its last 4x improvement, and when
        int rec_mark_it = 0;
        // Peeks: once the recursion has met the tolerance, the finishing
        // check waits for it to fall 4x below the last true residual, which
        // near the FP64 floor takes hundreds of iterations per check while
        // each true residual scatters around the floor by rounding.  Every
        // kPeekEvery iterations in that phase a true residual is computed
        // without replacing r (about one apply's cost): the solve ends on
        // the first that meets the tolerance, and the recursion is left as
        // it was.
        constexpr int kPeekEvery = 32;
        int last_peek_it = 0;
        // Tail: the launches queued past the finish run as no-ops, but each
        // still costs its launch and its share of the final synchronisation
        // (a chunk of 32 queued two polls late left 39 of them after a
        // 61-iteration two-level solve).  The finish is predicted from the
        // average fall of the polled r.r per iteration since the start, and
        // a chunk ends short of it (kAheadFrac of the predicted distance;
        // never below kMinChunk, which keeps the device ahead of the host's
        // polls), so that a poll falls just before the finish.
        constexpr int kMinChunk = 2;
        constexpr double kAheadFrac = 0.85;
        const double rr_start = cur.s[0].rr;
        // Each pass: decide on the newest polled state `cur`, enqueue one
        // chunk of iterations (4 growing to 32; 8 near the tolerance) and a
        // state copy, and block on the older of two outstanding copies.  Iterations past the finish are no-ops on
        // the device, so running one or two chunks ahead costs little.
        for (;;) {
            const cg_slot cs = cur.s[cur_launched & 1];
            if (cur.err || cs.done) break;      // done is final: later iterations are no-ops
            // a replacement is decided only once the previous one shows in
            // the polled state, so each is based on an up-to-date rr_ref
            const bool rep_visible = cur.nrep == nrep_enq;
            if (cur.nrep > seen_rep) {
                if (restart == 2 && cur.rr_ref > 0.25 * last_ref) restarting = true;
                last_ref = cur.rr_ref;
                seen_rep = cur.nrep;
            }
            if (rep_visible && check_pending) {
                // the requested true residual is in and the solve is not
                // done; the recursion fell 16x or more (in r.r) since the
                // previous one
                check_pending = false;
                if (floor_watch && floor_in_reach(cur.rr_ref, cur.bb) && cur.rr_ref > 0.25 * ref_at_rep) {
                    if (lag_count++ == 0) lag_it = cs.it;
                    if (lag_count >= (pcg ? kStallChecksPcg : kStallChecks) && cs.it - lag_it >= stall_window(cs.it))
                        break;
                } else {
                    lag_count = 0;
                }
            }
            // a recursion that does not fall 4x (in r.r) below the last true
            // residual within stall_window iterations has stalled at it
            // (preconditioned solvers: the residual of plain CG on this
            // operator can swing above its start for that long)
            if (floor_watch && pcg && seen_rep > 0 && !check_pending) {
                rec_min = rec_min < 0.0 ? cs.rr : std::min(rec_min, cs.rr);
                if (floor_in_reach(cur.rr_ref, cur.bb) && rec_min > 0.25 * cur.rr_ref &&
                    cs.it - rep_it >= stall_window(cs.it))
                    break;
                // nor one that fell that far but has hovered since: its
                // lowest r.r not 4x lower for stall_window iterations, short
                // of the next replacement's drop, so no true residual comes
                // for the lag rule to judge
                if (rec_mark < 0.0 || rec_min <= 0.25 * rec_mark) {
                    rec_mark = rec_min;
                    rec_mark_it = cs.it;
                }
                if (floor_in_reach(cur.rr_ref, cur.bb) && cs.it - rec_mark_it >= stall_window(cs.it) &&
                    cs.it - rep_it >= stall_window(cs.it))
                    break;
            }
            bool want_rep = false, restart_p = restarting;
            if (rep_visible) {
                // the recursion met the tolerance: check it on a true
                // residual, but not before it has fallen 4x below the last
                // one -- near the FP64 floor the true residual lags the
                // recursion, and replacing at every small drop stalls it
                if (true_stop && (cs.flags & kSlotWantCheck) && cs.rr <= 0.0625 * cur.rr_ref) want_rep = true;
                if (replace_factor > 0.0 && cs.rr <= replace_factor * replace_factor * cur.rr_ref)
                    want_rep = true;
                // divergence guard: a recursion that grew 100x past the last
                // true residual is recomputed and restarted from p = r
                // (FP64 PCG: 10^6x; its residual norm swings up to ~100x
                // early in a solve while the error's energy norm falls)
                const double guard = pcg && floor_watch ? 1e12 : 1e4;
                if ((replace_factor > 0.0 || true_stop) && cs.rr > guard * cur.rr_ref) {
                    want_rep = true;
                    restart_p = true;
                }
            }
            if (xpk && rep_visible && !want_rep && !check_pending && (cs.flags & kSlotWantCheck) &&
                cs.it - last_peek_it >= kPeekEvery) {
                if ((rc = peek())) break;
                last_peek_it = cs.it;
            }
            // near the tolerance, short chunks keep the host's lag small
            const bool near = true_stop && cs.rr <= 1e4 * cur.stop * cur.stop;
            int n_it = near ? std::min(chunk, 8) : chunk;
            const double goal = cur.stop * cur.stop;
            // drain: this pass waits for the newest state copy (the device
            // idles for one host round trip), not the older one
            bool drain = false;
            if (cs.it > 0 && rr_start > 0.0 && cs.rr < rr_start) {
                const int64_t queued = launched - cur_launched;      // launched past the polled state
                double left = 0.0;                                  // predicted iterations to the finish
                if (cs.rr > goal) left = std::ceil(std::log(cs.rr / goal) * cs.it / std::log(rr_start / cs.rr));
                left = std::min(left, 1e9);
                const int64_t cap = std::max<int64_t>(kMinChunk, (int64_t)(kAheadFrac * left) - queued);
                n_it = (int)std::min<int64_t>(n_it, cap);
                // the recursion is predicted to meet the tolerance within
                // what is queued and this chunk: the chunk ends there and
                // its state is waited for, so that the confirming true
                // residual follows the recursion's flag at once instead of
                // two chunks later
                if (true_stop && floor_watch && cs.rr > goal && !(cs.flags & kSlotWantCheck) && !check_pending &&
                    left <= (double)(queued + n_it)) {
                    drain = true;
                    n_it = (int)std::max<int64_t>(1, (int64_t)left - queued);
                }
            }
            // the final check alone, its outcome waited for
            if (want_rep && true_stop && floor_watch && (cs.flags & kSlotWantCheck)) {
                drain = true;
                n_it = 1;
            }
            for (int c = 0; c < n_it; ++c)
                if ((rc = iteration(want_rep && c == 0, restart_p))) break;
            if (rc) break;
            if (want_rep) {
                ++nrep_enq;
                check_pending = true;
                ref_at_rep = cur.rr_ref;
                rep_it = cs.it;
                rec_min = -1.0;
                rec_mark = -1.0;
            }
            if (chunk < 32) chunk *= 2;
            SIPG_HIP_TRY(hipMemcpyAsync(&hst[buf], st, sizeof(cg_state), hipMemcpyDeviceToHost, s));
            SIPG_HIP_TRY(hipEventRecord(ev[buf], s));
            pending_launched[buf] = launched;
            ++inflight;
            buf ^= 1;
            if (inflight == 2) {
                // the older copy sits in `buf` after the flip
                SIPG_HIP_TRY(hipEventSynchronize(ev[buf]));
                cur = hst[buf];
                cur_launched = pending_launched[buf];
                --inflight;
            }
            if (drain && inflight == 1) {
                // and the newest in the other
                SIPG_HIP_TRY(hipEventSynchronize(ev[buf ^ 1]));
                cur = hst[buf ^ 1];
                cur_launched = pending_launched[buf ^ 1];
                --inflight;
            }
        }
        // fold whatever the last cycle accumulated (a no-op after a final
        // true residual, which already folded it)
        if (!rc && mixed) rc = launch_accumulate(opO->dtype, opI->dtype, xO, xI, opI->n, s, xL);
        SIPG_HIP_TRY(hipStreamSynchronize(s));
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        if (rc) return rc;
        SIPG_HIP_TRY(hipMemcpy(&hst[0], st, sizeof(cg_state), hipMemcpyDeviceToHost));
        const cg_state &f = hst[0];
        const cg_slot &fs = f.s[launched & 1];
        int status = SIPG_OK;
        if (f.err == 2) status = SIPG_ERR_BREAKDOWN;
        else if (f.err == 3) status = SIPG_ERR_NONFINITE;
        else if (!(std::sqrt(fs.rr) <= f.stop) || (true_stop && !fs.checked)) status = SIPG_NOT_CONVERGED;
        if (info) {
            info->status = status;
            info->iters = fs.it;
            info->replacements = f.nrep;
            info->res_norm = std::sqrt(fs.rr);
            info->true_res_norm = f.nrep > 0 || fs.it == 0 || fs.checked ? std::sqrt(f.rr_ref) : -1.0;
            info->b_norm = std::sqrt(f.bb);
        }
        return status;
    }
};

int check_opts(const sipg_cg_opts *o) {
    if (!o) return SIPG_ERR_INVALID;
    if (!(o->tol >= 0.0) || (o->tol_mode != SIPG_TOL_ABS && o->tol_mode != SIPG_TOL_REL))
        return SIPG_ERR_INVALID;
    if (!(o->replace_factor >= 0.0 && o->replace_factor < 1.0)) return SIPG_ERR_INVALID;
    if (o->restart < 0 || o->restart > 2) return SIPG_ERR_INVALID;
    return SIPG_OK;
}

// Block-Jacobi tables of the op (its dtype), built and uploaded once.
int ensure_prec(sipg_op *op) {
    if (op->d_prec) return SIPG_OK;
    std::vector<double> tab;
    int rc = build_block_jacobi(op->p, op->pen, tab);
    if (rc) return rc;
    const size_t es = dtype_size(op->dtype);
    std::vector<float> tf;
    const void *src = tab.data();
    if (op->dtype == SIPG_F32) {
        tf.assign(tab.begin(), tab.end());
        src = tf.data();
    }
    void *d = nullptr;
    SIPG_HIP_TRY(hipMalloc(&d, tab.size() * es));
    const hipError_t e = hipMemcpy(d, src, tab.size() * es, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        set_last_hip_error(e);
        return SIPG_ERR_HIP;
    }
    op->d_prec = d;
    return SIPG_OK;
}

// Coarse-level tables and work arrays of the op (FP64), built once.
int ensure_coarse(sipg_op *op) {
    if (op->d_coarse) return SIPG_OK;
    const int dx = op->dimx, dy = op->dimy, n1 = op->p + 1;
    const int nvx = std::max(dx - 1, 0), nvy = std::max(dy - 1, 0);
    const int64_t noe = (int64_t)dx * dy, nv = (int64_t)nvx * nvy;
    std::vector<double> Qx, Qy, kx, mx, ky, my;
    coarse_dst(dx, Qx, kx, mx, op->p);
    coarse_dst(dy, Qy, ky, my, op->p);
    std::vector<double> linv((size_t)nv);
    for (int a = 0; a < nvy; ++a)
        for (int b = 0; b < nvx; ++b) linv[(size_t)a * nvx + b] = 1.0 / (my[a] * kx[b] + ky[a] * mx[b]);
    std::vector<double> xi(n1), w(n1);
    gll(op->p, xi.data(), w.data());
    for (double &v : xi) v = 0.5 * (v + 1.0);
    const int npart = std::max(coarse_num_partials(dx, dy), 1);
    // Qx, Qy, Linv, xi, E4, S, T1, T2, Cpad, part
    const size_t len[10] = {(size_t)nvx * nvx, (size_t)nvy * nvy, (size_t)nv, (size_t)n1, (size_t)(4 * noe),
                            (size_t)nv, (size_t)nv, (size_t)nv, (size_t)(dx + 1) * (dy + 1), (size_t)npart};
    size_t total = 0;
    for (size_t l : len) total += (l + 31) & ~size_t(31);
    double *d = nullptr;
    SIPG_HIP_TRY(hipMalloc(&d, total * sizeof(double)));
    coarse_tables &c = op->coarse;
    double **slots[10] = {&c.Qx, &c.Qy, &c.Linv, &c.xi, &c.E4, &c.S, &c.T1, &c.T2, &c.Cpad, &c.part};
    size_t o = 0;
    for (int k = 0; k < 10; ++k) {
        *slots[k] = d + o;
        o += (len[k] + 31) & ~size_t(31);
    }
    hipError_t e = hipSuccess;
    const std::vector<double> *src[4] = {&Qx, &Qy, &linv, &xi};
    for (int k = 0; k < 4 && e == hipSuccess; ++k)
        if (!src[k]->empty()) e = hipMemcpy(*slots[k], src[k]->data(), src[k]->size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c.Cpad, 0, len[8] * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(d);
        c = coarse_tables{};
        set_last_hip_error(e);
        return SIPG_ERR_HIP;
    }
    op->d_coarse = d;
    return SIPG_OK;
}

int solve_plain(sipg_op *op, void *x, const void *b, const sipg_cg_opts *o, sipg_cg_info *info,
                bool pcg = false, bool two_level = false) {
    int rc = check_opts(o);
    if (rc) return rc;
    const size_t vb = (size_t)op->n * dtype_size(op->dtype);
    if (overlaps(x, b, vb)) return SIPG_ERR_ALIAS;
    if (pcg && (rc = ensure_prec(op))) return rc;
    if (two_level && (rc = ensure_coarse(op))) return rc;
    // With true residuals the iterate is group-updated (van der Vorst & Ye
    // 2000): the CG increments go to a separate vector z that is folded
    // into x only when r is recomputed, so the rounding of x += alpha p
    // does not accumulate at the scale of |x| between recomputations.
    const bool group = o->replace_factor > 0.0 || o->true_residual_stop != 0;
    // Two-level PCG keeps the rounding errors of its first increments (xL,
    // Engine::kCompIters): its first directions carry the coarse solution,
    // so plainly accumulated roundings (eps |x| per update, amplified by
    // |A|) leave b - A x about 1e-9 of |b| away from the recursion at
    // 256 x 256, p = 8, and every true residual then costs iterations;
    // compensated, they stay near one rounding of x.
    const bool peeks = group && o->true_residual_stop != 0;
    const bool comp = group && two_level;
    const int nvec = 3 + (group ? 1 : 0) + (pcg ? 1 : 0);
    if ((rc = ensure_ws(op, (nvec + (peeks ? 1 : 0) + (comp ? 1 : 0)) * vb))) return rc;
    char *ws = (char *)op->ws;
    Engine E;
    E.opI = E.opO = op;
    E.xO = x;
    E.xI = group ? (void *)(ws + 3 * vb) : x;
    E.mixed = group;
    E.r = ws;
    E.p = ws + vb;
    E.Ap = ws + 2 * vb;
    E.pcg = pcg;
    E.z = pcg ? ws + (size_t)(nvec - 1) * vb : nullptr;
    E.prec = op->d_prec;
    E.coarse = two_level ? &op->coarse : nullptr;
    E.ncp = two_level ? coarse_num_partials(op->dimx, op->dimy) : 0;
    E.AxO = E.Ap;        // the initial and the replacement A x use the Ap slot
    E.xpk = peeks ? ws + (size_t)nvec * vb : nullptr;
    E.xL = comp ? ws + (size_t)(nvec + (peeks ? 1 : 0)) * vb : nullptr;
    E.b = b;
    E.replace_factor = o->replace_factor;
    E.true_stop = o->true_residual_stop != 0;
    E.restart = o->restart;
    E.maxit = (int)std::min<int64_t>(o->maxit > 0 ? (int64_t)o->maxit : op->n, INT_MAX);
    E.s = op->stream;
    E.st = op->d_state;
    E.hst = op->h_state;
    E.P = part_layout(op);
    E.nb = vec_num_blocks(op->n);
    E.fold = E.P.napply > kMaxPartials;
    if ((rc = E.init(o->tol, o->tol_mode))) return rc;
    return E.run(info);
}

// FP32 inner (P)CG cycles on the op's FP32 twin inside FP64 refinement.
int solve_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info, bool pcg,
                bool two_level = false) {
    if (!op || !x || !b || op->dtype != SIPG_F64) return SIPG_ERR_INVALID;
    int rc = check_opts(opts);
    if (rc) return rc;
    if (overlaps(x, b, (size_t)op->n * 8)) return SIPG_ERR_ALIAS;
    if (!op->twin) {
        if ((rc = sipg_op_create(&op->twin, op->dimx, op->dimy, op->p, op->pen, SIPG_F32, op->stream)))
            return rc;
    }
    sipg_op *tw = op->twin;
    tw->stream = op->stream;
    tw->kernel = op->kernel;
    if (pcg && (rc = ensure_prec(tw))) return rc;
    if (two_level && (rc = ensure_coarse(tw))) return rc;       // coarse solve in FP64, vectors in FP32
    const size_t v32 = (size_t)op->n * 4, v64 = (size_t)op->n * 8;
    if ((rc = ensure_ws(tw, (pcg ? 5 : 4) * v32))) return rc;    // r^, p^, A p^, x^ (, z^)
    if ((rc = ensure_ws(op, v64))) return rc;           // A x in FP64
    char *w = (char *)tw->ws;
    Engine E;
    E.pcg = pcg;
    E.z = pcg ? w + 4 * v32 : nullptr;
    E.prec = tw->d_prec;
    E.coarse = two_level ? &tw->coarse : nullptr;
    E.ncp = two_level ? coarse_num_partials(tw->dimx, tw->dimy) : 0;
    E.opI = tw;
    E.opO = op;
    E.r = w;
    E.p = w + v32;
    E.Ap = w + 2 * v32;
    E.xI = w + 3 * v32;
    E.xO = x;
    E.b = b;
    E.AxO = op->ws;
    E.mixed = true;
    E.restart = opts->restart;
    E.replace_factor = opts->replace_factor > 0.0 ? opts->replace_factor : 0.1;
    E.true_stop = true;
    E.maxit = (int)std::min<int64_t>(opts->maxit > 0 ? (int64_t)opts->maxit : op->n, INT_MAX);
    E.s = op->stream;
    E.st = tw->d_state;
    E.hst = tw->h_state;
    E.P = part_layout(tw);
    E.nb = vec_num_blocks(op->n);
    E.fold = E.P.napply > kMaxPartials;
    if ((rc = E.init(opts->tol, opts->tol_mode))) return rc;
    return E.run(info);
}

}  // namespace

extern "C" {

int sipg_cg(sipg_op *op, void *x, const void *b, double tol, int tol_mode, int maxit, int *iters,
            double *res_norm) {
    if (!op || !x || !b) return SIPG_ERR_INVALID;
    sipg_cg_opts o{};
    o.tol = tol;
    o.tol_mode = tol_mode;
    o.maxit = maxit;
    o.replace_factor = 0.0;
    o.true_residual_stop = 0;
    sipg_cg_info info{};
    const int rc = solve_plain(op, x, b, &o, &info);
    if (iters) *iters = info.iters;
    if (res_norm) *res_norm = info.res_norm;
    return rc;
}

int sipg_cg_ex(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info) {
    if (!op || !x || !b) return SIPG_ERR_INVALID;
    return solve_plain(op, x, b, opts, info);
}

int sipg_pcg(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info) {
    if (!op || !x || !b) return SIPG_ERR_INVALID;
    return solve_plain(op, x, b, opts, info, true);
}

int sipg_prec_apply(sipg_op *op, const void *r, void *z) {
    if (!op || !r || !z) return SIPG_ERR_INVALID;
    const size_t vb = (size_t)op->n * dtype_size(op->dtype);
    if (overlaps(r, z, vb)) return SIPG_ERR_ALIAS;
    int rc = ensure_prec(op);
    if (rc) return rc;
    parts P = part_layout(op);
    return launch_pcg_rz(op->dtype, op->p, op->pen, 0, const_cast<void *>(r), nullptr, z, nullptr,
                         (int64_t)op->dimx * op->dimy, op->d_prec, nullptr, 0, P.rr, P.rz, nullptr, 0, op->stream);
}

int sipg_pcg2(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info) {
    if (!op || !x || !b) return SIPG_ERR_INVALID;
    return solve_plain(op, x, b, opts, info, true, true);
}

int sipg_prec2_apply(sipg_op *op, const void *r, void *z) {
    if (!op || !r || !z) return SIPG_ERR_INVALID;
    const size_t vb = (size_t)op->n * dtype_size(op->dtype);
    if (overlaps(r, z, vb)) return SIPG_ERR_ALIAS;
    int rc = ensure_prec(op);
    if (rc || (rc = ensure_coarse(op))) return rc;
    parts P = part_layout(op);
    const bool fused = pcg_rz_fuses_restrict(op->dtype, op->p, op->pen);
    if ((rc = launch_pcg_rz(op->dtype, op->p, op->pen, 0, const_cast<void *>(r), nullptr, z, nullptr,
                            (int64_t)op->dimx * op->dimy, op->d_prec, nullptr, 0, P.rr, P.rz, nullptr, 0, op->stream,
                            fused ? op->coarse.E4 : nullptr)) ||
        (rc = launch_coarse_solve(op->dtype, op->p, r, op->dimx, op->dimy, op->coarse, op->stream, fused)))
        return rc;
    return launch_coarse_add(op->dtype, op->p, z, op->dimx, op->n, op->coarse, vec_num_blocks(op->n), op->stream);
}

int sipg_cg_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info) {
    return solve_mixed(op, x, b, opts, info, false);
}

int sipg_pcg_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info) {
    return solve_mixed(op, x, b, opts, info, true);
}

int sipg_pcg2_mixed(sipg_op *op, void *x, const void *b, const sipg_cg_opts *opts, sipg_cg_info *info) {
    return solve_mixed(op, x, b, opts, info, true, true);
}

int sipg_true_relres(sipg_op *op, const void *x, const void *b, double *relres) {
    if (!op || !x || !b || !relres) return SIPG_ERR_INVALID;
    if (op->dtype != SIPG_F64) return SIPG_ERR_INVALID;
    const size_t vb = (size_t)op->n * 8;
    int rc = ensure_ws(op, 3 * vb);
    if (rc) return rc;
    char *ws = (char *)op->ws;
    void *Ax = ws + 2 * vb;
    void *r = ws;
    if ((rc = launch_apply(op->dtype, op->p, x, Ax, op->d_tab, op->dimx, op->dimy, op->pen, nullptr,
                           nullptr, 0, op->stream, op->kernel)))
        return rc;
    if ((rc = launch_residual(op->dtype, b, Ax, r, op->n, op->stream))) return rc;
    double rr = 0, bb = 0;
    if ((rc = sipg_dot(op, r, r, &rr))) return rc;
    if ((rc = sipg_dot(op, b, b, &bb))) return rc;
    *relres = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return SIPG_OK;
}

}  // extern "C"

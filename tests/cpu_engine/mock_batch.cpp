// CPU stand-ins for the batched launchers (kernels_batch.hip), for the CPU
// test build of batch.cpp.  Each mirrors its kernel's rules (skips of
// finished instances, carry-forward, per-chunk and per-item partial sums,
// check / resume); the operator is the oracle's, applied to whole instances
// (the engine always lists all blocks of an instance together).  Work is run
// through mock_run, so it is recorded while a stream is captured and
// replayed by the mock graph.  Test infrastructure only.
#include "../../cuda-sipg-matrix-free_amd/csrc/sipg_internal.hpp"

#include <cmath>
#include <set>
#include <vector>

extern "C" void orc_apply(int dimx, int dimy, int p, double pen, const double *x, double *y);

namespace sipg_internal {

const int kBatchShapeBX[kBatchShapes] = {16, 8, 8};
const int kBatchShapeBY[kBatchShapes] = {4, 4, 2};

bool batch_shape_supported(int p, int shape) {
    if (p < 1 || p > kMaxDegree || shape < 0 || shape >= kBatchShapes) return false;
    return shape != 0 || p <= 12;
}

namespace {

void apply_instance(const batch_inst &in, int p, const double *X, double *Y) {
    const int64_t noe = (int64_t)in.dimx * in.dimy, nq = (int64_t)(p + 1) * (p + 1);
    std::vector<double> xr(noe * nq), yr(noe * nq);
    const double *x = X + in.off;
    double *y = Y + in.off;
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) xr[q * noe + e] = x[e * nq + q];
    orc_apply(in.dimx, in.dimy, p, in.pen, xr.data(), yr.data());
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) y[e * nq + q] = yr[q * noe + e];
}

bool finished(const cg_state *st, int slot) { return st->s[slot].done | st->err; }

}  // namespace

int launch_batch_apply(int p, int shape, const double *X, double *Y, const batch_inst *inst,
                       const batch_item *items, const int *count, int max_items, double *item_part,
                       const cg_state *states, int slot, hipStream_t) {
    if (!batch_shape_supported(p, shape)) return SIPG_ERR_UNSUPPORTED;
    if (item_part && !states) return SIPG_ERR_INVALID;
    mock_run([=]() {
        const int n = *count;
        if (n > max_items) std::abort();
        const int BX = kBatchShapeBX[shape], BY = kBatchShapeBY[shape];
        const int64_t nq = (int64_t)(p + 1) * (p + 1);
        std::set<int> done_inst;
        for (int k = 0; k < n; ++k) {
            const batch_item it = items[k];
            if (item_part && finished(states + it.inst, slot)) continue;
            const batch_inst in = inst[it.inst];
            if (!done_inst.count(it.inst)) {
                apply_instance(in, p, X, Y);
                done_inst.insert(it.inst);
            }
            if (item_part) {
                double s = 0.0;
                for (int ey = it.y0; ey < it.y0 + BY && ey < in.dimy; ++ey)
                    for (int ex = it.x0; ex < it.x0 + BX && ex < in.dimx; ++ex) {
                        const int64_t b = in.off + ((int64_t)ey * in.dimx + ex) * nq;
                        for (int64_t q = 0; q < nq; ++q) s += X[b + q] * Y[b + q];
                    }
                item_part[k] = s;
            }
        }
    });
    return SIPG_OK;
}

int launch_batch_alpha(const int *list, const int *count, int, const int2 *item_range, const double *item_part,
                       cg_state *states, int slot, hipStream_t) {
    mock_run([=]() {
        for (int k = 0; k < *count; ++k) {
            cg_state *st = states + list[k];
            if (finished(st, slot)) continue;
            const int2 r = item_range[list[k]];
            double pAp = 0.0;
            for (int i = r.x; i < r.y; ++i) pAp += item_part[i];
            const double alpha = st->s[slot].rz / pAp;
            if (!(pAp > 0.0) || !std::isfinite(alpha)) {
                st->err = std::isfinite(pAp) ? 2 : 3;
            } else {
                st->s[slot].alpha = alpha;
                st->s[slot].pAp = pAp;
            }
        }
    });
    return SIPG_OK;
}

int launch_batch_update_xr(double *X, const double *Pv, double *Rv, const double *APv, const batch_chunk *chunks,
                           const int *count, int max_chunks, double *chunk_part, const cg_state *states,
                           int slot, hipStream_t) {
    mock_run([=]() {
        if (*count > max_chunks) std::abort();
        for (int k = 0; k < *count; ++k) {
            const batch_chunk c = chunks[k];
            const cg_state *st = states + c.inst;
            if (finished(st, slot)) continue;
            const double a = st->s[slot].alpha;
            double s = 0.0;
            for (int64_t i = c.begin; i < c.end; ++i) {
                X[i] += a * Pv[i];
                Rv[i] -= a * APv[i];
                s += Rv[i] * Rv[i];
            }
            chunk_part[k] = s;
        }
    });
    return SIPG_OK;
}

int launch_batch_beta(const int *list, const int *count, int, const int2 *chunk_range, const double *chunk_part,
                      const int2 *repl_range, const double *repl_part, cg_state *states, int slot, hipStream_t) {
    mock_run([=]() {
        for (int k = 0; k < *count; ++k) {
            cg_state *st = states + list[k];
            const cg_slot cur = st->s[slot];
            if (cur.done | st->err) {
                st->s[slot ^ 1] = cur;
                continue;
            }
            const bool replaced = (cur.flags & kSlotReplaced) != 0;
            const int2 r = replaced ? repl_range[list[k]] : chunk_range[list[k]];
            const double *part = replaced ? repl_part : chunk_part;
            double rr = 0.0;
            for (int i = r.x; i < r.y; ++i) rr += part[i];
            const bool met = !(std::sqrt(rr) > st->stop);
            const bool need_true = st->true_stop && !replaced;
            cg_slot nx{};
            nx.rr = rr;
            nx.rz = rr;
            nx.beta = rr / cur.rr;
            nx.it = cur.it + 1;
            nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
            nx.checked = replaced ? 1 : 0;
            nx.flags = (met && need_true) ? kSlotWantCheck : 0;
            st->s[slot ^ 1] = nx;
            if (replaced) st->rr_ref = rr;
            if (!std::isfinite(rr)) st->err = 3;
        }
    });
    return SIPG_OK;
}

int launch_batch_fold(double *X, double *Z, const batch_chunk *chunks, const int *count, int max_chunks,
                      hipStream_t) {
    mock_run([=]() {
        if (*count > max_chunks) std::abort();
        for (int k = 0; k < *count; ++k)
            for (int64_t i = chunks[k].begin; i < chunks[k].end; ++i) {
                X[i] += Z[i];
                Z[i] = 0.0;
            }
    });
    return SIPG_OK;
}

int launch_batch_mark(const int *list, const int *count, int, cg_state *states, int slot, hipStream_t) {
    mock_run([=]() {
        for (int k = 0; k < *count; ++k) {
            cg_state *st = states + list[k];
            cg_slot c = st->s[slot];
            if (c.done | st->err) continue;
            c.flags |= kSlotReplaced;
            st->s[slot] = c;
            st->nrep += 1;
        }
    });
    return SIPG_OK;
}

int launch_batch_update_p(const double *Rv, double *Pv, const batch_chunk *chunks, const int *count, int,
                          const cg_state *states, int slot, hipStream_t) {
    mock_run([=]() {
        for (int k = 0; k < *count; ++k) {
            const batch_chunk c = chunks[k];
            const cg_state *st = states + c.inst;
            if (finished(st, slot)) continue;
            const double beta = st->s[slot ^ 1].beta;
            for (int64_t i = c.begin; i < c.end; ++i) Pv[i] = Rv[i] + beta * Pv[i];
        }
    });
    return SIPG_OK;
}

int launch_batch_residual(const double *B, const double *AX, double *Rv, double *Pv, const batch_chunk *chunks,
                          const int *count, int max_chunks, double *part_rr, double *part_bb, hipStream_t) {
    mock_run([=]() {
        if (*count > max_chunks) std::abort();
        for (int k = 0; k < *count; ++k) {
            const batch_chunk c = chunks[k];
            double srr = 0.0, sbb = 0.0;
            for (int64_t i = c.begin; i < c.end; ++i) {
                const double r = B[i] - AX[i];
                Rv[i] = r;
                if (Pv) Pv[i] = r;
                srr += r * r;
                sbb += B[i] * B[i];
            }
            part_rr[k] = srr;
            part_bb[k] = sbb;
        }
    });
    return SIPG_OK;
}

int launch_batch_init(const int *list, const int *count, int, const int2 *chunk_range, const double *part_rr,
                      const double *part_bb, cg_state *states, const batch_ctl *ctl, const int2 *item_range,
                      const double *item_rz, hipStream_t) {
    mock_run([=]() {
        for (int k = 0; k < *count; ++k) {
            const int i = list[k];
            const int2 r = chunk_range[i];
            double rr = 0.0, bb = 0.0;
            for (int c = r.x; c < r.y; ++c) {
                rr += part_rr[c];
                bb += part_bb[c];
            }
            cg_state *st = states + i;
            const double tol = ctl[i].tol;
            const int maxit_i = ctl[i].maxit;
            const double stop = ctl[i].tol_mode == SIPG_TOL_REL ? tol * std::sqrt(bb) : tol;
            cg_slot s0{};
            s0.rr = rr;
            s0.rz = rr;
            if (item_rz) {
                double rz = 0.0;
                for (int c = item_range[i].x; c < item_range[i].y; ++c) rz += item_rz[c];
                s0.rz = rz;
            }
            s0.done = (std::sqrt(rr) > stop && maxit_i > 0) ? 0 : 1;
            s0.checked = 1;
            st->s[0] = s0;
            st->s[1] = s0;
            st->bb = bb;
            st->stop = stop;
            st->stop_true = stop;
            st->rr0 = rr;
            st->rr_ref = rr;
            st->maxit = maxit_i;
            st->nrep = 0;
            st->err = (!std::isfinite(rr) || !std::isfinite(bb)) ? 3 : 0;
            st->true_stop = 1;
        }
    });
    return SIPG_OK;
}

// ---- batched PCG (kernels_batch_pcg.hip) ---------------------------------
bool batch_prec_supported(int p, double pen, int shape) {
    return batch_shape_supported(p, shape) && pen == 5.0 * (p + 1) * (p + 1);
}

// z = M^-1 r on each element of every listed block, with the op tables of
// build_block_jacobi; per-item r.r and r.z over the block's elements
int launch_batch_prec(int p, int shape, const double *R, double *Zq, double *Pv, const batch_inst *inst,
                      const batch_item *items, const int *count, int max_items, double *item_rr, double *item_rz,
                      const cg_state *states, int slot, hipStream_t) {
    if (!batch_shape_supported(p, shape)) return SIPG_ERR_UNSUPPORTED;
    std::vector<double> tab;
    const double pen = 5.0 * (p + 1) * (p + 1);
    if (build_block_jacobi(p, pen, tab)) return SIPG_ERR_INVALID;
    mock_run([=]() {
        const int n = *count;
        if (n > max_items) std::abort();
        const int BX = kBatchShapeBX[shape], BY = kBatchShapeBY[shape];
        const int n1 = p + 1, nq = n1 * n1;
        const double *S = tab.data(), *L = tab.data() + nq;
        std::vector<double> t1(nq), t2(nq);
        for (int k = 0; k < n; ++k) {
            const batch_item it = items[k];
            if (states && finished(states + it.inst, slot)) continue;
            const batch_inst in = inst[it.inst];
            double srr = 0, srz = 0;
            for (int ey = it.y0; ey < it.y0 + BY && ey < in.dimy; ++ey)
                for (int ex = it.x0; ex < it.x0 + BX && ex < in.dimx; ++ex) {
                    const int64_t b = in.off + ((int64_t)ey * in.dimx + ex) * nq;
                    const double *Re = R + b;
                    for (int a = 0; a < n1; ++a)
                        for (int j = 0; j < n1; ++j) {
                            double acc = 0;
                            for (int i = 0; i < n1; ++i) acc += S[i * n1 + a] * Re[i * n1 + j];
                            t1[a * n1 + j] = acc;
                        }
                    for (int a = 0; a < n1; ++a)
                        for (int c = 0; c < n1; ++c) {
                            double acc = 0;
                            for (int bb = 0; bb < n1; ++bb) acc += t1[a * n1 + bb] * S[bb * n1 + c];
                            t2[a * n1 + c] = acc * L[a * n1 + c];
                        }
                    for (int a = 0; a < n1; ++a)
                        for (int d = 0; d < n1; ++d) {
                            double acc = 0;
                            for (int c = 0; c < n1; ++c) acc += t2[a * n1 + c] * S[d * n1 + c];
                            t1[a * n1 + d] = acc;
                        }
                    for (int a = 0; a < n1; ++a)
                        for (int j = 0; j < n1; ++j) {
                            double acc = 0;
                            for (int i = 0; i < n1; ++i) acc += S[a * n1 + i] * t1[i * n1 + j];
                            Zq[b + a * n1 + j] = acc;
                            if (Pv) Pv[b + a * n1 + j] = acc;
                        }
                    for (int q = 0; q < nq; ++q) {
                        srr += Re[q] * Re[q];
                        srz += Re[q] * Zq[b + q];
                    }
                }
            item_rr[k] = srr;
            item_rz[k] = srz;
        }
    });
    return SIPG_OK;
}

int launch_batch_beta_pcg(const int *list, const int *count, int, const int2 *item_range, const double *item_rr,
                          const double *item_rz, cg_state *states, int slot, hipStream_t) {
    mock_run([=]() {
        for (int k = 0; k < *count; ++k) {
            cg_state *st = states + list[k];
            const cg_slot cur = st->s[slot];
            if (cur.done | st->err) {
                st->s[slot ^ 1] = cur;
                continue;
            }
            const int2 r = item_range[list[k]];
            double rr = 0.0, rz = 0.0;
            for (int i = r.x; i < r.y; ++i) {
                rr += item_rr[i];
                rz += item_rz[i];
            }
            const bool replaced = (cur.flags & kSlotReplaced) != 0;
            const bool met = !(std::sqrt(rr) > st->stop);
            const bool need_true = st->true_stop && !replaced;
            cg_slot nx{};
            nx.rr = rr;
            nx.rz = rz;
            nx.beta = rz / cur.rz;
            nx.it = cur.it + 1;
            nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
            nx.checked = replaced ? 1 : 0;
            nx.flags = (met && need_true) ? kSlotWantCheck : 0;
            st->s[slot ^ 1] = nx;
            if (replaced) st->rr_ref = rr;
            if (!std::isfinite(rr) || !std::isfinite(rz)) st->err = 3;
            else if (!(rz > 0.0) && rr > 0.0) st->err = 2;
        }
    });
    return SIPG_OK;
}

}  // namespace sipg_internal

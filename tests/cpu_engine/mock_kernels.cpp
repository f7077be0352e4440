// CPU stand-ins for libsipg_hip's kernel launchers, for the CPU test build
// of the host orchestration (abi.cpp, solvers.cpp).  Each function mirrors
// the semantics of the kernel it replaces (kernels_apply.hip,
// kernels_vec.hip) as one "block": the state-machine rules (slots,
// carry-forward, err, replacement finalize) are restated line for line so
// the engine's protocol can be exercised without a GPU.  The operator apply
// itself is the oracle's.  Test infrastructure only.
#include "../../cuda-sipg-matrix-free_amd/csrc/sipg_internal.hpp"

#include <cmath>
#include <cstring>
#include <vector>

#include "../../oracle/sipg_oracle.h"

extern "C" {
void orc_apply(int dimx, int dimy, int p, double pen, const double *x, double *y);
void orc_apply_f32(int dimx, int dimy, int p, float pen, const float *x, float *y);
}

namespace sipg_internal {

static const int kMockApplyBlocks = 3;
static const int kMockVecBlocks = 2;

int apply_num_blocks(int, int p, int, int, int) { return p <= kMaxDegree ? kMockApplyBlocks : -1; }
// the CPU stand-in builds no run-time-table operator (op creation above the
// compiled degrees then fails on the block count above)
size_t generic_table_len(int, int p, int, int) { return (size_t)table_len(p); }
// the CPU stand-in has one contraction path; the setting is only recorded
int resolve_kernel(int, int, int) { return SIPG_KERNEL_VALU; }
int vec_num_blocks(int64_t) { return kMockVecBlocks; }

template <typename T> static T *tp(void *p) { return (T *)p; }
template <typename T> static const T *tp(const void *p) { return (const T *)p; }

// x_h += a p with the rounding errors kept in x_l (TwoProd by fma, TwoSum),
// as comp_axpy of the device kernels
template <typename T> static void comp_axpy_h(T &xh, T &xl, T a, T pv) {
    const T t = a * pv;
    const T tp_ = std::fma(a, pv, -t);
    const T s = xh + t;
    const T bb = s - xh;
    const T e = (xh - (s - bb)) + (t - bb);
    xh = s;
    xl = xl + (e + tp_);
}

static double dotv(int dtype, const void *a, const void *b, int64_t n) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i)
        s += dtype == SIPG_F64 ? tp<double>(a)[i] * tp<double>(b)[i]
                               : (double)tp<float>(a)[i] * (double)tp<float>(b)[i];
    return s;
}

static void put_partials(double *part, int np, double v) {
    part[0] = v;
    for (int i = 1; i < np; ++i) part[i] = 0.0;
}

static double sum_partials(const double *part, int np) {
    double s = 0;
    for (int i = 0; i < np; ++i) s += part[i];
    return s;
}

// the device layout is element-major; the oracle takes the reference order
template <typename T>
static void apply_aos(int dimx, int dimy, int p, double pen, const T *x, T *y) {
    const int64_t noe = (int64_t)dimx * dimy, nq = (int64_t)(p + 1) * (p + 1);
    std::vector<T> xr(noe * nq), yr(noe * nq);
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) xr[q * noe + e] = x[e * nq + q];
    if constexpr (sizeof(T) == 8) orc_apply(dimx, dimy, p, pen, xr.data(), yr.data());
    else orc_apply_f32(dimx, dimy, p, (float)pen, xr.data(), yr.data());
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) y[e * nq + q] = yr[q * noe + e];
}

int launch_apply(int dtype, int p, const void *x, void *y, const void *, int dimx, int dimy,
                 double pen, double *partials, const cg_state *st, int slot, hipStream_t, int) {
    if (partials && (st->s[slot].done | st->err)) return SIPG_OK;
    const int64_t n = (int64_t)dimx * dimy * (p + 1) * (p + 1);
    if (dtype == SIPG_F64) apply_aos<double>(dimx, dimy, p, pen, tp<double>(x), tp<double>(y));
    else apply_aos<float>(dimx, dimy, p, pen, tp<float>(x), tp<float>(y));
    if (partials) put_partials(partials, kMockApplyBlocks, dotv(dtype, x, y, n));
    return SIPG_OK;
}

int launch_dot_partials(int dtype, const void *a, const void *b, int64_t n, double *partials, int nblk,
                        hipStream_t) {
    put_partials(partials, nblk, dotv(dtype, a, b, n));
    return SIPG_OK;
}

int launch_reduce_partials(const double *partials, int np, double *out, hipStream_t) {
    out[0] = sum_partials(partials, np);
    return SIPG_OK;
}

template <typename T> static void axpy_t(T a, const T *x, T *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
}
int launch_copy(int dtype, const void *x, void *y, int64_t n, hipStream_t) {
    std::memcpy(y, x, (size_t)n * (dtype == SIPG_F64 ? 8 : 4));
    return SIPG_OK;
}

int launch_axpy(int dtype, double alpha, const void *x, void *y, int64_t n, hipStream_t) {
    if (dtype == SIPG_F64) axpy_t<double>(alpha, tp<double>(x), tp<double>(y), n);
    else axpy_t<float>((float)alpha, tp<float>(x), tp<float>(y), n);
    return SIPG_OK;
}

template <typename T> static void xpby_t(const T *r, T b, T *p, int64_t n) {
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + b * p[i];
}
int launch_xpby(int dtype, const void *r, double beta, void *p, int64_t n, hipStream_t) {
    if (dtype == SIPG_F64) xpby_t<double>(tp<double>(r), beta, tp<double>(p), n);
    else xpby_t<float>(tp<float>(r), (float)beta, tp<float>(p), n);
    return SIPG_OK;
}

template <typename T> static void alpha_t(T a, T *x, const T *p, T *r, const T *Ap, int64_t n) {
    for (int64_t i = 0; i < n; ++i) { x[i] = x[i] + a * p[i]; r[i] = r[i] - a * Ap[i]; }
}
int launch_alpha_update(int dtype, double alpha, void *x, const void *p, void *r, const void *Ap,
                        int64_t n, hipStream_t) {
    if (dtype == SIPG_F64) alpha_t<double>(alpha, tp<double>(x), tp<double>(p), tp<double>(r), tp<double>(Ap), n);
    else alpha_t<float>((float)alpha, tp<float>(x), tp<float>(p), tp<float>(r), tp<float>(Ap), n);
    return SIPG_OK;
}

int launch_cg_init_finalize(const double *prr, const double *pbb, int np, cg_state *st, double tol,
                            int tol_mode, int maxit, int true_stop, hipStream_t) {
    const double rr = sum_partials(prr, np), bb = sum_partials(pbb, np);
    const double stop = tol_mode == SIPG_TOL_REL ? tol * std::sqrt(bb) : tol;
    cg_slot s0{};
    s0.rr = rr;
    s0.rz = rr;
    s0.done = (std::sqrt(rr) > stop && maxit > 0) ? 0 : 1;
    s0.checked = 1;
    st->s[0] = s0;
    st->s[1] = s0;
    st->rr0 = rr;
    st->stop_true = stop;
    st->rr_ref = rr;
    st->bb = bb;
    st->stop = stop;
    st->maxit = maxit;
    st->nrep = 0;
    st->err = (!std::isfinite(rr) || !std::isfinite(bb)) ? 3 : 0;
    st->true_stop = true_stop;
    return SIPG_OK;
}

template <typename T>
static void update_xr_t(T *x, const T *p, T *r, const T *Ap, int64_t n, const double *ppAp, int np,
                        double *prr, int nblk, cg_state *st, int slot, T *xlo) {
    if (st->s[slot].done | st->err) return;
    const double pAp = sum_partials(ppAp, np);
    const double rr = st->s[slot].rr;
    const double alpha_d = rr / pAp;
    if (!(pAp > 0.0) || !std::isfinite(alpha_d)) { st->err = std::isfinite(pAp) ? 2 : 3; return; }
    const T a = (T)alpha_d;
    double s = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (xlo) comp_axpy_h(x[i], xlo[i], a, p[i]);
        else x[i] = x[i] + a * p[i];
        const T ri = r[i] - a * Ap[i];
        r[i] = ri;
        s += (double)ri * (double)ri;
    }
    put_partials(prr, nblk, s);
}
int &mock_p_updates();

int launch_cg_update_xr(int dtype, void *x, const void *p, void *r, const void *Ap, int64_t n,
                        const double *part_pAp, int np_pAp, double *part_rr, int nblk, cg_state *st,
                        int slot, hipStream_t, void *xlo) {
    if (dtype == SIPG_F64)
        update_xr_t<double>(tp<double>(x), tp<double>(p), tp<double>(r), tp<double>(Ap), n, part_pAp, np_pAp, part_rr, nblk, st, slot, tp<double>(xlo));
    else
        update_xr_t<float>(tp<float>(x), tp<float>(p), tp<float>(r), tp<float>(Ap), n, part_pAp, np_pAp, part_rr, nblk, st, slot, tp<float>(xlo));
    return SIPG_OK;
}

template <typename T>
static void update_p_t(const T *r, T *p, int64_t n, const double *prr, int np, cg_state *st, int slot) {
    const cg_slot cur = st->s[slot];
    if (cur.done | st->err) { st->s[slot ^ 1] = cur; return; }
    const double rr_new = sum_partials(prr, np);
    const bool replaced = (cur.flags & kSlotReplaced) != 0;
    const double beta_d = (cur.flags & kSlotRestart) ? 0.0 : rr_new / cur.rr;
    const bool met = !(std::sqrt(rr_new) > st->stop);
    const bool need_true = st->true_stop && !replaced;
    cg_slot nx{};
    nx.rr = rr_new;
    nx.rz = rr_new;
    nx.beta = beta_d;
    nx.it = cur.it + 1;
    nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
    nx.checked = replaced ? 1 : 0;
    nx.flags = (met && need_true) ? kSlotWantCheck : 0;
    st->s[slot ^ 1] = nx;
    if (replaced) st->rr_ref = rr_new;
    if (!std::isfinite(rr_new)) st->err = 3;
    const T b = (T)beta_d;
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + b * p[i];
}
int launch_cg_update_p(int dtype, const void *r, void *p, int64_t n, const double *part_rr, int np_rr,
                       int, cg_state *st, int slot, hipStream_t) {
    ++mock_p_updates();
    if (dtype == SIPG_F64) update_p_t<double>(tp<double>(r), tp<double>(p), n, part_rr, np_rr, st, slot);
    else update_p_t<float>(tp<float>(r), tp<float>(p), n, part_rr, np_rr, st, slot);
    return SIPG_OK;
}

int launch_residual(int dtype, const void *b, const void *Ax, void *r, int64_t n, hipStream_t) {
    for (int64_t i = 0; i < n; ++i) {
        if (dtype == SIPG_F64) tp<double>(r)[i] = tp<double>(b)[i] - tp<double>(Ax)[i];
        else tp<float>(r)[i] = tp<float>(b)[i] - tp<float>(Ax)[i];
    }
    return SIPG_OK;
}

template <typename TI, typename TO>
static void true_res_t(const TI *b, const TI *Ax, TO *out, TO *pc, int64_t n, double *prr, double *pbb,
                       int nblk) {
    double srr = 0, sbb = 0;
    for (int64_t i = 0; i < n; ++i) {
        const TI ri_ = b[i] - Ax[i];
        const double ri = (double)ri_, bi = (double)b[i];
        const TO ro = (TO)ri;
        if (out) out[i] = ro;
        if (pc) pc[i] = ro;
        srr += ri * ri;
        sbb += bi * bi;
    }
    put_partials(prr, nblk, srr);
    if (pbb) put_partials(pbb, nblk, sbb);
}
int launch_true_residual(int in_dtype, int out_dtype, const void *b, const void *Ax, void *out,
                         void *p_copy, int64_t n, double *prr, double *pbb, int nblk, hipStream_t) {
    if (in_dtype == SIPG_F64 && out_dtype == SIPG_F64)
        true_res_t<double, double>(tp<double>(b), tp<double>(Ax), tp<double>(out), tp<double>(p_copy), n, prr, pbb, nblk);
    else if (in_dtype == SIPG_F64 && out_dtype == SIPG_F32)
        true_res_t<double, float>(tp<double>(b), tp<double>(Ax), tp<float>(out), tp<float>(p_copy), n, prr, pbb, nblk);
    else if (in_dtype == SIPG_F32 && out_dtype == SIPG_F32)
        true_res_t<float, float>(tp<float>(b), tp<float>(Ax), tp<float>(out), tp<float>(p_copy), n, prr, pbb, nblk);
    else
        return SIPG_ERR_INVALID;
    return SIPG_OK;
}

int launch_accumulate(int x_dtype, int xl_dtype, void *x, void *xl, int64_t n, hipStream_t, void *xlo) {
    if (x_dtype == SIPG_F32 && xl_dtype == SIPG_F64) return SIPG_ERR_INVALID;
    for (int64_t i = 0; i < n; ++i) {
        if (xl_dtype == SIPG_F64) {
            const double v = xlo ? tp<double>(xl)[i] + tp<double>(xlo)[i] : tp<double>(xl)[i];
            tp<double>(x)[i] += v;
            tp<double>(xl)[i] = 0;
            if (xlo) tp<double>(xlo)[i] = 0;
        } else {
            const float v = xlo ? tp<float>(xl)[i] + tp<float>(xlo)[i] : tp<float>(xl)[i];
            if (x_dtype == SIPG_F64) tp<double>(x)[i] += (double)v;
            else tp<float>(x)[i] += v;
            tp<float>(xl)[i] = 0;
            if (xlo) tp<float>(xlo)[i] = 0;
        }
    }
    return SIPG_OK;
}

int launch_comp_fold(int dtype, void *x, void *xlo, void *xi, int64_t n, hipStream_t) {
    for (int64_t i = 0; i < n; ++i) {
        if (dtype == SIPG_F64) {
            comp_axpy_h(tp<double>(x)[i], tp<double>(xlo)[i], 1.0, tp<double>(xi)[i]);
            tp<double>(xi)[i] = 0;
        } else {
            comp_axpy_h(tp<float>(x)[i], tp<float>(xlo)[i], 1.0f, tp<float>(xi)[i]);
            tp<float>(xi)[i] = 0;
        }
    }
    return SIPG_OK;
}

// solves the peek finished (read and reset by the engine driver)
int &mock_peeks() {
    static int n = 0;
    return n;
}

// p-update launches, one per iteration the engine enqueues, including the
// no-op ones past a solve's finish (read and reset by the engine driver)
int &mock_p_updates() {
    static int n = 0;
    return n;
}

int launch_sum(int dtype, const void *x, const void *xl, void *out, int64_t n, hipStream_t, const void *xlo) {
    for (int64_t i = 0; i < n; ++i) {
        if (dtype == SIPG_F64)
            tp<double>(out)[i] = tp<double>(x)[i] + (xlo ? tp<double>(xl)[i] + tp<double>(xlo)[i] : tp<double>(xl)[i]);
        else
            tp<float>(out)[i] = tp<float>(x)[i] + (xlo ? tp<float>(xl)[i] + tp<float>(xlo)[i] : tp<float>(xl)[i]);
    }
    return SIPG_OK;
}

int launch_peek_finalize(const double *prr, int np, cg_state *st, int slot, hipStream_t) {
    const double rr = sum_partials(prr, np);
    cg_slot c = st->s[slot];
    if (c.done | st->err) return SIPG_OK;
    if (!std::isfinite(rr) || std::sqrt(rr) > st->stop_true) return SIPG_OK;
    c.rr = rr;
    c.done = 1;
    c.checked = 1;
    c.flags = 0;
    st->s[slot] = c;
    st->rr_ref = rr;
    ++mock_peeks();
    return SIPG_OK;
}

int launch_mark_replaced(cg_state *st, int slot, int restart, hipStream_t) {
    cg_slot c = st->s[slot];
    if (c.done | st->err) return SIPG_OK;
    c.flags |= kSlotReplaced | (restart ? kSlotRestart : 0);
    st->s[slot] = c;
    st->nrep += 1;
    return SIPG_OK;
}

int launch_convert(int from, int to, const void *src, void *dst, int64_t n, hipStream_t) {
    for (int64_t i = 0; i < n; ++i) {
        const double v = from == SIPG_F64 ? tp<double>(src)[i] : (double)tp<float>(src)[i];
        if (to == SIPG_F64) tp<double>(dst)[i] = v;
        else tp<float>(dst)[i] = (float)v;
    }
    return SIPG_OK;
}

int launch_layout(int dtype, int to_device, const void *src, void *dst, int64_t noe, int nq, hipStream_t) {
    const size_t sz = dtype == SIPG_F64 ? 8 : 4;
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) {
            const int64_t r = q * noe + e, d = e * nq + q;
            if (to_device) std::memcpy((char *)dst + d * sz, (const char *)src + r * sz, sz);
            else std::memcpy((char *)dst + r * sz, (const char *)src + d * sz, sz);
        }
    return SIPG_OK;
}

// the oracle's right-hand side of the Fourier forcing (g = 0), transposed
// to the element-major order; FP64 only
int launch_rhs_fourier(int dtype, int p, int dim, const void *, int nmodes, const double *a, const int *m,
                       const int *n, void *b, hipStream_t) {
    if (dtype != SIPG_F64) return SIPG_ERR_UNSUPPORTED;
    mock_drain();      // runs now: the work queued before it (the modes' copies) first
    orc_problem pb{0, nmodes, a, m, n};
    const int64_t noe = (int64_t)dim * dim, nq = (int64_t)(p + 1) * (p + 1);
    std::vector<double> br(noe * nq);
    orc_rhs(dim, p, 5.0 * (p + 1) * (p + 1), &pb, br.data());   // pen does not enter with g = 0
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) tp<double>(b)[e * nq + q] = br[q * noe + e];
    return SIPG_OK;
}

// error norms by the oracle, as one block partial (max, L2^2, 0, 0)
int launch_err_norms(int dtype, int p, int dim, const double *, const double *, int nmodes, const double *a,
                     const int *m, const int *n, const void *u, double *part, int nb, hipStream_t) {
    if (dtype != SIPG_F64 || nb < 1) return SIPG_ERR_UNSUPPORTED;
    orc_problem pb{0, nmodes, a, m, n};
    const int64_t noe = (int64_t)dim * dim, nq = (int64_t)(p + 1) * (p + 1);
    std::vector<double> ur(noe * nq);
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) ur[q * noe + e] = tp<const double>(u)[e * nq + q];
    double out[3];
    orc_err_norms(dim, p, ur.data(), &pb, out);
    for (int k = 0; k < 4; ++k)
        for (int b = 0; b < nb; ++b) part[(size_t)k * nb + b] = k == 0 ? -1.0 : 0.0;
    part[0] = out[0];
    part[nb] = out[1] * out[1];
    part[2 * nb] = out[2] * out[2] - out[1] * out[1];
    return SIPG_OK;
}

// error norms against nodal samples, as one block partial: the norms'
// definition written out per node (FP64 only)
int launch_err_norms_nodal(int dtype, int p, int dim, const double *, const double *D, const double *ue,
                           const double *uxe, const double *uye, const void *u, double *part, int nb, hipStream_t) {
    if (dtype != SIPG_F64 || nb < 1) return SIPG_ERR_UNSUPPORTED;
    const int n1 = p + 1, nq = n1 * n1;
    const int64_t noe = (int64_t)dim * dim;
    std::vector<double> z(n1), w(n1);
    orc_gll(p, z.data(), w.data());
    const double *U = tp<const double>(u);
    double emax = -1.0, l2 = 0.0, sx = 0.0, sy = 0.0;
    for (int64_t e = 0; e < noe; ++e)
        for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n1; ++i2) {
                const int64_t k = e * nq + i1 * n1 + i2;
                double dxu = 0.0, dyu = 0.0;
                for (int i = 0; i < n1; ++i) {
                    dxu += U[e * nq + i * n1 + i2] * 2.0 * dim * D[i * n1 + i1];
                    dyu += U[e * nq + i1 * n1 + i] * 2.0 * dim * D[i * n1 + i2];
                }
                const double wgt = (w[i1] * .5 / dim) * (w[i2] * .5 / dim), err = std::fabs(U[k] - ue[k]);
                emax = std::fmax(emax, err);
                l2 += wgt * err * err;
                sx += wgt * (dxu - uxe[k]) * (dxu - uxe[k]);
                sy += wgt * (dyu - uye[k]) * (dyu - uye[k]);
            }
    for (int k = 0; k < 4; ++k)
        for (int b = 0; b < nb; ++b) part[(size_t)k * nb + b] = k == 0 ? -1.0 : 0.0;
    part[0] = emax;
    part[nb] = l2;
    part[2 * nb] = sx;
    part[3 * nb] = sy;
    return SIPG_OK;
}

// ---- PCG (kernels_pcg.hip) ----------------------------------------------
// z = M^-1 r per element with the op's tables [S, Linv]:
// Z = S [(S^T R S) .* Linv] S^T on each (p+1) x (p+1) tile.
int pcg_num_blocks(int, int p, double, int, int64_t) { return p <= kMaxDegree ? kMockApplyBlocks : -1; }

template <typename T>
static void pcg_rz_t(int p, int mode, T *r, const T *Ap, T *z, T *pcopy, int64_t noe, const T *prec,
                     const double *ppAp, int np, double *prr, double *prz, cg_state *st, int slot) {
    if (mode == 1) {
        if (st->s[slot].done | st->err) return;
        const double pAp = sum_partials(ppAp, np);
        const double a = st->s[slot].rz / pAp;
        if (!(pAp > 0.0) || !std::isfinite(a)) { st->err = std::isfinite(pAp) ? 2 : 3; return; }
        const int64_t n = noe * (p + 1) * (p + 1);
        for (int64_t i = 0; i < n; ++i) r[i] = r[i] - (T)a * Ap[i];
    }
    const int n1 = p + 1, nq = n1 * n1;
    const T *S = prec, *L = prec + nq;
    std::vector<double> t1(nq), t2(nq);
    double srr = 0, srz = 0;
    for (int64_t e = 0; e < noe; ++e) {
        const T *R = r + e * nq;
        for (int a = 0; a < n1; ++a)                 // t1 = S^T R
            for (int j = 0; j < n1; ++j) {
                double acc = 0;
                for (int i = 0; i < n1; ++i) acc += (double)S[i * n1 + a] * R[i * n1 + j];
                t1[a * n1 + j] = acc;
            }
        for (int a = 0; a < n1; ++a)                 // t2 = (t1 S) .* L
            for (int c = 0; c < n1; ++c) {
                double acc = 0;
                for (int b = 0; b < n1; ++b) acc += t1[a * n1 + b] * S[b * n1 + c];
                t2[a * n1 + c] = acc * L[a * n1 + c];
            }
        for (int a = 0; a < n1; ++a)                 // t1 = t2 S^T
            for (int d = 0; d < n1; ++d) {
                double acc = 0;
                for (int c = 0; c < n1; ++c) acc += t2[a * n1 + c] * S[d * n1 + c];
                t1[a * n1 + d] = acc;
            }
        for (int a = 0; a < n1; ++a)                 // Z = S t1
            for (int j = 0; j < n1; ++j) {
                double acc = 0;
                for (int i = 0; i < n1; ++i) acc += (double)S[a * n1 + i] * t1[i * n1 + j];
                const T zv = (T)acc;
                z[e * nq + a * n1 + j] = zv;
                if (mode == 2) pcopy[e * nq + a * n1 + j] = zv;
            }
        for (int q = 0; q < nq; ++q) {
            srr += (double)R[q] * (double)R[q];
            srz += (double)R[q] * (double)z[e * nq + q];
        }
    }
    put_partials(prr, kMockApplyBlocks, srr);
    put_partials(prz, kMockApplyBlocks, srz);
}

int launch_pcg_rz(int dtype, int p, double, int mode, void *r, const void *Ap, void *z, void *pcopy,
                  int64_t noe, const void *prec, const double *part_pAp, int np_pAp, double *part_rr,
                  double *part_rz, cg_state *st, int slot, hipStream_t, double *) {
    if (p < 1 || p > kMaxDegree || !prec) return SIPG_ERR_INVALID;
    if (dtype == SIPG_F64)
        pcg_rz_t<double>(p, mode, tp<double>(r), tp<double>(Ap), tp<double>(z), tp<double>(pcopy), noe,
                         tp<double>(prec), part_pAp, np_pAp, part_rr, part_rz, st, slot);
    else
        pcg_rz_t<float>(p, mode, tp<float>(r), tp<float>(Ap), tp<float>(z), tp<float>(pcopy), noe,
                        tp<float>(prec), part_pAp, np_pAp, part_rr, part_rz, st, slot);
    return SIPG_OK;
}

template <typename T>
static void pcg_update_p_t(T *x, T *p, const T *z, int64_t n, const double *ppAp, int np_pAp,
                           const double *prr, const double *prz, int np, cg_state *st, int slot, T *xlo) {
    const cg_slot cur = st->s[slot];
    if (cur.done | st->err) { st->s[slot ^ 1] = cur; return; }
    const double rr_new = sum_partials(prr, np), rz_new = sum_partials(prz, np);
    const double pAp = sum_partials(ppAp, np_pAp);
    const bool replaced = (cur.flags & kSlotReplaced) != 0;
    const double beta_d = (cur.flags & kSlotRestart) ? 0.0 : rz_new / cur.rz;
    const bool met = !(std::sqrt(rr_new) > st->stop);
    const bool need_true = st->true_stop && !replaced;
    cg_slot nx{};
    nx.rr = rr_new;
    nx.rz = rz_new;
    nx.beta = beta_d;
    nx.it = cur.it + 1;
    nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
    nx.checked = replaced ? 1 : 0;
    nx.flags = (met && need_true) ? kSlotWantCheck : 0;
    st->s[slot ^ 1] = nx;
    if (replaced) st->rr_ref = rr_new;
    if (!std::isfinite(rr_new) || !std::isfinite(rz_new)) st->err = 3;
    else if (!(rz_new > 0.0) && rr_new > 0.0) st->err = 2;
    const T a = (T)(cur.rz / pAp), b = (T)beta_d;
    for (int64_t i = 0; i < n; ++i) {
        const T pi = p[i];
        if (!replaced) {
            if (xlo) comp_axpy_h(x[i], xlo[i], a, pi);
            else x[i] = x[i] + a * pi;
        }
        p[i] = z[i] + b * pi;
    }
}

int launch_pcg_update_p(int dtype, void *x, void *p, const void *z, int64_t n, const double *part_pAp,
                        int np_pAp, const double *part_rr, const double *part_rz, int np, int, cg_state *st,
                        int slot, hipStream_t, void *xlo) {
    ++mock_p_updates();
    if (dtype == SIPG_F64)
        pcg_update_p_t<double>(tp<double>(x), tp<double>(p), tp<double>(z), n, part_pAp, np_pAp, part_rr, part_rz,
                               np, st, slot, tp<double>(xlo));
    else
        pcg_update_p_t<float>(tp<float>(x), tp<float>(p), tp<float>(z), n, part_pAp, np_pAp, part_rr, part_rz, np,
                              st, slot, tp<float>(xlo));
    return SIPG_OK;
}

bool pcg_rz_fuses_restrict(int, int, double) { return true; }

// ---- coarse level (kernels_coarse.hip), FP64: the same DST solve in loops -
int coarse_num_partials(int dimx, int dimy) { return (dimx > 1 && dimy > 1) ? 1 : 0; }

static double corner_weight(int c, double xa, double yb) {
    return ((c & 1) ? xa : 1.0 - xa) * ((c & 2) ? yb : 1.0 - yb);
}

int launch_coarse_solve(int dtype, int p, const void *r, int dimx, int dimy, const coarse_tables &c, hipStream_t,
                        bool, const cg_state *st, int slot) {
    if (dtype != SIPG_F64) return SIPG_ERR_UNSUPPORTED;
    if (st && (st->s[slot].done | st->err)) return SIPG_OK;
    const int nvx = dimx - 1, nvy = dimy - 1, n1 = p + 1, nq = n1 * n1;
    if (nvx < 1 || nvy < 1) return SIPG_OK;
    const double *R = tp<const double>(r);
    std::vector<double> S((size_t)nvx * nvy, 0.0);
    // P0 r: each element adds its corner-weighted sums to its interior corners
    for (int row = 0; row < dimy; ++row)
        for (int col = 0; col < dimx; ++col) {
            const int64_t e = (int64_t)row * dimx + col;
            for (int cc = 0; cc < 4; ++cc) {
                const int a = row + ((cc & 2) ? 0 : 1), b = col + ((cc & 1) ? 1 : 0);   // vertex of corner cc
                if (a < 1 || a > nvy || b < 1 || b > nvx) continue;
                double acc = 0.0;
                for (int i1 = 0; i1 < n1; ++i1)
                    for (int i2 = 0; i2 < n1; ++i2)
                        acc += R[e * nq + i1 * n1 + i2] * corner_weight(cc, c.xi[i1], c.xi[i2]);
                S[(size_t)(a - 1) * nvx + (b - 1)] += acc;
            }
        }
    auto mul = [](const double *L, const double *Rm, int M, int K, int N, double *out) {
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < N; ++j) {
                double acc = 0.0;
                for (int k = 0; k < K; ++k) acc += L[(size_t)i * K + k] * Rm[(size_t)k * N + j];
                out[(size_t)i * N + j] = acc;
            }
    };
    std::vector<double> T1(S.size()), T2(S.size()), C(S.size());
    mul(S.data(), c.Qx, nvy, nvx, nvx, T1.data());
    mul(c.Qy, T1.data(), nvy, nvy, nvx, T2.data());
    for (size_t k = 0; k < T2.size(); ++k) T2[k] *= c.Linv[k];
    mul(T2.data(), c.Qx, nvy, nvx, nvx, T1.data());
    mul(c.Qy, T1.data(), nvy, nvy, nvx, C.data());
    double dot = 0.0;
    for (int a = 0; a < nvy; ++a)
        for (int b = 0; b < nvx; ++b) {
            c.Cpad[(size_t)(a + 1) * (dimx + 1) + b + 1] = C[(size_t)a * nvx + b];
            dot += C[(size_t)a * nvx + b] * S[(size_t)a * nvx + b];
        }
    c.part[0] = dot;
    return SIPG_OK;
}

static double mock_prolong(const coarse_tables &c, int dimx, int n1, int64_t idx) {
    const int nq = n1 * n1;
    const int64_t e = idx / nq;
    const int node = (int)(idx - e * nq), i1 = node / n1, i2 = node % n1;
    const int row = (int)(e / dimx), col = (int)(e % dimx), ld = dimx + 1;
    double v = 0.0;
    for (int cc = 0; cc < 4; ++cc) {
        const int a = row + ((cc & 2) ? 0 : 1), b = col + ((cc & 1) ? 1 : 0);
        v += c.Cpad[(size_t)a * ld + b] * corner_weight(cc, c.xi[i1], c.xi[i2]);
    }
    return v;
}

int launch_coarse_add(int dtype, int p, void *v, int dimx, int64_t n, const coarse_tables &c, int, hipStream_t) {
    if (dtype != SIPG_F64) return SIPG_ERR_UNSUPPORTED;
    for (int64_t i = 0; i < n; ++i) tp<double>(v)[i] += mock_prolong(c, dimx, p + 1, i);
    return SIPG_OK;
}

int launch_pcg2_update_p(int dtype, int p, void *x, void *pv, const void *z, int64_t n, const double *part_pAp,
                         int np_pAp, const double *part_rr, const double *part_rz, int np, int npc, int dimx,
                         const coarse_tables &c, int, cg_state *st, int slot, hipStream_t, void *xlo) {
    ++mock_p_updates();
    if (dtype != SIPG_F64) return SIPG_ERR_UNSUPPORTED;
    // z + P0^T C through a scratch copy, then the plain update with r.z + S.C
    std::vector<double> zc(tp<const double>(z), tp<const double>(z) + n);
    for (int64_t i = 0; i < n; ++i) zc[i] += mock_prolong(c, dimx, p + 1, i);
    std::vector<double> prz(part_rz, part_rz + np);
    if (np > 0) prz[np - 1] += sum_partials(c.part, npc);
    pcg_update_p_t<double>(tp<double>(x), tp<double>(pv), zc.data(), n, part_pAp, np_pAp, part_rr, prz.data(), np,
                           st, slot, tp<double>(xlo));
    return SIPG_OK;
}

int launch_pcg_x(int dtype, void *x, const void *p, int64_t n, const double *part_pAp, int np_pAp, int,
                 const cg_state *st, int slot, hipStream_t, void *xlo) {
    if (st->s[slot].done | st->err) return SIPG_OK;
    const double a = st->s[slot].rz / sum_partials(part_pAp, np_pAp);
    for (int64_t i = 0; i < n; ++i) {
        if (dtype == SIPG_F64) {
            if (xlo) comp_axpy_h(tp<double>(x)[i], tp<double>(xlo)[i], a, tp<double>(p)[i]);
            else tp<double>(x)[i] += (double)a * tp<double>(p)[i];
        } else {
            if (xlo) comp_axpy_h(tp<float>(x)[i], tp<float>(xlo)[i], (float)a, tp<float>(p)[i]);
            else tp<float>(x)[i] += (float)a * tp<float>(p)[i];
        }
    }
    return SIPG_OK;
}

int launch_pcg_init_finalize(const double *prr, const double *prz, int np, const double *pbb, int npb,
                             cg_state *st, double tol, int tol_mode, int maxit, int true_stop, hipStream_t,
                             const double *pc, int npc) {
    const double rr = sum_partials(prr, np), rz = sum_partials(prz, np) + sum_partials(pc, npc),
                 bb = sum_partials(pbb, npb);
    const double stop = tol_mode == SIPG_TOL_REL ? tol * std::sqrt(bb) : tol;
    cg_slot s0{};
    s0.rr = rr;
    s0.rz = rz;
    s0.done = (std::sqrt(rr) > stop && maxit > 0) ? 0 : 1;
    s0.checked = 1;
    st->s[0] = s0;
    st->s[1] = s0;
    st->rr0 = rr;
    st->stop_true = stop;
    st->rr_ref = rr;
    st->bb = bb;
    st->stop = stop;
    st->maxit = maxit;
    st->nrep = 0;
    st->err = (!std::isfinite(rr) || !std::isfinite(bb) || !std::isfinite(rz)) ? 3 : 0;
    st->true_stop = true_stop;
    return SIPG_OK;
}

}  // namespace sipg_internal

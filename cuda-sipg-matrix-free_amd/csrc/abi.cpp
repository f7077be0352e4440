// abi.cpp -- the C-ABI of libsipg_hip (declared in include/sipg_hip.h).
//
// Host orchestration only: argument checks, table set-up, workspace, and
// the device-resident CG loop.  All arithmetic on vectors runs in the HIP
// kernels of kernels_apply.hip / kernels_vec.hip; there is no CPU fallback.
#include "sipg_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <random>

using namespace sipg_internal;

namespace {
thread_local hipError_t g_last_hip = hipSuccess;

size_t dtype_size(int dtype) { return dtype == SIPG_F64 ? 8 : 4; }

bool overlaps(const void *a, const void *b, size_t bytes) {
    const char *pa = (const char *)a, *pb = (const char *)b;
    return pa < pb + bytes && pb < pa + bytes;
}

int ensure_ws(sipg_op *op, size_t bytes) {
    if (op->ws_bytes >= bytes) return SIPG_OK;
    if (op->ws) {
        SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
        SIPG_HIP_TRY(hipFree(op->ws));
        op->ws = nullptr;
        op->ws_bytes = 0;
    }
    SIPG_HIP_TRY(hipMalloc(&op->ws, bytes));
    op->ws_bytes = bytes;
    return SIPG_OK;
}

// Partial-buffer layout inside op->d_part
struct parts {
    double *pAp, *rr, *bb, *misc;
    int napply;
};
parts part_layout(sipg_op *op) {
    parts p;
    p.napply = apply_num_blocks(op->dtype, op->p, op->dimx, op->dimy);
    p.pAp = op->d_part;
    p.rr = op->d_part + std::max(p.napply, kMaxPartials);
    p.bb = p.rr + kMaxPartials;
    p.misc = p.bb + kMaxPartials;
    return p;
}
}  // namespace

void sipg_internal::set_last_hip_error(hipError_t e) { g_last_hip = e; }

extern "C" {

const char *sipg_strerror(int s) {
    switch (s) {
    case SIPG_OK: return "ok";
    case SIPG_ERR_INVALID: return "invalid argument";
    case SIPG_ERR_ALIAS: return "input and output vectors overlap";
    case SIPG_ERR_UNSUPPORTED: return "degree or dtype not supported by the built kernels";
    case SIPG_ERR_HIP: return g_last_hip != hipSuccess ? hipGetErrorString(g_last_hip) : "HIP runtime error";
    case SIPG_ERR_NOMEM: return "device allocation failed";
    case SIPG_NOT_CONVERGED: return "iteration cap reached before the tolerance";
    case SIPG_ERR_NONFINITE: return "non-finite value in a residual or dot product";
    case SIPG_ERR_BREAKDOWN: return "CG breakdown: p.Ap <= 0 (operator not positive definite)";
    case SIPG_ERR_BUDGET: return "request exceeds the memory budget";
    case SIPG_ERR_CANCELLED: return "request cancelled or timed out";
    case SIPG_ERR_NOT_FOUND: return "unknown request or handle";
    default: return "unknown status";
    }
}

int sipg_max_degree(void) { return kMaxDegree; }

int sipg_device_count(int *count) {
    if (!count) return SIPG_ERR_INVALID;
    SIPG_HIP_TRY(hipGetDeviceCount(count));
    return SIPG_OK;
}

int sipg_set_device(int device) {
    SIPG_HIP_TRY(hipSetDevice(device));
    return SIPG_OK;
}

int sipg_device_synchronize(void) {
    SIPG_HIP_TRY(hipDeviceSynchronize());
    return SIPG_OK;
}

int sipg_malloc(void **ptr, size_t bytes) {
    if (!ptr) return SIPG_ERR_INVALID;
    *ptr = nullptr;
    if (bytes == 0) return SIPG_OK;
    SIPG_HIP_TRY(hipMalloc(ptr, bytes));
    return SIPG_OK;
}

int sipg_free(void *ptr) {
    if (ptr) SIPG_HIP_TRY(hipFree(ptr));
    return SIPG_OK;
}

int sipg_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    SIPG_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SIPG_OK;
}
int sipg_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    SIPG_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SIPG_OK;
}
int sipg_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    SIPG_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return SIPG_OK;
}
int sipg_memset_zero(void *dst, size_t bytes) {
    SIPG_HIP_TRY(hipMemset(dst, 0, bytes));
    return SIPG_OK;
}

int sipg_op_create(sipg_op **out, int dimx, int dimy, int p, double pen, int dtype, void *stream) {
    if (!out) return SIPG_ERR_INVALID;
    *out = nullptr;
    if (dimx < 1 || dimy < 1 || p < 1 || !std::isfinite(pen)) return SIPG_ERR_INVALID;
    if (dtype != SIPG_F64 && dtype != SIPG_F32) return SIPG_ERR_INVALID;
    if (p > kMaxDegree) return SIPG_ERR_UNSUPPORTED;
    const int64_t noe = (int64_t)dimx * dimy;
    if (noe > (int64_t)INT_MAX) return SIPG_ERR_INVALID;
    sipg_op *op = new sipg_op();
    op->dimx = dimx;
    op->dimy = dimy;
    op->p = p;
    op->pen = pen;
    op->dtype = dtype;
    op->n = noe * (p + 1) * (p + 1);
    op->stream = (hipStream_t)stream;
    build_tables(p, op->h_tab);
    const int tl = table_len(p);
    int rc = SIPG_OK;
    auto fail = [&](int code) { sipg_op_destroy(op); return code; };
    hipError_t e;
    if ((e = hipMalloc(&op->d_tab64, tl * sizeof(double))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
    if ((e = hipMemcpy(op->d_tab64, op->h_tab.data(), tl * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_HIP); }
    if (dtype == SIPG_F64) {
        op->d_tab = op->d_tab64;
    } else {
        std::vector<float> tf(op->h_tab.begin(), op->h_tab.end());
        if ((e = hipMalloc(&op->d_tab, tl * sizeof(float))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
        if ((e = hipMemcpy(op->d_tab, tf.data(), tl * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_HIP); }
    }
    const int napply = apply_num_blocks(dtype, p, dimx, dimy);
    if (napply < 1) return fail(SIPG_ERR_UNSUPPORTED);
    const size_t npart = (size_t)std::max(napply, kMaxPartials) + 3 * (size_t)kMaxPartials;
    if ((e = hipMalloc(&op->d_part, npart * sizeof(double))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
    if ((e = hipMalloc(&op->d_scalar, 256 * sizeof(double))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
    if ((e = hipMalloc(&op->d_state, sizeof(cg_state))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
    if ((e = hipHostMalloc((void **)&op->h_state, 2 * sizeof(cg_state))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
    (void)rc;
    *out = op;
    return SIPG_OK;
}

int sipg_op_destroy(sipg_op *op) {
    if (!op) return SIPG_OK;
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    else (void)hipDeviceSynchronize();
    if (op->d_tab && op->d_tab != op->d_tab64) (void)hipFree(op->d_tab);
    if (op->d_tab64) (void)hipFree(op->d_tab64);
    if (op->d_part) (void)hipFree(op->d_part);
    if (op->d_scalar) (void)hipFree(op->d_scalar);
    if (op->d_state) (void)hipFree(op->d_state);
    if (op->h_state) (void)hipHostFree(op->h_state);
    if (op->ws) (void)hipFree(op->ws);
    delete op;
    return SIPG_OK;
}

int sipg_op_size(const sipg_op *op, int64_t *n) {
    if (!op || !n) return SIPG_ERR_INVALID;
    *n = op->n;
    return SIPG_OK;
}

int sipg_op_set_stream(sipg_op *op, void *stream) {
    if (!op) return SIPG_ERR_INVALID;
    op->stream = (hipStream_t)stream;
    return SIPG_OK;
}

int sipg_op_tables(const sipg_op *op, double *nodes, double *weights, double *D, double *K1) {
    if (!op) return SIPG_ERR_INVALID;
    const int n1 = op->p + 1;
    std::vector<double> z(n1), w(n1), Dm(n1 * n1), K(n1 * n1);
    gll(op->p, z.data(), w.data());
    dphi(op->p, z.data(), Dm.data());
    stiffness_1d(op->p, w.data(), Dm.data(), K.data());
    if (nodes) std::memcpy(nodes, z.data(), n1 * sizeof(double));
    if (weights) std::memcpy(weights, w.data(), n1 * sizeof(double));
    if (D) std::memcpy(D, Dm.data(), n1 * n1 * sizeof(double));
    if (K1) std::memcpy(K1, K.data(), n1 * n1 * sizeof(double));
    return SIPG_OK;
}

int sipg_apply(sipg_op *op, const void *x, void *y) {
    if (!op || !x || !y) return SIPG_ERR_INVALID;
    if (overlaps(x, y, (size_t)op->n * dtype_size(op->dtype))) return SIPG_ERR_ALIAS;
    return launch_apply(op->dtype, op->p, x, y, op->d_tab, op->dimx, op->dimy, op->pen, nullptr,
                        nullptr, op->stream);
}

int sipg_dot(sipg_op *op, const void *a, const void *b, double *result) {
    if (!op || !a || !b || !result) return SIPG_ERR_INVALID;
    parts P = part_layout(op);
    const int nb = vec_num_blocks(op->n);
    int rc = launch_dot_partials(op->dtype, a, b, op->n, P.misc, nb, op->stream);
    if (rc) return rc;
    rc = launch_reduce_partials(P.misc, nb, op->d_scalar, op->stream);
    if (rc) return rc;
    SIPG_HIP_TRY(hipMemcpyAsync(result, op->d_scalar, sizeof(double), hipMemcpyDeviceToHost, op->stream));
    SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    return SIPG_OK;
}

int sipg_axpy(sipg_op *op, double alpha, const void *x, void *y) {
    if (!op || !x || !y) return SIPG_ERR_INVALID;
    return launch_axpy(op->dtype, alpha, x, y, op->n, op->stream);
}

int sipg_xpby(sipg_op *op, const void *r, double beta, void *p) {
    if (!op || !r || !p) return SIPG_ERR_INVALID;
    return launch_xpby(op->dtype, r, beta, p, op->n, op->stream);
}

int sipg_alpha_update(sipg_op *op, double alpha, void *x, const void *p, void *r, const void *Ap) {
    if (!op || !x || !p || !r || !Ap) return SIPG_ERR_INVALID;
    return launch_alpha_update(op->dtype, alpha, x, p, r, Ap, op->n, op->stream);
}

int sipg_cg(sipg_op *op, void *x, const void *b, double tol, int tol_mode, int maxit, int *iters,
            double *res_norm) {
    if (!op || !x || !b) return SIPG_ERR_INVALID;
    if (!(tol >= 0.0) || (tol_mode != SIPG_TOL_ABS && tol_mode != SIPG_TOL_REL)) return SIPG_ERR_INVALID;
    const size_t sz = dtype_size(op->dtype);
    const size_t vb = (size_t)op->n * sz;
    if (overlaps(x, b, vb)) return SIPG_ERR_ALIAS;
    int rc = ensure_ws(op, 3 * vb);
    if (rc) return rc;
    char *ws = (char *)op->ws;
    void *r = ws, *p = ws + vb, *Ap = ws + 2 * vb;
    const int64_t cap = maxit > 0 ? (int64_t)maxit : op->n;
    const int imax = (int)std::min<int64_t>(cap, INT_MAX);
    parts P = part_layout(op);
    const int nb = vec_num_blocks(op->n);
    hipStream_t s = op->stream;
    cg_state *st = op->d_state;

    // r = b - A x ; p = r ; r.r, b.b   (conjugate_gradient_gpu.hpp:40-60)
    if ((rc = launch_apply(op->dtype, op->p, x, r, op->d_tab, op->dimx, op->dimy, op->pen, nullptr,
                           nullptr, s)))
        return rc;
    if ((rc = launch_cg_init_residual(op->dtype, b, r, p, op->n, P.rr, P.bb, nb, s))) return rc;
    if ((rc = launch_cg_init_finalize(P.rr, P.bb, nb, st, tol, tol_mode, imax, s))) return rc;

    // p.Ap block sums: re-reduced in every consumer block when few, else
    // folded to one value by a one-block kernel first.
    const bool fold = P.napply > kMaxPartials;
    const double *pAp_src = fold ? P.misc : P.pAp;
    const int np_pAp = fold ? 1 : P.napply;

    auto launch_iter = [&](int64_t k) -> int {
        const int slot = (int)(k & 1);
        int r2;
        if ((r2 = launch_apply(op->dtype, op->p, p, Ap, op->d_tab, op->dimx, op->dimy, op->pen,
                               P.pAp, st, s)))
            return r2;
        if (fold && (r2 = launch_reduce_partials(P.pAp, P.napply, P.misc, s))) return r2;
        if ((r2 = launch_cg_update_xr(op->dtype, x, p, r, Ap, op->n, pAp_src, np_pAp, P.rr, nb, st,
                                      slot, s)))
            return r2;
        return launch_cg_update_p(op->dtype, r, p, op->n, P.rr, nb, nb, st, slot, s);
    };

    // Host loop: keep two chunks of iterations queued; poll the state copy
    // of the older chunk.  Kernels become no-ops once the state says done.
    hipEvent_t ev[2];
    SIPG_HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    SIPG_HIP_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int64_t launched = 0;
    int chunk = 4;
    int inflight = 0;
    int buf = 0;
    rc = SIPG_OK;
    // initial state check
    SIPG_HIP_TRY(hipMemcpyAsync(&op->h_state[0], st, sizeof(cg_state), hipMemcpyDeviceToHost, s));
    SIPG_HIP_TRY(hipStreamSynchronize(s));
    bool finished = op->h_state[0].done != 0;
    while (!finished) {
        // enqueue one chunk
        for (int c = 0; c < chunk && launched < imax; ++c, ++launched)
            if ((rc = launch_iter(launched))) break;
        if (rc) break;
        SIPG_HIP_TRY(hipMemcpyAsync(&op->h_state[buf], st, sizeof(cg_state), hipMemcpyDeviceToHost, s));
        SIPG_HIP_TRY(hipEventRecord(ev[buf], s));
        ++inflight;
        buf ^= 1;
        if (chunk < 64) chunk *= 2;
        if (inflight == 2 || launched >= imax) {
            // wait for the older chunk
            SIPG_HIP_TRY(hipEventSynchronize(ev[buf]));
            --inflight;
            if (op->h_state[buf].done) finished = true;
            if (launched >= imax) {
                SIPG_HIP_TRY(hipStreamSynchronize(s));
                finished = true;
            }
        }
    }
    SIPG_HIP_TRY(hipStreamSynchronize(s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
    SIPG_HIP_TRY(hipMemcpy(&op->h_state[0], st, sizeof(cg_state), hipMemcpyDeviceToHost));
    const cg_state &hs = op->h_state[0];
    if (iters) *iters = hs.it;
    if (res_norm) *res_norm = std::sqrt(hs.s[hs.it & 1].rr);
    if (hs.done == 2) return SIPG_ERR_BREAKDOWN;
    if (hs.done == 3) return SIPG_ERR_NONFINITE;
    if (!(std::sqrt(hs.s[hs.it & 1].rr) <= hs.stop)) return SIPG_NOT_CONVERGED;
    return SIPG_OK;
}

int sipg_true_relres(sipg_op *op, const void *x, const void *b, double *relres) {
    if (!op || !x || !b || !relres) return SIPG_ERR_INVALID;
    if (op->dtype != SIPG_F64) return SIPG_ERR_INVALID;
    const size_t vb = (size_t)op->n * 8;
    int rc = ensure_ws(op, 3 * vb);
    if (rc) return rc;
    char *ws = (char *)op->ws;
    void *Ax = ws + 2 * vb;   // reuse the Ap slot
    void *r = ws;
    if ((rc = launch_apply(op->dtype, op->p, x, Ax, op->d_tab, op->dimx, op->dimy, op->pen, nullptr,
                           nullptr, op->stream)))
        return rc;
    if ((rc = launch_residual(op->dtype, b, Ax, r, op->n, op->stream))) return rc;
    double rr = 0, bb = 0;
    if ((rc = sipg_dot(op, r, r, &rr))) return rc;
    if ((rc = sipg_dot(op, b, b, &bb))) return rc;
    *relres = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return SIPG_OK;
}

int sipg_forcing_draw(uint64_t seed, int kmin, int kmax, int mmax, int *nmodes, double *a, int *m,
                      int *n) {
    if (!nmodes || !a || !m || !n || kmin < 1 || kmax < kmin || mmax < 1) return SIPG_ERR_INVALID;
    std::mt19937_64 gen(seed);
    const int K = kmin + (int)(gen() % (uint64_t)(kmax - kmin + 1));
    for (int k = 0; k < K; ++k) {
        a[k] = -1.0 + 2.0 * ((double)(gen() >> 11) * 0x1.0p-53);
        m[k] = 1 + (int)(gen() % (uint64_t)mmax);
        n[k] = 1 + (int)(gen() % (uint64_t)mmax);
    }
    *nmodes = K;
    return SIPG_OK;
}

int sipg_rhs_fourier(sipg_op *op, int nmodes, const double *a, const int *m, const int *n, void *b) {
    if (!op || !b || nmodes < 0 || nmodes > 64 || (nmodes > 0 && (!a || !m || !n)))
        return SIPG_ERR_INVALID;
    if (op->dimx != op->dimy) return SIPG_ERR_INVALID;
    // modes live in the op's scalar scratch: [0,64) a, then m, n as ints
    double *da = op->d_scalar + 64;
    int *dm = (int *)(op->d_scalar + 128);
    int *dn = dm + 64;
    if (nmodes > 0) {
        SIPG_HIP_TRY(hipMemcpyAsync(da, a, nmodes * sizeof(double), hipMemcpyHostToDevice, op->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(dm, m, nmodes * sizeof(int), hipMemcpyHostToDevice, op->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(dn, n, nmodes * sizeof(int), hipMemcpyHostToDevice, op->stream));
    }
    int rc = launch_rhs_fourier(op->dtype, op->p, op->dimx, op->d_tab64, nmodes, da, dm, dn, b, op->stream);
    if (rc) return rc;
    SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    return SIPG_OK;
}

int sipg_time_apply(sipg_op *op, const void *x, void *y, int reps, double *ms_per_apply) {
    if (!op || !x || !y || !ms_per_apply || reps < 1) return SIPG_ERR_INVALID;
    if (overlaps(x, y, (size_t)op->n * dtype_size(op->dtype))) return SIPG_ERR_ALIAS;
    hipEvent_t a, b;
    SIPG_HIP_TRY(hipEventCreate(&a));
    SIPG_HIP_TRY(hipEventCreate(&b));
    int rc = SIPG_OK;
    SIPG_HIP_TRY(hipEventRecord(a, op->stream));
    for (int i = 0; i < reps && rc == SIPG_OK; ++i)
        rc = launch_apply(op->dtype, op->p, x, y, op->d_tab, op->dimx, op->dimy, op->pen, nullptr,
                          nullptr, op->stream);
    SIPG_HIP_TRY(hipEventRecord(b, op->stream));
    SIPG_HIP_TRY(hipEventSynchronize(b));
    float ms = 0.f;
    SIPG_HIP_TRY(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_apply = (double)ms / reps;
    return rc;
}

int sipg_upload(sipg_op *op, const double *host, void *dev) {
    if (!op || !host || !dev) return SIPG_ERR_INVALID;
    if (op->dtype == SIPG_F64) {
        SIPG_HIP_TRY(hipMemcpy(dev, host, (size_t)op->n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(host, host + op->n);
        SIPG_HIP_TRY(hipMemcpy(dev, tmp.data(), (size_t)op->n * 4, hipMemcpyHostToDevice));
    }
    return SIPG_OK;
}

int sipg_download(sipg_op *op, const void *dev, double *host) {
    if (!op || !host || !dev) return SIPG_ERR_INVALID;
    SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    if (op->dtype == SIPG_F64) {
        SIPG_HIP_TRY(hipMemcpy(host, dev, (size_t)op->n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(op->n);
        SIPG_HIP_TRY(hipMemcpy(tmp.data(), dev, (size_t)op->n * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < op->n; ++i) host[i] = tmp[i];
    }
    return SIPG_OK;
}

}  // extern "C"

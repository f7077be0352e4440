// abi.cpp -- the C-ABI of libsipg_hip (declared in include/sipg_hip.h).
//
// Host orchestration only: argument checks, table set-up, workspace, and
// the device-resident CG loop.  All arithmetic on vectors runs in the HIP
// kernels of kernels_apply.hip / kernels_vec.hip; there is no CPU fallback.
#include "sipg_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

using namespace sipg_internal;

namespace {
thread_local hipError_t g_last_hip = hipSuccess;
}  // namespace

int sipg_internal::ensure_ws(sipg_op *op, size_t bytes) {
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

sipg_internal::parts sipg_internal::part_layout(sipg_op *op) {
    parts p;
    p.napply = apply_num_blocks(op->dtype, op->p, op->dimx, op->dimy, op->kernel);
    p.pAp = op->d_part;
    p.rr = op->d_part + std::max(p.napply, kMaxPartials);
    p.bb = p.rr + kMaxPartials;
    p.misc = p.bb + kMaxPartials;
    p.rz = p.misc + kMaxPartials;
    return p;
}

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

int sipg_max_degree(void) { return kMaxGenericDegree; }
int sipg_max_fused_degree(void) { return kMaxDegree; }

int sipg_basis_tables(int p, double *nodes, double *weights, double *D, double *K1) {
    if (p < 1 || p > 63) return SIPG_ERR_INVALID;
    const int n1 = p + 1;
    std::vector<double> z(n1), w(n1), Dm(n1 * n1), K(n1 * n1);
    gll(p, z.data(), w.data());
    dphi(p, z.data(), Dm.data());
    stiffness_1d(p, w.data(), Dm.data(), K.data());
    if (nodes) std::memcpy(nodes, z.data(), n1 * sizeof(double));
    if (weights) std::memcpy(weights, w.data(), n1 * sizeof(double));
    if (D) std::memcpy(D, Dm.data(), n1 * n1 * sizeof(double));
    if (K1) std::memcpy(K1, K.data(), n1 * n1 * sizeof(double));
    return SIPG_OK;
}

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
    if (p > kMaxGenericDegree) return SIPG_ERR_UNSUPPORTED;
    const int64_t noe = (int64_t)dimx * dimy;
    if (noe > (int64_t)INT_MAX) return SIPG_ERR_INVALID;
    // the fused operator kernels address a block's footprint (up to 6
    // element rows) with 32-bit offsets
    if (p <= kMaxDegree && (int64_t)6 * dimx * (p + 1) * (p + 1) > (int64_t)INT_MAX) return SIPG_ERR_UNSUPPORTED;
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
    // above the compiled degrees the operator's face buffer follows the
    // tables in the op's dtype allocation
    const bool generic = p > kMaxDegree;
    const size_t tl64 = generic && dtype == SIPG_F64 ? generic_table_len(SIPG_F64, p, dimx, dimy) : (size_t)tl;
    const size_t tl32 = generic ? generic_table_len(SIPG_F32, p, dimx, dimy) : (size_t)tl;
    int rc = SIPG_OK;
    auto fail = [&](int code) { sipg_op_destroy(op); return code; };
    hipError_t e;
    if ((e = hipMalloc(&op->d_tab64, tl64 * sizeof(double))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
    if ((e = hipMemcpy(op->d_tab64, op->h_tab.data(), tl * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_HIP); }
    if (dtype == SIPG_F64) {
        op->d_tab = op->d_tab64;
    } else {
        std::vector<float> tf(op->h_tab.begin(), op->h_tab.end());
        if ((e = hipMalloc(&op->d_tab, tl32 * sizeof(float))) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_NOMEM); }
        if ((e = hipMemcpy(op->d_tab, tf.data(), tl * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) { set_last_hip_error(e); return fail(SIPG_ERR_HIP); }
    }
    // room for the block sums of any contraction kernel (sipg_op_set_kernel)
    const int napply = std::max({apply_num_blocks(dtype, p, dimx, dimy, SIPG_KERNEL_VALU),
                                 apply_num_blocks(dtype, p, dimx, dimy, SIPG_KERNEL_MFMA),
                                 apply_num_blocks(dtype, p, dimx, dimy, SIPG_KERNEL_WAVE)});
    if (napply < 1) return fail(SIPG_ERR_UNSUPPORTED);
    const size_t npart = (size_t)std::max(napply, kMaxPartials) + 4 * (size_t)kMaxPartials;
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
    if (op->d_prec) (void)hipFree(op->d_prec);
    if (op->d_coarse) (void)hipFree(op->d_coarse);
    if (op->twin) sipg_op_destroy(op->twin);
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
    return sipg_basis_tables(op->p, nodes, weights, D, K1);
}

int sipg_apply(sipg_op *op, const void *x, void *y) {
    if (!op || !x || !y) return SIPG_ERR_INVALID;
    if (overlaps(x, y, (size_t)op->n * dtype_size(op->dtype))) return SIPG_ERR_ALIAS;
    return launch_apply(op->dtype, op->p, x, y, op->d_tab, op->dimx, op->dimy, op->pen, nullptr,
                        nullptr, 0, op->stream, op->kernel);
}

int sipg_op_set_kernel(sipg_op *op, int kernel) {
    if (!op || kernel < SIPG_KERNEL_AUTO || kernel > SIPG_KERNEL_WAVE)
        return SIPG_ERR_INVALID;
    op->kernel = kernel;
    return SIPG_OK;
}

int sipg_op_kernel(const sipg_op *op, int dtype, int *kernel) {
    if (!op || !kernel || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    *kernel = resolve_kernel(dtype, op->p, op->kernel);
    return SIPG_OK;
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

int sipg_vec_dot(int dtype, int64_t n, const void *a, const void *b, double *result, void *stream) {
    if (!a || !b || !result || n < 0 || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    const int nb = vec_num_blocks(n);
    double *tmp = nullptr;
    SIPG_HIP_TRY(hipMallocAsync((void **)&tmp, (nb + 1) * sizeof(double), s));
    int rc = launch_dot_partials(dtype, a, b, n, tmp + 1, nb, s);
    if (!rc) rc = launch_reduce_partials(tmp + 1, nb, tmp, s);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(result, tmp, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_last_hip_error(e); rc = SIPG_ERR_HIP; }
    }
    (void)hipFreeAsync(tmp, s);
    return rc;
}

int sipg_vec_axpy(int dtype, int64_t n, double alpha, const void *x, void *y, void *stream) {
    if (!x || !y || n < 0 || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    return launch_axpy(dtype, alpha, x, y, n, (hipStream_t)stream);
}

int sipg_vec_copy(int dtype, int64_t n, const void *x, void *y, void *stream) {
    if (!x || !y || n < 0 || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    if (overlaps(x, y, (size_t)n * dtype_size(dtype))) return SIPG_ERR_ALIAS;
    return launch_copy(dtype, x, y, n, (hipStream_t)stream);
}

int sipg_vec_xpby(int dtype, int64_t n, const void *r, double beta, void *p, void *stream) {
    if (!r || !p || n < 0 || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    return launch_xpby(dtype, r, beta, p, n, (hipStream_t)stream);
}

int sipg_vec_alpha_update(int dtype, int64_t n, double alpha, void *x, const void *p, void *r,
                          const void *Ap, void *stream) {
    if (!x || !p || !r || !Ap || n < 0 || (dtype != SIPG_F64 && dtype != SIPG_F32)) return SIPG_ERR_INVALID;
    return launch_alpha_update(dtype, alpha, x, p, r, Ap, n, (hipStream_t)stream);
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

int sipg_err_norms_fourier(sipg_op *op, const void *x, int nmodes, const double *a, const int *m, const int *n,
                           double *out) {
    if (!op || !x || !out || nmodes < 0 || nmodes > 64 || (nmodes > 0 && (!a || !m || !n)))
        return SIPG_ERR_INVALID;
    if (op->dimx != op->dimy) return SIPG_ERR_INVALID;
    const int p = op->p, n1 = p + 1;
    std::vector<double> z(n1), w(n1), D((size_t)n1 * n1);
    gll(p, z.data(), w.data());
    dphi(p, z.data(), D.data());
    // modes in the scalar scratch (as sipg_rhs_fourier); D in the head of
    // the p.Ap partials, the block partials in the four regions after it
    const parts P = part_layout(op);
    double *da = op->d_scalar + 64;
    int *dm = (int *)(op->d_scalar + 128);
    int *dn = dm + 64;
    if (nmodes > 0) {
        SIPG_HIP_TRY(hipMemcpyAsync(da, a, nmodes * sizeof(double), hipMemcpyHostToDevice, op->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(dm, m, nmodes * sizeof(int), hipMemcpyHostToDevice, op->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(dn, n, nmodes * sizeof(int), hipMemcpyHostToDevice, op->stream));
    }
    SIPG_HIP_TRY(hipMemcpyAsync(P.pAp, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice, op->stream));
    const int nb = norms_num_blocks(op);
    int rc = launch_err_norms(op->dtype, p, op->dimx, (const double *)op->d_tab64, P.pAp, nmodes, da, dm, dn, x,
                              P.rr, nb, op->stream);
    if (rc) return rc;
    return norms_finish(op, nb, out);
}

}  // extern "C"

int sipg_internal::norms_num_blocks(sipg_op *op) {
    return (int)std::min<int64_t>(kMaxPartials, (op->n + kVecThreads - 1) / kVecThreads);
}

int sipg_internal::norms_finish(sipg_op *op, int nb, double *out) {
    const parts P = part_layout(op);
    std::vector<double> h((size_t)4 * nb);
    SIPG_HIP_TRY(hipMemcpyAsync(h.data(), P.rr, h.size() * sizeof(double), hipMemcpyDeviceToHost, op->stream));
    SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    double emax = -1.0, l2 = 0.0, sx = 0.0, sy = 0.0;
    for (int b = 0; b < nb; ++b) {
        emax = std::max(emax, h[b]);
        l2 += h[(size_t)nb + b];
        sx += h[(size_t)2 * nb + b];
        sy += h[(size_t)3 * nb + b];
    }
    out[0] = emax;
    out[1] = std::sqrt(l2);
    out[2] = std::sqrt(l2 + sx + sy);
    return SIPG_OK;
}

extern "C" {

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
                          nullptr, 0, op->stream, op->kernel);
    SIPG_HIP_TRY(hipEventRecord(b, op->stream));
    SIPG_HIP_TRY(hipEventSynchronize(b));
    float ms = 0.f;
    SIPG_HIP_TRY(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_apply = (double)ms / reps;
    return rc;
}

// Host data in the reference order ((i1*(p+1)+i2)*noe + e) <-> the
// element-major device order (e*(p+1)^2 + i1*(p+1)+i2).
int sipg_upload(sipg_op *op, const double *host, void *dev) {
    if (!op || !host || !dev) return SIPG_ERR_INVALID;
    const int64_t noe = (int64_t)op->dimx * op->dimy, nq = (int64_t)(op->p + 1) * (op->p + 1);
    if (op->dtype == SIPG_F64) {
        std::vector<double> tmp(op->n);
        for (int64_t e = 0; e < noe; ++e)
            for (int64_t q = 0; q < nq; ++q) tmp[e * nq + q] = host[q * noe + e];
        SIPG_HIP_TRY(hipMemcpy(dev, tmp.data(), (size_t)op->n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(op->n);
        for (int64_t e = 0; e < noe; ++e)
            for (int64_t q = 0; q < nq; ++q) tmp[e * nq + q] = (float)host[q * noe + e];
        SIPG_HIP_TRY(hipMemcpy(dev, tmp.data(), (size_t)op->n * 4, hipMemcpyHostToDevice));
    }
    return SIPG_OK;
}

int sipg_download(sipg_op *op, const void *dev, double *host) {
    if (!op || !host || !dev) return SIPG_ERR_INVALID;
    SIPG_HIP_TRY(hipStreamSynchronize(op->stream));
    const int64_t noe = (int64_t)op->dimx * op->dimy, nq = (int64_t)(op->p + 1) * (op->p + 1);
    if (op->dtype == SIPG_F64) {
        std::vector<double> tmp(op->n);
        SIPG_HIP_TRY(hipMemcpy(tmp.data(), dev, (size_t)op->n * 8, hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < noe; ++e)
            for (int64_t q = 0; q < nq; ++q) host[q * noe + e] = tmp[e * nq + q];
    } else {
        std::vector<float> tmp(op->n);
        SIPG_HIP_TRY(hipMemcpy(tmp.data(), dev, (size_t)op->n * 4, hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < noe; ++e)
            for (int64_t q = 0; q < nq; ++q) host[q * noe + e] = tmp[e * nq + q];
    }
    return SIPG_OK;
}

int sipg_write_result(sipg_op *op, const void *x, const char *path) {
    if (!op || !x) return SIPG_ERR_INVALID;
    if (op->dimx != op->dimy) return SIPG_ERR_INVALID;
    std::vector<double> u(op->n);
    int rc = sipg_download(op, x, u.data());
    if (rc) return rc;
    const int dim = op->dimx, p = op->p, n1 = p + 1;
    const int64_t noe = (int64_t)dim * dim;
    std::vector<double> z(n1), w(n1), nodes(n1);
    gll(p, z.data(), w.data());
    for (int i = 0; i < n1; ++i) nodes[i] = (z[i] * .5 + .5) / dim;
    FILE *f = (path && path[0]) ? std::fopen(path, "w") : stderr;
    if (!f) return SIPG_ERR_INVALID;
    // element corners as the mesh builder places them, so the printed
    // coordinates carry the same rounding
    std::vector<double> xe, ye;
    mesh_corners(dim, xe, ye);
    bool ok = true;
    for (int64_t e = 0; e < noe && ok; ++e) {
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n1; ++j)
                ok = ok && std::fprintf(f, "%10g\t%10g\t%10g\n", xe[e] + nodes[i], ye[e] + nodes[j],
                                        u[((int64_t)i * n1 + j) * noe + e]) > 0;
        ok = ok && std::fputc('\n', f) != EOF;
    }
    if (f != stderr) ok = (std::fclose(f) == 0) && ok;
    return ok ? SIPG_OK : SIPG_ERR_INVALID;
}

int sipg_to_device_layout(sipg_op *op, const void *ref, void *dev) {
    if (!op || !ref || !dev) return SIPG_ERR_INVALID;
    if (overlaps(ref, dev, (size_t)op->n * dtype_size(op->dtype))) return SIPG_ERR_ALIAS;
    return launch_layout(op->dtype, 1, ref, dev, (int64_t)op->dimx * op->dimy, (op->p + 1) * (op->p + 1),
                         op->stream);
}

int sipg_to_reference_layout(sipg_op *op, const void *dev, void *ref) {
    if (!op || !ref || !dev) return SIPG_ERR_INVALID;
    if (overlaps(ref, dev, (size_t)op->n * dtype_size(op->dtype))) return SIPG_ERR_ALIAS;
    return launch_layout(op->dtype, 0, dev, ref, (int64_t)op->dimx * op->dimy, (op->p + 1) * (op->p + 1),
                         op->stream);
}

}  // extern "C"

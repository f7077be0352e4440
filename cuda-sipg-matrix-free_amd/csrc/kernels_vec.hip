// kernels_vec.hip -- CG vector kernels with fused, deterministic reductions.
//
// Reference counterparts: alpha_kernel / beta_kernel
// (conjugate_gradient_kernels.hpp:5-38) and cublas_dot
// (cublas_wrapper.hpp:62-84, called at conjugate_gradient_gpu.hpp:42,60,75,85).
// Differences by design:
//   * the dots are fused into the kernels that already stream the vectors
//     (r.r into the x/r update, b.b and r.r into the initial residual,
//     p.Ap into the operator apply), accumulated in double;
//   * every reduction is two-stage with a fixed order (per-thread sums in
//     index order, a wave butterfly, waves summed in order, block sums in
//     block order), so results are bitwise reproducible run to run;
//   * the CG scalars (alpha, beta, r.r, the stop flag) never leave the
//     device: each consumer kernel re-reduces the previous kernel's block
//     sums itself, and block 0 records them in a cg_state.
#include "sipg_internal.hpp"
#include "partials.hpp"

#include <cmath>

namespace sipg_internal {

namespace {

constexpr int NT = kVecThreads;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block; the result is returned to every thread.
__device__ __forceinline__ double block_sum_all(double v) {
    __shared__ double red[NT / 64];
    __shared__ double total;
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();   // protect red/total from a previous call
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s += red[i];
        total = s;
    }
    __syncthreads();
    return total;
}

// Fixed-order sum of np block partials, identical in every block.
__device__ __forceinline__ double reduce_partials_dev(const double *__restrict__ part, int np) {
    return block_sum_all(strided_sum<NT>(part, 0, np));
}

template <typename T>
__global__ __launch_bounds__(NT) void dot_kernel(const T *__restrict__ a, const T *__restrict__ b,
                                                 int64_t n, double *__restrict__ part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        s += (double)a[i] * (double)b[i];
    s = block_sum_all(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(NT) void reduce_kernel(const double *__restrict__ part, int np,
                                                    double *__restrict__ out) {
    const double s = reduce_partials_dev(part, np);
    if (threadIdx.x == 0) out[0] = s;
}

template <typename T>
__global__ __launch_bounds__(NT) void axpy_kernel(T alpha, const T *__restrict__ x, T *__restrict__ y,
                                                  int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        y[i] = y[i] + alpha * x[i];
}

// device copy of an existing vector (copy(), mode_vector.hpp:314-359,
// without its reallocation)
template <typename T>
__global__ __launch_bounds__(NT) void copy_kernel(const T *__restrict__ x, T *__restrict__ y, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        y[i] = x[i];
}

// beta_kernel (conjugate_gradient_kernels.hpp:26-38): p = r + beta p
template <typename T>
__global__ __launch_bounds__(NT) void xpby_kernel(const T *__restrict__ r, T beta, T *__restrict__ p,
                                                  int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        p[i] = r[i] + beta * p[i];
}

// alpha_kernel (conjugate_gradient_kernels.hpp:5-23)
template <typename T>
__global__ __launch_bounds__(NT) void alpha_kernel(T alpha, T *__restrict__ x, const T *__restrict__ p,
                                                   T *__restrict__ r, const T *__restrict__ Ap,
                                                   int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        x[i] = x[i] + alpha * p[i];
        r[i] = r[i] - alpha * Ap[i];
    }
}

// Reads the initial r.r and b.b block sums (from true_residual_kernel, the
// r = b - A x, p = r step of conjugate_gradient_gpu.hpp:40-60) into slot 0.
__global__ __launch_bounds__(NT) void cg_init_finalize_kernel(const double *__restrict__ prr,
                                                              const double *__restrict__ pbb,
                                                              int np, cg_state *st, double tol,
                                                              int tol_mode, int maxit, int true_stop) {
    const double rr = reduce_partials_dev(prr, np);
    const double bb = reduce_partials_dev(pbb, np);
    if (threadIdx.x == 0) {
        const double stop = tol_mode == SIPG_TOL_REL ? tol * sqrt(bb) : tol;
        cg_slot s0;
        s0.rr = rr;
        s0.alpha = s0.beta = s0.pAp = 0.0;
        s0.rz = rr;
        s0.it = 0;
        s0.done = (sqrt(rr) > stop && maxit > 0) ? 0 : 1;   // while-condition of :69
        s0.checked = 1;                                     // r is the true residual here
        s0.flags = 0;
        st->s[0] = s0;
        st->s[1] = s0;
        st->rr0 = rr;
        st->stop_true = stop;
        st->rr_ref = rr;
        st->bb = bb;
        st->stop = stop;
        st->maxit = maxit;
        st->nrep = 0;
        st->err = (!isfinite(rr) || !isfinite(bb)) ? 3 : 0;
        st->true_stop = true_stop;
    }
}

// x += alpha p; r -= alpha Ap with alpha = r.r / p.Ap from the apply's
// block sums (alpha_kernel + the host-side alpha of
// conjugate_gradient_gpu.hpp:75-83); writes block sums of the new r.r.
template <typename T>
__global__ __launch_bounds__(NT) void cg_update_xr_kernel(T *__restrict__ x, const T *__restrict__ p,
                                                          T *__restrict__ r, const T *__restrict__ Ap,
                                                          int64_t n, const double *__restrict__ ppAp,
                                                          int np, double *__restrict__ prr,
                                                          cg_state *st, int slot, T *__restrict__ xlo) {
    if (st->s[slot].done | st->err) return;
    const double pAp = reduce_partials_dev(ppAp, np);
    const double rr = st->s[slot].rr;
    const double alpha_d = rr / pAp;
    if (!(pAp > 0.0) || !isfinite(alpha_d)) {
        // every block takes this branch (same pAp); block 0 records it in a
        // word no kernel of this iteration reads at entry
        if (blockIdx.x == 0 && threadIdx.x == 0) st->err = isfinite(pAp) ? 2 : 3;
        return;
    }
    const T alpha = (T)alpha_d;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        if (xlo) {
            T xh = x[i], xl = xlo[i];
            comp_axpy(xh, xl, alpha, p[i]);
            x[i] = xh;
            xlo[i] = xl;
        } else {
            x[i] = x[i] + alpha * p[i];
        }
        const T ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        s += (double)ri * (double)ri;
    }
    s = block_sum_all(s);
    if (threadIdx.x == 0) prr[blockIdx.x] = s;
}

// p = r + beta p, beta = r.r(new) / r.r(old) (beta_kernel and
// conjugate_gradient_gpu.hpp:85-95); block 0 writes the next slot.  When r
// was replaced by the true residual in this iteration (kSlotReplaced on the
// slot read), r.r(new) is the true residual's: the reliable update of van
// der Vorst & Ye keeps the direction, p = r_true + beta p, unless
// kSlotRestart asks for p = r.  With st->true_stop the iteration finishes
// only on a true residual below the tolerance; a recursive one raises
// kSlotWantCheck instead and the iteration goes on until the host inserts
// the true residual.
template <typename T>
__global__ __launch_bounds__(NT) void cg_update_p_kernel(const T *__restrict__ r, T *__restrict__ p,
                                                         int64_t n, const double *__restrict__ prr,
                                                         int np, cg_state *st, int slot) {
    const cg_slot cur = st->s[slot];
    if (cur.done | st->err) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->s[slot ^ 1] = cur;   // carry forward
        return;
    }
    const double rr_new = reduce_partials_dev(prr, np);
    const bool replaced = (cur.flags & kSlotReplaced) != 0;
    const double beta_d = (cur.flags & kSlotRestart) ? 0.0 : rr_new / cur.rr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool met = !(sqrt(rr_new) > st->stop);
        const bool need_true = st->true_stop && !replaced;
        cg_slot nx;
        nx.rr = rr_new;
        nx.alpha = 0.0;
        nx.beta = beta_d;
        nx.pAp = 0.0;
        nx.rz = rr_new;
        nx.it = cur.it + 1;
        nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
        nx.checked = replaced ? 1 : 0;
        nx.flags = (met && need_true) ? kSlotWantCheck : 0;
        st->s[slot ^ 1] = nx;
        if (replaced) st->rr_ref = rr_new;
        if (!isfinite(rr_new)) st->err = 3;
    }
    const T beta = (T)beta_d;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        p[i] = r[i] + beta * p[i];
}

// out = x + xl (+ its compensation xlo): the iterate that folding the
// pending increments into x would give (the same rounded sums as
// accumulate_kernel), without folding.
template <typename T>
__global__ __launch_bounds__(NT) void sum_kernel(const T *__restrict__ x, const T *__restrict__ xl,
                                                 const T *__restrict__ xlo, T *__restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        out[i] = x[i] + (xlo ? xl[i] + xlo[i] : xl[i]);
}

// A true residual that leaves r alone (prr: the block sums of its r.r),
// between two iterations: when it meets the tolerance, the slot the next
// iteration reads is finished on it (true residual, checked); otherwise
// nothing changes.  The single-problem solvers' threshold (stop_true = stop).
__global__ __launch_bounds__(NT) void peek_finalize_kernel(const double *__restrict__ prr, int np, cg_state *st,
                                                           int slot) {
    const double rr = reduce_partials_dev(prr, np);
    if (threadIdx.x != 0) return;
    cg_slot c = st->s[slot];
    if (c.done | st->err) return;
    if (!isfinite(rr) || sqrt(rr) > st->stop_true) return;
    c.rr = rr;
    c.done = 1;
    c.checked = 1;
    c.flags = 0;
    st->s[slot] = c;
    st->rr_ref = rr;
}

// Runs between an iteration's x/r update and its p update, after r has
// been overwritten with b - A x and the rr block sums with its own.
__global__ void mark_replaced_kernel(cg_state *st, int slot, int restart) {
    if (threadIdx.x != 0) return;
    cg_slot c = st->s[slot];
    if (c.done | st->err) return;
    c.flags |= kSlotReplaced | (restart ? kSlotRestart : 0);
    st->s[slot] = c;
    st->nrep += 1;
}

// out = (TO)(b - Ax); optional p = out; block sums of (b-Ax)^2 and b^2.
template <typename TI, typename TO>
__global__ __launch_bounds__(NT) void true_residual_kernel(const TI *__restrict__ b,
                                                           const TI *__restrict__ Ax,
                                                           TO *__restrict__ out, TO *__restrict__ pc,
                                                           int64_t n, double *__restrict__ prr,
                                                           double *__restrict__ pbb) {
    double srr = 0.0, sbb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        const TI bi_ = b[i];
        const TI ri_ = bi_ - Ax[i];
        const double bi = (double)bi_;
        const double ri = (double)ri_;
        const TO ro = (TO)ri;
        if (out) out[i] = ro;
        if (pc) pc[i] = ro;
        srr += ri * ri;
        sbb += bi * bi;
    }
    srr = block_sum_all(srr);
    if (pbb) sbb = block_sum_all(sbb);
    if (threadIdx.x == 0) {
        prr[blockIdx.x] = srr;
        if (pbb) pbb[blockIdx.x] = sbb;
    }
}

template <typename TO, typename TL>
__global__ __launch_bounds__(NT) void accumulate_kernel(TO *__restrict__ x, TL *__restrict__ xl,
                                                        TL *__restrict__ xlo, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        x[i] += (TO)(xlo ? xl[i] + xlo[i] : xl[i]);
        xl[i] = TL(0);
        if (xlo) xlo[i] = TL(0);
    }
}

// x += xi exactly as x + xlo (the sum's rounding error joins xlo), xi = 0
template <typename T>
__global__ __launch_bounds__(NT) void comp_fold_kernel(T *__restrict__ x, T *__restrict__ xlo, T *__restrict__ xi,
                                                       int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        T xh = x[i], xl = xlo[i];
        comp_axpy(xh, xl, T(1), xi[i]);
        x[i] = xh;
        xlo[i] = xl;
        xi[i] = T(0);
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void residual_kernel(const T *__restrict__ b, const T *__restrict__ Ax,
                                                      T *__restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        r[i] = b[i] - Ax[i];
}

template <typename S, typename D>
__global__ __launch_bounds__(NT) void convert_kernel(const S *__restrict__ src, D *__restrict__ dst,
                                                     int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
        dst[i] = (D)src[i];
}

// Layout conversion between the reference order ((i1*(p+1)+i2)*noe + e) and
// the element-major device order (e*(p+1)^2 + i1*(p+1)+i2).
template <typename T>
__global__ __launch_bounds__(NT) void to_device_layout_kernel(const T *__restrict__ ref, T *__restrict__ dev,
                                                              int64_t noe, int nq) {
    const int64_t n = noe * nq;
    for (int64_t idx = (int64_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * NT) {
        const int64_t e = idx / nq;
        const int q = (int)(idx - e * nq);
        dev[idx] = ref[(int64_t)q * noe + e];
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void to_reference_layout_kernel(const T *__restrict__ dev, T *__restrict__ ref,
                                                                 int64_t noe, int nq) {
    const int64_t n = noe * nq;
    for (int64_t idx = (int64_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * NT) {
        const int q = (int)(idx / noe);
        const int64_t e = idx - (int64_t)q * noe;
        ref[idx] = dev[e * nq + q];
    }
}

// Load vector of the Fourier problem with g = 0 (sipg_sem_2d.hpp:248-262):
// b(i1,i2,e) = (w_i1 h/2)(w_i2 h/2) f(x_e + h(z_i1+1)/2, y_e + h(z_i2+1)/2),
// x_e = col*h, y_e = 1 - (row+1)*h.
template <typename T>
__global__ __launch_bounds__(NT) void rhs_fourier_kernel(int p, int dim, const double *__restrict__ tab,
                                                         int nmodes, const double *__restrict__ am,
                                                         const int *__restrict__ mm,
                                                         const int *__restrict__ nn, T *__restrict__ b) {
    const int n1 = p + 1;
    const int64_t noe = (int64_t)dim * dim;
    const int64_t n = noe * n1 * n1;
    const double *w = tab;
    const double *z = tab + 3 * n1 + n1 * n1;
    const double h = 1.0 / dim;
    const int nq = n1 * n1;
    for (int64_t idx = (int64_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * NT) {
        // element-major device layout: idx = e * (p+1)^2 + (i1*(p+1) + i2)
        const int64_t e = idx / nq;
        const int node = (int)(idx - e * nq);
        const int i1 = node / n1, i2 = node % n1;
        const int col = (int)(e % dim), row = (int)(e / dim);
        const double xe = col * h, ye = 1.0 - (row + 1) * h;
        const double X = xe + (z[i1] * .5 + .5) / dim;
        const double Y = ye + (z[i2] * .5 + .5) / dim;
        double f = 0.0;
        for (int k = 0; k < nmodes; ++k) {
            const double m = mm[k], q = nn[k];
            f += am[k] * M_PI * M_PI * (m * m + q * q) * sin(m * M_PI * X) * sin(q * M_PI * Y);
        }
        b[idx] = (T)((w[i1] * .5 / dim) * (w[i2] * .5 / dim) * f);
    }
}

// The exact solution at a node for the error norms: the Fourier solution
// u_ex = sum_k a_k sin(m_k pi x) sin(n_k pi y) evaluated in place, or nodal
// samples of u_ex and its two derivatives (element-major, one per node).
struct FourierExact {
    int nmodes;
    const double *am;
    const int *mm, *nn;
    __device__ void at(int64_t, double X, double Y, double &u, double &dx, double &dy) const {
        u = dx = dy = 0.0;
        for (int k = 0; k < nmodes; ++k) {
            const double m = mm[k] * M_PI, q = nn[k] * M_PI;
            const double sxm = sin(m * X), cxm = cos(m * X), syq = sin(q * Y), cyq = cos(q * Y);
            u += am[k] * sxm * syq;
            dx += am[k] * m * cxm * syq;
            dy += am[k] * q * sxm * cyq;
        }
    }
};
struct NodalExact {
    const double *u, *ux, *uy;
    __device__ void at(int64_t idx, double, double, double &a, double &dx, double &dy) const {
        a = u[idx];
        dx = ux[idx];
        dy = uy[idx];
    }
};

// Error norms of u against an exact solution (sipg_sem_2d::err_norms,
// sipg_sem_2d.hpp:311-370), one thread per GLL node: the nodal error, and
// the x and y derivatives of u at the node from the element's line through
// it, u_x = 2 dim sum_i u(i,q) D_i(p) (the basis is nodal, so phi_j(z_q) =
// [j = q]), weighted by the GLL quadrature weights (w_p/2dim)(w_q/2dim).
// Block partials: max |e|, sum w e^2, sum w (u_x - u_ex,x)^2, sum w (u_y -
// u_ex,y)^2.  D[i*(p+1)+k] = phi_i'(z_k).
template <typename T, class Exact>
__global__ __launch_bounds__(NT) void err_norms_kernel(int p, int dim, const double *__restrict__ tab,
                                                       const double *__restrict__ D, Exact exact,
                                                       const T *__restrict__ u, double *__restrict__ part) {
    const int n1 = p + 1, nq = n1 * n1;
    const int64_t noe = (int64_t)dim * dim, n = noe * nq;
    const double *w = tab;
    const double *z = tab + 3 * n1 + n1 * n1;
    const double h = 1.0 / dim, two_dim = 2.0 * dim;
    double emax = -1.0, l2 = 0.0, sx = 0.0, sy = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * NT) {
        const int64_t e = idx / nq;
        const int node = (int)(idx - e * nq);
        const int i1 = node / n1, i2 = node - (node / n1) * n1;
        const T *ue = u + e * nq;
        double dxu = 0.0, dyu = 0.0;
        for (int i = 0; i < n1; ++i) {
            dxu += (double)ue[i * n1 + i2] * two_dim * D[i * n1 + i1];
            dyu += (double)ue[i1 * n1 + i] * two_dim * D[i * n1 + i2];
        }
        const int col = (int)(e % dim), row = (int)(e / dim);
        const double X = col * h + (z[i1] * .5 + .5) / dim;
        const double Y = 1.0 - (row + 1) * h + (z[i2] * .5 + .5) / dim;
        double ux, dx, dy;
        exact.at(idx, X, Y, ux, dx, dy);
        const double wgt = (w[i1] * .5 / dim) * (w[i2] * .5 / dim);
        const double err = fabs((double)ue[node] - ux);
        emax = fmax(emax, err);
        l2 += wgt * err * err;
        sx += wgt * (dxu - dx) * (dxu - dx);
        sy += wgt * (dyu - dy) * (dyu - dy);
    }
    __shared__ double red[4][NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) emax = fmax(emax, __shfl_xor(emax, off, 64));
    l2 = wave_sum(l2);
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) {
        red[0][wid] = emax;
        red[1][wid] = l2;
        red[2][wid] = sx;
        red[3][wid] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[4] = {-1.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < NT / 64; ++i) {
            r[0] = fmax(r[0], red[0][i]);
            r[1] += red[1][i];
            r[2] += red[2][i];
            r[3] += red[3][i];
        }
        for (int k = 0; k < 4; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = r[k];
    }
}

}  // namespace

int vec_num_blocks(int64_t n) {
    int64_t nb = (n + NT * 4 - 1) / (NT * 4);
    if (nb < 1) nb = 1;
    if (nb > 1024) nb = 1024;
    return (int)nb;
}

#define DISPATCH_T(dtype, KERNEL_CALL_F64, KERNEL_CALL_F32) \
    do {                                                   \
        if ((dtype) == SIPG_F64) { KERNEL_CALL_F64; }      \
        else { KERNEL_CALL_F32; }                          \
    } while (0)

int launch_dot_partials(int dtype, const void *a, const void *b, int64_t n, double *partials,
                        int nblk, hipStream_t s) {
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(dot_kernel<double>, dim3(nblk), dim3(NT), 0, s, (const double *)a, (const double *)b, n, partials),
        hipLaunchKernelGGL(dot_kernel<float>, dim3(nblk), dim3(NT), 0, s, (const float *)a, (const float *)b, n, partials));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_reduce_partials(const double *partials, int np, double *out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(NT), 0, s, partials, np, out);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_copy(int dtype, const void *x, void *y, int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(copy_kernel<double>, dim3(nb), dim3(NT), 0, s, (const double *)x, (double *)y, n),
        hipLaunchKernelGGL(copy_kernel<float>, dim3(nb), dim3(NT), 0, s, (const float *)x, (float *)y, n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_axpy(int dtype, double alpha, const void *x, void *y, int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(axpy_kernel<double>, dim3(nb), dim3(NT), 0, s, alpha, (const double *)x, (double *)y, n),
        hipLaunchKernelGGL(axpy_kernel<float>, dim3(nb), dim3(NT), 0, s, (float)alpha, (const float *)x, (float *)y, n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_xpby(int dtype, const void *r, double beta, void *p, int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(xpby_kernel<double>, dim3(nb), dim3(NT), 0, s, (const double *)r, beta, (double *)p, n),
        hipLaunchKernelGGL(xpby_kernel<float>, dim3(nb), dim3(NT), 0, s, (const float *)r, (float)beta, (float *)p, n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_alpha_update(int dtype, double alpha, void *x, const void *p, void *r, const void *Ap,
                        int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(alpha_kernel<double>, dim3(nb), dim3(NT), 0, s, alpha, (double *)x, (const double *)p, (double *)r, (const double *)Ap, n),
        hipLaunchKernelGGL(alpha_kernel<float>, dim3(nb), dim3(NT), 0, s, (float)alpha, (float *)x, (const float *)p, (float *)r, (const float *)Ap, n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_cg_init_finalize(const double *part_rr, const double *part_bb, int np, cg_state *st,
                            double tol, int tol_mode, int maxit, int true_stop, hipStream_t s) {
    hipLaunchKernelGGL(cg_init_finalize_kernel, dim3(1), dim3(NT), 0, s, part_rr, part_bb, np, st, tol,
                       tol_mode, maxit, true_stop);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_cg_update_xr(int dtype, void *x, const void *p, void *r, const void *Ap, int64_t n,
                        const double *part_pAp, int np_pAp, double *part_rr, int nblk, cg_state *st,
                        int slot, hipStream_t s, void *xlo) {
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(cg_update_xr_kernel<double>, dim3(nblk), dim3(NT), 0, s, (double *)x, (const double *)p, (double *)r, (const double *)Ap, n, part_pAp, np_pAp, part_rr, st, slot, (double *)xlo),
        hipLaunchKernelGGL(cg_update_xr_kernel<float>, dim3(nblk), dim3(NT), 0, s, (float *)x, (const float *)p, (float *)r, (const float *)Ap, n, part_pAp, np_pAp, part_rr, st, slot, (float *)xlo));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_cg_update_p(int dtype, const void *r, void *p, int64_t n, const double *part_rr, int np_rr,
                       int nblk, cg_state *st, int slot, hipStream_t s) {
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(cg_update_p_kernel<double>, dim3(nblk), dim3(NT), 0, s, (const double *)r, (double *)p, n, part_rr, np_rr, st, slot),
        hipLaunchKernelGGL(cg_update_p_kernel<float>, dim3(nblk), dim3(NT), 0, s, (const float *)r, (float *)p, n, part_rr, np_rr, st, slot));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_residual(int dtype, const void *b, const void *Ax, void *r, int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(residual_kernel<double>, dim3(nb), dim3(NT), 0, s, (const double *)b, (const double *)Ax, (double *)r, n),
        hipLaunchKernelGGL(residual_kernel<float>, dim3(nb), dim3(NT), 0, s, (const float *)b, (const float *)Ax, (float *)r, n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_true_residual(int in_dtype, int out_dtype, const void *b, const void *Ax, void *out,
                         void *p_copy, int64_t n, double *prr, double *pbb, int nblk, hipStream_t s) {
    if (in_dtype == SIPG_F64 && out_dtype == SIPG_F64)
        hipLaunchKernelGGL((true_residual_kernel<double, double>), dim3(nblk), dim3(NT), 0, s, (const double *)b, (const double *)Ax, (double *)out, (double *)p_copy, n, prr, pbb);
    else if (in_dtype == SIPG_F64 && out_dtype == SIPG_F32)
        hipLaunchKernelGGL((true_residual_kernel<double, float>), dim3(nblk), dim3(NT), 0, s, (const double *)b, (const double *)Ax, (float *)out, (float *)p_copy, n, prr, pbb);
    else if (in_dtype == SIPG_F32 && out_dtype == SIPG_F32)
        hipLaunchKernelGGL((true_residual_kernel<float, float>), dim3(nblk), dim3(NT), 0, s, (const float *)b, (const float *)Ax, (float *)out, (float *)p_copy, n, prr, pbb);
    else
        return SIPG_ERR_INVALID;
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_accumulate(int x_dtype, int xl_dtype, void *x, void *xl, int64_t n, hipStream_t s, void *xlo) {
    const int nb = vec_num_blocks(n);
    if (x_dtype == SIPG_F64 && xl_dtype == SIPG_F64)
        hipLaunchKernelGGL((accumulate_kernel<double, double>), dim3(nb), dim3(NT), 0, s, (double *)x, (double *)xl, (double *)xlo, n);
    else if (x_dtype == SIPG_F64 && xl_dtype == SIPG_F32)
        hipLaunchKernelGGL((accumulate_kernel<double, float>), dim3(nb), dim3(NT), 0, s, (double *)x, (float *)xl, (float *)xlo, n);
    else if (x_dtype == SIPG_F32 && xl_dtype == SIPG_F32)
        hipLaunchKernelGGL((accumulate_kernel<float, float>), dim3(nb), dim3(NT), 0, s, (float *)x, (float *)xl, (float *)xlo, n);
    else
        return SIPG_ERR_INVALID;
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_comp_fold(int dtype, void *x, void *xlo, void *xi, int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(comp_fold_kernel<double>, dim3(nb), dim3(NT), 0, s, (double *)x, (double *)xlo,
                           (double *)xi, n),
        hipLaunchKernelGGL(comp_fold_kernel<float>, dim3(nb), dim3(NT), 0, s, (float *)x, (float *)xlo, (float *)xi,
                           n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_mark_replaced(cg_state *st, int slot, int restart, hipStream_t s) {
    hipLaunchKernelGGL(mark_replaced_kernel, dim3(1), dim3(64), 0, s, st, slot, restart);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_sum(int dtype, const void *x, const void *xl, void *out, int64_t n, hipStream_t s, const void *xlo) {
    const int nb = vec_num_blocks(n);
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(sum_kernel<double>, dim3(nb), dim3(NT), 0, s, (const double *)x, (const double *)xl,
                           (const double *)xlo, (double *)out, n),
        hipLaunchKernelGGL(sum_kernel<float>, dim3(nb), dim3(NT), 0, s, (const float *)x, (const float *)xl,
                           (const float *)xlo, (float *)out, n));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_peek_finalize(const double *prr, int np, cg_state *st, int slot, hipStream_t s) {
    hipLaunchKernelGGL(peek_finalize_kernel, dim3(1), dim3(NT), 0, s, prr, np, st, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_convert(int from_dtype, int to_dtype, const void *src, void *dst, int64_t n, hipStream_t s) {
    const int nb = vec_num_blocks(n);
    if (from_dtype == SIPG_F64 && to_dtype == SIPG_F32)
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nb), dim3(NT), 0, s, (const double *)src, (float *)dst, n);
    else if (from_dtype == SIPG_F32 && to_dtype == SIPG_F64)
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nb), dim3(NT), 0, s, (const float *)src, (double *)dst, n);
    else if (from_dtype == SIPG_F64)
        hipLaunchKernelGGL((convert_kernel<double, double>), dim3(nb), dim3(NT), 0, s, (const double *)src, (double *)dst, n);
    else
        hipLaunchKernelGGL((convert_kernel<float, float>), dim3(nb), dim3(NT), 0, s, (const float *)src, (float *)dst, n);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_layout(int dtype, int to_device, const void *src, void *dst, int64_t noe, int nq, hipStream_t s) {
    int64_t nb = (noe * nq + NT - 1) / NT;
    if (nb > 4096) nb = 4096;
    if (to_device) {
        DISPATCH_T(dtype,
            hipLaunchKernelGGL(to_device_layout_kernel<double>, dim3((int)nb), dim3(NT), 0, s, (const double *)src, (double *)dst, noe, nq),
            hipLaunchKernelGGL(to_device_layout_kernel<float>, dim3((int)nb), dim3(NT), 0, s, (const float *)src, (float *)dst, noe, nq));
    } else {
        DISPATCH_T(dtype,
            hipLaunchKernelGGL(to_reference_layout_kernel<double>, dim3((int)nb), dim3(NT), 0, s, (const double *)src, (double *)dst, noe, nq),
            hipLaunchKernelGGL(to_reference_layout_kernel<float>, dim3((int)nb), dim3(NT), 0, s, (const float *)src, (float *)dst, noe, nq));
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_err_norms(int dtype, int p, int dim, const double *tab_f64, const double *D_dev, int nmodes,
                     const double *a_dev, const int *m_dev, const int *n_dev, const void *u, double *part, int nb,
                     hipStream_t s) {
    const FourierExact ex{nmodes, a_dev, m_dev, n_dev};
    DISPATCH_T(dtype,
        hipLaunchKernelGGL((err_norms_kernel<double, FourierExact>), dim3(nb), dim3(NT), 0, s, p, dim, tab_f64, D_dev, ex, (const double *)u, part),
        hipLaunchKernelGGL((err_norms_kernel<float, FourierExact>), dim3(nb), dim3(NT), 0, s, p, dim, tab_f64, D_dev, ex, (const float *)u, part));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_err_norms_nodal(int dtype, int p, int dim, const double *tab_f64, const double *D_dev,
                           const double *ue, const double *uxe, const double *uye, const void *u, double *part,
                           int nb, hipStream_t s) {
    const NodalExact ex{ue, uxe, uye};
    DISPATCH_T(dtype,
        hipLaunchKernelGGL((err_norms_kernel<double, NodalExact>), dim3(nb), dim3(NT), 0, s, p, dim, tab_f64, D_dev, ex, (const double *)u, part),
        hipLaunchKernelGGL((err_norms_kernel<float, NodalExact>), dim3(nb), dim3(NT), 0, s, p, dim, tab_f64, D_dev, ex, (const float *)u, part));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_rhs_fourier(int dtype, int p, int dim, const void *tab_f64, int nmodes, const double *a_dev,
                       const int *m_dev, const int *n_dev, void *b, hipStream_t s) {
    const int64_t n = (int64_t)dim * dim * (p + 1) * (p + 1);
    int64_t nb = (n + NT - 1) / NT;
    if (nb > 4096) nb = 4096;
    DISPATCH_T(dtype,
        hipLaunchKernelGGL(rhs_fourier_kernel<double>, dim3((int)nb), dim3(NT), 0, s, p, dim, (const double *)tab_f64, nmodes, a_dev, m_dev, n_dev, (double *)b),
        hipLaunchKernelGGL(rhs_fourier_kernel<float>, dim3((int)nb), dim3(NT), 0, s, p, dim, (const double *)tab_f64, nmodes, a_dev, m_dev, n_dev, (float *)b));
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace sipg_internal

// kernels_coarse.hip -- the coarse level of the two-level preconditioner.
//
// M2^-1 r = M_BJ^-1 r + P0^T A0^-1 P0 r: the block-Jacobi preconditioner of
// kernels_pcg.hip plus an exact solve on V0, the continuous bilinear (Q1)
// functions on the element vertices that vanish on the boundary.  P0^T
// interpolates vertex values bilinearly into every node of the adjacent
// elements; P0 is its transpose.  A continuous function has no jumps, so on
// V0 every face term of the SIPG form vanishes and A0 = P0 A P0^T is the Q1
// stiffness matrix, with the 1-D factors K = tridiag(-1, 2, -1) and the GLL
// quadrature mass M = tridiag(m1, 2 m0, m1) (exact for p >= 2: 1/6, 1/3;
// lumped for p = 1): A0 = Kx (x) My + Mx (x) Ky on the (dimy-1) x (dimx-1)
// interior vertices.  Both factors are diagonalised by the discrete sine
// transform Q_jk = sqrt(2/dim) sin(j k pi / dim), so
//
//   A0^-1 S = Qy [ (Qy S Qx) ./ Lambda ] Qx,
//   Lambda_ab = lamKy_a lamMx_b + lamMy_a lamKx_b
//
// four dense products of (dim-1)^2 matrices.  One coarse step per
// preconditioner application:
//   restrict  -- per element the four corner-weighted sums of r (E4); for
//                the standard penalty the block-Jacobi kernel forms them
//                from its own read of r (pcg_rz_kernel with RESTR), and
//                restrict_kernel runs only elsewhere
//   4 x gemm  -- the DST solve (dst_gemm_kernel, each product over half
//                the k range); the first gathers S(a,b), the sum of the
//                four elements meeting at interior vertex (a,b) (vertex
//                (a,b) sits at x = b hx, y = 1 - a hy; element (row, col)
//                has its top-left corner at vertex (row, col)), as it loads
//                it; the last writes C into the zero-bordered vertex grid
//                Cpad and the block partials of S.C
// and the p update interpolates Cpad into every node (p = z + P0^T C + beta p)
// with r.z = r.z_BJ + S.C.
//
// Large vertex grids (kBlasVertices and up: from about 271 x 271 elements)
// take the four products as rocBLAS FP64 GEMMs instead, with the
// gather, the 1/Lambda scaling and the S.C partials as small kernels of
// their own: the dense products grow as the cube of the mesh side, and the
// library's matrix-core GEMM runs them several times faster than the
// folded VALU kernel, which is latency-bound and wins below that size.
// rocBLAS is opened on first use (dlopen); without it the folded kernel
// runs at every size.
#include "sipg_internal.hpp"
#include "partials.hpp"

#include <dlfcn.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <map>

namespace sipg_internal {

namespace {

constexpr int VT = kVecThreads;
// elements per restrict workgroup: their tiles fill at most 32 KiB of LDS
template <int P>
constexpr int rne() { return 4096 / ((P + 1) * (P + 1)) < 32 ? 4096 / ((P + 1) * (P + 1)) : 32; }
constexpr int GT = 16;           // gemm output tile (GT x GT)
constexpr int KC2 = 128;         // folded k chunk staged in LDS
constexpr int KS = 4;            // gemm k slices per output
constexpr int GEMM_T = GT * GT * KS;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// fixed-order sum over the workgroup of VT threads, returned to all
__device__ __forceinline__ double vsum_all(double v) {
    __shared__ double red[VT / 64];
    __shared__ double total;
    v = wsum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < VT / 64; ++i) s += red[i];
        total = s;
    }
    __syncthreads();
    return total;
}

__device__ __forceinline__ double vreduce(const double *__restrict__ part, int np) {
    return vsum_all(strided_sum<VT>(part, 0, np));
}

// E4[c * noe + e] = sum_nodes r(e, i1, i2) wx_c(i1) wy_c(i2); corner c = 0
// bottom-left, 1 bottom-right, 2 top-left, 3 top-right; wx = 1 - xi or xi,
// wy = 1 - xi or xi (xi = (z+1)/2: x from the element's left, y from its
// bottom).
template <typename T, int P>
__global__ __launch_bounds__(VT) void restrict_kernel(const T *__restrict__ r, const double *__restrict__ xi,
                                                      int64_t noe, double *__restrict__ E4) {
    constexpr int N1 = P + 1, NQ = N1 * N1, RNE = rne<P>();
    __shared__ double tile[RNE * NQ];
    __shared__ double sx[N1];
    if (threadIdx.x < N1) sx[threadIdx.x] = xi[threadIdx.x];
    const int64_t e0 = (int64_t)blockIdx.x * RNE;
    const int ne = (int)min<int64_t>(RNE, noe - e0);
    const T *src = r + e0 * NQ;
    for (int i = threadIdx.x; i < ne * NQ; i += VT) tile[i] = (double)src[i];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 4 * ne) {
        const int el = t >> 2, c = t & 3;
        const double *u = tile + el * NQ;
        double acc = 0.0;
        for (int i1 = 0; i1 < N1; ++i1) {
            double line = 0.0;
#pragma unroll
            for (int i2 = 0; i2 < N1; ++i2) line += u[i1 * N1 + i2] * ((c & 2) ? sx[i2] : 1.0 - sx[i2]);
            acc += line * ((c & 1) ? sx[i1] : 1.0 - sx[i1]);
        }
        E4[(int64_t)c * noe + e0 + el] = acc;
    }
}

// The sine matrix Q (n x n, symmetric) has Q[j][n-1-k] = (-1)^j Q[j][k],
// so each product folds its k range in half: with h = n / 2,
//   (Q X)[j][m] = sum_{k<h} Q[j][k] (X[k][m] +- X[n-1-k][m]) (+ Q[j][h] X[h][m], n odd)
//   (X Q)[m][j] = sum_{k<h} (X[m][k] +- X[m][n-1-k]) Q[k][j]  (+ X[m][h] Q[h][j], n odd)
// with + for even j and - for odd j.  One workgroup per GT x GT output tile
// (GT = 16, so a 255 x 255 product has 256 of them, one per CU); the folded
// strips (KC2 values of the half range at a time, one chunk up to n = 256)
// are staged in LDS, and the 1024 threads split that range four ways (four
// waves per SIMD), the partial sums added in LDS in a fixed order.
//   SIDE 0: out (n x N) = Q X (X: n x N);  SIDE 1: out (M x n) = X Q (X: M x n)
//   MODE 0: out = the product                (SIDE 1)
//   MODE 1: out = product .* W               (SIDE 0; W the 1/Lambda table)
//   MODE 2: Cpad[(i+1)*(N+2) + j+1] = product, and per-workgroup partials of
//           sum product .* W into part[blockIdx.y * gridDim.x + blockIdx.x]   (SIDE 0)
//   MODE 3: as MODE 0 with X = S gathered from the corner sums E4 (X is E4,
//           M + 1 by n + 1 elements; the first tile column also stores S
//           into W)
template <int SIDE, int MODE>
__global__ __launch_bounds__(GEMM_T) void dst_gemm_kernel(const double *__restrict__ X,
                                                          const double *__restrict__ Q, int M, int n, int N,
                                                          double *__restrict__ W, double *__restrict__ out,
                                                          double *__restrict__ part, const cg_state *st, int slot) {
    // inside a solve: nothing to do once it has finished (the launches
    // queued past the finish run as no-ops)
    if (st && (st->s[slot].done | st->err)) return;
    // SIDE 0: A = Q rows (plain), B0/B1 = folded X (+/-); SIDE 1: A0/A1 =
    // folded X rows (+/-), B = Q columns
    __shared__ double Aq[SIDE == 1 ? 2 : 1][GT][KC2 + 1];
    __shared__ double Bq[SIDE == 0 ? 2 : 1][KC2][GT];
    __shared__ double ps[KS][GT][GT + 1];
    const int o = threadIdx.x % (GT * GT), ks = threadIdx.x / (GT * GT);
    const int tx = o % GT, ty = o / GT;
    const int row0 = blockIdx.y * GT, col0 = blockIdx.x * GT;
    const int rows = SIDE == 0 ? n : M, cols = SIDE == 0 ? N : n;
    const int h = n / 2, kh = h + (n & 1);          // folded length (with the middle)
    // X element (r, k) of the (rows x n) or (n x N) operand
    auto xval = [&](int r, int k) -> double {
        if (MODE == 3) {
            const int a = r + 1, b = k + 1, dimx = n + 1;
            const int64_t noe = (int64_t)dimx * (M + 1);
            return (X[1 * noe + (int64_t)(a - 1) * dimx + (b - 1)] + X[0 * noe + (int64_t)(a - 1) * dimx + b]) +
                   (X[3 * noe + (int64_t)a * dimx + (b - 1)] + X[2 * noe + (int64_t)a * dimx + b]);
        }
        return SIDE == 0 ? X[(int64_t)k * N + r] : X[(int64_t)r * n + k];
    };
    // output (gr, gc): its parity picks the folded operand
    const int gr = row0 + ty, gc = col0 + tx;
    const bool odd = SIDE == 0 ? (gr & 1) : (gc & 1);
    double acc[2] = {0.0, 0.0};
    for (int k0 = 0; k0 < kh; k0 += KC2) {
        const int kc = min(KC2, kh - k0);
        for (int i = threadIdx.x; i < GT * KC2; i += GEMM_T) {
            const int r = i / KC2, kk = i - (i / KC2) * KC2, k = k0 + kk;
            if (SIDE == 0) {
                const int gq = row0 + r;
                Aq[0][r][kk] = (gq < rows && kk < kc) ? Q[(int64_t)gq * n + k] : 0.0;
            } else {
                const int gx = row0 + r;
                double p = 0.0, q = 0.0;
                if (gx < rows && kk < kc) {
                    const double a = xval(gx, k);
                    if (k < h) {
                        const double b = xval(gx, n - 1 - k);
                        p = a + b;
                        q = a - b;
                        if (MODE == 3 && blockIdx.x == 0) {
                            W[(int64_t)gx * n + k] = a;
                            W[(int64_t)gx * n + n - 1 - k] = b;
                        }
                    } else {                                  // the middle (n odd)
                        p = q = a;
                        if (MODE == 3 && blockIdx.x == 0) W[(int64_t)gx * n + k] = a;
                    }
                }
                Aq[0][r][kk] = p;
                Aq[SIDE == 1 ? 1 : 0][r][kk] = q;
            }
        }
        for (int i = threadIdx.x; i < KC2 * GT; i += GEMM_T) {
            const int kk = i / GT, c = i % GT, k = k0 + kk;
            if (SIDE == 0) {
                const int gx = col0 + c;
                double p = 0.0, q = 0.0;
                if (kk < kc && gx < cols) {
                    const double a = xval(gx, k);
                    if (k < h) {
                        const double b = xval(gx, n - 1 - k);
                        p = a + b;
                        q = a - b;
                    } else {
                        p = q = a;
                    }
                }
                Bq[0][kk][c] = p;
                Bq[SIDE == 0 ? 1 : 0][kk][c] = q;
            } else {
                const int gq = col0 + c;
                Bq[0][kk][c] = (kk < kc && gq < cols) ? Q[(int64_t)k * n + gq] : 0.0;
            }
        }
        __syncthreads();
        const int kb = ks * (KC2 / KS);
        const int ia = SIDE == 1 ? (int)odd : 0, ib = SIDE == 0 ? (int)odd : 0;
#pragma unroll 8
        for (int kk = 0; kk < KC2 / KS; kk += 2) {
            acc[0] += Aq[ia][ty][kb + kk] * Bq[ib][kb + kk][tx];
            acc[1] += Aq[ia][ty][kb + kk + 1] * Bq[ib][kb + kk + 1][tx];
        }
        __syncthreads();
    }
    ps[ks][ty][tx] = acc[0] + acc[1];
    __syncthreads();
    double dot = 0.0;
    if (ks == 0) {
        const double v = (ps[0][ty][tx] + ps[1][ty][tx]) + (ps[2][ty][tx] + ps[3][ty][tx]);
        if (gr < rows && gc < cols) {
            const int64_t oo = (int64_t)gr * cols + gc;
            if (MODE == 0 || MODE == 3) out[oo] = v;
            if (MODE == 1) out[oo] = v * W[oo];
            if (MODE == 2) {
                out[(int64_t)(gr + 1) * (cols + 2) + gc + 1] = v;
                dot = v * W[oo];
            }
        }
    }
    if (MODE == 2) {
        __shared__ double red[GT * GT / 64];
        dot = wsum(dot);
        if (ks == 0 && (threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
        __syncthreads();
        if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

// bilinear interpolation of the zero-bordered vertex grid Cpad
// ((dimy+1) x (dimx+1)) at node (i1, i2) of element e
template <int P>
__device__ __forceinline__ double prolong(const double *__restrict__ Cpad, const double *sx, int dimx, int e,
                                          int i1, int i2) {
    const int row = e / dimx, col = e - (e / dimx) * dimx;
    const int ld = dimx + 1;
    const double tl = Cpad[row * ld + col], tr = Cpad[row * ld + col + 1];
    const double bl = Cpad[(row + 1) * ld + col], br = Cpad[(row + 1) * ld + col + 1];
    const double x = sx[i1], y = sx[i2];
    return (1.0 - y) * ((1.0 - x) * bl + x * br) + y * ((1.0 - x) * tl + x * tr);
}

// p += P0^T C (the start of a solve: p = z_BJ was written by the
// preconditioner), or with z given: z += P0^T C (a plain application)
template <typename T, int P>
__global__ __launch_bounds__(VT) void add_coarse_kernel(T *__restrict__ v, const double *__restrict__ Cpad,
                                                        const double *__restrict__ xi, int dimx, int64_t n) {
    constexpr int N1 = P + 1, NQ = N1 * N1;
    __shared__ double sx[N1];
    if (threadIdx.x < N1) sx[threadIdx.x] = xi[threadIdx.x];
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VT) {
        const int e = (int)(i / NQ), node = (int)(i - (int64_t)e * NQ);
        const int i1 = node / N1, i2 = node - (node / N1) * N1;
        v[i] = (T)((double)v[i] + prolong<P>(Cpad, sx, dimx, e, i1, i2));
    }
}

// pcg_update_p_kernel with the coarse level: r.z = sum(prz) + sum(pc)
// (the block-Jacobi part and S.C) and p = z + P0^T C + beta p.
template <typename T, int P>
__global__ __launch_bounds__(VT) void pcg2_update_p_kernel(T *__restrict__ x, T *__restrict__ p,
                                                           const T *__restrict__ z, int64_t n,
                                                           const double *__restrict__ ppAp, int np_pAp,
                                                           const double *__restrict__ prr,
                                                           const double *__restrict__ prz, int np,
                                                           const double *__restrict__ pc, int npc,
                                                           const double *__restrict__ Cpad,
                                                           const double *__restrict__ xi, int dimx, cg_state *st,
                                                           int slot, T *__restrict__ xlo) {
    constexpr int N1 = P + 1, NQ = N1 * N1;
    __shared__ double sx[N1];
    const cg_slot cur = st->s[slot];
    if (cur.done | st->err) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->s[slot ^ 1] = cur;   // carry forward
        return;
    }
    if (threadIdx.x < N1) sx[threadIdx.x] = xi[threadIdx.x];
    const double rr_new = vreduce(prr, np);
    const double rz_new = vreduce(prz, np) + vreduce(pc, npc);
    const double pAp = vreduce(ppAp, np_pAp);
    const bool replaced = (cur.flags & kSlotReplaced) != 0;
    const double beta_d = (cur.flags & kSlotRestart) ? 0.0 : rz_new / cur.rz;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool met = !(sqrt(rr_new) > st->stop);
        const bool need_true = st->true_stop && !replaced;
        cg_slot nx;
        nx.rr = rr_new;
        nx.alpha = 0.0;
        nx.beta = beta_d;
        nx.pAp = 0.0;
        nx.rz = rz_new;
        nx.it = cur.it + 1;
        nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
        nx.checked = replaced ? 1 : 0;
        nx.flags = (met && need_true) ? kSlotWantCheck : 0;
        st->s[slot ^ 1] = nx;
        if (replaced) st->rr_ref = rr_new;
        if (!isfinite(rr_new) || !isfinite(rz_new)) st->err = 3;
        else if (!(rz_new > 0.0) && rr_new > 0.0) st->err = 2;   // M not positive on r
    }
    const T alpha = (T)(cur.rz / pAp);
    const T beta = (T)beta_d;
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VT) {
        const int e = (int)(i / NQ), node = (int)(i - (int64_t)e * NQ);
        const int i1 = node / N1, i2 = node - (node / N1) * N1;
        const T pi = p[i];
        if (!replaced) {
            if (xlo) {
                T xh = x[i], xl = xlo[i];
                comp_axpy(xh, xl, alpha, pi);
                x[i] = xh;
                xlo[i] = xl;
            } else {
                x[i] = x[i] + alpha * pi;
            }
        }
        p[i] = (T)((double)z[i] + prolong<P>(Cpad, sx, dimx, e, i1, i2)) + beta * pi;
    }
}

template <int P>
struct CoarseFns {
    template <typename T>
    static void restrict_(const void *r, const double *xi, int64_t noe, double *E4, hipStream_t s) {
        const int nb = (int)((noe + rne<P>() - 1) / rne<P>());
        hipLaunchKernelGGL((restrict_kernel<T, P>), dim3(nb), dim3(VT), 0, s, (const T *)r, xi, noe, E4);
    }
    template <typename T>
    static void add_(void *v, const double *Cpad, const double *xi, int dimx, int64_t n, int nblk, hipStream_t s) {
        hipLaunchKernelGGL((add_coarse_kernel<T, P>), dim3(nblk), dim3(VT), 0, s, (T *)v, Cpad, xi, dimx, n);
    }
    template <typename T>
    static void update_(void *x, void *p, const void *z, int64_t n, const double *ppAp, int np_pAp, const double *prr,
                        const double *prz, int np, const double *pc, int npc, const double *Cpad, const double *xi,
                        int dimx, int nblk, cg_state *st, int slot, hipStream_t s, void *xlo) {
        hipLaunchKernelGGL((pcg2_update_p_kernel<T, P>), dim3(nblk), dim3(VT), 0, s, (T *)x, (T *)p, (const T *)z, n,
                           ppAp, np_pAp, prr, prz, np, pc, npc, Cpad, xi, dimx, st, slot, (T *)xlo);
    }
};

struct coarse_args {
    const void *r = nullptr;
    void *v = nullptr, *x = nullptr, *xlo = nullptr;
    const void *z = nullptr;
    const double *xi = nullptr, *Cpad = nullptr, *ppAp = nullptr, *prr = nullptr, *prz = nullptr, *pc = nullptr;
    double *E4 = nullptr;
    int64_t noe = 0, n = 0;
    int dimx = 0, nblk = 0, np_pAp = 0, np = 0, npc = 0, slot = 0;
    cg_state *st = nullptr;
    hipStream_t s = nullptr;
};

template <int P>
int dispatch_p(int p, int dtype, int what, const coarse_args &a) {
    if (p != P) {
        if constexpr (P < kMaxDegree) return dispatch_p<P + 1>(p, dtype, what, a);
        else return SIPG_ERR_UNSUPPORTED;
    }
    using F = CoarseFns<P>;
    const bool f64 = dtype == SIPG_F64;
    if (what == 0) {
        if (f64) F::template restrict_<double>(a.r, a.xi, a.noe, a.E4, a.s);
        else F::template restrict_<float>(a.r, a.xi, a.noe, a.E4, a.s);
    } else if (what == 1) {
        if (f64) F::template add_<double>(a.v, a.Cpad, a.xi, a.dimx, a.n, a.nblk, a.s);
        else F::template add_<float>(a.v, a.Cpad, a.xi, a.dimx, a.n, a.nblk, a.s);
    } else {
        if (f64)
            F::template update_<double>(a.x, a.v, a.z, a.n, a.ppAp, a.np_pAp, a.prr, a.prz, a.np, a.pc, a.npc, a.Cpad,
                                        a.xi, a.dimx, a.nblk, a.st, a.slot, a.s, a.xlo);
        else
            F::template update_<float>(a.x, a.v, a.z, a.n, a.ppAp, a.np_pAp, a.prr, a.prz, a.np, a.pc, a.npc, a.Cpad,
                                       a.xi, a.dimx, a.nblk, a.st, a.slot, a.s, a.xlo);
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

dim3 gemm_grid(int M, int N) { return dim3((N + GT - 1) / GT, (M + GT - 1) / GT); }

// ---- the rocBLAS route for large vertex grids -------------------------------
// Two-level PCG at p = 8, us per iteration, folded kernel / rocBLAS:
// 256^2 elements 135.5 / 198.8, 288^2 217.3 / 191.9, 384^2 377.4 / 324.9.
constexpr int64_t kBlasVertices = 270 * 270;

// S(a, b): the sum of the four elements' corner sums meeting at interior
// vertex (a, b) (the gather of dst_gemm_kernel MODE 3), into S (nvy x nvx)
__global__ __launch_bounds__(VT) void gather_vertex_kernel(const double *__restrict__ E4, int nvy, int nvx,
                                                           double *__restrict__ S) {
    const int64_t nv = (int64_t)nvy * nvx;
    const int dimx = nvx + 1;
    const int64_t noe = (int64_t)dimx * (nvy + 1);
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nv; i += (int64_t)gridDim.x * VT) {
        const int a = (int)(i / nvx) + 1, b = (int)(i - (i / nvx) * nvx) + 1;
        S[i] = (E4[1 * noe + (int64_t)(a - 1) * dimx + (b - 1)] + E4[0 * noe + (int64_t)(a - 1) * dimx + b]) +
               (E4[3 * noe + (int64_t)a * dimx + (b - 1)] + E4[2 * noe + (int64_t)a * dimx + b]);
    }
}

// T .*= W (the 1/Lambda table), nv values
__global__ __launch_bounds__(VT) void scale_kernel(double *__restrict__ T, const double *__restrict__ W, int64_t nv) {
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nv; i += (int64_t)gridDim.x * VT) T[i] *= W[i];
}

// per-tile sums of S .* C (C the interior of the zero-bordered grid Cpad),
// on the tiles and in the partial order of dst_gemm_kernel MODE 2
__global__ __launch_bounds__(GT * GT) void sc_partials_kernel(const double *__restrict__ S,
                                                              const double *__restrict__ Cpad, int nvy, int nvx,
                                                              double *__restrict__ part) {
    __shared__ double red[GT * GT / 64];
    const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;
    const int gr = blockIdx.y * GT + ty, gc = blockIdx.x * GT + tx;
    double dot = 0.0;
    if (gr < nvy && gc < nvx) dot = S[(int64_t)gr * nvx + gc] * Cpad[(int64_t)(gr + 1) * (nvx + 2) + gc + 1];
    dot = wsum(dot);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// The three rocBLAS entry points used, resolved once from the installed
// library.
struct BlasFns {
    using create_t = rocblas_status (*)(rocblas_handle *);
    using stream_t = rocblas_status (*)(rocblas_handle, hipStream_t);
    using dgemm_t = rocblas_status (*)(rocblas_handle, rocblas_operation, rocblas_operation, rocblas_int,
                                       rocblas_int, rocblas_int, const double *, const double *, rocblas_int,
                                       const double *, rocblas_int, const double *, double *, rocblas_int);
    create_t create = nullptr;
    stream_t set_stream = nullptr;
    dgemm_t dgemm = nullptr;
    bool ok() const { return create && set_stream && dgemm; }
};

const BlasFns &blas() {
    static const BlasFns fns = [] {
        BlasFns f;
        void *so = nullptr;
        for (const char *name : {"librocblas.so.5", "librocblas.so", "/opt/rocm/lib/librocblas.so.5"})
            if ((so = dlopen(name, RTLD_NOW | RTLD_LOCAL))) break;
        if (so) {
            f.create = reinterpret_cast<BlasFns::create_t>(dlsym(so, "rocblas_create_handle"));
            f.set_stream = reinterpret_cast<BlasFns::stream_t>(dlsym(so, "rocblas_set_stream"));
            f.dgemm = reinterpret_cast<BlasFns::dgemm_t>(dlsym(so, "rocblas_dgemm"));
        }
        return f;
    }();
    return fns;
}

// One handle per host thread and device, kept for the life of the process
// (a rocBLAS handle must not be used by two threads at once; the solver
// sets its stream before each use).
rocblas_handle blas_handle() {
    thread_local std::map<int, rocblas_handle> handles;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto it = handles.find(dev);
    if (it != handles.end()) return it->second;
    rocblas_handle h = nullptr;
    if (blas().create(&h) != rocblas_status_success) return nullptr;
    handles[dev] = h;
    return h;
}

// row-major C (m x n, ldc) = A (m x k, lda) B (k x n, ldb): the column-major
// product C^T = B^T A^T on the same storage
bool rm_gemm(rocblas_handle h, int m, int n, int k, const double *A, int lda, const double *B, int ldb, double *C,
             int ldc) {
    const double one = 1.0, zero = 0.0;
    return blas().dgemm(h, rocblas_operation_none, rocblas_operation_none, n, m, k, &one, B, ldb, A, lda, &zero, C,
                        ldc) == rocblas_status_success;
}

int coarse_solve_blas(int nvy, int nvx, const coarse_tables &c, hipStream_t s) {
    rocblas_handle h = blas_handle();
    if (!h || blas().set_stream(h, s) != rocblas_status_success) {
        set_last_hip_error(hipErrorNotInitialized);
        return SIPG_ERR_HIP;
    }
    const int64_t nv = (int64_t)nvy * nvx;
    const int nb = (int)std::min<int64_t>((nv + VT - 1) / VT, 4096);
    hipLaunchKernelGGL(gather_vertex_kernel, dim3(nb), dim3(VT), 0, s, c.E4, nvy, nvx, c.S);
    SIPG_HIP_TRY(hipGetLastError());
    // T1 = S Qx; T2 = (Qy T1) ./ Lambda; T1 = T2 Qx; C = Qy T1 into Cpad
    bool ok = rm_gemm(h, nvy, nvx, nvx, c.S, nvx, c.Qx, nvx, c.T1, nvx) &&
              rm_gemm(h, nvy, nvx, nvy, c.Qy, nvy, c.T1, nvx, c.T2, nvx);
    if (ok) {
        hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(VT), 0, s, c.T2, c.Linv, nv);
        SIPG_HIP_TRY(hipGetLastError());
        ok = rm_gemm(h, nvy, nvx, nvx, c.T2, nvx, c.Qx, nvx, c.T1, nvx) &&
             rm_gemm(h, nvy, nvx, nvy, c.Qy, nvy, c.T1, nvx, c.Cpad + (nvx + 2) + 1, nvx + 2);
    }
    if (!ok) {
        set_last_hip_error(hipErrorLaunchFailure);
        return SIPG_ERR_HIP;
    }
    hipLaunchKernelGGL(sc_partials_kernel, gemm_grid(nvy, nvx), dim3(GT * GT), 0, s, c.S, c.Cpad, nvy, nvx, c.part);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace

int coarse_num_partials(int dimx, int dimy) {
    const int nvx = dimx - 1, nvy = dimy - 1;
    if (nvx < 1 || nvy < 1) return 0;
    const dim3 g = gemm_grid(nvy, nvx);
    return (int)(g.x * g.y);
}

int launch_coarse_solve(int dtype, int p, const void *r, int dimx, int dimy, const coarse_tables &c, hipStream_t s,
                        bool have_e4, const cg_state *st, int slot) {
    const int nvx = dimx - 1, nvy = dimy - 1;
    const int64_t noe = (int64_t)dimx * dimy;
    if (nvx < 1 || nvy < 1) return SIPG_OK;          // no interior vertex: C = 0, no partials
    coarse_args a{};
    a.r = r;
    a.xi = c.xi;
    a.noe = noe;
    a.E4 = c.E4;
    a.s = s;
    int rc = have_e4 ? SIPG_OK : dispatch_p<1>(p, dtype, 0, a);
    if (rc) return rc;
    if ((int64_t)nvx * nvy >= kBlasVertices && blas().ok()) return coarse_solve_blas(nvy, nvx, c, s);
    // T1 = S Qx (S gathered from E4 on the fly, and stored); T2 = (Qy T1) ./
    // Lambda; T1 = T2 Qx; C = Qy T1 (into Cpad), partials of S.C
    const dim3 g = gemm_grid(nvy, nvx), b(GEMM_T);
    hipLaunchKernelGGL((dst_gemm_kernel<1, 3>), g, b, 0, s, c.E4, c.Qx, nvy, nvx, nvx, c.S, c.T1, nullptr, st, slot);
    hipLaunchKernelGGL((dst_gemm_kernel<0, 1>), g, b, 0, s, c.T1, c.Qy, nvy, nvy, nvx, c.Linv, c.T2, nullptr, st,
                       slot);
    hipLaunchKernelGGL((dst_gemm_kernel<1, 0>), g, b, 0, s, c.T2, c.Qx, nvy, nvx, nvx, nullptr, c.T1, nullptr, st,
                       slot);
    hipLaunchKernelGGL((dst_gemm_kernel<0, 2>), g, b, 0, s, c.T1, c.Qy, nvy, nvy, nvx, c.S, c.Cpad, c.part, st,
                       slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_coarse_add(int dtype, int p, void *v, int dimx, int64_t n, const coarse_tables &c, int nblk,
                      hipStream_t s) {
    coarse_args a{};
    a.v = v;
    a.Cpad = c.Cpad;
    a.xi = c.xi;
    a.dimx = dimx;
    a.n = n;
    a.nblk = nblk;
    a.s = s;
    return dispatch_p<1>(p, dtype, 1, a);
}

int launch_pcg2_update_p(int dtype, int p, void *x, void *pv, const void *z, int64_t n, const double *part_pAp,
                         int np_pAp, const double *part_rr, const double *part_rz, int np, int npc, int dimx,
                         const coarse_tables &c, int nblk, cg_state *st, int slot, hipStream_t s, void *xlo) {
    coarse_args a{};
    a.x = x;
    a.xlo = xlo;
    a.v = pv;
    a.z = z;
    a.n = n;
    a.ppAp = part_pAp;
    a.np_pAp = np_pAp;
    a.prr = part_rr;
    a.prz = part_rz;
    a.np = np;
    a.pc = c.part;
    a.npc = npc;
    a.Cpad = c.Cpad;
    a.xi = c.xi;
    a.dimx = dimx;
    a.nblk = nblk;
    a.st = st;
    a.slot = slot;
    a.s = s;
    return dispatch_p<1>(p, dtype, 2, a);
}

}  // namespace sipg_internal

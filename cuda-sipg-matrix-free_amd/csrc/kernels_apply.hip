// kernels_apply.hip -- the fused SIPG operator apply y = A x for gfx950.
//
// The reference applies A with three launches (sipg_sem_2d.hpp:167-201):
// `volume` (volume_gpu_kernels.hpp:9-86, one thread per dof, O(p^2) work per
// dof with the 1-D stiffness rebuilt in the inner loop), then `flux_term6a`
// and `flux_term6b` (sipg_sem_2d_gpu_kernels.hpp:189-316, one thread per
// element, read-modify-write of the output).  Here one kernel does all of
// it in a single pass over the input and a single write of the output:
//
//   * a workgroup owns a BX x BY block of elements and stages the block's
//     (p+1)^2 tiles plus a one-element halo ring (the four side neighbours'
//     tiles) in LDS, node-plane major: lds[node][slot];
//   * thread (slot, j) keeps column j of its element in registers and does
//     the x-direction work for it: the sum-factorised contraction
//     w_j * K1 * U[:,j] plus the left/right face terms; it then does the
//     y-direction work for row j from LDS (K1 * U[j,:] plus down/up faces);
//   * the row results go back through LDS (over the element's own tile) to
//     be added to the column results, and each thread stores one column:
//     for a fixed node the lanes of a wave write 16 consecutive elements.
//
// The discrete operator is exactly the reference's (see the oracle and the
// tests): Y = K1 U W + W U K1 + face terms of flux_term6a/6b, with the
// same interior (1/2-average, 2*pen) and Dirichlet (one-sided, pen) rules.
//
// Block -> tile mapping is XCD-aware: the eight XCDs of an MI355X take
// consecutive workgroup ids round-robin, so the ids are remapped such that
// each XCD walks a contiguous band of block rows and the halo rows it reads
// are mostly in its own L2.
#include "sipg_internal.hpp"

namespace sipg_internal {

template <typename T, int P>
struct ApplyCfg;   // block shape per (type, degree); see select below

// BX x BY elements per workgroup, NE = BX*BY, threads = NE*(P+1) <= 1024,
// LDS = (P+1)^2 * (NE + 2BX + 2BY) * sizeof(T) <= 160 KiB.
template <typename T, int P>
struct ApplyShape {
    static constexpr int N1 = P + 1;
    static constexpr bool big64 = sizeof(T) == 8 && P >= 14;
    static constexpr bool small = N1 > 16;
    static constexpr int BX = (big64 || small) ? 8 : 16;
    static constexpr int BY = 4;
};

__device__ __forceinline__ int xcd_remap(int b, int nblk) {
    // workgroup id b runs on XCD (b % 8); give XCD k the k-th contiguous
    // share of the logical block range.
    const int xcd = b & 7;
    const int idx = b >> 3;
    const int q = nblk >> 3, r = nblk & 7;
    const int start = xcd * q + (xcd < r ? xcd : r);
    return start + idx;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double *red) {
    // wave64 butterfly, then a fixed-order sum over waves: deterministic.
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    constexpr int NW = (NT + 63) / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NW; ++i) s += red[i];
    }
    return s;   // valid in thread 0
}

template <typename T, int P, int BX, int BY, bool DOT>
__global__ __launch_bounds__(BX * BY * (P + 1))
void apply_kernel(const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ tab,
                  int dimx, int dimy, int nbx, int nblk, T pen,
                  double *__restrict__ partials, const cg_state *__restrict__ st, int slot)
{
    constexpr int N1 = P + 1;
    constexpr int NQ = N1 * N1;
    constexpr int NE = BX * BY;
    constexpr int NH = 2 * BX + 2 * BY;
    constexpr int NS = NE + NH;
    constexpr int NT = NE * N1;
    // halo slot groups
    constexpr int H_UP = NE, H_DN = NE + BX, H_LF = NE + 2 * BX, H_RT = NE + 2 * BX + BY;

    __shared__ T lds[NQ * NS];
    __shared__ double red[(NT + 63) / 64];

    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }

    const T *w = tab;
    const T *K1 = tab + N1;
    const T *D0 = tab + N1 + NQ;
    const T *DP = tab + 2 * N1 + NQ;

    const int tid = threadIdx.x;
    const int lb = xcd_remap(blockIdx.x, nblk);
    const int bx = lb % nbx, by = lb / nbx;
    const int x0 = bx * BX, y0 = by * BY;
    const size_t noe = (size_t)dimx * (size_t)dimy;

    const int s = tid % NE;
    int j = tid / NE;
    if (NE % 64 == 0) j = __builtin_amdgcn_readfirstlane(j);
    const int ex = s % BX, ey = s / BX;
    const int gx = x0 + ex, gy = y0 + ey;
    const bool valid = (gx < dimx) && (gy < dimy);
    const size_t e = (size_t)gy * dimx + gx;

    // ---- stage: own column j into registers and LDS ----------------------
    T u[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
        u[i] = valid ? x[(size_t)(i * N1 + j) * noe + e] : T(0);
        lds[(i * N1 + j) * NS + s] = u[i];
    }
    // ---- stage: halo ring (full tiles of the side neighbours) ------------
    for (int k = tid; k < NH * NQ; k += NT) {
        const int h = k % NH;
        const int q = k / NH;
        int hx, hy;
        if (h < BX)               { hx = x0 + h;              hy = y0 - 1;  }
        else if (h < 2 * BX)      { hx = x0 + h - BX;         hy = y0 + BY; }
        else if (h < 2 * BX + BY) { hx = x0 - 1;              hy = y0 + h - 2 * BX; }
        else                      { hx = x0 + BX;             hy = y0 + h - 2 * BX - BY; }
        T v = T(0);
        if (hx >= 0 && hx < dimx && hy >= 0 && hy < dimy)
            v = x[(size_t)q * noe + (size_t)hy * dimx + hx];
        lds[q * NS + NE + h] = v;
    }
    __syncthreads();

    const bool Lb = (gx == 0), Rb = (gx == dimx - 1);
    const bool Ub = (gy == 0), Db = (gy == dimy - 1);   // y grows downwards
    const T wj = w[j];
    const T pen2 = pen + pen;

    // ---- x-direction, column j (flux_term6a + first half of volume) ------
    T Yx[N1];
    {
        T nL = 0, tL = 0, nR = 0, tR = 0;
        const int sL = ex > 0 ? s - 1 : H_LF + ey;
        const int sR = ex < BX - 1 ? s + 1 : H_RT + ey;
        if (!Lb) {
#pragma unroll
            for (int i = 0; i < N1; ++i) {
                const T v = lds[(i * N1 + j) * NS + sL];
                nL += DP[i] * v;
                if (i == P) tL = v;
            }
        }
        if (!Rb) {
#pragma unroll
            for (int i = 0; i < N1; ++i) {
                const T v = lds[(i * N1 + j) * NS + sR];
                nR += D0[i] * v;
                if (i == 0) tR = v;
            }
        }
        T dL = 0, dR = 0;
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            dL += D0[i] * u[i];
            dR += DP[i] * u[i];
        }
        const T ev = wj * (Lb ? u[0] : T(0.5) * (u[0] - tL));
        const T fv = wj * (Rb ? -u[P] : T(0.5) * (tR - u[P]));
        const T A = Lb ? dL : T(0.5) * (dL + nL);
        const T B = Rb ? -dR : T(-0.5) * (dR + nR);
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            T acc = 0;
#pragma unroll
            for (int i = 0; i < N1; ++i) acc += K1[a * N1 + i] * u[i];
            Yx[a] = wj * acc + D0[a] * ev + DP[a] * fv;
        }
        Yx[0] += wj * A + (Lb ? pen : pen2) * ev;
        Yx[P] += wj * B - (Rb ? pen : pen2) * fv;
    }

    // ---- y-direction, row j (flux_term6b + second half of volume) --------
    T Yy[N1];
    {
        T v[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) v[k] = lds[(j * N1 + k) * NS + s];
        T nD = 0, tD = 0, nU = 0, tU = 0;
        const int sD = ey < BY - 1 ? s + BX : H_DN + ex;   // element below (y+1)
        const int sU = ey > 0 ? s - BX : H_UP + ex;        // element above (y-1)
        if (!Db) {
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                const T t = lds[(j * N1 + k) * NS + sD];
                nD += DP[k] * t;
                if (k == P) tD = t;
            }
        }
        if (!Ub) {
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                const T t = lds[(j * N1 + k) * NS + sU];
                nU += D0[k] * t;
                if (k == 0) tU = t;
            }
        }
        T dD = 0, dU = 0;
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            dD += D0[k] * v[k];
            dU += DP[k] * v[k];
        }
        const T g = wj * (Db ? v[0] : T(0.5) * (v[0] - tD));
        const T h = wj * (Ub ? -v[P] : T(0.5) * (tU - v[P]));
        const T C = Db ? dD : T(0.5) * (dD + nD);
        const T Dd = Ub ? -dU : T(-0.5) * (dU + nU);
#pragma unroll
        for (int b = 0; b < N1; ++b) {
            T acc = 0;
#pragma unroll
            for (int k = 0; k < N1; ++k) acc += K1[b * N1 + k] * v[k];
            Yy[b] = wj * acc + D0[b] * g + DP[b] * h;
        }
        Yy[0] += wj * C + (Db ? pen : pen2) * g;
        Yy[P] += wj * Dd - (Ub ? pen : pen2) * h;
    }

    // ---- transpose the row results through the element's own tile --------
    __syncthreads();
#pragma unroll
    for (int b = 0; b < N1; ++b) lds[(j * N1 + b) * NS + s] = Yy[b];
    __syncthreads();

    double dot = 0.0;
#pragma unroll
    for (int a = 0; a < N1; ++a) {
        const T ya = Yx[a] + lds[(a * N1 + j) * NS + s];
        if (valid) y[(size_t)(a * N1 + j) * noe + e] = ya;
        if (DOT) dot += (double)u[a] * (double)ya;
    }
    if (DOT) {
        const double bs = block_sum<NT>(dot, red);
        if (tid == 0) partials[blockIdx.x] = bs;
    }
}

namespace {

template <typename T, int P>
int launch_one(const void *x, void *y, const void *tab, int dimx, int dimy, double pen,
               double *partials, const cg_state *st, int slot, hipStream_t s)
{
    using S = ApplyShape<T, P>;
    constexpr int BX = S::BX, BY = S::BY;
    const int nbx = (dimx + BX - 1) / BX, nby = (dimy + BY - 1) / BY;
    const int nblk = nbx * nby;
    dim3 grid(nblk), block(BX * BY * (P + 1));
    if (partials) {
        hipLaunchKernelGGL((apply_kernel<T, P, BX, BY, true>), grid, block, 0, s,
                           (const T *)x, (T *)y, (const T *)tab, dimx, dimy, nbx, nblk, (T)pen,
                           partials, st, slot);
    } else {
        hipLaunchKernelGGL((apply_kernel<T, P, BX, BY, false>), grid, block, 0, s,
                           (const T *)x, (T *)y, (const T *)tab, dimx, dimy, nbx, nblk, (T)pen,
                           partials, st, slot);
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

template <typename T, int P>
int nblocks_one(int dimx, int dimy) {
    using S = ApplyShape<T, P>;
    return ((dimx + S::BX - 1) / S::BX) * ((dimy + S::BY - 1) / S::BY);
}

template <typename T, int... Ps>
struct Dispatch;

template <typename T>
struct Dispatch<T> {
    static int launch(int, const void *, void *, const void *, int, int, double, double *,
                      const cg_state *, int, hipStream_t) { return SIPG_ERR_UNSUPPORTED; }
    static int nblocks(int, int, int) { return -1; }
};

template <typename T, int P0, int... Ps>
struct Dispatch<T, P0, Ps...> {
    static int launch(int p, const void *x, void *y, const void *tab, int dimx, int dimy,
                      double pen, double *partials, const cg_state *st, int slot, hipStream_t s) {
        if (p == P0) return launch_one<T, P0>(x, y, tab, dimx, dimy, pen, partials, st, slot, s);
        return Dispatch<T, Ps...>::launch(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
    }
    static int nblocks(int p, int dimx, int dimy) {
        if (p == P0) return nblocks_one<T, P0>(dimx, dimy);
        return Dispatch<T, Ps...>::nblocks(p, dimx, dimy);
    }
};

template <typename T>
using AllDegrees = Dispatch<T, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>;

}  // namespace

int apply_num_blocks(int dtype, int p, int dimx, int dimy) {
    return dtype == SIPG_F64 ? AllDegrees<double>::nblocks(p, dimx, dimy)
                             : AllDegrees<float>::nblocks(p, dimx, dimy);
}

int launch_apply(int dtype, int p, const void *x, void *y, const void *tab, int dimx, int dimy,
                 double pen, double *partials, const cg_state *st, int slot, hipStream_t s)
{
    if (dtype == SIPG_F64)
        return AllDegrees<double>::launch(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
    return AllDegrees<float>::launch(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
}

}  // namespace sipg_internal

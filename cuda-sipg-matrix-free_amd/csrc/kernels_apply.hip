// kernels_apply.hip -- the fused SIPG operator apply y = A x for gfx950.
//
// The reference applies A with three launches (sipg_sem_2d.hpp:167-201):
// `volume` (volume_gpu_kernels.hpp:9-86, one thread per dof, O(p^2) work per
// dof with the 1-D stiffness rebuilt in the inner loop), then `flux_term6a`
// and `flux_term6b` (sipg_sem_2d_gpu_kernels.hpp:189-316, one thread per
// element, read-modify-write of the output).  Here one kernel does all of
// it in a single pass over the input and a single write of the output:
//
//   * vectors are element-major on the device: the (p+1)^2 values of element
//     e = dimx*y + x are contiguous at e*(p+1)^2 + i1*(p+1) + i2;
//   * a workgroup owns a BX x BY block of elements and stages the block's
//     tiles plus a one-element halo ring (the four side neighbours' tiles)
//     in LDS, one tile per slot with an odd row stride.  Each block row and
//     each halo row is one contiguous run of BX*(p+1)^2 values in memory and
//     each halo-column tile a run of (p+1)^2, so the loads are full 64-lane
//     runs of consecutive words;
//   * thread (slot, j) reads column j of its element and does the
//     x-direction work for it: the sum-factorised contraction
//     w_j * K1 * U[:,j] plus the left/right face terms; it also reads row j
//     and does the y-direction work (K1 * U[j,:] plus down/up faces);
//   * the column results are written over the element's own LDS tile, the
//     row results are added to them, and the block's rows are stored back
//     as contiguous runs.

// The discrete operator is exactly the reference's (see the oracle and the
// tests): Y = K1 U W + W U K1 + face terms of flux_term6a/6b, with the
// same interior (1/2-average, 2*pen) and Dirichlet (one-sided, pen) rules.
//
// Block -> tile mapping is XCD-aware: the eight XCDs of an MI355X take
// consecutive workgroup ids round-robin, so the ids are remapped such that
// each XCD walks a contiguous band of block rows and the halo rows it reads
// are mostly in its own L2.
#include "sipg_internal.hpp"
#include "basis_constexpr.hpp"

#include <algorithm>
#include <atomic>

namespace sipg_internal {

// BX x BY elements per workgroup, NE = BX*BY, threads = NE*(P+1) <= 1024,
// LDS = (P+1)^2 * (NE + 2BX + 2BY) * sizeof(T) <= 160 KiB.
template <typename T, int P>
struct ApplyShape {
    static constexpr int N1 = P + 1;
    static constexpr bool big64 = sizeof(T) == 8 && P >= 13;
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

// Odd LDS row stride (in elements of T) for a (p+1)^2 tile: lanes that read
// the same node of consecutive tiles then hit distinct banks.
__host__ __device__ constexpr int tile_stride(int nq) { return (nq % 2) ? nq : nq + 1; }

// A copy of v that the compiler cannot see through.  The block loop of the
// apply recomputes the staging index arithmetic from it in every pass, which
// is cheaper than holding ~20 loop-invariant offsets in registers across
// the compute phase.  (A register move; no memory access.)
__device__ __forceinline__ int opaque(int v) {
    int r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// LDS position of the idx-th value of consecutive tiles stored with row
// stride TS >= NQ.
template <int NQ, int TS>
__device__ __forceinline__ int lds_index(int idx) {
    if constexpr (TS == NQ) return idx;
    else return idx + (TS - NQ) * (idx / NQ);
}

template <typename T, int P, int BX, int BY, bool DOT>
__global__ __launch_bounds__(BX * BY * (P + 1))
void apply_kernel(const T *__restrict__ x, T *__restrict__ y, int dimx, int dimy, int nbx,
                  int nblk, T pen, double *__restrict__ partials,
                  const cg_state *__restrict__ st, int slot)
{
    constexpr int N1 = P + 1;
    constexpr int NQ = N1 * N1;
    constexpr int TS = tile_stride(NQ);
    constexpr int NE = BX * BY;
    // LDS slots: (BY+2) rows of BX tiles for the element rows y0-1 .. y0+BY
    // (slot r*BX + t holds element (x0+t, y0-1+r)), then the left halo
    // column (x0-1, y0+k) at SL_COL+k and the right one (x0+BX, y0+k) at
    // SL_COL+BY+k.  A block's own element (ex, ey) sits in slot (ey+1)*BX+ex,
    // so its up/down neighbours are always slot -/+ BX.
    constexpr int RUN = BX * NQ;                 // values of one row run
    constexpr int NROWV = (BY + 2) * RUN;
    constexpr int SL_COL = (BY + 2) * BX;
    constexpr int NCOLV = 2 * BY * NQ;
    constexpr int NS = SL_COL + 2 * BY;
    constexpr int NT = NE * N1;
    constexpr int KR = (NROWV + NT - 1) / NT;    // row values per thread
    constexpr int KC = (NCOLV + NT - 1) / NT;    // halo-column values per thread
    constexpr int NOUT = BY * RUN;
    constexpr int KO = (NOUT + NT - 1) / NT;

    __shared__ T lds[NS * TS];
    __shared__ double red[(NT + 63) / 64];

    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }

    // degree-p tables as compile-time constants (see basis_constexpr.hpp)
    constexpr cx::Tables<P> CT = cx::make_tables<P>();
    struct CW { __device__ T operator[](int i) const { return (T)CT.w[i]; } } w;
    struct C0 { __device__ T operator[](int i) const { return (T)CT.D[i][0]; } } D0;
    struct CP { __device__ T operator[](int i) const { return (T)CT.D[i][P]; } } DP;

    const int tid = threadIdx.x;
    // offsets are relative to a block's first own tile, in int: a block's
    // footprint spans (BY+2) rows of dimx*(p+1)^2 values
    const int RS = dimx * NQ;                    // row stride
    const int GAP = RS - RUN;

    // Persistent walk: the workgroups split the blocks into contiguous
    // ranges (consecutive blocks are horizontal neighbours), and each
    // workgroup loads block b+1 into registers while it computes block b.
    const int G = gridDim.x;
    const int lw = xcd_remap(blockIdx.x, G);
    const int bbeg = (int)((int64_t)lw * nblk / G);
    const int bend = (int)((int64_t)(lw + 1) * nblk / G);

    // Row value idx lies in run r = idx / RUN at offset idx + r*GAP - RS
    // from the block's first tile and in LDS at lds_index(idx) (the row
    // tiles are consecutive slots); halo-column value idx is node idx % NQ of
    // column tile idx / NQ.
    T buf[KR], cbuf[KC];
    auto prefetch = [&](int b) {
        const int tid = opaque(threadIdx.x);
        const int x0 = (b % nbx) * BX, y0 = (b / nbx) * BY;
        const T *xb = x + ((size_t)y0 * dimx + x0) * NQ;
        const bool interior = x0 >= 1 && x0 + BX < dimx && y0 >= 1 && y0 + BY < dimy;
        if (interior) {
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int idx = tid + k * NT;
                if ((k + 1) * NT <= NROWV || idx < NROWV) {
                    const int r = idx / RUN;
                    buf[k] = xb[idx + r * GAP - RS];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int idx = tid + k * NT;
                const int r = idx / RUN;
                const int t = (idx - r * RUN) / NQ;
                const int gy = y0 - 1 + r;
                const bool ok = idx < NROWV && gy >= 0 && gy < dimy && x0 + t < dimx;
                buf[k] = ok ? xb[idx + r * GAP - RS] : T(0);
            }
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int idx = tid + k * NT;
            const int kk = idx / NQ, q = idx - kk * NQ;
            const bool right = kk >= BY;
            const int row = right ? kk - BY : kk;
            const bool ok = idx < NCOLV && y0 + row < dimy && (right ? x0 + BX < dimx : x0 >= 1);
            cbuf[k] = ok ? xb[row * RS + (right ? BX * NQ : -NQ) + q] : T(0);
        }
    };
    auto stage = [&]() {
        const int tid = opaque(threadIdx.x);
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int idx = tid + k * NT;
            if ((k + 1) * NT <= NROWV || idx < NROWV) lds[lds_index<NQ, TS>(idx)] = buf[k];
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int idx = tid + k * NT;
            const int kk = idx / NQ, q = idx - kk * NQ;
            if (idx < NCOLV) lds[(SL_COL + kk) * TS + q] = cbuf[k];
        }
    };

    // ---- thread (s, j): element s of a block, line j ---------------------
    const int s = tid % NE;
    int j = tid / NE;
    if (NE % 64 == 0) j = __builtin_amdgcn_readfirstlane(j);
    const int ex = s % BX, ey = s / BX;
    const T wj = w[j];
    const T pen2 = pen + pen;
    const int sl = (ey + 1) * BX + ex;                  // own LDS slot
    const T *tile = lds + sl * TS;
    const T *tlL = lds + (ex > 0 ? sl - 1 : SL_COL + ey) * TS;
    const T *tlR = lds + (ex < BX - 1 ? sl + 1 : SL_COL + BY + ey) * TS;
    const T *tlD = lds + (sl + BX) * TS;                // element below (y+1)
    const T *tlU = lds + (sl - BX) * TS;                // element above (y-1)
    T *otile = lds + sl * TS;

    double dot = 0.0;
    if (bbeg < bend) prefetch(bbeg);
    for (int b = bbeg; b < bend; ++b) {
        const int x0 = (b % nbx) * BX, y0 = (b / nbx) * BY;
        stage();
        __syncthreads();
        if (b + 1 < bend) prefetch(b + 1);

        const int gx = x0 + ex, gy = y0 + ey;
        const bool Lb = (gx == 0), Rb = (gx == dimx - 1);
        const bool Ub = (gy == 0), Db = (gy == dimy - 1);   // y grows downwards

        // column j (x-direction, flux_term6a) and row j (y-direction, flux_term6b)
        T u[N1], v[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            u[i] = tile[i * N1 + j];
            v[i] = tile[j * N1 + i];
        }
        T ev, fv, Ax, Bx, g, h, Cy, Dy;
        {
            T nL = 0, tL = 0, nR = 0, tR = 0;
            if (!Lb) {
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    const T t = tlL[i * N1 + j];
                    nL += DP[i] * t;
                    if (i == P) tL = t;
                }
            }
            if (!Rb) {
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    const T t = tlR[i * N1 + j];
                    nR += D0[i] * t;
                    if (i == 0) tR = t;
                }
            }
            T dL = 0, dR = 0;
#pragma unroll
            for (int i = 0; i < N1; ++i) {
                dL += D0[i] * u[i];
                dR += DP[i] * u[i];
            }
            ev = wj * (Lb ? u[0] : T(0.5) * (u[0] - tL));
            fv = wj * (Rb ? -u[P] : T(0.5) * (tR - u[P]));
            Ax = Lb ? dL : T(0.5) * (dL + nL);
            Bx = Rb ? -dR : T(-0.5) * (dR + nR);
        }
        {
            T nD = 0, tD = 0, nU = 0, tU = 0;
            if (!Db) {
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    const T t = tlD[j * N1 + k];
                    nD += DP[k] * t;
                    if (k == P) tD = t;
                }
            }
            if (!Ub) {
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    const T t = tlU[j * N1 + k];
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
            g = wj * (Db ? v[0] : T(0.5) * (v[0] - tD));
            h = wj * (Ub ? -v[P] : T(0.5) * (tU - v[P]));
            Cy = Db ? dD : T(0.5) * (dD + nD);
            Dy = Ub ? -dU : T(-0.5) * (dU + nU);
        }

        // ---- both sum-factorised contractions, Yx[a] = (K1 U)[a][j] and
        // Yy[a] = (U K1)[j][a] (K1 symmetric), in even-odd form: each constant
        // of the half-size E/O blocks serves both directions while it is live.
        T Yx[N1], Yy[N1];
        {
            constexpr cx::EvenOdd<P> EO = cx::make_even_odd<P>();
            constexpr int H = cx::EvenOdd<P>::H, ME = cx::EvenOdd<P>::ME;
            T ue[ME], uo[H > 0 ? H : 1], ve[ME], vo[H > 0 ? H : 1];
#pragma unroll
            for (int i = 0; i < H; ++i) {
                ue[i] = u[i] + u[P - i];
                uo[i] = u[i] - u[P - i];
                ve[i] = v[i] + v[P - i];
                vo[i] = v[i] - v[P - i];
            }
            if (ME > H) { ue[ME - 1] = u[H]; ve[ME - 1] = v[H]; }
#pragma unroll
            for (int a = 0; a < ME; ++a) {
                T ex_ = 0, ey_ = 0, ox_ = 0, oy_ = 0;
#pragma unroll
                for (int i = 0; i < ME; ++i) {
                    const T c = (T)EO.E[a][i];
                    ex_ += c * ue[i];
                    ey_ += c * ve[i];
                }
                if (a < H) {
#pragma unroll
                    for (int i = 0; i < H; ++i) {
                        const T c = (T)EO.O[a][i];
                        ox_ += c * uo[i];
                        oy_ += c * vo[i];
                    }
                    Yx[a] = ex_ + ox_;
                    Yx[P - a] = ex_ - ox_;
                    Yy[a] = ey_ + oy_;
                    Yy[P - a] = ey_ - oy_;
                } else {
                    Yx[a] = ex_;
                    Yy[a] = ey_;
                }
            }
        }
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            Yx[a] = wj * Yx[a] + D0[a] * ev + DP[a] * fv;
            Yy[a] = wj * Yy[a] + D0[a] * g + DP[a] * h;
        }
        Yx[0] += wj * Ax + (Lb ? pen : pen2) * ev;
        Yx[P] += wj * Bx - (Rb ? pen : pen2) * fv;
        Yy[0] += wj * Cy + (Db ? pen : pen2) * g;
        Yy[P] += wj * Dy - (Ub ? pen : pen2) * h;

        // ---- Y[a][b] = Yx (column j) + Yy (row j), assembled in the
        // element's own LDS tile, then copied out as contiguous row runs.
        __syncthreads();                                   // all tile reads done
#pragma unroll
        for (int a = 0; a < N1; ++a) otile[a * N1 + j] = Yx[a];
        __syncthreads();
        const bool valid = gx < dimx && gy < dimy;
#pragma unroll
        for (int c = 0; c < N1; ++c) {
            const T yv = otile[j * N1 + c] + Yy[c];
            otile[j * N1 + c] = yv;
            if (DOT && valid) dot += (double)v[c] * (double)yv;
        }
        __syncthreads();
        {
            const int tid = opaque(threadIdx.x);
            T *yb = y + ((size_t)y0 * dimx + x0) * NQ;
            const bool full = x0 + BX <= dimx && y0 + BY <= dimy;
#pragma unroll
            for (int k = 0; k < KO; ++k) {
                const int idx = tid + k * NT;
                if ((k + 1) * NT <= NOUT || idx < NOUT) {
                    const int r = idx / RUN;
                    bool ok = true;
                    if (!full) ok = y0 + r < dimy && x0 + (idx - r * RUN) / NQ < dimx;
                    if (ok) yb[idx + r * GAP] = lds[lds_index<NQ, TS>(idx + RUN)];
                }
            }
        }
        __syncthreads();                                   // LDS free for the next block
    }
    if (DOT) {
        const double bs = block_sum<NT>(dot, red);
        if (tid == 0) partials[blockIdx.x] = bs;
    }
}

namespace {

// Workgroups of one kernel that the device holds at once: the persistent
// grid.  Cached per device (occupancy depends only on the kernel).
template <typename T, int P, bool DOT>
int resident_groups() {
    using S = ApplyShape<T, P>;
    constexpr int kMaxDev = 64;
    static std::atomic<int> cache[kMaxDev];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) dev = 0;
    int g = cache[dev].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&apply_kernel<T, P, S::BX, S::BY, DOT>),
            S::BX * S::BY * (P + 1), 0) != hipSuccess)
        per = 0;
    g = (cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    cache[dev].store(g, std::memory_order_relaxed);
    return g;
}

template <typename T, int P>
int launch_one(const void *x, void *y, const void *tab, int dimx, int dimy, double pen,
               double *partials, const cg_state *st, int slot, hipStream_t s)
{
    using S = ApplyShape<T, P>;
    constexpr int BX = S::BX, BY = S::BY;
    const int nbx = (dimx + BX - 1) / BX, nby = (dimy + BY - 1) / BY;
    const int nblk = nbx * nby;
    dim3 block(BX * BY * (P + 1));
    if (partials) {
        dim3 grid(std::min(nblk, resident_groups<T, P, true>()));
        hipLaunchKernelGGL((apply_kernel<T, P, BX, BY, true>), grid, block, 0, s,
                           (const T *)x, (T *)y, dimx, dimy, nbx, nblk, (T)pen,
                           partials, st, slot);
    } else {
        dim3 grid(std::min(nblk, resident_groups<T, P, false>()));
        hipLaunchKernelGGL((apply_kernel<T, P, BX, BY, false>), grid, block, 0, s,
                           (const T *)x, (T *)y, dimx, dimy, nbx, nblk, (T)pen,
                           partials, st, slot);
    }
    (void)tab;
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

// Grid size of the fused apply + p.Ap launch = number of p.Ap partials.
template <typename T, int P>
int nblocks_one(int dimx, int dimy) {
    using S = ApplyShape<T, P>;
    const int nblk = ((dimx + S::BX - 1) / S::BX) * ((dimy + S::BY - 1) / S::BY);
    return std::min(nblk, resident_groups<T, P, true>());
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

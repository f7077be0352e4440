// kernels_pcg.hip -- block-Jacobi preconditioner and the PCG pieces.
//
// The reference's preconditioned CG (conjugate_gradient_gpu.hpp:116-213)
// applies a block-diagonal preconditioner built from the element Laplacian
// (laplacian_operator.hpp:95-186, mode_matrix.hpp:474-521): a dense block
// per element, with pen truncated to an integer and only the Laplacian-
// pattern entries of the inverse kept.  Here the preconditioner is the
// exact inverse of an interior element's diagonal block of A, applied to
// every element:
//
//   A_ee U = K U W + W U K,   K = K1 + (the element's own face terms in one
//                             direction; the same matrix in x and y for an
//                             interior element),  W = diag(w)
//
// a Kronecker sum, so with K S = W S diag(lambda), S^T W S = I (host side,
// precond.cpp) its inverse is
//
//   U = S [ (S^T R S) ./ (lambda_a + lambda_c) ] S^T
//
// four (p+1)-long contractions per node -- the cost of a volume apply, with
// no neighbour data.  Elements on the mesh boundary get the interior block
// too (still SPD; about 11 % more iterations than their own blocks, in
// exchange for coefficients that are uniform across every wave).
#include "apply_block.hpp"
#include "prec_block.hpp"
#include "partials.hpp"

#include <algorithm>
#include <atomic>

namespace sipg_internal {

namespace {

// elements per workgroup group; thread (s, j) owns line j of element s
template <int P>
struct PcgShape {
    static constexpr int N1 = P + 1;
    static constexpr int NE = N1 > 16 ? 32 : 64;
    static constexpr int NT = NE * N1;
    static constexpr int NQ = N1 * N1;
    static constexpr int TS = tile_stride(NQ);
    static constexpr int NV = NE * NQ;                  // values per group
    static constexpr int KR = (NV + NT - 1) / NT;       // = N1
};

// The restriction's line shares (4 (p+1) doubles per element) are summed in
// the element tiles' LDS once z has left it, where they fit: FP64 from
// p = 3, FP32 from p = 7; elsewhere the separate restriction kernel runs.
// es: bytes per value.
__host__ __device__ constexpr bool rz_fuses_restrict(int p, int es) {
    return 4 * (p + 1) * 8 <= tile_stride((p + 1) * (p + 1)) * es;
}

// Sum over the workgroup, returned to every thread (fixed order).
template <int NT>
__device__ __forceinline__ double group_sum_all(double v, double *red, double *total) {
    const double s = block_sum<NT>(v, red);
    if (threadIdx.x == 0) *total = s;
    __syncthreads();
    const double out = *total;
    __syncthreads();
    return out;
}

// z = M^-1 r over every element, with the block sums of r.r and r.z.
//   UPDATE: first r -= alpha Ap with alpha = slot.rz / p.Ap (the p.Ap block
//   sums of this iteration's apply), r written back -- the r half of the
//   reference's alpha step; returns at once when the slot is finished.
//   COPYP:  also p = z (the start of the iteration).
// prec = [S (row-major, S[i][a] = eigenvector a at node i), Linv[a][c] =
// 1 / (lambda_a + lambda_c)], in T.
//   CX: the coefficients are the compile-time tables of the standard penalty
//   pen = 5 (p+1)^2 (prec is then unused); otherwise they are read from prec.
//   RESTR: with E4, also the two-level restriction's corner sums (its own
//   instantiation, so that the plain kernel keeps its registers).
template <typename T, int P, bool UPDATE, bool COPYP, bool CX, bool RESTR = false>
__global__ __launch_bounds__(PcgShape<P>::NT)
void pcg_rz_kernel(T *__restrict__ r, const T *__restrict__ Ap, T *__restrict__ z,
                   T *__restrict__ pcopy, int64_t noe, const T *__restrict__ prec,
                   const double *__restrict__ ppAp, int np, double *__restrict__ prr,
                   double *__restrict__ prz, cg_state *st, int slot, double *__restrict__ E4)
{
    using G = PcgShape<P>;
    constexpr int N1 = G::N1, NQ = G::NQ, TS = G::TS, NE = G::NE, NT = G::NT, KR = G::KR;
    __shared__ __attribute__((aligned(16))) T lds[NE * TS];
    __shared__ double red[(NT + 63) / 64];
    __shared__ double total;
    // with E4 (the two-level preconditioner's restriction): each line's
    // share of its element's four corner-weighted sums, per corner c
    // (0 bottom-left, 1 bottom-right, 2 top-left, 3 top-right), held in
    // registers and summed over the lines in the tile LDS at the end
    constexpr bool FUSE = RESTR && rz_fuses_restrict(P, (int)sizeof(T));
    constexpr cx::Tables<P> TB = cx::make_tables<P>();

    T alpha = T(0);
    if (UPDATE) {
        if (st->s[slot].done | st->err) return;
        const double pAp = group_sum_all<NT>(strided_sum<NT>(ppAp, 0, np), red, &total);
        const double a = st->s[slot].rz / pAp;
        if (!(pAp > 0.0) || !isfinite(a)) {
            // every workgroup takes this branch (same p.Ap); the first
            // records it in a word that the p update reads afterwards
            if (blockIdx.x == 0 && threadIdx.x == 0) st->err = isfinite(pAp) ? 2 : 3;
            return;
        }
        alpha = (T)a;
    }
    constexpr cx::PrecTables<P> PT = cx::make_prec<P>();
    struct CS {
        const T *p;
        __device__ T operator()(int i, int a) const { return CX ? (T)PT.S[i][a] : p[i * N1 + a]; }
    } S{prec};
    struct CL {
        const T *p;
        __device__ T operator()(int a, int c) const { return CX ? (T)PT.Linv[a][c] : p[a * N1 + c]; }
    } Linv{prec + NQ};
    const int tid = threadIdx.x;
    const int s_el = tid % NE;
    int j = tid / NE;
    if (NE % 64 == 0) j = __builtin_amdgcn_readfirstlane(j);
    const int64_t nval = noe * NQ;
    const int64_t ngroups = (noe + NE - 1) / NE;
    double srr = 0.0, srz = 0.0;

    for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int64_t base = g * (int64_t)(NE * NQ);
        T rv[KR];
        double cv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int idx = tid + k * NT;
            const int64_t gi = base + idx;
            T v = T(0);
            if (idx < G::NV && gi < nval) {
                v = r[gi];
                if (UPDATE) {
                    v = v - alpha * Ap[gi];
                    r[gi] = v;
                }
                srr += (double)v * (double)v;
            }
            rv[k] = v;
            if (idx < G::NV) lds[lds_index<NQ, TS>(idx)] = v;
        }
        __syncthreads();
        T *tile = lds + s_el * TS;
        // 1: column j of S^T R
        {
            T c[N1], t[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) c[i] = tile[i * N1 + j];
            if (FUSE && E4) {
                // column j is the x line r(., j) at y node j: x weights here,
                // the y weight of line j, summed over the lines below
                double a0 = 0.0, a1 = 0.0;
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    const double xi = 0.5 * (TB.z[i] + 1.0);
                    a0 += (1.0 - xi) * (double)c[i];
                    a1 += xi * (double)c[i];
                }
                const double eta = 0.5 * (TB.z[j] + 1.0);
                cv[0] = (1.0 - eta) * a0;
                cv[1] = (1.0 - eta) * a1;
                cv[2] = eta * a0;
                cv[3] = eta * a1;
            }
#pragma unroll
            for (int a = 0; a < N1; ++a) {
                T acc = T(0);
#pragma unroll
                for (int i = 0; i < N1; ++i) acc += S(i, a) * c[i];
                t[a] = acc;
            }
#pragma unroll
            for (int a = 0; a < N1; ++a) tile[a * N1 + j] = t[a];
        }
        __syncthreads();
        // 2: row j of ((S^T R S) ./ L) S^T
        {
            T w[N1], v[N1];
#pragma unroll
            for (int b = 0; b < N1; ++b) w[b] = tile[j * N1 + b];
#pragma unroll
            for (int c = 0; c < N1; ++c) {
                T acc = T(0);
#pragma unroll
                for (int b = 0; b < N1; ++b) acc += w[b] * S(b, c);
                v[c] = acc * Linv(j, c);
            }
#pragma unroll
            for (int d = 0; d < N1; ++d) {
                T acc = T(0);
#pragma unroll
                for (int c = 0; c < N1; ++c) acc += v[c] * S(d, c);
                w[d] = acc;
            }
#pragma unroll
            for (int d = 0; d < N1; ++d) tile[j * N1 + d] = w[d];
        }
        __syncthreads();
        // 3: column j of S (...)
        {
            T c[N1], t[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) c[i] = tile[i * N1 + j];
#pragma unroll
            for (int a = 0; a < N1; ++a) {
                T acc = T(0);
#pragma unroll
                for (int i = 0; i < N1; ++i) acc += S(a, i) * c[i];
                t[a] = acc;
            }
#pragma unroll
            for (int a = 0; a < N1; ++a) tile[a * N1 + j] = t[a];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int idx = tid + k * NT;
            const int64_t gi = base + idx;
            if (idx < G::NV && gi < nval) {
                const T zv = lds[lds_index<NQ, TS>(idx)];
                z[gi] = zv;
                if (COPYP) pcopy[gi] = zv;
                srz += (double)rv[k] * (double)zv;
            }
        }
        __syncthreads();                               // LDS free for the next group
        if (FUSE && E4) {
            // E4[c * noe + e]: the lines' shares summed in a fixed order
            double *cs = reinterpret_cast<double *>(lds);
#pragma unroll
            for (int c = 0; c < 4; ++c) cs[(c * N1 + j) * NE + s_el] = cv[c];
            __syncthreads();
            for (int q = tid; q < 4 * NE; q += NT) {
                const int c = q / NE, el = q - (q / NE) * NE;
                const int64_t e = g * NE + el;
                double sum = 0.0;
#pragma unroll
                for (int jj = 0; jj < N1; ++jj) sum += cs[(c * N1 + jj) * NE + el];
                if (e < noe) E4[(int64_t)c * noe + e] = sum;
            }
            __syncthreads();
        }
    }
    const double a = block_sum<NT>(srr, red);
    __syncthreads();
    const double b = block_sum<NT>(srz, red);
    if (threadIdx.x == 0) {
        prr[blockIdx.x] = a;
        prz[blockIdx.x] = b;
    }
}

// ---- z = M^-1 r with wave-owned elements ----------------------------------
// The same preconditioner and pass structure with another work split: a
// wave holds EPW = 64 / (p+1) whole elements (7 at p = 8), lane (s, j) =
// s (p+1) + j on line j of element s, and streams runs of EPW consecutive
// elements on its own.  The transposes between the three contraction
// passes stay inside the wave, so the element loop has no workgroup barrier;
// with the standard penalty the contractions take the even-odd form of S
// (cx::PrecEvenOdd), half the multiplies of the plain one.
template <int P>
struct RzWave {
    static constexpr int N1 = P + 1, NQ = N1 * N1;
    static constexpr int EPW = 64 / N1;
    static constexpr int TS = tile_stride(NQ);
    static constexpr int NV = EPW * NQ;                 // values per run
    static constexpr int KR = (NV + 63) / 64;
    static constexpr int W = 4;                         // waves per workgroup
    static constexpr int NT = 64 * W;
};

// LDS writes of one pass are visible to every lane of the wave in the next
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Arguments and flags as pcg_rz_kernel; RESTR: with E4, the restriction's
// corner sums (for every degree: they are summed in an LDS array of their
// own).
template <typename T, int P, bool UPDATE, bool COPYP, bool CX, bool RESTR = false>
__global__ __launch_bounds__(RzWave<P>::NT)
void pcg_rz_wave_kernel(T *__restrict__ r, const T *__restrict__ Ap, T *__restrict__ z,
                        T *__restrict__ pcopy, int64_t noe, const T *__restrict__ prec,
                        const double *__restrict__ ppAp, int np, double *__restrict__ prr,
                        double *__restrict__ prz, cg_state *st, int slot, double *__restrict__ E4)
{
    using G = RzWave<P>;
    constexpr int N1 = G::N1, NQ = G::NQ, TS = G::TS, EPW = G::EPW, NV = G::NV, KR = G::KR, NT = G::NT;
    __shared__ __attribute__((aligned(16))) T lds[G::W * EPW * TS];
    __shared__ T linv[NQ];
    __shared__ double csum[RESTR ? G::W * 4 * N1 * EPW : 1];
    __shared__ double red[G::W];
    __shared__ double total;
    using TB = cx::Basis<P>;                           // TB::T.z: the nodes (a lane's own j: a load)

    T alpha = T(0);
    if (UPDATE) {
        if (st->s[slot].done | st->err) return;
        const double pAp = group_sum_all<NT>(strided_sum<NT>(ppAp, 0, np), red, &total);
        const double a = st->s[slot].rz / pAp;
        if (!(pAp > 0.0) || !isfinite(a)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) st->err = isfinite(pAp) ? 2 : 3;
            return;
        }
        alpha = (T)a;
    }
    // 1 / (lambda_a + lambda_c), read per lane (row j) from LDS
    for (int q = threadIdx.x; q < NQ; q += NT) {
        if constexpr (CX) linv[q] = (T)PrecEO<P>::V.Linv[q / N1][q % N1];
        else linv[q] = prec[NQ + q];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = lane / N1, j = lane - s * N1;
    const bool act = s < EPW;                          // lanes past the last element: copies, no writes
    T *wl = lds + w * (EPW * TS);
    T *tile = wl + (act ? s : EPW - 1) * TS;
    const T *lrow = linv + j * N1;
    const int64_t nval = noe * NQ;
    const int64_t nrun = (noe + EPW - 1) / EPW;
    double srr = 0.0, srz = 0.0;

    for (int64_t run = (int64_t)blockIdx.x * G::W + w; run < nrun; run += (int64_t)gridDim.x * G::W) {
        const int64_t base = run * NV;
        T rv[KR];
        double cv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int idx = lane + 64 * k;
            const int64_t gi = base + idx;
            T v = T(0);
            if (idx < NV && gi < nval) {
                v = r[gi];
                if (UPDATE) {
                    v = v - alpha * Ap[gi];
                    r[gi] = v;
                }
                srr += (double)v * (double)v;
            }
            rv[k] = v;
            if (idx < NV) wl[lds_index<NQ, TS>(idx)] = v;
        }
        wave_sync();
        // 1: column j of S^T R
        {
            T c[N1], t[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) c[i] = tile[i * N1 + j];
            if (RESTR && E4) {
                // column j is the x line r(., j) at y node j: x weights here,
                // the y weight of line j, summed over the lines below
                double a0 = 0.0, a1 = 0.0;
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    const double xi = 0.5 * (TB::T.z[i] + 1.0);
                    a0 += (1.0 - xi) * (double)c[i];
                    a1 += xi * (double)c[i];
                }
                const double eta = 0.5 * (TB::T.z[j] + 1.0);
                cv[0] = (1.0 - eta) * a0;
                cv[1] = (1.0 - eta) * a1;
                cv[2] = eta * a0;
                cv[3] = eta * a1;
            }
            contract_st<T, P, CX>(prec, c, t);
            if (act) {
#pragma unroll
                for (int a = 0; a < N1; ++a) tile[a * N1 + j] = t[a];
            }
        }
        wave_sync();
        // 2: row j of ((S^T R S) ./ L) S^T
        {
            T wv[N1], v[N1];
#pragma unroll
            for (int b = 0; b < N1; ++b) wv[b] = tile[j * N1 + b];
            contract_st<T, P, CX>(prec, wv, v);
#pragma unroll
            for (int c = 0; c < N1; ++c) v[c] *= lrow[c];
            contract_s<T, P, CX>(prec, v, wv);
            if (act) {
#pragma unroll
                for (int d = 0; d < N1; ++d) tile[j * N1 + d] = wv[d];
            }
        }
        wave_sync();
        // 3: column j of S (...)
        {
            T c[N1], t[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) c[i] = tile[i * N1 + j];
            contract_s<T, P, CX>(prec, c, t);
            if (act) {
#pragma unroll
                for (int a = 0; a < N1; ++a) tile[a * N1 + j] = t[a];
            }
        }
        wave_sync();
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int idx = lane + 64 * k;
            const int64_t gi = base + idx;
            if (idx < NV && gi < nval) {
                const T zv = wl[lds_index<NQ, TS>(idx)];
                z[gi] = zv;
                if (COPYP) pcopy[gi] = zv;
                srz += (double)rv[k] * (double)zv;
            }
        }
        if (RESTR && E4) {
            // E4[c * noe + e]: the lines' shares summed in a fixed order
            double *cs = csum + w * (4 * N1 * EPW);
            if (act) {
#pragma unroll
                for (int c = 0; c < 4; ++c) cs[(c * N1 + j) * EPW + s] = cv[c];
            }
            wave_sync();
            for (int q = lane; q < 4 * EPW; q += 64) {
                const int c = q / EPW, el = q - c * EPW;
                const int64_t e = run * EPW + el;
                double sum = 0.0;
#pragma unroll
                for (int jj = 0; jj < N1; ++jj) sum += cs[(c * N1 + jj) * EPW + el];
                if (e < noe) E4[(int64_t)c * noe + e] = sum;
            }
        }
        wave_sync();                                   // the wave's LDS is free for its next run
    }
    const double a = block_sum<NT>(srr, red);
    __syncthreads();
    const double b = block_sum<NT>(srz, red);
    if (threadIdx.x == 0) {
        prr[blockIdx.x] = a;
        prz[blockIdx.x] = b;
    }
}

// Which split z = M^-1 r takes at degree p with es bytes per value: the
// wave-owned one where it measured faster (z = M^-1 r over ~5.3M values on
// MI355X, FP64: 23.9 against 28.5 us at p = 8, 31.3 against 68.5 at
// p = 16; FP32 likewise at p = 4, 6 and from 8), the workgroup-wide one
// below (up to 15 % faster there: p <= 3, and FP32 at p = 5, 7).
__host__ __device__ constexpr bool rz_by_wave(int p, int es) {
    return es == 8 ? p >= 4 : (p == 4 || p == 6 || p >= 8);
}

// Workgroups of one pcg_rz kernel that the device holds at once, cached per
// device; the grid is min(groups, that, kMaxPartials).
template <typename T, int P, bool UPDATE, bool COPYP, bool CX>
int pcg_resident() {
    constexpr int kMaxDev = 64;
    static std::atomic<int> cache[kMaxDev];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) dev = 0;
    int g = cache[dev].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    hipError_t e;
    if constexpr (rz_by_wave(P, (int)sizeof(T)))
        // the instantiation that also restricts (more registers) where it exists
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&pcg_rz_wave_kernel<T, P, UPDATE, COPYP, CX, CX>), RzWave<P>::NT, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&pcg_rz_kernel<T, P, UPDATE, COPYP, CX>), PcgShape<P>::NT, 0);
    if (e != hipSuccess) per = 0;
    g = (cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    cache[dev].store(g, std::memory_order_relaxed);
    return g;
}

// The wave-owned split: W runs of EPW elements per workgroup and pass, in
// as many passes as it takes for the grid to fit on the device at once
// (at most kMaxPartials workgroups), with the runs spread evenly over the
// passes.  Fewer workgroups also means fewer block sums for the kernels
// that reduce them afterwards (each of their workgroups reads all of them).
template <typename T, int P, bool UPDATE, bool COPYP, bool CX>
int pcg_grid(int64_t noe) {
    if constexpr (rz_by_wave(P, (int)sizeof(T))) {
        using G = RzWave<P>;
        const int64_t runs = (noe + G::EPW - 1) / G::EPW;
        const int64_t wgs = std::max<int64_t>(1, (runs + G::W - 1) / G::W);
        const int64_t cap = std::min<int64_t>(pcg_resident<T, P, UPDATE, COPYP, CX>(), kMaxPartials);
        const int64_t rounds = (wgs + cap - 1) / cap;
        return (int)((wgs + rounds - 1) / rounds);
    } else {
        const int64_t groups = (noe + PcgShape<P>::NE - 1) / PcgShape<P>::NE;
        return (int)std::min<int64_t>(std::min<int64_t>(groups, pcg_resident<T, P, UPDATE, COPYP, CX>()),
                                      kMaxPartials);
    }
}

// grid: that of the plain kernel (the block-sum count the solver expects);
// with E4 on a degree where it fits, the restricting instantiation
template <typename T, int P, bool UPDATE, bool COPYP, bool CX>
void rz_launch(dim3 grid, void *r, const void *Ap, void *z, void *pcopy, int64_t noe, const void *prec,
               const double *ppAp, int np, double *prr, double *prz, cg_state *st, int slot, double *E4,
               hipStream_t s) {
    if constexpr (rz_by_wave(P, (int)sizeof(T))) {
        if (CX && E4)
            hipLaunchKernelGGL((pcg_rz_wave_kernel<T, P, UPDATE, COPYP, CX, CX>), grid, dim3(RzWave<P>::NT), 0, s,
                               (T *)r, (const T *)Ap, (T *)z, (T *)pcopy, noe, (const T *)prec, ppAp, np, prr, prz,
                               st, slot, E4);
        else
            hipLaunchKernelGGL((pcg_rz_wave_kernel<T, P, UPDATE, COPYP, CX>), grid, dim3(RzWave<P>::NT), 0, s,
                               (T *)r, (const T *)Ap, (T *)z, (T *)pcopy, noe, (const T *)prec, ppAp, np, prr, prz,
                               st, slot, nullptr);
        return;
    }
    if (CX && E4 && rz_fuses_restrict(P, (int)sizeof(T)))
        hipLaunchKernelGGL((pcg_rz_kernel<T, P, UPDATE, COPYP, CX, CX>), grid, dim3(PcgShape<P>::NT), 0, s, (T *)r,
                           (const T *)Ap, (T *)z, (T *)pcopy, noe, (const T *)prec, ppAp, np, prr, prz, st, slot, E4);
    else
        hipLaunchKernelGGL((pcg_rz_kernel<T, P, UPDATE, COPYP, CX>), grid, dim3(PcgShape<P>::NT), 0, s, (T *)r,
                           (const T *)Ap, (T *)z, (T *)pcopy, noe, (const T *)prec, ppAp, np, prr, prz, st, slot,
                           nullptr);
}

// The standard penalty 5 (p+1)^2 takes the kernel with compile-time
// coefficients; any other the one that reads the op's tables.
inline bool standard_pen(int p, double pen) { return pen == 5.0 * (p + 1) * (p + 1); }

template <typename T, int P>
int launch_rz_one(int mode, bool cx, void *r, const void *Ap, void *z, void *pcopy, int64_t noe,
                  const void *prec, const double *ppAp, int np, double *prr, double *prz, cg_state *st,
                  int slot, double *E4, hipStream_t s)
{
    if (cx) {
        if (mode == 1)
            rz_launch<T, P, true, false, true>(dim3(pcg_grid<T, P, true, false, true>(noe)), r, Ap, z, pcopy,
                                               noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
        else if (mode == 2)
            rz_launch<T, P, false, true, true>(dim3(pcg_grid<T, P, false, true, true>(noe)), r, Ap, z, pcopy,
                                               noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
        else
            rz_launch<T, P, false, false, true>(dim3(pcg_grid<T, P, false, false, true>(noe)), r, Ap, z,
                                                pcopy, noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
    } else {
        if (mode == 1)
            rz_launch<T, P, true, false, false>(dim3(pcg_grid<T, P, true, false, false>(noe)), r, Ap, z,
                                                pcopy, noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
        else if (mode == 2)
            rz_launch<T, P, false, true, false>(dim3(pcg_grid<T, P, false, true, false>(noe)), r, Ap, z,
                                                pcopy, noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
        else
            rz_launch<T, P, false, false, false>(dim3(pcg_grid<T, P, false, false, false>(noe)), r, Ap, z,
                                                 pcopy, noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

template <typename T, int P>
int grid_one(int mode, bool cx, int64_t noe) {
    if (cx)
        return mode == 1 ? pcg_grid<T, P, true, false, true>(noe)
                         : mode == 2 ? pcg_grid<T, P, false, true, true>(noe)
                                     : pcg_grid<T, P, false, false, true>(noe);
    return mode == 1 ? pcg_grid<T, P, true, false, false>(noe)
                     : mode == 2 ? pcg_grid<T, P, false, true, false>(noe)
                                 : pcg_grid<T, P, false, false, false>(noe);
}

template <typename T, int... Ps>
struct PcgDispatch;

template <typename T>
struct PcgDispatch<T> {
    static int launch(int, int, bool, void *, const void *, void *, void *, int64_t, const void *,
                      const double *, int, double *, double *, cg_state *, int, double *, hipStream_t) {
        return SIPG_ERR_UNSUPPORTED;
    }
    static int grid(int, int, bool, int64_t) { return -1; }
};

template <typename T, int P0, int... Ps>
struct PcgDispatch<T, P0, Ps...> {
    static int launch(int p, int mode, bool cx, void *r, const void *Ap, void *z, void *pcopy, int64_t noe,
                      const void *prec, const double *ppAp, int np, double *prr, double *prz, cg_state *st,
                      int slot, double *E4, hipStream_t s) {
        if (p == P0)
            return launch_rz_one<T, P0>(mode, cx, r, Ap, z, pcopy, noe, prec, ppAp, np, prr, prz, st, slot, E4, s);
        return PcgDispatch<T, Ps...>::launch(p, mode, cx, r, Ap, z, pcopy, noe, prec, ppAp, np, prr, prz, st,
                                             slot, E4, s);
    }
    static int grid(int p, int mode, bool cx, int64_t noe) {
        if (p == P0) return grid_one<T, P0>(mode, cx, noe);
        return PcgDispatch<T, Ps...>::grid(p, mode, cx, noe);
    }
};

template <typename T>
using PcgDegrees = PcgDispatch<T, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>;

constexpr int VT = kVecThreads;

__device__ __forceinline__ double vsum_all(double v) {
    __shared__ double red[VT / 64];
    __shared__ double total;
    return group_sum_all<VT>(v, red, &total);
}

__device__ __forceinline__ double vreduce(const double *__restrict__ part, int np) {
    return vsum_all(strided_sum<VT>(part, 0, np));
}

// The step after z = M^-1 r: beta = r.z(new) / r.z(old) and p = z + beta p;
// unless this iteration's r was replaced by the true residual (whose x
// update ran before it), also x += alpha p with the old p.  Block 0 writes
// the next slot: rr = r.r (the stopping test and the host's decisions),
// rz = r.z.  Flag semantics are those of the CG p update.
template <typename T>
__global__ __launch_bounds__(VT) void pcg_update_p_kernel(T *__restrict__ x, T *__restrict__ p,
                                                          const T *__restrict__ z, int64_t n,
                                                          const double *__restrict__ ppAp, int np_pAp,
                                                          const double *__restrict__ prr,
                                                          const double *__restrict__ prz, int np,
                                                          cg_state *st, int slot, T *__restrict__ xlo) {
    const cg_slot cur = st->s[slot];
    if (cur.done | st->err) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->s[slot ^ 1] = cur;   // carry forward
        return;
    }
    const double rr_new = vreduce(prr, np);
    const double rz_new = vreduce(prz, np);
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
        p[i] = z[i] + beta * pi;
    }
}

// x += alpha p (alpha = slot.rz / p.Ap): the x half of the step, run before
// a true residual replaces r in the same iteration.
template <typename T>
__global__ __launch_bounds__(VT) void pcg_x_kernel(T *__restrict__ x, const T *__restrict__ p, int64_t n,
                                                   const double *__restrict__ ppAp, int np,
                                                   const cg_state *st, int slot, T *__restrict__ xlo) {
    if (st->s[slot].done | st->err) return;
    const double pAp = vreduce(ppAp, np);
    const T alpha = (T)(st->s[slot].rz / pAp);
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VT) {
        if (xlo) {
            T xh = x[i], xl = xlo[i];
            comp_axpy(xh, xl, alpha, p[i]);
            x[i] = xh;
            xlo[i] = xl;
        } else {
            x[i] = x[i] + alpha * p[i];
        }
    }
}

// Initial state: r.r and r.z from the preconditioner's block sums, b.b
// from the true residual's.
__global__ __launch_bounds__(VT) void pcg_init_finalize_kernel(const double *__restrict__ prr,
                                                               const double *__restrict__ prz, int np,
                                                               const double *__restrict__ pbb, int npb,
                                                               cg_state *st, double tol, int tol_mode,
                                                               int maxit, int true_stop,
                                                               const double *__restrict__ pc, int npc) {
    const double rr = vreduce(prr, np);
    const double rz = vreduce(prz, np) + vreduce(pc, npc);     // + S.C of a coarse level
    const double bb = vreduce(pbb, npb);
    if (threadIdx.x == 0) {
        const double stop = tol_mode == SIPG_TOL_REL ? tol * sqrt(bb) : tol;
        cg_slot s0;
        s0.rr = rr;
        s0.alpha = s0.beta = s0.pAp = 0.0;
        s0.rz = rz;
        s0.it = 0;
        s0.done = (sqrt(rr) > stop && maxit > 0) ? 0 : 1;
        s0.checked = 1;
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
        st->err = (!isfinite(rr) || !isfinite(bb) || !isfinite(rz)) ? 3 : 0;
        st->true_stop = true_stop;
    }
}

}  // namespace

int pcg_num_blocks(int dtype, int p, double pen, int mode, int64_t noe) {
    const bool cx = standard_pen(p, pen);
    return dtype == SIPG_F64 ? PcgDegrees<double>::grid(p, mode, cx, noe)
                             : PcgDegrees<float>::grid(p, mode, cx, noe);
}

bool pcg_rz_fuses_restrict(int dtype, int p, double pen) {
    const int es = dtype == SIPG_F64 ? 8 : 4;
    return p >= 1 && p <= kMaxDegree && standard_pen(p, pen) && (rz_by_wave(p, es) || rz_fuses_restrict(p, es));
}

int launch_pcg_rz(int dtype, int p, double pen, int mode, void *r, const void *Ap, void *z, void *pcopy,
                  int64_t noe, const void *prec, const double *part_pAp, int np_pAp, double *part_rr,
                  double *part_rz, cg_state *st, int slot, hipStream_t s, double *E4) {
    const bool cx = standard_pen(p, pen);
    if (dtype == SIPG_F64)
        return PcgDegrees<double>::launch(p, mode, cx, r, Ap, z, pcopy, noe, prec, part_pAp, np_pAp, part_rr,
                                          part_rz, st, slot, E4, s);
    return PcgDegrees<float>::launch(p, mode, cx, r, Ap, z, pcopy, noe, prec, part_pAp, np_pAp, part_rr,
                                     part_rz, st, slot, E4, s);
}

int launch_pcg_update_p(int dtype, void *x, void *p, const void *z, int64_t n, const double *part_pAp,
                        int np_pAp, const double *part_rr, const double *part_rz, int np, int nblk,
                        cg_state *st, int slot, hipStream_t s, void *xlo) {
    if (dtype == SIPG_F64)
        hipLaunchKernelGGL(pcg_update_p_kernel<double>, dim3(nblk), dim3(VT), 0, s, (double *)x, (double *)p,
                           (const double *)z, n, part_pAp, np_pAp, part_rr, part_rz, np, st, slot, (double *)xlo);
    else
        hipLaunchKernelGGL(pcg_update_p_kernel<float>, dim3(nblk), dim3(VT), 0, s, (float *)x, (float *)p,
                           (const float *)z, n, part_pAp, np_pAp, part_rr, part_rz, np, st, slot, (float *)xlo);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_pcg_x(int dtype, void *x, const void *p, int64_t n, const double *part_pAp, int np_pAp, int nblk,
                 const cg_state *st, int slot, hipStream_t s, void *xlo) {
    if (dtype == SIPG_F64)
        hipLaunchKernelGGL(pcg_x_kernel<double>, dim3(nblk), dim3(VT), 0, s, (double *)x, (const double *)p, n,
                           part_pAp, np_pAp, st, slot, (double *)xlo);
    else
        hipLaunchKernelGGL(pcg_x_kernel<float>, dim3(nblk), dim3(VT), 0, s, (float *)x, (const float *)p, n,
                           part_pAp, np_pAp, st, slot, (float *)xlo);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_pcg_init_finalize(const double *part_rr, const double *part_rz, int np, const double *part_bb,
                             int npb, cg_state *st, double tol, int tol_mode, int maxit, int true_stop,
                             hipStream_t s, const double *part_c, int npc) {
    hipLaunchKernelGGL(pcg_init_finalize_kernel, dim3(1), dim3(VT), 0, s, part_rr, part_rz, np, part_bb, npb,
                       st, tol, tol_mode, maxit, true_stop, part_c, npc);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace sipg_internal

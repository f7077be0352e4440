// kernels_generic.hip -- the SIPG operator y = A x for any degree up to
// kMaxGenericDegree (63, the reference's 64-point LGL tables,
// device_const.hpp:10-18), on run-time tables: the degrees above the
// compiled fused kernels (volume_gpu_kernels.hpp:9-86 + flux_term6a/6b,
// sipg_sem_2d_gpu_kernels.hpp:189-316).
//
// Two launches per apply.  (1) face_derivs: per element and line j the
// four normal derivatives on its sides, dx0[j] = sum_i D0[i] u(i,j),
// dxN[j] = sum_i DN[i] u(i,j), dy0[j] = sum_i D0[i] u(j,i), dyN[j] = sum_i
// DN[i] u(j,i) (D0[i] = phi_i'(-1), DN[i] = phi_i'(+1)), into a face buffer
// of 4 (p+1) values per element.  (2) apply: per element, the tile in LDS,
// the neighbours' traces from x and their derivatives from the face buffer,
// then every output
//   Y(i1,i2) = sum_i w[i2] K1[i][i1] U(i,i2) + w[i1] K1[i][i2] U(i1,i)
//            + D0[i1] ev[i2] + DN[i1] fv[i2] + D0[i2] g[i1] + DN[i2] h[i1]
//            + side corrections on the element's own boundary lines,
// with ev/fv/g/h the weighted jumps on the left/right/lower/upper sides
// and a/b/c/d the averaged normal derivatives (the reference's flux terms).
// One workgroup per element in a grid-stride loop over a capped grid, so
// that the fused p.Ap partials stay within the partial buffer.
#include "apply_block.hpp"

namespace sipg_internal {

namespace {

constexpr int kGenThreads = 256;
constexpr int kGenMaxBlocks = 2048;

// run-time table views (the op's table layout: w, K1, D0, DP, z)
template <typename T>
struct RtTab {
    const T *w, *K1, *D0, *DN;
    __device__ RtTab(const T *tab, int n1)
        : w(tab), K1(tab + n1), D0(tab + n1 + n1 * n1), DN(tab + 2 * n1 + n1 * n1) {}
};

template <typename T>
__global__ __launch_bounds__(kGenThreads) void face_derivs_kernel(int n1, const T *__restrict__ x,
                                                                  const T *__restrict__ tab, T *__restrict__ face,
                                                                  int64_t noe, const cg_state *__restrict__ st,
                                                                  int slot)
{
    if (st && (st->s[slot].done | st->err)) return;
    extern __shared__ char smem[];
    T *U = reinterpret_cast<T *>(smem);
    const RtTab<T> t(tab, n1);
    const int nq = n1 * n1;
    for (int64_t e = blockIdx.x; e < noe; e += gridDim.x) {
        const T *ue = x + e * nq;
        for (int k = threadIdx.x; k < nq; k += kGenThreads) U[k] = ue[k];
        __syncthreads();
        T *fe = face + e * 4 * n1;
        for (int k = threadIdx.x; k < 4 * n1; k += kGenThreads) {
            const int side = k / n1, j = k - side * n1;
            const T *d = (side & 1) ? t.DN : t.D0;
            T s = 0;
            if (side < 2)
                for (int i = 0; i < n1; ++i) s += d[i] * U[i * n1 + j];     // along x: line j
            else
                for (int i = 0; i < n1; ++i) s += d[i] * U[j * n1 + i];     // along y: line j
            fe[k] = s;
        }
        __syncthreads();
    }
}

// face buffer of element e: [dx0 | dxN | dy0 | dyN], n1 each
template <typename T, bool DOT>
__global__ __launch_bounds__(kGenThreads) void apply_generic_kernel(int n1, const T *__restrict__ x,
                                                                    T *__restrict__ y, const T *__restrict__ tab,
                                                                    const T *__restrict__ face, int dimx,
                                                                    int dimy, T pen, double *__restrict__ partials,
                                                                    const cg_state *__restrict__ st, int slot)
{
    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }
    extern __shared__ char smem[];
    const int nq = n1 * n1, N = n1 - 1;
    T *U = reinterpret_cast<T *>(smem);
    T *ev = U + nq, *fv = ev + n1, *g = fv + n1, *h = g + n1;
    T *a = h + n1, *b = a + n1, *c = b + n1, *d = c + n1;
    __shared__ double red[kGenThreads / 64];
    const RtTab<T> t(tab, n1);
    const int64_t noe = (int64_t)dimx * dimy;
    double dot = 0.0;
    for (int64_t e = blockIdx.x; e < noe; e += gridDim.x) {
        const int col = (int)(e % dimx), row = (int)(e / dimx);
        const bool Lb = col == 0, Rb = col == dimx - 1, Db = row == dimy - 1, Ub = row == 0;
        const T *ue = x + e * nq;
        for (int k = threadIdx.x; k < nq; k += kGenThreads) U[k] = ue[k];
        __syncthreads();
        // per line j and side: the weighted jump of the traces (ev, fv, g,
        // h) and the averaged normal derivative (a, b, c, d); a boundary
        // side takes the one-sided Dirichlet form
        const T half = T(0.5);
        const T *fe = face + e * 4 * n1;
        for (int k = threadIdx.x; k < 4 * n1; k += kGenThreads) {
            const int side = k / n1, j = k - side * n1;
            const T wj = t.w[j];
            if (side == 0) {                          // left, x = 0 side of the element
                const T own = U[j], dn = fe[j];
                if (Lb) {
                    ev[j] = wj * own;
                    a[j] = dn;
                } else {
                    const T *fl = face + (e - 1) * 4 * n1;
                    ev[j] = wj * (half * (own - x[(e - 1) * nq + N * n1 + j]));
                    a[j] = half * (dn + fl[n1 + j]);
                }
            } else if (side == 1) {                   // right
                const T own = U[N * n1 + j], dn = fe[n1 + j];
                if (Rb) {
                    fv[j] = wj * -own;
                    b[j] = -dn;
                } else {
                    const T *fr = face + (e + 1) * 4 * n1;
                    fv[j] = wj * (half * (x[(e + 1) * nq + j] - own));
                    b[j] = -half * (dn + fr[j]);
                }
            } else if (side == 2) {                   // lower (y = 0 side): the next row down
                const T own = U[j * n1], dn = fe[2 * n1 + j];
                if (Db) {
                    g[j] = wj * own;
                    c[j] = dn;
                } else {
                    const int64_t en = e + dimx;
                    const T *fd = face + en * 4 * n1;
                    g[j] = wj * (half * (own - x[en * nq + j * n1 + N]));
                    c[j] = half * (dn + fd[3 * n1 + j]);
                }
            } else {                                  // upper: the row above
                const T own = U[j * n1 + N], dn = fe[3 * n1 + j];
                if (Ub) {
                    h[j] = wj * -own;
                    d[j] = -dn;
                } else {
                    const int64_t eu = e - dimx;
                    const T *fu = face + eu * 4 * n1;
                    h[j] = wj * (half * (x[eu * nq + j * n1] - own));
                    d[j] = -half * (dn + fu[2 * n1 + j]);
                }
            }
        }
        __syncthreads();
        const T penL = Lb ? pen : 2 * pen, penR = Rb ? pen : 2 * pen;
        const T penD = Db ? pen : 2 * pen, penU = Ub ? pen : 2 * pen;
        T *ye = y + e * nq;
        for (int k = threadIdx.x; k < nq; k += kGenThreads) {
            const int i1 = k / n1, i2 = k - i1 * n1;
            T vx = 0, vy = 0;
            for (int i = 0; i < n1; ++i) {
                vx += t.K1[i * n1 + i1] * U[i * n1 + i2];
                vy += t.K1[i * n1 + i2] * U[i1 * n1 + i];
            }
            T r = t.w[i2] * vx + t.w[i1] * vy;
            r += t.D0[i1] * ev[i2] + t.DN[i1] * fv[i2] + t.D0[i2] * g[i1] + t.DN[i2] * h[i1];
            if (i1 == 0) r += t.w[i2] * a[i2] + penL * ev[i2];
            if (i1 == N) r += t.w[i2] * b[i2] - penR * fv[i2];
            if (i2 == 0) r += t.w[i1] * c[i1] + penD * g[i1];
            if (i2 == N) r += t.w[i1] * d[i1] - penU * h[i1];
            ye[k] = r;
            if (DOT) dot += (double)U[k] * (double)r;
        }
        __syncthreads();
    }
    if (DOT) {
        const double s = block_sum<kGenThreads>(dot, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = s;
    }
}

int grid_of(int64_t noe) { return (int)std::min<int64_t>(noe, kGenMaxBlocks); }

template <typename T>
int launch_t(int p, const void *x, void *y, const void *tab, int dimx, int dimy, double pen, double *partials,
             const cg_state *st, int slot, hipStream_t s)
{
    const int n1 = p + 1, nq = n1 * n1;
    const int64_t noe = (int64_t)dimx * dimy;
    const T *tb = (const T *)tab;
    // the face buffer follows the tables in the op's table allocation
    T *face = const_cast<T *>(tb) + generic_face_offset(p, sizeof(T));
    const int grid = grid_of(noe);
    hipLaunchKernelGGL(face_derivs_kernel<T>, dim3(grid), dim3(kGenThreads), (size_t)nq * sizeof(T), s, n1,
                       (const T *)x, tb, face, noe, partials ? st : nullptr, slot);
    const size_t lds = (size_t)(nq + 8 * n1) * sizeof(T);
    if (partials)
        hipLaunchKernelGGL((apply_generic_kernel<T, true>), dim3(grid), dim3(kGenThreads), lds, s, n1,
                           (const T *)x, (T *)y, tb, (const T *)face, dimx, dimy, (T)pen, partials, st, slot);
    else
        hipLaunchKernelGGL((apply_generic_kernel<T, false>), dim3(grid), dim3(kGenThreads), lds, s, n1,
                           (const T *)x, (T *)y, tb, (const T *)face, dimx, dimy, (T)pen, partials, st, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace

size_t generic_face_offset(int p, size_t elem) { return ((size_t)table_len(p) * elem + 255) / 256 * 256 / elem; }

size_t generic_table_len(int dtype, int p, int dimx, int dimy) {
    const size_t elem = dtype == SIPG_F64 ? 8 : 4;
    return generic_face_offset(p, elem) + (size_t)4 * (p + 1) * dimx * dimy;
}

int generic_apply_num_blocks(int dimx, int dimy) { return grid_of((int64_t)dimx * dimy); }

int launch_apply_generic(int dtype, int p, const void *x, void *y, const void *tab, int dimx, int dimy, double pen,
                         double *partials, const cg_state *st, int slot, hipStream_t s)
{
    if (p <= kMaxDegree || p > kMaxGenericDegree) return SIPG_ERR_UNSUPPORTED;
    return dtype == SIPG_F64 ? launch_t<double>(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s)
                             : launch_t<float>(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
}

}  // namespace sipg_internal

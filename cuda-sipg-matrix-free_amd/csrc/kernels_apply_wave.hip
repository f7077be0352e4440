// kernels_apply_wave.hip -- the fused SIPG operator apply y = A x with
// wave-owned elements (volume + flux_term6a + flux_term6b of the reference,
// volume_gpu_kernels.hpp:9-86, sipg_sem_2d_gpu_kernels.hpp:189-316).
//
// Same operator and same per-line arithmetic as apply_kernel / _lds
// (line_operator in apply_block.hpp), another work split: a wave holds
// EPW = 64 / (p+1) whole elements (7 at p = 8), one block row, with lane
// (s, j) = s (p+1) + j on line j of element s.  A workgroup of BY waves
// owns an EPW x BY block, staged in LDS with its halo ring by direct
// global -> LDS loads.  Because each element's lines sit in one wave, the
// column-to-row assembly of the output tile and the copy-out of the block
// row are wave-local: a block costs two workgroup barriers (the staged
// block is complete; every wave is done reading it), against four when an
// element's lines are spread over the workgroup's waves.
//
// Launch: one block per workgroup, one LDS image per workgroup, so several
// workgroups share a CU and one computes while another streams its block
// in.  Block -> workgroup mapping is XCD-aware as in the other operator
// kernels.  Even degrees (unpadded tiles, which the direct loads need).
#include "apply_block.hpp"

#include <algorithm>

namespace sipg_internal {

template <typename T, int P, int BY, bool DOT, class C>
__global__ __launch_bounds__(64 * BY)
void apply_kernel_wave(const T *__restrict__ x, T *__restrict__ y, int dimx, int dimy, int nbx, int nblk, T pen,
                       double *__restrict__ partials, const cg_state *__restrict__ st, int slot)
{
    constexpr int N1 = P + 1, NQ = N1 * N1, BX = 64 / N1;
    using G = BlockGeo<T, P, BX, BY>;
    static_assert(G::TS == NQ, "wave-owned elements need unpadded tiles (even p)");
    __shared__ T lds[G::LDS];
    __shared__ double red[BY];

    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }
    const int b = xcd_remap(blockIdx.x, nblk);
    const int x0 = (b % nbx) * BX, y0 = (b / nbx) * BY;
    block_issue<T, P, BX, BY, BY>(lds, x, dimx, dimy, x0, y0);      // every wave issues

    const C coef{};
    const int lane = threadIdx.x & 63;
    const int ey = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = lane / N1, j = lane - s * N1;
    const int ex = s < BX ? s : BX - 1;                 // lanes past the last element: copies of it
    const int gx = x0 + ex, gy = y0 + ey;
    const bool valid = s < BX && gx < dimx && gy < dimy;
    const bool Lb = gx == 0, Rb = gx == dimx - 1;
    const bool Ub = gy == 0, Db = gy == dimy - 1;       // y grows downwards
    const int sl = (ey + 1) * BX + ex;
    T *tile = lds + sl * NQ;
    const T *tlL = lds + (ex > 0 ? sl - 1 : G::SL_COL + ey) * NQ;
    const T *tlR = lds + (ex < BX - 1 ? sl + 1 : G::SL_COL + BY + ey) * NQ;
    const T *tlD = lds + (sl + BX) * NQ;
    const T *tlU = lds + (sl - BX) * NQ;

    block_issue_wait();
    __syncthreads();                                    // the staged block is complete
    T v[N1], Yx[N1], Yy[N1];
    line_operator<T, P, C>(tile, tlL, tlR, tlD, tlU, j, Lb, Rb, Db, Ub, pen, Yx, Yy, v, coef);
    __syncthreads();                                    // every wave is done reading the block

    // wave-local from here: the element's tile takes the column results,
    // then each lane adds its row results to its row
    if (valid) {
#pragma unroll
        for (int a = 0; a < N1; ++a) tile[a * N1 + j] = Yx[a];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double dot = 0.0;
    if (valid) {
#pragma unroll
        for (int c = 0; c < N1; ++c) {
            const T yv = tile[j * N1 + c] + Yy[c];
            tile[j * N1 + c] = yv;
            if (DOT) dot += (double)v[c] * (double)yv;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // this wave's block row is one contiguous run of ncol tiles in y
    if (gy < dimy) {
        const int ncol = dimx - x0 < BX ? dimx - x0 : BX;
        const int n = ncol * NQ;
        const T *src = lds + (ey + 1) * BX * NQ;
        T *dst = y + ((size_t)gy * dimx + x0) * NQ;
#pragma unroll
        for (int k = 0; k < (BX * NQ + 63) / 64; ++k) {
            const int idx = lane + 64 * k;
            if (idx < n) dst[idx] = src[idx];
        }
    }
    if (DOT) {
        const double bs = block_sum<64 * BY>(dot, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    }
}

namespace {

// Rows of elements per workgroup (waves), one LDS image of (BY + 2) * BX +
// 2 BY tiles per workgroup, and where the coefficients come from.
// Measured at 256 x 256 and 1024 x 1024, p = 8, FP64 (us per apply):
// BY = 2: 22.2 / 281, 3: 21.7 / 277, 4: 21.1 / 277, 6: 24.3 / 301,
// 8: 22.8 / 297, 10: 26.9 / 374.
template <typename T, int P>
struct WaveShape {
    static constexpr int BY = 4;
    using Coef = CxCoef<T, P>;
};

template <typename T, int P, bool DOT>
void launch_wave(const void *x, void *y, int dimx, int dimy, double pen, double *partials, const cg_state *st,
                 int slot, hipStream_t s)
{
    constexpr int BX = 64 / (P + 1), BY = WaveShape<T, P>::BY;
    const int nbx = (dimx + BX - 1) / BX, nby = (dimy + BY - 1) / BY;
    const int nblk = nbx * nby;
    hipLaunchKernelGGL((apply_kernel_wave<T, P, BY, DOT, typename WaveShape<T, P>::Coef>), dim3(nblk),
                       dim3(64 * BY), 0, s, (const T *)x, (T *)y, dimx, dimy, nbx, nblk, (T)pen, partials, st, slot);
}

template <typename T, int P>
int wave_blocks(int dimx, int dimy) {
    constexpr int BX = 64 / (P + 1), BY = WaveShape<T, P>::BY;
    return ((dimx + BX - 1) / BX) * ((dimy + BY - 1) / BY);
}

template <typename T>
int dispatch(int p, const void *x, void *y, int dimx, int dimy, double pen, double *partials, const cg_state *st,
             int slot, hipStream_t s)
{
    switch (p) {
#define SIPG_WAVE_CASE(PP)                                                                 \
    case PP:                                                                               \
        if (partials) launch_wave<T, PP, true>(x, y, dimx, dimy, pen, partials, st, slot, s); \
        else launch_wave<T, PP, false>(x, y, dimx, dimy, pen, partials, st, slot, s);         \
        return SIPG_OK;
        SIPG_WAVE_CASE(2)
        SIPG_WAVE_CASE(4)
        SIPG_WAVE_CASE(6)
        SIPG_WAVE_CASE(8)
#undef SIPG_WAVE_CASE
    default:
        return SIPG_ERR_UNSUPPORTED;
    }
}

template <typename T>
int blocks(int p, int dimx, int dimy) {
    switch (p) {
    case 2: return wave_blocks<T, 2>(dimx, dimy);
    case 4: return wave_blocks<T, 4>(dimx, dimy);
    case 6: return wave_blocks<T, 6>(dimx, dimy);
    case 8: return wave_blocks<T, 8>(dimx, dimy);
    default: return -1;
    }
}

}  // namespace

bool wave_apply_available(int dtype, int p) {
    return (dtype == SIPG_F64 || dtype == SIPG_F32) && (p == 2 || p == 4 || p == 6 || p == 8);
}

int wave_apply_num_blocks(int dtype, int p, int dimx, int dimy) {
    if (!wave_apply_available(dtype, p)) return -1;
    return dtype == SIPG_F64 ? blocks<double>(p, dimx, dimy) : blocks<float>(p, dimx, dimy);
}

int launch_apply_wave(int dtype, int p, const void *x, void *y, int dimx, int dimy, double pen, double *partials,
                      const cg_state *st, int slot, hipStream_t s)
{
    if (!wave_apply_available(dtype, p)) return SIPG_ERR_UNSUPPORTED;
    const int rc = dtype == SIPG_F64 ? dispatch<double>(p, x, y, dimx, dimy, pen, partials, st, slot, s)
                                     : dispatch<float>(p, x, y, dimx, dimy, pen, partials, st, slot, s);
    if (rc) return rc;
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace sipg_internal

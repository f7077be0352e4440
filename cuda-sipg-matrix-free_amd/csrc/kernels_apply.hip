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
#include "apply_block.hpp"

#include <algorithm>
#include <atomic>

namespace sipg_internal {

// Persistent walk over the blocks of one mesh: the workgroups split the
// blocks into contiguous ranges (consecutive blocks are horizontal
// neighbours), and each workgroup loads block b+1 into registers while it
// computes block b.
template <typename T, int P, int BX, int BY, bool DOT>
__global__ __launch_bounds__(BX * BY * (P + 1))
void apply_kernel(const T *__restrict__ x, T *__restrict__ y, int dimx, int dimy, int nbx,
                  int nblk, T pen, double *__restrict__ partials,
                  const cg_state *__restrict__ st, int slot)
{
    using G = BlockGeo<T, P, BX, BY>;
    __shared__ T lds[G::LDS];
    __shared__ double red[(G::NT + 63) / 64];

    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }
    const int NG = gridDim.x;
    const int lw = xcd_remap(blockIdx.x, NG);
    const int bbeg = (int)((int64_t)lw * nblk / NG);
    const int bend = (int)((int64_t)(lw + 1) * nblk / NG);

    BlockRegs<T, P, BX, BY> R;
    double dot = 0.0;
    if (bbeg < bend)
        block_prefetch<T, P, BX, BY>(R, x, dimx, dimy, (bbeg % nbx) * BX, (bbeg / nbx) * BY);
    for (int b = bbeg; b < bend; ++b) {
        const int x0 = (b % nbx) * BX, y0 = (b / nbx) * BY;
        block_stage<T, P, BX, BY>(lds, R);
        __syncthreads();
        if (b + 1 < bend)
            block_prefetch<T, P, BX, BY>(R, x, dimx, dimy, ((b + 1) % nbx) * BX, ((b + 1) / nbx) * BY);
        dot += block_compute<T, P, BX, BY, DOT>(lds, y, dimx, dimy, x0, y0, pen);
    }
    if (DOT) {
        const double bs = block_sum<G::NT>(dot, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    }
}

// The same walk with direct global -> LDS staging into two LDS images:
// block b+1 streams into one while block b is computed from the other, with
// no registers held for it.  The barrier at the top of each pass (after the
// wait for this wave's loads) both publishes the staged block and tells the
// issuing waves that every wave is done with the image they refill next.
template <typename T, int P, int BX, int BY, bool DOT>
__global__ __launch_bounds__(BX * BY * (P + 1))
void apply_kernel_lds(const T *__restrict__ x, T *__restrict__ y, int dimx, int dimy, int nbx,
                      int nblk, T pen, double *__restrict__ partials,
                      const cg_state *__restrict__ st, int slot)
{
    using G = BlockGeo<T, P, BX, BY>;
    __shared__ T lds[2 * G::LDS];
    __shared__ double red[(G::NT + 63) / 64];

    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }
    const int NG = gridDim.x;
    const int lw = xcd_remap(blockIdx.x, NG);
    const int bbeg = (int)((int64_t)lw * nblk / NG);
    const int bend = (int)((int64_t)(lw + 1) * nblk / NG);

    double dot = 0.0;
    if (bbeg < bend)
        block_issue<T, P, BX, BY>(lds, x, dimx, dimy, (bbeg % nbx) * BX, (bbeg / nbx) * BY);
    for (int b = bbeg; b < bend; ++b) {
        const int x0 = (b % nbx) * BX, y0 = (b / nbx) * BY;
        const int cur = (b - bbeg) & 1;
        block_issue_wait();
        __syncthreads();
        if (b + 1 < bend)
            block_issue<T, P, BX, BY>(lds + (cur ^ 1) * G::LDS, x, dimx, dimy, ((b + 1) % nbx) * BX,
                                      ((b + 1) / nbx) * BY);
        dot += block_compute<T, P, BX, BY, DOT, false>(lds + cur * G::LDS, y, dimx, dimy, x0, y0, pen);
    }
    if (DOT) {
        const double bs = block_sum<G::NT>(dot, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    }
}

namespace {

// The apply kernel of one degree: direct LDS staging where the block's LDS
// image is unpadded and two images fit, register staging otherwise.
template <typename T, int P, bool DOT>
struct ApplyKernel {
    using S = ApplyShape<T, P>;
    static constexpr bool lds_staged = BlockDma<T, P, S::BX, S::BY>::ok;
    using Fn = void (*)(const T *, T *, int, int, int, int, T, double *, const cg_state *, int);
    static Fn fn() {
        if constexpr (lds_staged) return &apply_kernel_lds<T, P, S::BX, S::BY, DOT>;
        else return &apply_kernel<T, P, S::BX, S::BY, DOT>;
    }
};

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
            &per, reinterpret_cast<const void *>(ApplyKernel<T, P, DOT>::fn()),
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
        hipLaunchKernelGGL((ApplyKernel<T, P, true>::fn()), grid, block, 0, s,
                           (const T *)x, (T *)y, dimx, dimy, nbx, nblk, (T)pen,
                           partials, st, slot);
    } else {
        dim3 grid(std::min(nblk, resident_groups<T, P, false>()));
        hipLaunchKernelGGL((ApplyKernel<T, P, false>::fn()), grid, block, 0, s,
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

int resolve_kernel(int dtype, int p, int kernel) {
    if (kernel == SIPG_KERNEL_WAVE) return wave_apply_available(dtype, p) ? SIPG_KERNEL_WAVE : SIPG_KERNEL_VALU;
    if (kernel == SIPG_KERNEL_MFMA) return mfma_apply_available(dtype, p) ? SIPG_KERNEL_MFMA : SIPG_KERNEL_VALU;
    if (kernel == SIPG_KERNEL_VALU) return SIPG_KERNEL_VALU;
    // AUTO: the fastest measured -- the matrix cores at p = 15 in FP32, the
    // wave-owned elements at even p <= 8 (1.2-1.6x the block kernel there)
    if (mfma_apply_available(dtype, p)) return SIPG_KERNEL_MFMA;
    if (wave_apply_available(dtype, p)) return SIPG_KERNEL_WAVE;
    return SIPG_KERNEL_VALU;
}

int apply_num_blocks(int dtype, int p, int dimx, int dimy, int kernel) {
    if (p > kMaxDegree) return p <= kMaxGenericDegree ? generic_apply_num_blocks(dimx, dimy) : -1;
    const int k = resolve_kernel(dtype, p, kernel);
    if (k == SIPG_KERNEL_MFMA) return mfma_apply_num_blocks(dimx, dimy);
    if (k == SIPG_KERNEL_WAVE) return wave_apply_num_blocks(dtype, p, dimx, dimy);
    return dtype == SIPG_F64 ? AllDegrees<double>::nblocks(p, dimx, dimy)
                             : AllDegrees<float>::nblocks(p, dimx, dimy);
}

int launch_apply(int dtype, int p, const void *x, void *y, const void *tab, int dimx, int dimy,
                 double pen, double *partials, const cg_state *st, int slot, hipStream_t s, int kernel)
{
    // above the compiled degrees: the run-time-table operator, whatever the
    // setting (its only implementation)
    if (p > kMaxDegree)
        return launch_apply_generic(dtype, p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
    const int k = resolve_kernel(dtype, p, kernel);
    if (k == SIPG_KERNEL_MFMA)
        return launch_apply_mfma(dtype, p, x, y, dimx, dimy, pen, partials, st, slot, s);
    if (k == SIPG_KERNEL_WAVE)
        return launch_apply_wave(dtype, p, x, y, dimx, dimy, pen, partials, st, slot, s);
    if (dtype == SIPG_F64)
        return AllDegrees<double>::launch(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
    return AllDegrees<float>::launch(p, x, y, tab, dimx, dimy, pen, partials, st, slot, s);
}

}  // namespace sipg_internal

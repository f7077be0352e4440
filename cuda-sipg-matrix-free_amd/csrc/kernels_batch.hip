// kernels_batch.hip -- many independent SIPG solves in the same launches.
//
// Instances of one degree share a launch of the operator kernel: the
// workgroups walk a device-resident list of (instance, block) items, so the
// list can shrink as instances finish without relaunching anything (the CG
// iterations are replayed from a captured hipGraph).  Vector updates walk a
// list of (instance, range) chunks.  Each instance has its own cg_state;
// its scalars are reduced per instance in a fixed order (deterministic, and
// independent of how the work is split).  All vectors of all instances are
// concatenated, element-major within each instance.
#include "apply_block.hpp"

#include <algorithm>
#include <atomic>

namespace sipg_internal {

namespace {

constexpr int NT = kVecThreads;

// y = A x for every listed block; with DOT also this block's share of x.y
// into item_part[k], and blocks of instances already finished (per
// states[inst].s[slot].done) are skipped.  Persistent: workgroup ranges of
// the item list are contiguous, so consecutive items are neighbouring blocks
// of one instance; the next live block is loaded while the current one is
// computed.
template <typename T, int P, int BX, int BY, bool DOT>
__global__ __launch_bounds__(BX * BY * (P + 1))
void batch_apply_kernel(const T *__restrict__ X, T *__restrict__ Y, const batch_inst *__restrict__ inst,
                        const batch_item *__restrict__ items, const int *__restrict__ count,
                        double *__restrict__ item_part, const cg_state *__restrict__ states, int slot)
{
    using G = BlockGeo<T, P, BX, BY>;
    __shared__ T lds[G::LDS];
    __shared__ double red[(G::NT + 63) / 64];
    const int n = *count;
    const int NG = gridDim.x;
    const int lw = xcd_remap(blockIdx.x, NG);
    const int kbeg = (int)((int64_t)lw * n / NG);
    const int kend = (int)((int64_t)(lw + 1) * n / NG);
    // first item >= k of an unfinished instance: SCAN items per step, one
    // per lane, then the lowest set bit of the ballot.  SCAN is the lane
    // count of the smallest wave (the last one is partial when NT is not a
    // multiple of 64), so every wave reaches the same answer.
    constexpr int SCAN = (G::NT % 64) ? (G::NT % 64) : 64;
    auto next_live = [&](int k) -> int {
        if (!DOT) return k;
        const int lane = threadIdx.x & 63;
        while (k < kend) {
            const int kk = k + lane;
            bool lv = false;
            if (lane < SCAN && kk < kend) {
                const cg_state *st = states + items[kk].inst;
                lv = !(st->s[slot].done | st->err);
            }
            const unsigned long long m = __ballot(lv);
            if (m) return __builtin_amdgcn_readfirstlane(k + __ffsll((long long)m) - 1);
            k += SCAN;
        }
        return kend;
    };

    BlockRegs<T, P, BX, BY> R;
    int k = next_live(kbeg);
    if (k < kend) {
        const batch_item it = items[k];
        const batch_inst in = inst[it.inst];
        block_prefetch<T, P, BX, BY>(R, X + in.off, in.dimx, in.dimy, it.x0, it.y0);
    }
    while (k < kend) {
        const batch_item it = items[k];
        const batch_inst in = inst[it.inst];
        block_stage<T, P, BX, BY>(lds, R);
        __syncthreads();
        const int kn = next_live(k + 1);
        if (kn < kend) {
            const batch_item nx = items[kn];
            const batch_inst ni = inst[nx.inst];
            block_prefetch<T, P, BX, BY>(R, X + ni.off, ni.dimx, ni.dimy, nx.x0, nx.y0);
        }
        const double d = block_compute<T, P, BX, BY, DOT>(lds, Y + in.off, in.dimx, in.dimy, it.x0,
                                                          it.y0, (T)in.pen);
        if (DOT) {
            const double bs = block_sum<G::NT>(d, red);
            if (threadIdx.x == 0) item_part[k] = bs;
        }
        k = kn;
    }
}

__device__ __forceinline__ double block_reduce(double v, double *red) {
    // fixed-order block sum, valid in every thread
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    return s;
}

// sum of part[b, e) in a fixed order (strided per thread, then the block)
__device__ __forceinline__ double range_sum(const double *__restrict__ part, int b, int e, double *red) {
    double s = 0.0;
    for (int i = b + (int)threadIdx.x; i < e; i += NT) s += part[i];
    return block_reduce(s, red);
}

// alpha = r.z / p.Ap per active instance (r.z = r.r for plain CG; the host-side alpha of
// conjugate_gradient_gpu.hpp:75-83); breakdown and non-finite values end
// that instance only.
__global__ __launch_bounds__(NT) void batch_alpha_kernel(const int *__restrict__ list, const int *__restrict__ count,
                                                         const int2 *__restrict__ item_range,
                                                         const double *__restrict__ item_part,
                                                         cg_state *states, int slot) {
    __shared__ double red[NT / 64];
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int i = list[k];
        cg_state *st = states + i;
        if (st->s[slot].done | st->err) continue;                 // uniform per block
        const int2 r = item_range[i];
        const double pAp = range_sum(item_part, r.x, r.y, red);
        if (threadIdx.x == 0) {
            const double alpha = st->s[slot].rz / pAp;
            if (!(pAp > 0.0) || !isfinite(alpha)) {
                st->err = isfinite(pAp) ? 2 : 3;
            } else {
                st->s[slot].alpha = alpha;
                st->s[slot].pAp = pAp;
            }
        }
    }
}

// x += alpha p, r -= alpha Ap over the listed chunks; chunk sums of r.r
template <typename T>
__global__ __launch_bounds__(NT) void batch_update_xr_kernel(T *__restrict__ X, const T *__restrict__ Pv,
                                                             T *__restrict__ Rv, const T *__restrict__ APv,
                                                             const batch_chunk *__restrict__ chunks,
                                                             const int *__restrict__ count,
                                                             double *__restrict__ chunk_part,
                                                             const cg_state *__restrict__ states, int slot) {
    __shared__ double red[NT / 64];
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const batch_chunk c = chunks[k];
        const cg_state *st = states + c.inst;
        if (st->s[slot].done | st->err) continue;
        const T alpha = (T)st->s[slot].alpha;
        double s = 0.0;
        for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT) {
            X[i] = X[i] + alpha * Pv[i];
            const T ri = Rv[i] - alpha * APv[i];
            Rv[i] = ri;
            s += (double)ri * (double)ri;
        }
        s = block_reduce(s, red);
        if (threadIdx.x == 0) chunk_part[k] = s;
    }
}

// beta = r.r(new) / r.r(old) per active instance; writes the next slot
// (conjugate_gradient_gpu.hpp:85-95 and the while-condition of :69).  An
// instance whose r was replaced by b - A x in this iteration (kSlotReplaced,
// set by batch_mark_kernel) takes r.r(new) from the replacement's chunk
// sums and keeps its direction (reliable update); with st->true_stop it
// finishes only on such a true residual, and a recursion that meets the
// tolerance raises kSlotWantCheck for the host to schedule one.
__global__ __launch_bounds__(NT) void batch_beta_kernel(const int *__restrict__ list, const int *__restrict__ count,
                                                        const int2 *__restrict__ chunk_range,
                                                        const double *__restrict__ chunk_part,
                                                        const int2 *__restrict__ repl_range,
                                                        const double *__restrict__ repl_part,
                                                        cg_state *states, int slot) {
    __shared__ double red[NT / 64];
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int i = list[k];
        cg_state *st = states + i;
        const cg_slot cur = st->s[slot];
        if (cur.done | st->err) {
            if (threadIdx.x == 0) st->s[slot ^ 1] = cur;           // carry forward
            continue;
        }
        const bool replaced = (cur.flags & kSlotReplaced) != 0;
        const int2 r = replaced ? repl_range[i] : chunk_range[i];
        const double rr_new = range_sum(replaced ? repl_part : chunk_part, r.x, r.y, red);
        if (threadIdx.x == 0) {
            const bool met = !(sqrt(rr_new) > st->stop);
            const bool need_true = st->true_stop && !replaced;
            cg_slot nx{};
            nx.rr = rr_new;
            nx.rz = rr_new;
            nx.alpha = 0.0;
            nx.beta = rr_new / cur.rr;
            nx.pAp = 0.0;
            nx.it = cur.it + 1;
            nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
            nx.checked = replaced ? 1 : 0;
            nx.flags = (met && need_true) ? kSlotWantCheck : 0;
            st->s[slot ^ 1] = nx;
            if (replaced) st->rr_ref = rr_new;
            if (!isfinite(rr_new)) st->err = 3;
        }
    }
}

// marks the listed instances' slot as replaced (r is b - A x now and the
// replacement chunk sums hold its r.r); a no-op for finished instances
__global__ __launch_bounds__(NT) void batch_mark_kernel(const int *__restrict__ list, const int *__restrict__ count,
                                                        cg_state *states, int slot) {
    const int n = *count;
    for (int k = blockIdx.x * NT + threadIdx.x; k < n; k += gridDim.x * NT) {
        cg_state *st = states + list[k];
        cg_slot c = st->s[slot];
        if (c.done | st->err) continue;
        c.flags |= kSlotReplaced;
        st->s[slot] = c;
        st->nrep += 1;
    }
}

// x += z; z = 0 over the listed chunks: the group update (van der Vorst &
// Ye 2000) folds the increments accumulated since the last true residual
template <typename T>
__global__ __launch_bounds__(NT) void batch_fold_kernel(T *__restrict__ X, T *__restrict__ Z,
                                                        const batch_chunk *__restrict__ chunks,
                                                        const int *__restrict__ count) {
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const batch_chunk c = chunks[k];
        for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT) {
            X[i] += Z[i];
            Z[i] = T(0);
        }
    }
}

// p = r + beta p over the listed chunks of instances still iterating
template <typename T>
__global__ __launch_bounds__(NT) void batch_update_p_kernel(const T *__restrict__ Rv, T *__restrict__ Pv,
                                                            const batch_chunk *__restrict__ chunks,
                                                            const int *__restrict__ count,
                                                            const cg_state *__restrict__ states, int slot) {
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const batch_chunk c = chunks[k];
        const cg_state *st = states + c.inst;
        if (st->s[slot].done | st->err) continue;
        const T beta = (T)st->s[slot ^ 1].beta;
        for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT) Pv[i] = Rv[i] + beta * Pv[i];
    }
}

// r = b - Ax over the listed chunks (optionally p = r); chunk sums of r.r
// and b.b
template <typename T>
__global__ __launch_bounds__(NT) void batch_residual_kernel(const T *__restrict__ B, const T *__restrict__ AX,
                                                            T *__restrict__ Rv, T *__restrict__ Pv,
                                                            const batch_chunk *__restrict__ chunks,
                                                            const int *__restrict__ count,
                                                            double *__restrict__ part_rr,
                                                            double *__restrict__ part_bb) {
    __shared__ double red[NT / 64];
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const batch_chunk c = chunks[k];
        double srr = 0.0, sbb = 0.0;
        for (int64_t i = c.begin + threadIdx.x; i < c.end; i += NT) {
            const T bi = B[i];
            const T ri = bi - AX[i];
            Rv[i] = ri;
            if (Pv) Pv[i] = ri;
            srr += (double)ri * (double)ri;
            sbb += (double)bi * (double)bi;
        }
        srr = block_reduce(srr, red);
        sbb = block_reduce(sbb, red);
        if (threadIdx.x == 0) {
            part_rr[k] = srr;
            part_bb[k] = sbb;
        }
    }
}

// initial state of each listed instance from its r.r and b.b chunk sums
// (and, for PCG, its r.z item sums; otherwise r.z = r.r)
__global__ __launch_bounds__(NT) void batch_init_kernel(const int *__restrict__ list, const int *__restrict__ count,
                                                        const int2 *__restrict__ chunk_range,
                                                        const double *__restrict__ part_rr,
                                                        const double *__restrict__ part_bb, cg_state *states,
                                                        const batch_ctl *__restrict__ ctl,
                                                        const int2 *__restrict__ item_range,
                                                        const double *__restrict__ item_rz) {
    __shared__ double red[NT / 64];
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int i = list[k];
        const int2 r = chunk_range[i];
        const double rr = range_sum(part_rr, r.x, r.y, red);
        const double bb = range_sum(part_bb, r.x, r.y, red);
        double rz = rr;
        if (item_rz) {
            const int2 q = item_range[i];
            rz = range_sum(item_rz, q.x, q.y, red);
        }
        if (threadIdx.x == 0) {
            cg_state *st = states + i;
            const batch_ctl c = ctl[i];
            const double stop = c.tol_mode == SIPG_TOL_REL ? c.tol * sqrt(bb) : c.tol;
            cg_slot s0{};
            s0.rr = rr;
            s0.rz = rz;
            s0.alpha = s0.beta = s0.pAp = 0.0;
            s0.it = 0;
            s0.done = (sqrt(rr) > stop && c.maxit > 0) ? 0 : 1;
            s0.checked = 1;
            s0.flags = 0;
            st->s[0] = s0;
            st->s[1] = s0;
            st->bb = bb;
            st->stop = stop;
            st->stop_true = stop;
            st->rr0 = rr;
            st->rr_ref = rr;
            st->maxit = c.maxit;
            st->nrep = 0;
            st->err = (!isfinite(rr) || !isfinite(bb) || !isfinite(rz)) ? 3 : 0;
            st->true_stop = 1;
        }
    }
}

// ---- per-degree operator launch table --------------------------------------

template <int P, int BX, int BY, bool DOT>
int resident_batch_groups() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int g = cache[dev].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&batch_apply_kernel<double, P, BX, BY, DOT>),
            BX * BY * (P + 1), 0) != hipSuccess)
        per = 0;
    g = (cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    cache[dev].store(g, std::memory_order_relaxed);
    return g;
}

template <int P, int BX, int BY>
int launch_batch_apply_one(const double *X, double *Y, const batch_inst *inst, const batch_item *items,
                           const int *count, int max_items, double *item_part, const cg_state *states,
                           int slot, hipStream_t s) {
    if (max_items <= 0) return SIPG_OK;
    dim3 block(BX * BY * (P + 1));
    if (item_part) {
        dim3 grid(std::min(max_items, resident_batch_groups<P, BX, BY, true>()));
        hipLaunchKernelGGL((batch_apply_kernel<double, P, BX, BY, true>), grid, block, 0, s, X, Y, inst, items,
                           count, item_part, states, slot);
    } else {
        dim3 grid(std::min(max_items, resident_batch_groups<P, BX, BY, false>()));
        hipLaunchKernelGGL((batch_apply_kernel<double, P, BX, BY, false>), grid, block, 0, s, X, Y, inst, items,
                           count, item_part, states, slot);
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

using BatchApplyFn = int (*)(const double *, double *, const batch_inst *, const batch_item *, const int *, int,
                             double *, const cg_state *, int, hipStream_t);

template <int P>
constexpr BatchApplyFn wide_shape() {
    // 16 x 4 blocks need (p+1)*64 <= 1024 threads and fit LDS in FP64 up to p = 12
    if constexpr (P <= 12) return &launch_batch_apply_one<P, 16, 4>;
    else return nullptr;
}

template <int... Ps>
struct BatchTable {
    static constexpr BatchApplyFn fn[kBatchShapes][sizeof...(Ps)] = {
        {wide_shape<Ps>()...},
        {&launch_batch_apply_one<Ps, 8, 4>...},
        {&launch_batch_apply_one<Ps, 8, 2>...},
    };
};
using Table = BatchTable<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>;

int vec_blocks(int nchunks) { return std::max(1, std::min(nchunks, 2048)); }
int inst_blocks(int ninst) { return std::max(1, std::min(ninst, 2048)); }

}  // namespace

const int kBatchShapeBX[kBatchShapes] = {16, 8, 8};
const int kBatchShapeBY[kBatchShapes] = {4, 4, 2};

bool batch_shape_supported(int p, int shape) {
    return p >= 1 && p <= kMaxDegree && shape >= 0 && shape < kBatchShapes && Table::fn[shape][p - 1] != nullptr;
}

int launch_batch_apply(int p, int shape, const double *X, double *Y, const batch_inst *inst,
                       const batch_item *items, const int *count, int max_items, double *item_part,
                       const cg_state *states, int slot, hipStream_t s) {
    if (!batch_shape_supported(p, shape)) return SIPG_ERR_UNSUPPORTED;
    if (item_part && !states) return SIPG_ERR_INVALID;
    return Table::fn[shape][p - 1](X, Y, inst, items, count, max_items, item_part, states, slot, s);
}

int launch_batch_alpha(const int *list, const int *count, int max_inst, const int2 *item_range,
                       const double *item_part, cg_state *states, int slot, hipStream_t s) {
    hipLaunchKernelGGL(batch_alpha_kernel, dim3(inst_blocks(max_inst)), dim3(NT), 0, s, list, count, item_range,
                       item_part, states, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_update_xr(double *X, const double *Pv, double *Rv, const double *APv, const batch_chunk *chunks,
                           const int *count, int max_chunks, double *chunk_part, const cg_state *states, int slot,
                           hipStream_t s) {
    hipLaunchKernelGGL(batch_update_xr_kernel<double>, dim3(vec_blocks(max_chunks)), dim3(NT), 0, s, X, Pv, Rv,
                       APv, chunks, count, chunk_part, states, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_beta(const int *list, const int *count, int max_inst, const int2 *chunk_range,
                      const double *chunk_part, const int2 *repl_range, const double *repl_part,
                      cg_state *states, int slot, hipStream_t s) {
    hipLaunchKernelGGL(batch_beta_kernel, dim3(inst_blocks(max_inst)), dim3(NT), 0, s, list, count, chunk_range,
                       chunk_part, repl_range, repl_part, states, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_update_p(const double *Rv, double *Pv, const batch_chunk *chunks, const int *count,
                          int max_chunks, const cg_state *states, int slot, hipStream_t s) {
    hipLaunchKernelGGL(batch_update_p_kernel<double>, dim3(vec_blocks(max_chunks)), dim3(NT), 0, s, Rv, Pv, chunks,
                       count, states, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_residual(const double *B, const double *AX, double *Rv, double *Pv, const batch_chunk *chunks,
                          const int *count, int max_chunks, double *part_rr, double *part_bb, hipStream_t s) {
    hipLaunchKernelGGL(batch_residual_kernel<double>, dim3(vec_blocks(max_chunks)), dim3(NT), 0, s, B, AX, Rv, Pv,
                       chunks, count, part_rr, part_bb);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_init(const int *list, const int *count, int max_inst, const int2 *chunk_range,
                      const double *part_rr, const double *part_bb, cg_state *states, const batch_ctl *ctl,
                      const int2 *item_range, const double *item_rz, hipStream_t s) {
    hipLaunchKernelGGL(batch_init_kernel, dim3(inst_blocks(max_inst)), dim3(NT), 0, s, list, count, chunk_range,
                       part_rr, part_bb, states, ctl, item_range, item_rz);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_fold(double *X, double *Z, const batch_chunk *chunks, const int *count, int max_chunks,
                      hipStream_t s) {
    hipLaunchKernelGGL(batch_fold_kernel<double>, dim3(vec_blocks(max_chunks)), dim3(NT), 0, s, X, Z, chunks, count);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

int launch_batch_mark(const int *list, const int *count, int max_inst, cg_state *states, int slot,
                      hipStream_t s) {
    hipLaunchKernelGGL(batch_mark_kernel, dim3(std::max(1, std::min((max_inst + NT - 1) / NT, 64))), dim3(NT), 0,
                       s, list, count, states, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace sipg_internal

// kernels_batch_pcg.hip -- the block-Jacobi preconditioner of sipg_pcg for
// batched solves: z = M^-1 r over the (instance, block) items of the
// operator's lists, with per-item sums of r.r and r.z, and the PCG beta step
// per instance.  M is the exact inverse of the interior element block of
// the standard penalty 5 (p+1)^2 (kernels_pcg.hip), so every coefficient is
// a compile-time constant of the degree; batched PCG requires that penalty.
#include "apply_block.hpp"
#include "prec_block.hpp"

#include <algorithm>
#include <atomic>

namespace sipg_internal {

namespace {

constexpr int VT = kVecThreads;

// z = M^-1 r on every element of each listed block (COPYP: also p = z);
// item k's sums of r.r and r.z go to item_rr[k], item_rz[k].  With states,
// blocks of instances already finished at `slot` are skipped.
template <int P, int BX, int BY, bool COPYP>
__global__ __launch_bounds__(BX * BY * (P + 1))
void batch_prec_kernel(const double *__restrict__ R, double *__restrict__ Zq, double *__restrict__ Pv,
                       const batch_inst *__restrict__ inst, const batch_item *__restrict__ items,
                       const int *__restrict__ count, double *__restrict__ item_rr,
                       double *__restrict__ item_rz, const cg_state *__restrict__ states, int slot)
{
    constexpr int N1 = P + 1, NQ = N1 * N1, NE = BX * BY, NT = NE * N1, TS = tile_stride(NQ);
    constexpr int RUN = BX * NQ, KR = N1;               // NE*NQ values = KR per thread
    __shared__ double lds[NE * TS];
    __shared__ double red[(NT + 63) / 64];
    const int tid = threadIdx.x;
    const int s_el = tid % NE;
    int j = tid / NE;
    if (NE % 64 == 0) j = __builtin_amdgcn_readfirstlane(j);
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const batch_item it = items[k];
        if (states) {
            const cg_state *st = states + it.inst;
            if (st->s[slot].done | st->err) continue;     // uniform over the workgroup
        }
        const batch_inst in = inst[it.inst];
        const int64_t RS = (int64_t)in.dimx * NQ;
        const int64_t base = in.off + ((int64_t)it.y0 * in.dimx + it.x0) * NQ;
        const int ncol = min(BX, in.dimx - it.x0), nrow = min(BY, in.dimy - it.y0);
        double rv[KR];
        unsigned valid = 0;
        double srr = 0.0, srz = 0.0;
#pragma unroll
        for (int kk = 0; kk < KR; ++kk) {
            const int idx = tid + kk * NT;                 // < NE * NQ
            const int rr = idx / RUN, w = idx - rr * RUN;
            const bool ok = rr < nrow && w / NQ < ncol;
            double v = 0.0;
            if (ok) {
                v = R[base + rr * RS + w];
                valid |= 1u << kk;
                srr += v * v;
            }
            rv[kk] = v;
            lds[lds_index<NQ, TS>(idx)] = v;
        }
        __syncthreads();
        double *tile = lds + s_el * TS;
        {   // column j of S^T R
            double c[N1], t[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) c[i] = tile[i * N1 + j];
            contract_st<double, P, true>(nullptr, c, t);
#pragma unroll
            for (int a = 0; a < N1; ++a) tile[a * N1 + j] = t[a];
        }
        __syncthreads();
        {   // row j of ((S^T R S) ./ L) S^T
            double w[N1], v[N1];
#pragma unroll
            for (int b = 0; b < N1; ++b) w[b] = tile[j * N1 + b];
            contract_st<double, P, true>(nullptr, w, v);
#pragma unroll
            for (int c = 0; c < N1; ++c) v[c] *= PrecEO<P>::V.Linv[j][c];
            contract_s<double, P, true>(nullptr, v, w);
#pragma unroll
            for (int d = 0; d < N1; ++d) tile[j * N1 + d] = w[d];
        }
        __syncthreads();
        {   // column j of S (...)
            double c[N1], t[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) c[i] = tile[i * N1 + j];
            contract_s<double, P, true>(nullptr, c, t);
#pragma unroll
            for (int a = 0; a < N1; ++a) tile[a * N1 + j] = t[a];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KR; ++kk) {
            if (valid & (1u << kk)) {
                const int idx = tid + kk * NT;
                const int rr = idx / RUN, w = idx - rr * RUN;
                const int64_t gi = base + rr * RS + w;
                const double zv = lds[lds_index<NQ, TS>(idx)];
                Zq[gi] = zv;
                if (COPYP) Pv[gi] = zv;
                srz += rv[kk] * zv;
            }
        }
        const double a = block_sum<NT>(srr, red);
        __syncthreads();
        const double b = block_sum<NT>(srz, red);
        if (tid == 0) {
            item_rr[k] = a;
            item_rz[k] = b;
        }
        __syncthreads();                                   // LDS and red free for the next item
    }
}

template <int P, int BX, int BY, bool COPYP>
int resident_prec_groups() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int g = cache[dev].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&batch_prec_kernel<P, BX, BY, COPYP>), BX * BY * (P + 1), 0) !=
        hipSuccess)
        per = 0;
    g = (cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    cache[dev].store(g, std::memory_order_relaxed);
    return g;
}

template <int P, int BX, int BY>
int launch_prec_one(const double *R, double *Zq, double *Pv, const batch_inst *inst, const batch_item *items,
                    const int *count, int max_items, double *item_rr, double *item_rz, const cg_state *states,
                    int slot, hipStream_t s) {
    if (max_items <= 0) return SIPG_OK;
    dim3 block(BX * BY * (P + 1));
    if (Pv) {
        dim3 grid(std::min(max_items, resident_prec_groups<P, BX, BY, true>()));
        hipLaunchKernelGGL((batch_prec_kernel<P, BX, BY, true>), grid, block, 0, s, R, Zq, Pv, inst, items, count,
                           item_rr, item_rz, states, slot);
    } else {
        dim3 grid(std::min(max_items, resident_prec_groups<P, BX, BY, false>()));
        hipLaunchKernelGGL((batch_prec_kernel<P, BX, BY, false>), grid, block, 0, s, R, Zq, Pv, inst, items, count,
                           item_rr, item_rz, states, slot);
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

using PrecFn = int (*)(const double *, double *, double *, const batch_inst *, const batch_item *, const int *, int,
                       double *, double *, const cg_state *, int, hipStream_t);

template <int P>
constexpr PrecFn wide_prec() {
    if constexpr (P <= 12) return &launch_prec_one<P, 16, 4>;
    else return nullptr;
}

template <int... Ps>
struct PrecTable {
    static constexpr PrecFn fn[kBatchShapes][sizeof...(Ps)] = {
        {wide_prec<Ps>()...},
        {&launch_prec_one<Ps, 8, 4>...},
        {&launch_prec_one<Ps, 8, 2>...},
    };
};
using PTable = PrecTable<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>;

__device__ __forceinline__ double vreduce(double v, double *red) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < VT / 64; ++i) s += red[i];
    return s;
}

__device__ __forceinline__ double vrange(const double *__restrict__ part, int b, int e, double *red) {
    double s = 0.0;
    for (int i = b + (int)threadIdx.x; i < e; i += VT) s += part[i];
    return vreduce(s, red);
}

// beta = r.z(new) / r.z(old) per active instance from the preconditioner's
// item sums; writes the next slot with r.r (stopping test, host decisions)
// and r.z.  Flag rules as in batch_beta_kernel; the sums are those of the
// replaced residual when the instance's r was replaced this iteration (the
// preconditioner runs after the replacement).
__global__ __launch_bounds__(VT) void batch_beta_pcg_kernel(const int *__restrict__ list,
                                                            const int *__restrict__ count,
                                                            const int2 *__restrict__ item_range,
                                                            const double *__restrict__ item_rr,
                                                            const double *__restrict__ item_rz, cg_state *states,
                                                            int slot) {
    __shared__ double red[VT / 64];
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int i = list[k];
        cg_state *st = states + i;
        const cg_slot cur = st->s[slot];
        if (cur.done | st->err) {
            if (threadIdx.x == 0) st->s[slot ^ 1] = cur;          // carry forward
            continue;
        }
        const int2 r = item_range[i];
        const double rr_new = vrange(item_rr, r.x, r.y, red);
        const double rz_new = vrange(item_rz, r.x, r.y, red);
        if (threadIdx.x == 0) {
            const bool replaced = (cur.flags & kSlotReplaced) != 0;
            const bool met = !(sqrt(rr_new) > st->stop);
            const bool need_true = st->true_stop && !replaced;
            cg_slot nx{};
            nx.rr = rr_new;
            nx.rz = rz_new;
            nx.beta = rz_new / cur.rz;
            nx.it = cur.it + 1;
            nx.done = ((met && !need_true) || nx.it >= st->maxit) ? 1 : 0;
            nx.checked = replaced ? 1 : 0;
            nx.flags = (met && need_true) ? kSlotWantCheck : 0;
            st->s[slot ^ 1] = nx;
            if (replaced) st->rr_ref = rr_new;
            if (!isfinite(rr_new) || !isfinite(rz_new)) st->err = 3;
            else if (!(rz_new > 0.0) && rr_new > 0.0) st->err = 2;
        }
    }
}

}  // namespace

bool batch_prec_supported(int p, double pen, int shape) {
    return p >= 1 && p <= kMaxDegree && shape >= 0 && shape < kBatchShapes && PTable::fn[shape][p - 1] != nullptr &&
           pen == 5.0 * (p + 1) * (p + 1);
}

int launch_batch_prec(int p, int shape, const double *R, double *Zq, double *Pv, const batch_inst *inst,
                      const batch_item *items, const int *count, int max_items, double *item_rr, double *item_rz,
                      const cg_state *states, int slot, hipStream_t s) {
    if (p < 1 || p > kMaxDegree || shape < 0 || shape >= kBatchShapes || !PTable::fn[shape][p - 1])
        return SIPG_ERR_UNSUPPORTED;
    return PTable::fn[shape][p - 1](R, Zq, Pv, inst, items, count, max_items, item_rr, item_rz, states, slot, s);
}

int launch_batch_beta_pcg(const int *list, const int *count, int max_inst, const int2 *item_range,
                          const double *item_rr, const double *item_rz, cg_state *states, int slot, hipStream_t s) {
    hipLaunchKernelGGL(batch_beta_pcg_kernel, dim3(std::max(1, std::min(max_inst, 2048))), dim3(VT), 0, s, list,
                       count, item_range, item_rr, item_rz, states, slot);
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace sipg_internal

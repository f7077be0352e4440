// kernels_mfma.hip -- the fused SIPG operator apply y = A x on the matrix
// cores, for elements of 16 x 16 LGL nodes (p = 15), FP32.
//
// One wave computes one element.  The element's output tile, transposed,
// is one 16 x 16 accumulator of v_mfma_f32_16x16x4f32 (lane l = c + 16 q
// holds Y[c][4q + r] in register r, so a lane stores four consecutive
// values of row c and a wave one contiguous 1 KiB tile):
//
//   Y^T = K1 (U^T W) + (W U^T) K1          volume (volume_gpu_kernels.hpp:46-83)
//       + [e f alpha beta] [D0; DN; d0; dN]   x faces (flux_term6a, sipg_sem_2d_gpu_kernels.hpp:212-247)
//       + [D0 DN d0 dN] [g; h; gamma; zeta]   y faces (flux_term6b, :279-312)
//
// with d0, dN the unit vectors of nodes 0 and p.  The volume is eight
// products over k in blocks of four; each face set is one product with
// k = 4, whose four rank-one terms are the symmetry term (D0 e + DN f), and
// the consistency plus penalty terms that act on the face rows only.  The
// 16-long face vectors come from the element's own traces and face
// derivatives and those of its four neighbours: every lane forms the part
// of a face sum over its own four nodes, and a sum over the four lanes of a
// column (two cross-lane exchanges) completes it; the face vector each
// lane feeds to the product is then a few multiply-adds.
//
// Register layouts (lane l = c + 16 q, register kb):
//   natural  un[kb] = U[q + 4kb][c]   four coalesced dword loads of the tile
//   rows     ut[kb] = U[c][4q + kb]   one 16-byte load per lane
// The first gives the x-direction sums (over the first node index), the
// second the y-direction ones; as matrix-core operands each is read with k
// mapped accordingly (k = q + 4kb, resp. k = 4q + kb), and the constant
// operands (K1) are laid out to match.
//
// The kernel is persistent: each workgroup (four waves, one element each)
// walks a contiguous range of groups of four elements in row-major order,
// with the XCD-aware range split of the other operator kernels.
#include "apply_block.hpp"

#include <algorithm>
#include <atomic>

namespace sipg_internal {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

struct Mfma16Tab {
    float K1[16][16];
    float D0[16];      // D0[i] = phi_i'(-1)
    float DN[16];      // DN[i] = phi_i'(+1)
    float w[16];
};

constexpr Mfma16Tab make_mfma16_tab() {
    const cx::Tables<15> t = cx::make_tables<15>();
    Mfma16Tab r{};
    for (int a = 0; a < 16; ++a) {
        for (int b = 0; b < 16; ++b) r.K1[a][b] = (float)t.K1[a][b];
        r.D0[a] = (float)t.D[a][0];
        r.DN[a] = (float)t.D[a][15];
        r.w[a] = (float)t.w[a];
    }
    return r;
}

__constant__ Mfma16Tab kTab16 = make_mfma16_tab();

__device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Reduce-scatter of a pair over the four lanes of a column (lanes c + 16q):
// afterwards lanes with even q hold sum_q a, lanes with odd q sum_q b.
__device__ __forceinline__ float pair_sum(float a, float b, bool odd) {
    float keep = odd ? b : a;
    const float send = odd ? a : b;
    keep += __shfl_xor(send, 16, 64);
    keep += __shfl_xor(keep, 32, 64);
    return keep;
}

__device__ __forceinline__ f4 load4(const float *p) { return *reinterpret_cast<const f4 *>(p); }

}  // namespace

template <bool DOT>
__global__ __launch_bounds__(256)
void apply_mfma16_f32(const float *__restrict__ x, float *__restrict__ y, int dimx, int dimy, int ngroup,
                      float pen, double *__restrict__ partials, const cg_state *__restrict__ st, int slot)
{
    __shared__ double red[4];
    if (DOT) {
        if (st->s[slot].done | st->err) return;
    }
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int c = lane & 15, q = lane >> 4;
    const bool odd = q & 1;

    // per-lane constants of the degree-15 tables
    float KA[4], KB[4], d0n[4], dnn[4], d0t[4], dnt[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        KA[kb] = kTab16.K1[c][4 * q + kb];          // A of K1 (U^T W): K1[c][k], k = 4q + kb
        KB[kb] = kTab16.K1[q + 4 * kb][c];          // B of (W U^T) K1: K1[k][c], k = q + 4kb
        d0n[kb] = kTab16.D0[q + 4 * kb];            // face weights of the natural rows
        dnn[kb] = kTab16.DN[q + 4 * kb];
        d0t[kb] = kTab16.D0[4 * q + kb];            // face weights of the row layout
        dnt[kb] = kTab16.DN[4 * q + kb];
    }
    const float wc = kTab16.w[c];
    // operand of the face products: D0, DN, d0, dN for q = 0..3
    const float cf = q == 0 ? kTab16.D0[c] : q == 1 ? kTab16.DN[c] : q == 2 ? (c == 0 ? 1.f : 0.f)
                                                                          : (c == 15 ? 1.f : 0.f);
    const float pen2 = pen + pen;
    const int noe = dimx * dimy;

    const int NG = gridDim.x;
    const int lw = xcd_remap(blockIdx.x, NG);
    const int gbeg = (int)((int64_t)lw * ngroup / NG);
    const int gend = (int)((int64_t)(lw + 1) * ngroup / NG);
    double dot = 0.0;
    for (int g = gbeg; g < gend; ++g) {
        const int e = 4 * g + wid;
        if (e >= noe) continue;                      // wave-uniform
        const int gy = e / dimx, gx = e - gy * dimx;
        const bool Lb = gx == 0, Rb = gx == dimx - 1;
        const bool Ub = gy == 0, Db = gy == dimy - 1;   // y grows downwards
        const float *xe = x + (size_t)e * 256;

        float un[4], nl[4] = {0, 0, 0, 0}, nr[4] = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) un[kb] = xe[64 * kb + lane];
        const f4 ut = load4(xe + 16 * c + 4 * q);
        f4 dt = {0, 0, 0, 0}, uu = {0, 0, 0, 0};
        if (!Lb) {
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) nl[kb] = xe[64 * kb + lane - 256];
        }
        if (!Rb) {
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) nr[kb] = xe[64 * kb + lane + 256];
        }
        if (!Db) dt = load4(xe + (size_t)dimx * 256 + 16 * c + 4 * q);
        if (!Ub) uu = load4(xe - (size_t)dimx * 256 + 16 * c + 4 * q);

        // volume: K1 (U^T W) + (W U^T) K1
        f4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(KA[kb], ut[kb] * wc, acc);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(un[kb] * wc, KB[kb], acc);

        // x faces (column c of the tile, sums over the first index)
        float sdl = 0, sdr = 0, snl = 0, snr = 0;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            sdl += d0n[kb] * un[kb];
            sdr += dnn[kb] * un[kb];
            snl += dnn[kb] * nl[kb];
            snr += d0n[kb] * nr[kb];
        }
        const float tl = q == 0 ? un[0] : 0.f, tr = q == 3 ? un[3] : 0.f;
        const float tln = q == 3 ? nl[3] : 0.f, trn = q == 0 ? nr[0] : 0.f;
        // E_L, F_R: the face traces of e and f; A_L, B_R: the face derivatives
        const float pE = Lb ? tl : 0.5f * (tl - tln);
        const float pF = Rb ? -tr : 0.5f * (trn - tr);
        const float pA = Lb ? sdl : 0.5f * (sdl + snl);
        const float pB = Rb ? -sdr : -0.5f * (sdr + snr);
        const float EX = pair_sum(pE, pF, odd);      // q even: E_L, odd: F_R
        const float AX = pair_sum(pA, pB, odd);      // q even: A_L, odd: B_R
        const float sx = odd ? (Rb ? -pen : -pen2) : (Lb ? pen : pen2);
        const float vx = q < 2 ? wc * EX : wc * (AX + sx * EX);

        // y faces (row c of the tile, sums over the second index)
        float sdd = 0, sdu = 0, snd = 0, snu = 0;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            sdd += d0t[kb] * ut[kb];
            sdu += dnt[kb] * ut[kb];
            snd += dnt[kb] * dt[kb];
            snu += d0t[kb] * uu[kb];
        }
        const float td = q == 0 ? ut[0] : 0.f, tu = q == 3 ? ut[3] : 0.f;
        const float tdn = q == 3 ? dt[3] : 0.f, tun = q == 0 ? uu[0] : 0.f;
        const float pG = Db ? td : 0.5f * (td - tdn);
        const float pH = Ub ? -tu : 0.5f * (tun - tu);
        const float pC = Db ? sdd : 0.5f * (sdd + snd);
        const float pD = Ub ? -sdu : -0.5f * (sdu + snu);
        const float GY = pair_sum(pG, pH, odd);      // q even: G_D, odd: H_U
        const float CY = pair_sum(pC, pD, odd);      // q even: C_D, odd: D_U
        const float sy = odd ? (Ub ? -pen : -pen2) : (Db ? pen : pen2);
        const float vy = q < 2 ? wc * GY : wc * (CY + sy * GY);

        acc = mfma(vx, cf, acc);                     // x faces
        acc = mfma(cf, vy, acc);                     // y faces

        *reinterpret_cast<f4 *>(y + (size_t)e * 256 + 16 * c + 4 * q) = acc;
        if (DOT) {
#pragma unroll
            for (int r = 0; r < 4; ++r) dot += (double)ut[r] * (double)acc[r];
        }
    }
    if (DOT) {
        const double bs = block_sum<256>(dot, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    }
}

namespace {

template <bool DOT>
int mfma16_resident_groups() {
    constexpr int kMaxDev = 64;
    static std::atomic<int> cache[kMaxDev];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) dev = 0;
    int g = cache[dev].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void *>(&apply_mfma16_f32<DOT>),
                                                     256, 0) != hipSuccess)
        per = 0;
    g = (cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    cache[dev].store(g, std::memory_order_relaxed);
    return g;
}

}  // namespace

bool mfma_apply_available(int dtype, int p) { return dtype == SIPG_F32 && p == 15; }

int mfma_apply_num_blocks(int dimx, int dimy) {
    const int ngroup = (int)(((int64_t)dimx * dimy + 3) / 4);
    return std::min(ngroup, mfma16_resident_groups<true>());
}

int launch_apply_mfma(int dtype, int p, const void *x, void *y, int dimx, int dimy, double pen, double *partials,
                      const cg_state *st, int slot, hipStream_t s)
{
    if (!mfma_apply_available(dtype, p)) return SIPG_ERR_UNSUPPORTED;
    const int ngroup = (int)(((int64_t)dimx * dimy + 3) / 4);
    if (partials) {
        const int grid = std::min(ngroup, mfma16_resident_groups<true>());
        hipLaunchKernelGGL(apply_mfma16_f32<true>, dim3(grid), dim3(256), 0, s, (const float *)x, (float *)y, dimx,
                           dimy, ngroup, (float)pen, partials, st, slot);
    } else {
        const int grid = std::min(ngroup, mfma16_resident_groups<false>());
        hipLaunchKernelGGL(apply_mfma16_f32<false>, dim3(grid), dim3(256), 0, s, (const float *)x, (float *)y, dimx,
                           dimy, ngroup, (float)pen, partials, st, slot);
    }
    SIPG_HIP_TRY(hipGetLastError());
    return SIPG_OK;
}

}  // namespace sipg_internal

// apply_block.hpp -- the device pieces of the fused SIPG operator shared by
// the single-mesh kernel (kernels_apply.hip) and the batched kernel
// (kernels_batch.hip): block geometry, the register prefetch of a block and
// its halo ring, the LDS stage, and the operator on a staged block.
#pragma once
#include "sipg_internal.hpp"
#include "basis_constexpr.hpp"

#include <utility>

namespace sipg_internal {

// BX x BY elements per workgroup, NE = BX*BY, threads = NE*(P+1) <= 1024,
// LDS = (P+1)^2 * (NE + 2BX + 2BY) * sizeof(T) <= 160 KiB.  FP64 at p >= 9
// takes 8-wide blocks and at p >= 15 8 x 2: fewer waves per workgroup leave
// each thread the registers that its (p+1)-long lines need without spills.
template <typename T, int P>
struct ApplyShape {
    static constexpr int N1 = P + 1;
    static constexpr bool big64 = sizeof(T) == 8 && P >= 9;
    static constexpr bool small = N1 > 16;
    static constexpr int BX = (big64 || small) ? 8 : 16;
    static constexpr int BY = (sizeof(T) == 8 && P >= 15) ? 2 : 4;
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
    // When NT is not a multiple of 64 the last wave is partial: its lanes
    // >= NT % 64 do not exist and their registers hold nothing meaningful,
    // so a partner outside the workgroup contributes zero.
    constexpr int NW = (NT + 63) / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        if (NT % 64 != 0 && wid * 64 + (lane ^ off) >= NT) o = 0.0;
        v += o;
    }
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

// Compile-time geometry of a BX x BY block of degree-P elements.  LDS slots:
// (BY+2) rows of BX tiles for the element rows y0-1 .. y0+BY (slot r*BX + t
// holds element (x0+t, y0-1+r)), then the left halo column (x0-1, y0+k) at
// SL_COL+k and the right one (x0+BX, y0+k) at SL_COL+BY+k.  A block's own
// element (ex, ey) sits in slot (ey+1)*BX+ex, so its up/down neighbours are
// always slot -/+ BX.
template <typename T, int P, int BX, int BY>
struct BlockGeo {
    static constexpr int N1 = P + 1;
    static constexpr int NQ = N1 * N1;
    static constexpr int TS = tile_stride(NQ);
    static constexpr int NE = BX * BY;
    static constexpr int RUN = BX * NQ;                 // values of one row run
    static constexpr int NROWV = (BY + 2) * RUN;
    static constexpr int SL_COL = (BY + 2) * BX;
    static constexpr int NCOLV = 2 * BY * NQ;
    static constexpr int NS = SL_COL + 2 * BY;
    static constexpr int NT = NE * N1;                  // threads
    static constexpr int KR = (NROWV + NT - 1) / NT;    // row values per thread
    static constexpr int KC = (NCOLV + NT - 1) / NT;    // halo-column values per thread
    static constexpr int NOUT = BY * RUN;
    static constexpr int KO = (NOUT + NT - 1) / NT;
    static constexpr int LDS = NS * TS;                 // elements of T
};

// A block's values in flight between the global loads and the LDS stage.
template <typename T, int P, int BX, int BY>
struct BlockRegs {
    T buf[BlockGeo<T, P, BX, BY>::KR];
    T cbuf[BlockGeo<T, P, BX, BY>::KC];
};

// Loads block (x0, y0) of a dimx x dimy element-major vector x into
// registers: row value idx lies in run r = idx / RUN at offset
// idx + r*(RS - RUN) - RS from the block's first tile (RS = dimx*(p+1)^2);
// halo-column value idx is node idx % NQ of column tile idx / NQ.  Values
// outside the mesh are zero.  Interior blocks take a path without checks.
template <typename T, int P, int BX, int BY>
__device__ __forceinline__ void block_prefetch(BlockRegs<T, P, BX, BY> &R, const T *__restrict__ x,
                                               int dimx, int dimy, int x0, int y0)
{
    using G = BlockGeo<T, P, BX, BY>;
    T *buf = R.buf, *cbuf = R.cbuf;
    constexpr int NQ = G::NQ, NT = G::NT, RUN = G::RUN, NROWV = G::NROWV, NCOLV = G::NCOLV;
    const int tid = opaque(threadIdx.x);
    const int RS = dimx * NQ, GAP = RS - RUN;
    const T *xb = x + ((size_t)y0 * dimx + x0) * NQ;
    const bool interior = x0 >= 1 && x0 + BX < dimx && y0 >= 1 && y0 + BY < dimy;
    if (interior) {
#pragma unroll
        for (int k = 0; k < G::KR; ++k) {
            const int idx = tid + k * NT;
            if ((k + 1) * NT <= NROWV || idx < NROWV) {
                const int r = idx / RUN;
                buf[k] = xb[idx + r * GAP - RS];
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < G::KR; ++k) {
            const int idx = tid + k * NT;
            const int r = idx / RUN;
            const int t = (idx - r * RUN) / NQ;
            const int gy = y0 - 1 + r;
            const bool ok = idx < NROWV && gy >= 0 && gy < dimy && x0 + t < dimx;
            buf[k] = ok ? xb[idx + r * GAP - RS] : T(0);
        }
    }
#pragma unroll
    for (int k = 0; k < G::KC; ++k) {
        const int idx = tid + k * NT;
        const int kk = idx / NQ, q = idx - kk * NQ;
        const bool right = kk >= BY;
        const int row = right ? kk - BY : kk;
        const bool ok = idx < NCOLV && y0 + row < dimy && (right ? x0 + BX < dimx : x0 >= 1);
        cbuf[k] = ok ? xb[row * RS + (right ? BX * NQ : -NQ) + q] : T(0);
    }
}

// Direct global -> LDS staging (global_load_lds_dword): used where the LDS
// image of a block is the concatenation of its row runs and halo-column
// tiles (no row padding, TS == NQ) and two images fit in LDS, so the next
// block streams in while the current one is computed.
template <typename T, int P, int BX, int BY>
struct BlockDma {
    using G = BlockGeo<T, P, BX, BY>;
    static constexpr bool ok = G::TS == G::NQ && G::NT >= 64 && 2 * G::LDS * sizeof(T) <= 160 * 1024;
};

typedef __attribute__((address_space(3))) unsigned lds_dword;

// Direct loads of chunks C... of 64 dwords: lane l moves dword 64C + l of
// src to dword 64C + l of dst.  The chunk offset is the instruction's
// immediate (applied to both addresses), so a group costs one M0 write.
template <int... C>
__device__ __forceinline__ void dma_group(const unsigned *src, lds_dword *dst, std::integer_sequence<int, C...>)
{
    (__builtin_amdgcn_global_load_lds(src, dst, 4, C * 256, 0), ...);
}

// The first N full 64-dword chunks from src (per-lane pointer: base + lane)
// to dst (wave-uniform base), in groups of 16 (immediate offsets < 4 KiB).
template <int N>
__device__ __forceinline__ void dma_chunks(const unsigned *src, lds_dword *dst)
{
    if constexpr (N >= 16) {
        dma_group(src, dst, std::make_integer_sequence<int, 16>{});
        dma_chunks<N - 16>(src + 1024, dst + 1024);
    } else if constexpr (N > 0) {
        dma_group(src, dst, std::make_integer_sequence<int, N>{});
    }
}

// A run of exactly ND dwords: its full chunks, then the lanes of the last
// partial chunk that fall inside the run.
template <int ND>
__device__ __forceinline__ void dma_run(const unsigned *src, lds_dword *dst, int lane)
{
    dma_chunks<ND / 64>(src, dst);
    if constexpr (ND % 64 != 0) {
        if (lane < ND % 64)
            __builtin_amdgcn_global_load_lds(src + (ND / 64) * 64, dst + (ND / 64) * 64, 4, 0, 0);
    }
}

// Issues the loads of block (x0, y0) into the LDS image `lds`; returns
// without waiting.  Every lane of a load instruction writes the dword at
// (LDS base + 4 * lane), so an instruction moves 64 consecutive dwords of
// one row run or halo tile, and only full waves issue them.  Work items
// (BY + 2 row runs, then 2 BY halo tiles) go round-robin over the waves;
// runs and tiles outside the mesh are skipped (their slots are never read
// by a valid element), and a run cut by the right mesh edge loads only the
// elements that exist.
// NW: the number of (full) waves of the workgroup that issue.
template <typename T, int P, int BX, int BY, int NW = BlockGeo<T, P, BX, BY>::NT / 64>
__device__ __forceinline__ void block_issue(T *lds, const T *__restrict__ x, int dimx, int dimy, int x0, int y0)
{
    using G = BlockGeo<T, P, BX, BY>;
    static_assert(BlockDma<T, P, BX, BY>::ok, "direct staging needs unpadded tiles and full waves");
    constexpr int WD = sizeof(T) / 4;                   // dwords per value
    constexpr int NQD = G::NQ * WD, RUND = G::RUN * WD;
    constexpr int CHR = (RUND + 63) / 64;
    constexpr int NITEM = BY + 2 + 2 * BY;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= NW) return;
    const unsigned *xd = reinterpret_cast<const unsigned *>(x);
    lds_dword *ld = (lds_dword *)lds;
    const int ncol = dimx - x0 < BX ? dimx - x0 : BX;
    for (int it = wid; it < NITEM; it += NW) {
        if (it < BY + 2) {
            const int gy = y0 - 1 + it;
            if (gy < 0 || gy >= dimy) continue;
            const unsigned *src = xd + ((size_t)gy * dimx + x0) * NQD + lane;
            lds_dword *dst = ld + it * RUND;
            if (ncol == BX) {
                dma_run<RUND>(src, dst, lane);
            } else {
                const int nd = ncol * NQD;
#pragma unroll
                for (int c = 0; c < CHR; ++c)
                    if (c * 64 + lane < nd)
                        __builtin_amdgcn_global_load_lds(src + c * 64, dst + c * 64, 4, 0, 0);
            }
        } else {
            const int k = it - (BY + 2);                // halo tile: left k < BY, right k >= BY
            const bool right = k >= BY;
            const int gy = y0 + (right ? k - BY : k), gx = right ? x0 + BX : x0 - 1;
            if (gy >= dimy || gx < 0 || gx >= dimx) continue;
            dma_run<NQD>(xd + ((size_t)gy * dimx + gx) * NQD + lane, ld + (G::SL_COL + k) * NQD, lane);
        }
    }
}

// Waits for this wave's outstanding vector-memory operations, the direct
// LDS loads included; a barrier after it publishes the staged block.
__device__ __forceinline__ void block_issue_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <typename T, int P, int BX, int BY>
__device__ __forceinline__ void block_stage(T *lds, const BlockRegs<T, P, BX, BY> &R)
{
    using G = BlockGeo<T, P, BX, BY>;
    const T *buf = R.buf, *cbuf = R.cbuf;
    constexpr int NQ = G::NQ, TS = G::TS, NT = G::NT;
    const int tid = opaque(threadIdx.x);
#pragma unroll
    for (int k = 0; k < G::KR; ++k) {
        const int idx = tid + k * NT;
        if ((k + 1) * NT <= G::NROWV || idx < G::NROWV) lds[lds_index<NQ, TS>(idx)] = buf[k];
    }
#pragma unroll
    for (int k = 0; k < G::KC; ++k) {
        const int idx = tid + k * NT;
        const int kk = idx / NQ, q = idx - kk * NQ;
        if (idx < G::NCOLV) lds[(G::SL_COL + kk) * TS + q] = cbuf[k];
    }
}

// The operator for line j of one element staged in LDS, in three parts:
//   line_load        column j (x direction, flux_term6a) and row j (y
//                    direction, flux_term6b) of the element's tile, their
//                    even-odd split and the element's own face derivatives;
//   line_neighbours  the neighbours' face derivative and trace on line j,
//                    from their staged tiles (boundary sides are not read);
//   line_finish      the sum-factorised contractions in even-odd form and
//                    the face terms: Yx[a] is this line's contribution to
//                    Y[a][j], Yy[c] to Y[j][c].
// line_operator runs the three.
//
// Even-odd form: K1 is centrosymmetric and the derivative table
// antisymmetric (DP[i] = -D0[p-i], exactly), so the contraction, the
// element's own face derivatives and the face terms all act on the pair
// sums ue = u[i] + u[p-i] and differences uo = u[i] - u[p-i].
//
// The coefficients come from a policy object c; CxCoef has them as
// compile-time constants (immediates built in scalar registers at each use).
template <typename T, int P>
struct CxCoef {
    static constexpr cx::Tables<P> CT = cx::make_tables<P>();
    static constexpr cx::EvenOdd<P> EO = cx::make_even_odd<P>();
    // E and O are symmetric (K1 symmetric and centrosymmetric); one copy of
    // each pair is used, so the two index orders give the same value
    __device__ T E(int a, int i) const { return (T)(a <= i ? EO.E[a][i] : EO.E[i][a]); }
    __device__ T O(int a, int i) const { return (T)(a <= i ? EO.O[a][i] : EO.O[i][a]); }
    __device__ T DA(int i) const { return (T)EO.DA[i]; }
    __device__ T DB(int i) const { return (T)EO.DB[i]; }
    __device__ T DM() const { return (T)EO.DM; }
    __device__ T D0(int i) const { return (T)CT.D[i][0]; }
    // DP[i] = -D0[p-i] exactly (the derivative table is antisymmetric)
    __device__ T DP(int i) const { return -(T)CT.D[P - i][0]; }
    __device__ T w(int i) const { return (T)CT.w[i]; }
};

template <typename T, int P>
struct LineState {
    static constexpr int N1 = P + 1, H = N1 / 2, ME = (N1 + 1) / 2;
    T ue[ME], uo[H > 0 ? H : 1], ve[ME], vo[H > 0 ? H : 1];
    T u0, uP, v0, vP;          // traces of the line at the element's faces
    T dL, dR, dD, dU;          // the element's own face derivatives on the line
};

// Face derivative n* and trace t* of each neighbour on the shared face, line j.
template <typename T>
struct LineNbr {
    T nL = 0, tL = 0, nR = 0, tR = 0, nD = 0, tD = 0, nU = 0, tU = 0;
};

template <typename T, int P, class C = CxCoef<T, P>>
__device__ __forceinline__ void line_load(const T *tile, int j, LineState<T, P> &S, T (&v)[P + 1],
                                          const C &c = C())
{
    constexpr int N1 = P + 1;
    constexpr int H = cx::EvenOdd<P>::H, ME = cx::EvenOdd<P>::ME;
    T u[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
        u[i] = tile[i * N1 + j];
        v[i] = tile[j * N1 + i];
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        S.ue[i] = u[i] + u[P - i];
        S.uo[i] = u[i] - u[P - i];
        S.ve[i] = v[i] + v[P - i];
        S.vo[i] = v[i] - v[P - i];
    }
    if (ME > H) { S.ue[ME - 1] = u[H]; S.ve[ME - 1] = v[H]; }
    S.u0 = u[0];
    S.uP = u[P];
    S.v0 = v[0];
    S.vP = v[P];
    // own face derivatives: dL = sum D0[i] u[i] = S + B, dR = sum DP[i] u[i]
    // = B - S with S = sum DA[i] ue[i] (+ D0[h] u[h]), B = sum DB[i] uo[i]
    T su = 0, bu = 0, sv = 0, bv = 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        su += c.DA(i) * S.ue[i];
        bu += c.DB(i) * S.uo[i];
        sv += c.DA(i) * S.ve[i];
        bv += c.DB(i) * S.vo[i];
    }
    if (ME > H) {
        su += c.DM() * S.ue[ME - 1];
        sv += c.DM() * S.ve[ME - 1];
    }
    S.dL = su + bu;
    S.dR = bu - su;
    S.dD = sv + bv;
    S.dU = bv - sv;
}

template <typename T, int P, class C = CxCoef<T, P>>
__device__ __forceinline__ void line_neighbours(const T *tlL, const T *tlR, const T *tlD, const T *tlU, int j,
                                                bool Lb, bool Rb, bool Db, bool Ub, LineNbr<T> &N,
                                                const C &c = C())
{
    constexpr int N1 = P + 1;
    if (!Lb) {
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            const T t = tlL[i * N1 + j];
            N.nL += c.DP(i) * t;
            if (i == P) N.tL = t;
        }
    }
    if (!Rb) {
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            const T t = tlR[i * N1 + j];
            N.nR += c.D0(i) * t;
            if (i == 0) N.tR = t;
        }
    }
    if (!Db) {
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const T t = tlD[j * N1 + k];
            N.nD += c.DP(k) * t;
            if (k == P) N.tD = t;
        }
    }
    if (!Ub) {
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const T t = tlU[j * N1 + k];
            N.nU += c.D0(k) * t;
            if (k == 0) N.tU = t;
        }
    }
}

template <typename T, int P, class C = CxCoef<T, P>>
__device__ __forceinline__ void line_finish(const LineState<T, P> &S, const LineNbr<T> &N, int j, bool Lb, bool Rb,
                                            bool Db, bool Ub, T pen, T (&Yx)[P + 1], T (&Yy)[P + 1],
                                            const C &c = C())
{
    constexpr int H = cx::EvenOdd<P>::H, ME = cx::EvenOdd<P>::ME;
    const T wj = c.w(j);
    const T pen2 = pen + pen;
    const T ev = wj * (Lb ? S.u0 : T(0.5) * (S.u0 - N.tL));
    const T fv = wj * (Rb ? -S.uP : T(0.5) * (N.tR - S.uP));
    const T Ax = Lb ? S.dL : T(0.5) * (S.dL + N.nL);
    const T Bx = Rb ? -S.dR : T(-0.5) * (S.dR + N.nR);
    const T g = wj * (Db ? S.v0 : T(0.5) * (S.v0 - N.tD));
    const T h = wj * (Ub ? -S.vP : T(0.5) * (N.tU - S.vP));
    const T Cy = Db ? S.dD : T(0.5) * (S.dD + N.nD);
    const T Dy = Ub ? -S.dU : T(-0.5) * (S.dU + N.nU);

    // Yx[a] = wj (K1 u)[a] + D0[a] ev + DP[a] fv, and the same for Yy with
    // (v, g, h).  In even-odd form the face part of pair (a, p-a) is
    // DA[a] (ev - fv) +/- DB[a] (ev + fv), and D0[h] (ev - fv) at the middle
    // node; each constant of the E/O blocks serves both directions while live.
    {
        const T s0x = ev - fv, s1x = ev + fv, s0y = g - h, s1y = g + h;
#pragma unroll
        for (int a = 0; a < ME; ++a) {
            T ex_ = 0, ey_ = 0, ox_ = 0, oy_ = 0;
#pragma unroll
            for (int i = 0; i < ME; ++i) {
                const T k = c.E(a, i);
                ex_ += k * S.ue[i];
                ey_ += k * S.ve[i];
            }
            if (a < H) {
#pragma unroll
                for (int i = 0; i < H; ++i) {
                    const T k = c.O(a, i);
                    ox_ += k * S.uo[i];
                    oy_ += k * S.vo[i];
                }
                const T e1x = wj * ex_ + c.DA(a) * s0x, o1x = wj * ox_ + c.DB(a) * s1x;
                const T e1y = wj * ey_ + c.DA(a) * s0y, o1y = wj * oy_ + c.DB(a) * s1y;
                Yx[a] = e1x + o1x;
                Yx[P - a] = e1x - o1x;
                Yy[a] = e1y + o1y;
                Yy[P - a] = e1y - o1y;
            } else {
                Yx[a] = wj * ex_ + c.DM() * s0x;
                Yy[a] = wj * ey_ + c.DM() * s0y;
            }
        }
    }
    Yx[0] += wj * Ax + (Lb ? pen : pen2) * ev;
    Yx[P] += wj * Bx - (Rb ? pen : pen2) * fv;
    Yy[0] += wj * Cy + (Db ? pen : pen2) * g;
    Yy[P] += wj * Dy - (Ub ? pen : pen2) * h;
}

template <typename T, int P, class C = CxCoef<T, P>>
__device__ __forceinline__ void line_operator(const T *tile, const T *tlL, const T *tlR, const T *tlD, const T *tlU,
                                              int j, bool Lb, bool Rb, bool Db, bool Ub, T pen, T (&Yx)[P + 1],
                                              T (&Yy)[P + 1], T (&v)[P + 1], const C &c = C())
{
    LineState<T, P> S;
    LineNbr<T> N;
    line_load<T, P, C>(tile, j, S, v, c);
    line_neighbours<T, P, C>(tlL, tlR, tlD, tlU, j, Lb, Rb, Db, Ub, N, c);
    line_finish<T, P, C>(S, N, j, Lb, Rb, Db, Ub, pen, Yx, Yy, c);
}

// The operator on the staged block (x0, y0), result written to y; returns
// this thread's share of x.y over valid elements when DOT.  Contains the
// block's barriers; with TAIL it ends with one more, after which the LDS
// image may be overwritten (not needed when the next block is staged in
// another image and a barrier precedes the reuse of this one).
//
// Thread (s, j): element s of the block, line j.  Column j of the element
// (x-direction, flux_term6a) and row j (y-direction, flux_term6b) come from
// LDS, then both sum-factorised contractions Yx[a] = (K1 U)[a][j] and
// Yy[a] = (U K1)[j][a] (K1 symmetric) in even-odd form, the face terms, and
// the assembly Y[a][b] = Yx (column j) + Yy (row j) over the element's own
// tile before the block's rows are copied out as contiguous runs.
template <typename T, int P, int BX, int BY, bool DOT, bool TAIL = true>
__device__ __forceinline__ double block_compute(T *lds, T *__restrict__ y, int dimx, int dimy,
                                                int x0, int y0, T pen)
{
    using G = BlockGeo<T, P, BX, BY>;
    constexpr int N1 = G::N1, NQ = G::NQ, TS = G::TS, NE = G::NE, NT = G::NT;
    constexpr int SL_COL = G::SL_COL, RUN = G::RUN;

    const int tid = threadIdx.x;
    const int s = tid % NE;
    int j = tid / NE;
    if (NE % 64 == 0) j = __builtin_amdgcn_readfirstlane(j);
    const int ex = s % BX, ey = s / BX;
    const int sl = (ey + 1) * BX + ex;                  // own LDS slot
    const T *tile = lds + sl * TS;
    const T *tlL = lds + (ex > 0 ? sl - 1 : SL_COL + ey) * TS;
    const T *tlR = lds + (ex < BX - 1 ? sl + 1 : SL_COL + BY + ey) * TS;
    const T *tlD = lds + (sl + BX) * TS;                // element below (y+1)
    const T *tlU = lds + (sl - BX) * TS;                // element above (y-1)
    T *otile = lds + sl * TS;

    const int gx = x0 + ex, gy = y0 + ey;
    const bool Lb = (gx == 0), Rb = (gx == dimx - 1);
    const bool Ub = (gy == 0), Db = (gy == dimy - 1);   // y grows downwards

    T v[N1], Yx[N1], Yy[N1];
    line_operator<T, P>(tile, tlL, tlR, tlD, tlU, j, Lb, Rb, Db, Ub, pen, Yx, Yy, v);

    __syncthreads();                                   // all tile reads done
#pragma unroll
    for (int a = 0; a < N1; ++a) otile[a * N1 + j] = Yx[a];
    __syncthreads();
    const bool valid = gx < dimx && gy < dimy;
    double dot = 0.0;
#pragma unroll
    for (int c = 0; c < N1; ++c) {
        const T yv = otile[j * N1 + c] + Yy[c];
        otile[j * N1 + c] = yv;
        if (DOT && valid) dot += (double)v[c] * (double)yv;
    }
    __syncthreads();
    {
        const int tid2 = opaque(threadIdx.x);
        const int GAP = dimx * NQ - RUN;
        T *yb = y + ((size_t)y0 * dimx + x0) * NQ;
        const bool full = x0 + BX <= dimx && y0 + BY <= dimy;
#pragma unroll
        for (int k = 0; k < G::KO; ++k) {
            const int idx = tid2 + k * NT;
            if ((k + 1) * NT <= G::NOUT || idx < G::NOUT) {
                const int r = idx / RUN;
                bool ok = true;
                if (!full) ok = y0 + r < dimy && x0 + (idx - r * RUN) / NQ < dimx;
                if (ok) yb[idx + r * GAP] = lds[lds_index<NQ, TS>(idx + RUN)];
            }
        }
    }
    if (TAIL) __syncthreads();                         // LDS free for the next block
    return dot;
}

}  // namespace sipg_internal

// prec_block.hpp -- the contractions of the block-Jacobi preconditioner
// z = M^-1 r on one element line, shared by the single-mesh kernels
// (kernels_pcg.hip) and the batched ones (kernels_batch_pcg.hip).
//
// M^-1 R = S [ (S^T R S) ./ (lambda_a + lambda_c) ] S^T per element, S the
// W-orthonormal eigenvectors of the interior element block (precond.cpp,
// cx::make_prec): each pass contracts one line of the element's tile with
// S or S^T.
#pragma once
#include "basis_constexpr.hpp"

namespace sipg_internal {

// The standard penalty's tables as device constants, in even-odd form where
// the eigenvectors have parities (every degree but p = 1).
template <int P>
struct PrecEO {
    static constexpr cx::PrecEvenOdd<P> V = cx::make_prec_even_odd<P>();
    static constexpr bool EO = V.ok();
    static_assert(EO || P == 1, "block-Jacobi eigenvectors without a parity");
};

// out[a] = sum_i S[i][a] in[i] (S^T in) and out[d] = sum_c S[d][c] in[c]
// (S in).  The standard penalty's table in even-odd form works on the pair
// sums e[i] = in[i] + in[p-i] and differences o[i] = in[i] - in[p-i];
// otherwise the plain sums run over the compile-time Jacobi table (CX, a
// degree without parities) or a row-major table in memory.
template <typename T, int P, bool CX>
__device__ __forceinline__ T s_coef(const T *__restrict__ S, int i, int a) {
    if constexpr (CX) return (T)PrecEO<P>::V.S0[i][a];
    else return S[i * (P + 1) + a];
}

template <typename T, int P, bool CX>
__device__ __forceinline__ void contract_st(const T *__restrict__ S, const T (&in)[P + 1], T (&out)[P + 1]) {
    constexpr int N1 = P + 1, H = N1 / 2;
    if constexpr (CX && PrecEO<P>::EO) {
        using E = PrecEO<P>;
        T e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; ++i) {
            e[i] = in[i] + in[P - i];
            o[i] = in[i] - in[P - i];
        }
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            T acc = T(0);
            if (E::V.sig[a] > 0) {
#pragma unroll
                for (int i = 0; i < H; ++i) acc += (T)E::V.S[i][a] * e[i];
                if (P % 2 == 0) acc += (T)E::V.S[H][a] * in[H];
            } else {
#pragma unroll
                for (int i = 0; i < H; ++i) acc += (T)E::V.S[i][a] * o[i];
            }
            out[a] = acc;
        }
    } else {
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < N1; ++i) acc += s_coef<T, P, CX>(S, i, a) * in[i];
            out[a] = acc;
        }
    }
}

template <typename T, int P, bool CX>
__device__ __forceinline__ void contract_s(const T *__restrict__ S, const T (&in)[P + 1], T (&out)[P + 1]) {
    constexpr int N1 = P + 1, H = N1 / 2;
    if constexpr (CX && PrecEO<P>::EO) {
        using E = PrecEO<P>;
#pragma unroll
        for (int d = 0; d < H; ++d) {
            T ev = T(0), od = T(0);
#pragma unroll
            for (int c = 0; c < N1; ++c) {
                if (E::V.sig[c] > 0) ev += (T)E::V.S[d][c] * in[c];
                else od += (T)E::V.S[d][c] * in[c];
            }
            out[d] = ev + od;
            out[P - d] = ev - od;
        }
        if (P % 2 == 0) {
            T ev = T(0);
#pragma unroll
            for (int c = 0; c < N1; ++c)
                if (E::V.sig[c] > 0) ev += (T)E::V.S[H][c] * in[c];
            out[H] = ev;
        }
    } else {
#pragma unroll
        for (int d = 0; d < N1; ++d) {
            T acc = T(0);
#pragma unroll
            for (int c = 0; c < N1; ++c) acc += s_coef<T, P, CX>(S, d, c) * in[c];
            out[d] = acc;
        }
    }
}

}  // namespace sipg_internal

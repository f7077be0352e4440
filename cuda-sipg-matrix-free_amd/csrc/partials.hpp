// partials.hpp -- the first stage of every reduction of block sums, and the
// compensated update of an iterate.
#pragma once
#include <hip/hip_runtime.h>

namespace sipg_internal {

// Thread t's share of part[b] + ... + part[e-1]: the entries b + t,
// b + t + NT, b + t + 2 NT, ... in eight accumulators, so that eight loads
// are in flight at once (a loop with one accumulator waits for each load
// before issuing the next: (e - b) / NT dependent round trips to the cache,
// which for a few thousand block sums is several microseconds at the start
// of every workgroup that needs them).  Added in a fixed order.
template <int NT>
__device__ __forceinline__ double strided_sum(const double *__restrict__ part, int b, int e) {
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int i = b + (int)threadIdx.x;
    for (; i + 7 * NT < e; i += 8 * NT) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += part[i + u * NT];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (i + u * NT < e) a[u] += part[i + u * NT];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// xh += a p with the rounding errors kept in xl, so that xh + xl carries
// the running sum to about twice the working precision: the product's
// error from an fma, the sum's by Knuth's TwoSum (both exact).  Contraction
// is off so that the error terms are formed as written.  Used for the
// increments of the single-problem solvers: accumulated plainly, their
// rounding (eps |x| per step, amplified by |A|) leaves b - A x about 1e-9
// of |b| away from the recursion at 256 x 256, p = 8.
template <typename T>
__device__ __forceinline__ void comp_axpy(T &xh, T &xl, T a, T pv) {
#pragma clang fp contract(off)
    const T t = a * pv;
    const T tp = fma(a, pv, -t);
    const T s = xh + t;
    const T bb = s - xh;
    const T e = (xh - (s - bb)) + (t - bb);
    xh = s;
    xl = xl + (e + tp);
}

}  // namespace sipg_internal

// partials.hpp -- the first stage of every reduction of block sums.
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

}  // namespace sipg_internal

/*
 * thunderrw_device.h -- device-side definitions shared by the library's walk kernels and
 * by kernels compiled into an application with hipcc (the drop-in functor engine,
 * thunderrw_amd/dropin/functor_walk.h), so that both draw the same random words and make
 * the same comparisons:
 *   - the SFMT conversions real1 / real2 (SFMT.h:193-228),
 *   - Philox-4x32-10 (Salmon et al., SC'11) as the library keys it: key = the walk seed's
 *     two 32-bit halves, counter = (global walker id low, high, step, trial); a move's draws
 *     are words .x/.y of its trial's block, a PPR stop draw word .w of the step's trial-0 block,
 *   - Utility::upper_bound_branchless (util/utility.h:18-27).
 * HIP only (include after <hip/hip_runtime.h>).
 */
#ifndef THUNDERRW_DEVICE_H
#define THUNDERRW_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trw_dev {

constexpr double kToReal1 = 1.0 / 4294967295.0;  // SFMT.h:193-207
constexpr double kToReal2 = 1.0 / 4294967296.0;  // SFMT.h:214-228

__device__ __forceinline__ double real1(uint32_t v) { return static_cast<double>(v) * kToReal1; }
__device__ __forceinline__ double real2(uint32_t v) { return static_cast<double>(v) * kToReal2; }

// Philox-4x32-10: ten rounds of two 32x32->64 multiplies, Weyl-sequence key schedule.
__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return make_uint4(c0, c1, c2, c3);
}

// The block of (walker id, step, trial) under the walk seed.
__device__ __forceinline__ uint4 walk_block(uint64_t seed, uint64_t walker_id, uint32_t step, uint32_t trial) {
    return philox4x32_10(static_cast<uint32_t>(walker_id), static_cast<uint32_t>(walker_id >> 32), step, trial,
                         static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

// The branchless binary search of util/utility.h:18-27 over src[0, n): for a
// non-decreasing array, the number of entries <= target (std::upper_bound).
__device__ __forceinline__ int32_t upper_bound_branchless(const double *src, int32_t n, double target) {
    const double *base = src;
    while (n > 1) {
        const int32_t half = n >> 1;
        base = (base[half] <= target) ? base + half : base;
        n -= half;
    }
    return static_cast<int32_t>((*base <= target) + (base - src));
}

}  // namespace trw_dev

#endif /* THUNDERRW_DEVICE_H */

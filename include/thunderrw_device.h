/*
 * thunderrw_device.h -- device-side definitions shared by the library's walk kernels and
 * by kernels compiled into an application with hipcc (the drop-in functor engine,
 * thunderrw_amd/dropin/functor_walk.h), so that both draw the same random words and make
 * the same comparisons:
 *   - the SFMT conversions real1 / real2 (SFMT.h:193-228),
 *   - Philox-4x32-10 (Salmon et al., SC'11) as the library keys it: key = the walk seed's
 *     two 32-bit halves, counter = (global walker id low, high, step, trial); a move's draws
 *     are words .x/.y of its trial's block, a PPR stop draw word .w of the step's trial-0 block,
 *   - Utility::upper_bound_branchless (util/utility.h:18-27),
 *   - the membership tables Graph::is_neighbor (graph.h:180-199) is answered from on the device
 *     (trw_build_membership builds them; trw_graph_device_view gives their address).
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

// Membership table of one vertex's adjacency, for every vertex of degree d > kMemberMin (0: all
// with neighbours): a hash set of its neighbours in nb = 2^member_bucket_bits(d) buckets of 128
// bytes (32 int32 slots, up to 16 keys per bucket on average; one bucket up to degree 16).  Neighbour u lives in bucket member_hash(u) or, when that one
// is full, the next bucket (cyclically) with room; a bucket's keys fill its slots from the first,
// the other slots hold kMemberEmpty.  A vertex's table starts at line root[v] (kNoRoot: none).
constexpr int kMemberSlots = 32;
constexpr int kMemberMin = 0;
constexpr uint32_t kMemberEmpty = 0xffffffffu;
constexpr uint32_t kNoRoot = 0xffffffffu;

__host__ __device__ inline uint32_t member_bucket_bits(int32_t d) {
    const uint32_t need = static_cast<uint32_t>((d + 15) / 16);
    return need <= 1u ? 0u : 32u - static_cast<uint32_t>(__builtin_clz(need - 1u));
}

__host__ __device__ inline uint32_t member_hash(int32_t u, uint32_t bits) {
    return bits == 0u ? 0u : (static_cast<uint32_t>(u) * 0x9E3779B1u) >> (32u - bits);
}

// Half of a bucket: 16 keys in 4 16-byte loads issued together.  A search loads the second
// half only when the first cannot decide: a whole line in flight would take 32 registers.
struct MemberHalf {
    uint4 a0, a1, a2, a3;
    __host__ __device__ inline void load(const uint4 *p) {
        a0 = p[0];
        a1 = p[1];
        a2 = p[2];
        a3 = p[3];
    }
    static __host__ __device__ inline bool eq4(const uint4 &v, uint32_t x) {
        return (v.x == x) | (v.y == x) | (v.z == x) | (v.w == x);
    }
    __host__ __device__ inline bool contains(uint32_t x) const { return eq4(a0, x) | eq4(a1, x) | eq4(a2, x) | eq4(a3, x); }
    __host__ __device__ inline bool full() const { return a3.w != kMemberEmpty; }
};

// Whether x is among the d neighbours whose table starts at line root (lines: 8 uint4 each):
// x's bucket, its second half only when the first is full of other keys, the next bucket only
// when this one is full without x.  Almost always one 128-byte line.
__host__ __device__ inline bool member_contains(const uint4 *lines, uint32_t root, int32_t d, int32_t x) {
    const uint32_t bits = member_bucket_bits(d), nb = 1u << bits, y = static_cast<uint32_t>(x);
    MemberHalf h;
    uint32_t b = member_hash(x, bits);
    for (uint32_t i = 0; i < nb; ++i) {
        const uint4 *p = lines + static_cast<uint64_t>(root + b) * 8;
        h.load(p);
        if (h.contains(y)) return true;
        if (!h.full()) return false;
        h.load(p + 4);
        if (h.contains(y)) return true;
        if (!h.full()) return false;
        b = (b + 1u) & (nb - 1u);
    }
    return false;
}

}  // namespace trw_dev

#endif /* THUNDERRW_DEVICE_H */

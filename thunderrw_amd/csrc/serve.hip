// Batch assembly for the query server: expands a batch's segment table (one
// segment per query chunk) into the per-walker start vertices and Philox ids
// the walk kernels read, so a batch crosses PCIe as its segment table and the
// queries' source sets rather than as 12 bytes per walker.
#include "engine.h"

namespace trw {
namespace {

constexpr int kExpandBlock = 256;

__global__ __launch_bounds__(kExpandBlock) void serve_expand_kernel(const ServeSegment *__restrict__ segs,
                                                                    int32_t n_segs, const int32_t *__restrict__ pool,
                                                                    int32_t n, int32_t *__restrict__ sources,
                                                                    uint64_t *__restrict__ ids) {
    const int32_t i = static_cast<int32_t>(blockIdx.x) * kExpandBlock + static_cast<int32_t>(threadIdx.x);
    if (i >= n) return;
    // Last segment whose batch offset is <= i (segments are in batch order, first at 0).
    int32_t lo = 0, hi = n_segs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].batch_offset <= i) lo = mid;
        else hi = mid - 1;
    }
    const ServeSegment s = segs[lo];
    const int64_t j = static_cast<int64_t>(s.first_walker) + (i - s.batch_offset);
    sources[i] = pool[s.pool_offset + static_cast<int32_t>(j % s.n_sources)];
    ids[i] = s.id_base + static_cast<uint64_t>(j);
}

}  // namespace

hipError_t launch_serve_expand(const ServeSegment *segs, int32_t n_segs, const int32_t *pool, int32_t n,
                               int32_t *sources, uint64_t *ids, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int blocks = (n + kExpandBlock - 1) / kExpandBlock;
    hipLaunchKernelGGL(serve_expand_kernel, dim3(blocks), dim3(kExpandBlock), 0, stream, segs, n_segs, pool, n,
                       sources, ids);
    return hipGetLastError();
}

}  // namespace trw

// ---------------------------------------------------------------- PPR aggregation
//
// A PPR batch's result is, per query segment, the end vertices with their walker
// counts.  It is built on the device: each walker becomes the key (segment << 32 | end
// vertex), or (segment << 32 | 0xffffffff) when it ended in error; a radix sort over
// the low 32 + log2(segments) bits groups the keys by segment and vertex, a run-length
// encoding turns them into (key, count) runs, and a segmented sum gives each segment's
// steps (sum of length - 1).  Only runs and per-segment sums cross PCIe.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_segmented_reduce.hpp>

namespace trw {
namespace {

__global__ __launch_bounds__(kExpandBlock) void serve_keys_kernel(const ServeSegment *__restrict__ segs, int32_t n_segs,
                                                                  int32_t n, const int32_t *__restrict__ ends,
                                                                  const int32_t *__restrict__ lengths,
                                                                  const int32_t *__restrict__ status,
                                                                  uint64_t *__restrict__ keys,
                                                                  uint64_t *__restrict__ steps) {
    const int32_t i = static_cast<int32_t>(blockIdx.x) * kExpandBlock + static_cast<int32_t>(threadIdx.x);
    if (i >= n) return;
    int32_t lo = 0, hi = n_segs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].batch_offset <= i) lo = mid;
        else hi = mid - 1;
    }
    const bool ok = status[i] < 3;  // DONE, GATHER_TERMINATED, DEAD_END end at a vertex
    keys[i] = (static_cast<uint64_t>(lo) << 32) | (ok ? static_cast<uint32_t>(ends[i]) : 0xffffffffu);
    const int32_t len = lengths[i];
    steps[i] = static_cast<uint64_t>(len > 1 ? len - 1 : 0);
}

unsigned int key_bits(int32_t n_segs) {
    unsigned int b = 32;
    while ((int64_t(1) << (b - 32)) < n_segs) ++b;
    return b;
}

}  // namespace

hipError_t serve_ppr_temp_bytes(int32_t max_walkers, int32_t max_segments, size_t *bytes) {
    size_t a = 0, b = 0, c = 0;
    uint64_t *k = nullptr;
    uint32_t *cnt = nullptr;
    int32_t *off = nullptr;
    hipError_t e = rocprim::radix_sort_keys(nullptr, a, k, k, max_walkers, 0u, key_bits(max_segments), nullptr);
    if (e == hipSuccess) e = rocprim::run_length_encode(nullptr, b, k, static_cast<unsigned int>(max_walkers), k, cnt, cnt, nullptr);
    if (e == hipSuccess)
        e = rocprim::segmented_reduce(nullptr, c, k, k, static_cast<unsigned int>(max_segments), off, off + 1,
                                      rocprim::plus<uint64_t>(), uint64_t(0), nullptr);
    *bytes = std::max(a, std::max(b, c));
    return e;
}

hipError_t launch_serve_ppr_reduce(const ServeSegment *segs, const int32_t *seg_offsets, int32_t n_segs, int32_t n,
                                   const int32_t *ends, const int32_t *lengths, const int32_t *status,
                                   uint64_t *keys_in, uint64_t *keys_out, uint64_t *walker_steps, uint64_t *run_keys,
                                   uint32_t *run_counts, uint32_t *num_runs, uint64_t *seg_steps, void *temp,
                                   size_t temp_bytes, hipStream_t stream) {
    if (n <= 0 || n_segs <= 0) return hipSuccess;
    hipLaunchKernelGGL(serve_keys_kernel, dim3((n + kExpandBlock - 1) / kExpandBlock), dim3(kExpandBlock), 0, stream,
                       segs, n_segs, n, ends, lengths, status, keys_in, walker_steps);
    hipError_t e = hipGetLastError();
    size_t bytes = temp_bytes;
    if (e == hipSuccess)
        e = rocprim::radix_sort_keys(temp, bytes, keys_in, keys_out, n, 0u, key_bits(n_segs), stream);
    bytes = temp_bytes;
    if (e == hipSuccess)
        e = rocprim::run_length_encode(temp, bytes, keys_out, static_cast<unsigned int>(n), run_keys, run_counts,
                                       num_runs, stream);
    bytes = temp_bytes;
    if (e == hipSuccess)
        e = rocprim::segmented_reduce(temp, bytes, walker_steps, seg_steps, static_cast<unsigned int>(n_segs),
                                      seg_offsets, seg_offsets + 1, rocprim::plus<uint64_t>(), uint64_t(0), stream);
    return e;
}

}  // namespace trw

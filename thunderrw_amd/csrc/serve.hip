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

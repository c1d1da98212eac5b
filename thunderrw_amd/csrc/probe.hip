// Random-gather ceiling probe: the access pattern of a walk step without the walk.
//
// Every lane follows `chains` independent pointer chains through a table of
// 16-byte records: the next record index is a hash of the record just read, so
// each load depends on the previous one exactly as a walker's next vertex
// depends on the alias slot it read.  The result is the rate of dependent random
// 16-byte gathers the memory system sustains at full occupancy, which bounds any
// walk kernel whose step is one such gather.
#include "engine.h"

namespace trw {
namespace {

constexpr int kProbeBlock = 256;

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__global__ void probe_fill_kernel(uint4 *table, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint32_t h = mix32(static_cast<uint32_t>(i) ^ 0x9e3779b9U);
        table[i] = make_uint4(h, mix32(h), mix32(h + 1), mix32(h + 2));
    }
}

// RECORD 16-byte loads per gather, adjacent (a RECORD * 16-byte record in one line).
template <int CHAINS, int RECORD>
__global__ __launch_bounds__(kProbeBlock) void probe_gather_kernel(const uint4 *__restrict__ table, uint32_t n,
                                                                   int32_t steps, uint32_t *out) {
    const uint32_t lane = blockIdx.x * kProbeBlock + threadIdx.x;
    uint32_t idx[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) idx[c] = mix32(lane * CHAINS + c) % n;
    uint32_t acc = 0;
    for (int32_t s = 0; s < steps; ++s) {
        uint4 r[CHAINS][RECORD];
#pragma unroll
        for (int c = 0; c < CHAINS; ++c)
#pragma unroll
            for (int q = 0; q < RECORD; ++q) r[c][q] = table[static_cast<uint64_t>(idx[c]) * RECORD + q];
        // Every word of the record feeds the next index, so the compiler keeps each load a
        // whole 16-byte load (with words unused it narrows them, which is not a walk's read).
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) {
            uint32_t h = 0;
#pragma unroll
            for (int q = 0; q < RECORD; ++q) h ^= r[c][q].x ^ mix32(r[c][q].y) ^ (r[c][q].z * 3u) ^ (r[c][q].w * 5u);
            idx[c] = (h ^ static_cast<uint32_t>(s)) % n;
            acc += r[c][0].y;
        }
    }
    out[lane] = acc;
}

}  // namespace

template <int RECORD>
void launch_probe(int32_t chains, dim3 grid, dim3 block, hipStream_t stream, const uint4 *table, uint32_t n,
                  int32_t steps, uint32_t *out) {
    switch (chains) {
    case 1: hipLaunchKernelGGL((probe_gather_kernel<1, RECORD>), grid, block, 0, stream, table, n, steps, out); break;
    case 2: hipLaunchKernelGGL((probe_gather_kernel<2, RECORD>), grid, block, 0, stream, table, n, steps, out); break;
    case 4: hipLaunchKernelGGL((probe_gather_kernel<4, RECORD>), grid, block, 0, stream, table, n, steps, out); break;
    default: hipLaunchKernelGGL((probe_gather_kernel<8, RECORD>), grid, block, 0, stream, table, n, steps, out); break;
    }
}

hipError_t run_gather_probe(uint64_t table_bytes, int32_t record_bytes, int32_t chains, int32_t steps, int32_t cus,
                            hipStream_t stream, double *loads_per_s) {
    const int32_t record = record_bytes == 32 ? 2 : 1;
    const int64_t n = static_cast<int64_t>(table_bytes / (16 * record));  // records
    if (n < 1 || n > 0xffffffffLL) return hipErrorInvalidValue;
    uint4 *table = nullptr;
    uint32_t *out = nullptr;
    const int64_t lanes = static_cast<int64_t>(cus) * 2048;  // 32 wavefronts of 64 lanes per CU
    hipError_t e = hipMalloc(&table, static_cast<size_t>(n) * 16 * record);
    if (e != hipSuccess) return e;
    e = hipMalloc(&out, static_cast<size_t>(lanes) * 4);
    if (e != hipSuccess) {
        (void)hipFree(table);
        return e;
    }
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipLaunchKernelGGL(probe_fill_kernel, dim3(4096), dim3(256), 0, stream, table, n * record);
    const dim3 grid(static_cast<unsigned>(lanes / kProbeBlock)), block(kProbeBlock);
    for (int rep = 0; rep < 2; ++rep) {  // first launch warms the caches and the clocks
        (void)hipEventRecord(a, stream);
        if (record == 2) launch_probe<2>(chains, grid, block, stream, table, static_cast<uint32_t>(n), steps, out);
        else launch_probe<1>(chains, grid, block, stream, table, static_cast<uint32_t>(n), steps, out);
        (void)hipEventRecord(b, stream);
    }
    e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(table);
    (void)hipFree(out);
    const int c = (chains == 1 || chains == 2 || chains == 4) ? chains : 8;
    if (e == hipSuccess) *loads_per_s = static_cast<double>(lanes) * c * steps / (ms / 1e3);
    return e;
}

}  // namespace trw

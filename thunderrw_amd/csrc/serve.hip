// Batch assembly for the query server: expands a batch's segment table (one
// segment per query chunk) into the per-walker start vertices and Philox ids
// the walk kernels read, so a batch crosses PCIe as its segment table and the
// queries' source sets rather than as 12 bytes per walker.
#include <algorithm>

#include "engine.h"

namespace trw {
namespace {

constexpr int kExpandBlock = 256;

__global__ __launch_bounds__(kExpandBlock) void serve_expand_kernel(const ServeSegment *__restrict__ segs,
                                                                    int32_t n_segs, const int32_t *__restrict__ pool,
                                                                    int32_t n, int32_t *__restrict__ sources,
                                                                    uint64_t *__restrict__ ids,
                                                                    uint32_t *__restrict__ zero, int32_t n_zero) {
    const int32_t i = static_cast<int32_t>(blockIdx.x) * kExpandBlock + static_cast<int32_t>(threadIdx.x);
    if (i < n_zero) zero[i] = 0u;
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
                               int32_t *sources, uint64_t *ids, uint32_t *zero, int32_t n_zero, hipStream_t stream) {
    const int32_t m = std::max(n, n_zero);
    if (m <= 0) return hipSuccess;
    const int blocks = (m + kExpandBlock - 1) / kExpandBlock;
    hipLaunchKernelGGL(serve_expand_kernel, dim3(blocks), dim3(kExpandBlock), 0, stream, segs, n_segs, pool, n,
                       sources, ids, zero, n_zero);
    return hipGetLastError();
}

}  // namespace trw

// ---------------------------------------------------------------- PPR aggregation
//
// A PPR batch's result is, per query segment, its distinct end vertices in ascending
// order with their walker counts, plus the segment's steps and error walkers.  It is
// built with no sort and no global atomic (on this part every device-memory atomic is
// performed at the memory side, far below the HBM rate), as a partition of each
// segment's end vertices into chunks of 2^17 vertex ids, then LDS work per chunk:
//   hist    - per tile (2048 walkers of one segment): walkers per chunk (LDS counters),
//             the tile's steps and errors; the segment's last tile to finish (a ticket
//             per segment, taken with acquire-release ordering) then scans the segment: each
//             (tile, chunk) pair's first position in the segment's bucket order
//             (chunk-major, tile-minor), each bucket's extent, the segment's totals;
//   scatter - per tile: each walker's end vertex into its bucket (LDS cursors);
//   chunk   - per (segment, chunk) bucket: the bucket's vertices into a 16 KiB LDS
//             bitmap and their walker counts into LDS; reading the bitmap in order, with
//             a popcount prefix for each thread's first rank, writes the (vertex, count)
//             runs ascending at the bucket's first position, then copies them and the
//             bucket's extent into host memory with consecutive threads on consecutive
//             runs (whole lines over PCIe rather than one 8-byte write per run).
// Three launches per batch: each one's fixed cost (a few microseconds) is latency a
// lightly loaded server's queries pay in full.
// Runs of segment s lie in its own walker positions, bucket by bucket: buckets are in
// vertex order and runs within a bucket ascend, so the segment's runs read in bucket
// order are ascending.

namespace trw {
namespace {

constexpr int kTileBlock = 256;
constexpr int kTilePerThread = 8;
constexpr int kTileWalkers = kTileBlock * kTilePerThread;  // hist and scatter: one tile per block
constexpr int kChunkShift = 17;                            // 2^17 vertex ids per chunk
constexpr int kChunkWords = (1 << kChunkShift) / 64;       // 2048 bitmap words (16 KiB)
constexpr int kChunkBlock = 256;
constexpr int kWordsPerThread = kChunkWords / kChunkBlock;  // 8
constexpr int kAux = 4096;                                 // LDS words beside the bitmap (16 KiB)
constexpr int kHashMax = kAux / 2;                         // buckets up to this many walkers count in a hash
constexpr uint32_t kHashEmpty = 0xffffffffu;               // never a packed entry: counts stay <= kHashMax
constexpr int kRankCap = kAux - kChunkWords;               // run counters per pass of the large-bucket path

__device__ inline uint64_t block_sum_u64(uint64_t x, uint64_t *lds) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    const int tid = static_cast<int>(threadIdx.x);
    if ((tid & 63) == 0) lds[tid >> 6] = x;
    __syncthreads();
    uint64_t t = 0;
    for (int j = 0; j < static_cast<int>(blockDim.x) / 64; ++j) t += lds[j];
    __syncthreads();
    return t;
}

__device__ inline uint32_t wave_inclusive_u32(uint32_t x, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive prefix of x over the block (x of all threads summed into *total).
__device__ inline uint32_t block_exclusive_u32(uint32_t x, uint32_t *lds, uint32_t *total) {
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const uint32_t incl = wave_inclusive_u32(x, lane);
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int j = 0; j < static_cast<int>(blockDim.x) / 64; ++j) {
        if (j < wave) before += lds[j];
        all += lds[j];
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

// Segment of tile t: the last segment whose first tile is <= t.
__device__ inline int32_t tile_segment(const ServeSegment *__restrict__ segs, int32_t n_segs, int32_t t) {
    int32_t lo = 0, hi = n_segs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].first_tile <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The walker range [*i0, *i1) of tile t and its segment.
__device__ inline int32_t tile_range(const ServeSegment *__restrict__ segs, int32_t n_segs, int32_t t, int32_t *s_lds,
                                     int32_t *i0, int32_t *i1) {
    if (threadIdx.x == 0) *s_lds = tile_segment(segs, n_segs, t);
    __syncthreads();
    const int32_t s = *s_lds;
    const ServeSegment sg = segs[s];
    *i0 = sg.batch_offset + (t - sg.first_tile) * kTileWalkers;
    *i1 = min(*i0 + kTileWalkers, sg.batch_offset + sg.count);
    return s;
}

// A walker that ended at a vertex (DONE, GATHER_TERMINATED, DEAD_END) is counted at it.
__device__ inline bool counted(int32_t status, int32_t v, int32_t nv) { return status < 3 && v >= 0 && v < nv; }

// Segment s's scan, by one block of kTileBlock threads: tile_hist[t][c] becomes the first
// bucket position of tile t's walkers in chunk c; bucket[s][c] = {first position, walkers};
// the segment's totals go to host memory.
__device__ void scan_segment(const ServeSegment &sg, int32_t s, const PprScratch &p,
                             PprSegmentStats *__restrict__ host_stats, uint32_t *lds, uint64_t *red) {
    const int tid = static_cast<int>(threadIdx.x);
    const int32_t t0 = sg.first_tile, nt = (sg.count + kTileWalkers - 1) / kTileWalkers;
    uint32_t carry = static_cast<uint32_t>(sg.batch_offset);
    for (int32_t c0 = 0; c0 < p.C; c0 += kTileBlock) {
        const int32_t c = c0 + tid;
        uint32_t col = 0;
        if (c < p.C)
            for (int32_t t = 0; t < nt; ++t) col += p.tile_hist[static_cast<int64_t>(t0 + t) * p.C + c];
        uint32_t total;
        const uint32_t first = carry + block_exclusive_u32(col, lds, &total);
        if (c < p.C) {
            p.bucket[static_cast<int64_t>(s) * p.C + c] = make_uint2(first, col);
            uint32_t pos = first;
            for (int32_t t = 0; t < nt; ++t) {
                uint32_t *h = p.tile_hist + static_cast<int64_t>(t0 + t) * p.C + c;
                const uint32_t k = *h;
                *h = pos;
                pos += k;
            }
        }
        carry += total;
    }
    uint64_t steps = 0, errors = 0;
    for (int32_t t = tid; t < nt; t += kTileBlock) {
        steps += p.tile_stats[t0 + t].steps;
        errors += p.tile_stats[t0 + t].errors;
    }
    steps = block_sum_u64(steps, red);
    errors = block_sum_u64(errors, red);
    if (tid == 0) {
        PprSegmentStats o;
        o.steps = steps;
        o.errors = errors;
        host_stats[s] = o;
    }
}

// One block per tile.  p.tickets[s] (zeroed with the batch's control block) counts the
// segment's finished tiles; the block that takes the last ticket scans the segment.  Each
// block's counts are released (device scope) with its ticket and acquired by the scanning
// block with its own, so the scan sees every tile of the segment whichever XCD ran it.
__global__ __launch_bounds__(kTileBlock) void ppr_hist_kernel(const ServeSegment *__restrict__ segs, int32_t n_segs,
                                                              const int32_t *__restrict__ ends,
                                                              const int32_t *__restrict__ lengths,
                                                              const int32_t *__restrict__ status, PprScratch p,
                                                              PprSegmentStats *__restrict__ host_stats) {
    extern __shared__ uint32_t hist[];  // C counters
    __shared__ uint64_t red[kTileBlock / 64];
    __shared__ uint32_t lds[kTileBlock / 64];
    __shared__ int32_t s_lds;
    __shared__ int32_t last;
    const int tid = static_cast<int>(threadIdx.x);
    const int32_t t = static_cast<int32_t>(blockIdx.x);
    for (int32_t c = tid; c < p.C; c += kTileBlock) hist[c] = 0u;
    int32_t i0, i1;
    const int32_t s = tile_range(segs, n_segs, t, &s_lds, &i0, &i1);  // its barrier also orders the zeroing
    int32_t v[kTilePerThread], st[kTilePerThread], len[kTilePerThread];
#pragma unroll
    for (int k = 0; k < kTilePerThread; ++k) {
        const int32_t i = i0 + k * kTileBlock + tid;
        v[k] = i < i1 ? ends[i] : -1;
        st[k] = i < i1 ? status[i] : 0;
        len[k] = i < i1 ? lengths[i] : 0;
    }
    uint64_t steps = 0, errors = 0;
#pragma unroll
    for (int k = 0; k < kTilePerThread; ++k) {
        if (i0 + k * kTileBlock + tid >= i1) continue;
        steps += static_cast<uint64_t>(len[k] > 1 ? len[k] - 1 : 0);
        if (counted(st[k], v[k], p.nv)) atomicAdd(&hist[v[k] >> kChunkShift], 1u);
        else ++errors;
    }
    steps = block_sum_u64(steps, red);
    errors = block_sum_u64(errors, red);  // its barriers also publish hist
    if (tid == 0) {
        p.tile_stats[t].steps = steps;
        p.tile_stats[t].errors = errors;
    }
    for (int32_t c = tid; c < p.C; c += kTileBlock) p.tile_hist[static_cast<int64_t>(t) * p.C + c] = hist[c];
    const ServeSegment sg = segs[s];
    // Every wave's counts are written at the barrier; one thread's acquire-release ticket then
    // releases them (this XCD's L2 written back) and, for the last block, acquires the other
    // blocks' (stale lines invalidated).  A fence in every wave instead (each one a write-back
    // and an invalidate of the L2 under the other slots' walks) cost 9% of the saturated rate.
    __syncthreads();
    if (tid == 0) {
        const uint32_t nt = static_cast<uint32_t>((sg.count + kTileWalkers - 1) / kTileWalkers);
        last = __hip_atomic_fetch_add(&p.tickets[s], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == nt - 1u;
    }
    __syncthreads();
    if (!last) return;
    scan_segment(sg, s, p, host_stats, lds, red);
}

__global__ __launch_bounds__(kTileBlock) void ppr_scatter_kernel(const ServeSegment *__restrict__ segs,
                                                                 int32_t n_segs, const int32_t *__restrict__ ends,
                                                                 const int32_t *__restrict__ status, PprScratch p) {
    extern __shared__ uint32_t cursor[];  // C bucket positions
    __shared__ int32_t s_lds;
    const int tid = static_cast<int>(threadIdx.x);
    const int32_t t = static_cast<int32_t>(blockIdx.x);
    for (int32_t c = tid; c < p.C; c += kTileBlock) cursor[c] = p.tile_hist[static_cast<int64_t>(t) * p.C + c];
    int32_t i0, i1;
    tile_range(segs, n_segs, t, &s_lds, &i0, &i1);  // its barrier also publishes the cursors
    int32_t v[kTilePerThread], st[kTilePerThread];
#pragma unroll
    for (int k = 0; k < kTilePerThread; ++k) {
        const int32_t i = i0 + k * kTileBlock + tid;
        v[k] = i < i1 ? ends[i] : -1;
        st[k] = i < i1 ? status[i] : 3;
    }
#pragma unroll
    for (int k = 0; k < kTilePerThread; ++k)
        if (counted(st[k], v[k], p.nv)) p.bucketed[atomicAdd(&cursor[v[k] >> kChunkShift], 1u)] = v[k];
}

__device__ inline uint32_t hash_slot(uint32_t lv) { return (lv * 2654435761u) >> 20; }  // top 12 bits: kAux slots

// grid (C, segments): one bucket per block.  Its vertices go into an LDS bitmap of the
// chunk; the set bits, read in order with a popcount prefix giving each thread its first
// rank, are the bucket's runs in ascending order.  Counts: a bucket of at most kHashMax
// walkers counts in one pass, in an LDS hash of packed (vertex in chunk, count) words; a
// larger one ranks each walker (word prefix + popcount) and counts kRankCap ranks a pass.
__global__ __launch_bounds__(kChunkBlock) void ppr_chunk_kernel(PprScratch p, int2 *__restrict__ host_pairs,
                                                                uint2 *__restrict__ host_buckets) {
    __shared__ unsigned long long bm[kChunkWords];
    __shared__ uint32_t aux[kAux];
    __shared__ uint32_t lds[kChunkBlock / 64];
    const int tid = static_cast<int>(threadIdx.x);
    const int32_t c = static_cast<int32_t>(blockIdx.x), s = static_cast<int32_t>(blockIdx.y);
    const uint2 b = p.bucket[static_cast<int64_t>(s) * p.C + c];  // {first position, walkers}
    if (b.y == 0) {
        if (tid == 0) host_buckets[static_cast<int64_t>(s) * p.C + c] = make_uint2(b.x, 0u);
        return;
    }
    const int32_t vbase = c << kChunkShift;
    const bool hashed = b.y <= static_cast<uint32_t>(kHashMax);
    for (int j = tid; j < kChunkWords; j += kChunkBlock) bm[j] = 0ull;
    if (hashed)
        for (int j = tid; j < kAux; j += kChunkBlock) aux[j] = kHashEmpty;
    __syncthreads();
    for (uint32_t j = static_cast<uint32_t>(tid); j < b.y; j += kChunkBlock) {
        const uint32_t lv = static_cast<uint32_t>(p.bucketed[b.x + j] - vbase);
        if (lv >= (1u << kChunkShift)) continue;
        atomicOr(&bm[lv >> 6], 1ull << (lv & 63));
        if (!hashed) continue;
        for (uint32_t h = hash_slot(lv), probe = 0; probe < static_cast<uint32_t>(kAux); ++probe, h = (h + 1) & (kAux - 1)) {
            const uint32_t old = atomicCAS(&aux[h], kHashEmpty, lv | (1u << kChunkShift));
            if (old == kHashEmpty) break;
            if ((old & ((1u << kChunkShift) - 1)) == lv) {
                atomicAdd(&aux[h], 1u << kChunkShift);
                break;
            }
        }
    }
    __syncthreads();
    // Thread tid owns words [w0, w0 + kWordsPerThread): their ranks start at its exclusive prefix.
    const int w0 = tid * kWordsPerThread;
    uint32_t mine = 0;
#pragma unroll
    for (int j = 0; j < kWordsPerThread; ++j) mine += static_cast<uint32_t>(__popcll(bm[w0 + j]));
    uint32_t runs;
    const uint32_t first_rank = block_exclusive_u32(mine, lds, &runs);
    if (hashed) {
        uint32_t rank = first_rank;
        for (int j = 0; j < kWordsPerThread; ++j) {
            unsigned long long x = bm[w0 + j];
            while (x) {
                const int bit = __ffsll(x) - 1;
                x &= x - 1;
                const uint32_t lv = static_cast<uint32_t>((w0 + j) * 64 + bit);
                uint32_t n = 0;
                for (uint32_t h = hash_slot(lv), probe = 0; probe < static_cast<uint32_t>(kAux);
                     ++probe, h = (h + 1) & (kAux - 1)) {
                    const uint32_t e = aux[h];
                    if (e != kHashEmpty && (e & ((1u << kChunkShift) - 1)) == lv) {
                        n = e >> kChunkShift;
                        break;
                    }
                }
                p.pairs[b.x + rank] = make_int2(vbase + static_cast<int32_t>(lv), static_cast<int32_t>(n));
                ++rank;
            }
        }
    } else {
        uint32_t *wpre = aux, *cnt = aux + kChunkWords;
        {
            uint32_t r = first_rank;
            for (int j = 0; j < kWordsPerThread; ++j) {
                wpre[w0 + j] = r;
                r += static_cast<uint32_t>(__popcll(bm[w0 + j]));
            }
        }
        __syncthreads();
        for (uint32_t r0 = 0; r0 < runs; r0 += kRankCap) {
            const uint32_t m = min(static_cast<uint32_t>(kRankCap), runs - r0);
            for (uint32_t j = static_cast<uint32_t>(tid); j < m; j += kChunkBlock) cnt[j] = 0u;
            __syncthreads();
            for (uint32_t j = static_cast<uint32_t>(tid); j < b.y; j += kChunkBlock) {
                const uint32_t lv = static_cast<uint32_t>(p.bucketed[b.x + j] - vbase);
                if (lv >= (1u << kChunkShift)) continue;
                const uint32_t w = lv >> 6;
                const uint32_t rank = wpre[w] + static_cast<uint32_t>(__popcll(bm[w] & ((1ull << (lv & 63)) - 1ull)));
                if (rank >= r0 && rank < r0 + m) atomicAdd(&cnt[rank - r0], 1u);
            }
            __syncthreads();
            uint32_t rank = first_rank;
            for (int j = 0; j < kWordsPerThread && rank < r0 + m; ++j) {
                unsigned long long x = bm[w0 + j];
                while (x) {
                    const int bit = __ffsll(x) - 1;
                    x &= x - 1;
                    if (rank >= r0 && rank < r0 + m)
                        p.pairs[b.x + rank] =
                            make_int2(vbase + (w0 + j) * 64 + bit, static_cast<int32_t>(cnt[rank - r0]));
                    ++rank;
                }
            }
            __syncthreads();
        }
    }
    // The runs into host memory, consecutive threads on consecutive runs (the block's own
    // device writes, visible to it after the barrier).
    __syncthreads();
    for (uint32_t j = static_cast<uint32_t>(tid); j < runs; j += kChunkBlock) host_pairs[b.x + j] = p.pairs[b.x + j];
    if (tid == 0) host_buckets[static_cast<int64_t>(s) * p.C + c] = make_uint2(b.x, runs);
}

}  // namespace

int32_t ppr_chunks(int32_t nv) { return static_cast<int32_t>((static_cast<int64_t>(nv) + (1 << kChunkShift) - 1) >> kChunkShift); }

int32_t ppr_tiles(int32_t walkers) { return (walkers + kTileWalkers - 1) / kTileWalkers; }

hipError_t launch_serve_ppr_reduce(const ServeSegment *segs, int32_t n_segs, int32_t n_tiles, const int32_t *ends,
                                   const int32_t *lengths, const int32_t *status, const PprScratch &p,
                                   PprSegmentStats *host_stats, uint2 *host_buckets, int2 *host_pairs,
                                   hipStream_t stream) {
    if (n_tiles <= 0 || n_segs <= 0) return hipSuccess;
    const size_t per_chunk = sizeof(uint32_t) * static_cast<size_t>(p.C);
    const dim3 buckets(static_cast<unsigned>(p.C), static_cast<unsigned>(n_segs));
    hipLaunchKernelGGL(ppr_hist_kernel, dim3(n_tiles), dim3(kTileBlock), per_chunk, stream, segs, n_segs, ends,
                       lengths, status, p, host_stats);
    hipLaunchKernelGGL(ppr_scatter_kernel, dim3(n_tiles), dim3(kTileBlock), per_chunk, stream, segs, n_segs, ends,
                       status, p);
    hipLaunchKernelGGL(ppr_chunk_kernel, buckets, dim3(kChunkBlock), 0, stream, p, host_pairs, host_buckets);
    return hipGetLastError();
}

}  // namespace trw

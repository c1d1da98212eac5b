// Walk kernels for gfx950: the Gather-Move-Update step engine.
//
// The reference hides DRAM latency on a CPU with a software-pipelined ring of
// 64 walkers per thread (random_walk/uniform_sampling.h:11-45 and siblings).
// Here latency is hidden by occupancy instead: one walker per lane, tens of
// thousands of lanes in flight, every step a chain of dependent gathers.
//
// Three kernel forms share one per-step routine, advance():
//   fused  - thread per walker, the whole fixed-length walk in one launch; path
//            positions are staged in LDS and written as 64-byte row chunks.
//   queue  - persistent: a lane whose walker finished takes the next walker id
//            from a device counter (one atomic per wavefront per refill), the
//            GPU analogue of the reference's ring refill (creeper.h:321-332).
//   step   - the plain form: one launch per step over walker state kept in HBM.
//
// Exactness: the FP64 expressions are the reference's, evaluated in the same
// order with contraction off (-ffp-contract=off); u % d is 32-bit unsigned, which
// equals the reference's uint32 % int64 for 1 <= d < 2^31.
#include "engine.h"
#include "thunderrw_device.h"

#include <algorithm>

namespace trw {
namespace {

// Fused-form launch shapes: threads per block and the LDS bytes per block that stage path
// positions (see fused_tile).  Each form stages whole rows when they fit: the rows of a
// wavefront's 64 consecutive walkers are then one contiguous run of memory, written out with
// coalesced stores at the end of the walk; longer walks stage chunks of a multiple of 32
// positions (whole 128-byte lines per row).  Measured on the DeepWalk-alias headline (2^22
// vertices, length 80, one walker per lane): 32-position chunks 4.02e10 steps/s, 64-position
// chunks 4.60e10, whole rows 4.79e10, no path output at all 4.93e10 -- the one-gather Move
// tolerates the occupancy the tile costs, while the HBM write stream wants long runs.  Moves of
// several dependent reads need wavefronts in flight instead: Node2Vec loses 22% with
// 32-position chunks at 256 threads and 60% with whole rows, so it keeps 16-position chunks.
#ifndef TRW_MULTI_BLOCK1
#define TRW_MULTI_BLOCK1 128
#endif
#ifndef TRW_MULTI_BLOCK2
#define TRW_MULTI_BLOCK2 64
#endif
#ifndef TRW_MULTI_BLOCK4
#define TRW_MULTI_BLOCK4 64
#endif
constexpr int kMaxFusedBlock = 256;  // launch bound of both fused kernels
constexpr int kTileBudget = 65536;   // walk_fused_multi_kernel (uniform / alias walk records)
constexpr int kQueueBlock = 256;
constexpr int kStepBlock = 256;

using trw_dev::philox4x32_10;
using trw_dev::real1;
using trw_dev::real2;

// Free-running stream: one Philox block per (global walker id, step, trial).
// A step's move draws are words .x/.y of its trial's block; the PPR stop draw is
// word .w of the step's trial-0 block.  Callers hand the trial-0 block in, so it
// can be computed a step ahead, while the previous step's gather is in flight.
__device__ __forceinline__ uint4 philox_block(const BatchIO &io, const WalkerState &s, uint32_t step, uint32_t trial) {
    return philox4x32_10(s.gid_lo, s.gid_hi, step, trial, io.key0, io.key1);
}

// Holds v as computed at this point of the program (an empty volatile asm takes it as an
// operand), so the compiler cannot sink its computation below a following wait.
__device__ __forceinline__ void pin(uint4 &v) { asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w)); }

__device__ __forceinline__ bool replay_take(const BatchIO &io, int32_t w, WalkerState &s, uint32_t &u) {
    if (static_cast<int64_t>(s.k) >= io.K) {
        s.st = kReplayExhausted;
        return false;
    }
    u = io.draws[static_cast<int64_t>(s.k) * io.n + w];
    s.k += 1;
    return true;
}

template <bool REPLAY>
__device__ __forceinline__ bool draw_one(const BatchIO &io, int32_t w, WalkerState &s, const uint4 &blk, uint32_t &u) {
    if (REPLAY) return replay_take(io, w, s, u);
    u = blk.x;
    return true;
}

template <bool REPLAY>
__device__ __forceinline__ bool draw_pair(const BatchIO &io, int32_t w, WalkerState &s, uint32_t step, uint32_t trial,
                                          const uint4 &blk, uint32_t &a, uint32_t &b) {
    if (REPLAY) {
        if (!replay_take(io, w, s, a)) return false;
        return replay_take(io, w, s, b);
    }
    if (trial == 0) {
        a = blk.x;
        b = blk.y;
    } else {
        const uint4 x = philox_block(io, s, step, trial);
        a = x.x;
        b = x.y;
    }
    return true;
}

template <bool REPLAY>
__device__ __forceinline__ bool draw_stop(const BatchIO &io, int32_t w, WalkerState &s, const uint4 &blk, uint32_t &u) {
    if (REPLAY) {
        if (static_cast<int64_t>(s.kt) >= io.KT) {
            s.st = kReplayExhausted;
            return false;
        }
        u = io.tdraws[static_cast<int64_t>(s.kt) * io.n + w];
        s.kt += 1;
        return true;
    }
    u = blk.w;
    return true;
}

// Replay: the draws a walker's next move takes -- words s.k and s.k + 1 of its stream (0
// past the end) -- loaded a step ahead, while the current step's gather is in flight, so a
// replay move is one dependent read as a free-running one is.
template <int NDRAW>
__device__ __forceinline__ uint4 replay_prefetch(const BatchIO &io, int32_t w, const WalkerState &s) {
    uint4 b = make_uint4(0u, 0u, 0u, 0u);
    const int64_t k = s.k;
    if (k < io.K) b.x = io.draws[k * io.n + w];
    if (NDRAW > 1 && k + 1 < io.K) b.y = io.draws[(k + 1) * io.n + w];
    return b;
}

// Takes the prefetched draws (replay_take's accounting: the cursor advances per word taken).
template <int NDRAW>
__device__ __forceinline__ bool replay_take_prefetched(const BatchIO &io, WalkerState &s, const uint4 &b, uint32_t &u0,
                                                       uint32_t &u1) {
    const int64_t k = s.k;
    if (k >= io.K || (NDRAW > 1 && k + 1 >= io.K)) {
        if (k < io.K) s.k += 1;
        s.st = kReplayExhausted;
        return false;
    }
    u0 = b.x;
    u1 = b.y;
    s.k += NDRAW;
    return true;
}

// Utility::upper_bound_branchless (util/utility.h:18-27): the number of
// entries <= target in a non-decreasing array.
__device__ __forceinline__ int32_t count_le(const double *src, int32_t n, double target) {
    return trw_dev::upper_bound_branchless(src, n, target);
}

// Graph::is_neighbor (util/graph/graph.h:180-199) with both endpoints' ranges
// known: binary search of the larger-degree endpoint in the smaller-degree
// endpoint's sorted adjacency.
__device__ __forceinline__ bool is_neighbor(const DevGraph &g, int32_t u, int64_t uoff, int32_t udeg, int32_t v,
                                            int64_t voff, int32_t vdeg) {
    if (udeg < vdeg) {
        const int32_t t = u;
        u = v;
        v = t;
        voff = uoff;
        vdeg = udeg;
    }
    (void)v;
    const int32_t *nb = g.adj + voff;
    int32_t b = 0, e = vdeg - 1;
    while (b <= e) {
        const int32_t m = b + ((e - b) >> 1);
        const int32_t x = nb[m];
        if (x == u) return true;
        if (x > u) e = m - 1;
        else b = m + 1;
    }
    return false;
}

// Graph::is_neighbor (util/graph/graph.h:180-199) with both endpoints' ranges known: the
// larger-degree endpoint searched in the smaller-degree endpoint's sorted adjacency -- through
// that endpoint's membership table when it has one (root carried by the walker or the
// neighbour record, else looked up), by binary search otherwise.  The answer is the same.
__device__ __forceinline__ bool is_neighbor_member(const DevGraph &g, int32_t u, int64_t uoff, int32_t udeg,
                                                 uint32_t uroot, int32_t v, int64_t voff, int32_t vdeg,
                                                 uint32_t vroot) {
    if (udeg < vdeg) {
        const int32_t t = u;
        u = v;
        v = t;
        voff = uoff;
        vdeg = udeg;
        vroot = uroot;
    }
    if (g.mlines && vdeg > kMemberMin) {
        const uint32_t root = vroot != kNoRange ? vroot : g.mroot[v];
        if (root != kNoRange) return trw_dev::member_contains(g.mlines, root, vdeg, u);
    }
    return is_neighbor(g, u, voff, vdeg, v, voff, vdeg);  // equal degrees: searches u in v's list
}

__device__ __forceinline__ int32_t first_ge(const int32_t *a, int32_t n, int32_t x) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ __forceinline__ int32_t first_gt(const int32_t *a, int32_t lo, int32_t n, int32_t x) {
    int32_t hi = n;
    while (lo < hi) {
        const int32_t m = (lo + hi) >> 1;
        if (a[m] <= x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// First edge and degree of v from the graph: one 8-byte vertex record when the
// graph has them (num_edges < 2^32), else the two CSR offsets.
__device__ __forceinline__ void vertex_range(const DevGraph &g, int32_t v, int64_t &off, int32_t &d) {
    if (g.vrec) {
        const uint2 r = g.vrec[v];
        off = static_cast<int64_t>(r.x);
        d = static_cast<int32_t>(r.y);
    } else {
        off = g.offset[v];
        d = static_cast<int32_t>(g.offset[v + 1] - off);
    }
}

// Range of a vertex whose record may already be held by the walker.
__device__ __forceinline__ void known_range(const DevGraph &g, int32_t v, uint32_t roff, uint32_t rdeg, int64_t &off,
                                            int32_t &d) {
    if (rdeg != kNoRange) {
        off = static_cast<int64_t>(roff);
        d = static_cast<int32_t>(rdeg);
    } else {
        vertex_range(g, v, off, d);
    }
}

// Adjacency entry e; with neighbour records the same 16-byte read also yields
// the neighbour's range, which the walker carries into its next step.
__device__ __forceinline__ int32_t read_neighbor(const DevGraph &g, int64_t e, uint32_t &noff, uint32_t &ndeg) {
    if (g.erec) {
        const uint4 r = g.erec[e];
        noff = r.y;
        ndeg = r.z;
        return static_cast<int32_t>(r.x);
    }
    noff = 0u;
    ndeg = kNoRange;
    return g.adj[e];
}

// As read_neighbor, also returning the neighbour's ITS structure (record word .w).
__device__ __forceinline__ int32_t read_neighbor_its(const DevGraph &g, int64_t e, uint32_t &noff, uint32_t &ndeg,
                                                     uint32_t &nidx) {
    if (g.erec) {
        const uint4 r = g.erec[e];
        noff = r.y;
        ndeg = r.z;
        nidx = g.its_lines ? r.w : kNoRange;
        return static_cast<int32_t>(r.x);
    }
    noff = 0u;
    ndeg = kNoRange;
    nidx = kNoRange;
    return g.adj[e];
}

// As read_neighbor, also returning the neighbour's membership table root (Node2Vec).
__device__ __forceinline__ int32_t read_neighbor_member(const DevGraph &g, int64_t e, uint32_t &noff, uint32_t &ndeg,
                                                        uint32_t &nroot) {
    if (g.nrec) {
        const uint4 r = g.nrec[e];
        noff = r.y;
        ndeg = r.z;
        nroot = r.w;
        return static_cast<int32_t>(r.x);
    }
    nroot = kNoRange;
    return read_neighbor(g, e, noff, ndeg);
}

// Weighted Node2Vec: the candidate's edge weight with its neighbour record and membership root,
// one 32-byte record when the graph has them.
__device__ __forceinline__ int32_t read_neighbor_member_weighted(const DevGraph &g, int64_t e, uint32_t &noff,
                                                                 uint32_t &ndeg, uint32_t &nroot, double &ew) {
    if (g.wnrec) {
        const uint4 a = g.wnrec[2 * e], b = g.wnrec[2 * e + 1];
        ew = __hiloint2double(static_cast<int>(a.y), static_cast<int>(a.x));
        noff = a.w;
        ndeg = b.x;
        nroot = b.y;
        return static_cast<int32_t>(a.z);
    }
    ew = g.weight[e];
    return read_neighbor_member(g, e, noff, ndeg, nroot);
}

// ITS guide range of the walker's current vertex: {first bucket, bucket count}, read once
// per walk (the records carry the next vertex's).  A count of 0 is degree 0, or weights that
// sum to zero when the first bucket is kNoRange.
__device__ __forceinline__ void guide_range(const DevGraph &g, int32_t v, uint32_t roff, uint32_t rdeg, int64_t &off,
                                            int32_t &m) {
    if (rdeg == kNoRange) {
        const uint2 r = g.gvrec[v];
        roff = r.x;
        rdeg = r.y;
    }
    off = static_cast<int64_t>(roff);
    m = static_cast<int32_t>(rdeg);
}

// A guide bucket whose draws reach more than two edges: the Move's own search of the prefix
// sums (inverse_transformation_sampling.h:259-286), then the chosen neighbour's bucket range.
__device__ __forceinline__ int32_t guide_search(const DevGraph &g, int32_t cur, uint32_t u, uint32_t &noff,
                                                uint32_t &ndeg) {
    int64_t off;
    int32_t d;
    vertex_range(g, cur, off, d);
    const double *P_ = g.prefix + off;
    const int32_t pos = min(count_le(P_, d, real2(u) * P_[d - 1]), d - 1);
    const int32_t next = g.adj[off + pos];
    const uint2 r = g.gvrec[next];
    noff = r.x;
    ndeg = r.y;
    return next;
}

// One 128-byte line of an ITS structure, as 8 16-byte loads issued together.  Named
// members, not an array: a run-time choice among them stays a chain of selects, where
// an indexed array would be moved to scratch memory.
struct ItsLine {
    uint4 a0, a1, a2, a3, a4, a5, a6, a7;
    __device__ __forceinline__ void load(const uint4 *p) {
        a0 = p[0];
        a1 = p[1];
        a2 = p[2];
        a3 = p[3];
        a4 = p[4];
        a5 = p[5];
        a6 = p[6];
        a7 = p[7];
    }
    static __device__ __forceinline__ double lo(const uint4 &v) {
        return __hiloint2double(static_cast<int>(v.y), static_cast<int>(v.x));
    }
    static __device__ __forceinline__ double hi(const uint4 &v) {
        return __hiloint2double(static_cast<int>(v.w), static_cast<int>(v.z));
    }
    // Tree node: keys 0..15 are the 16 doubles of the line.
    __device__ __forceinline__ int32_t count_le(double t) const {
        return (lo(a0) <= t) + (hi(a0) <= t) + (lo(a1) <= t) + (hi(a1) <= t) + (lo(a2) <= t) + (hi(a2) <= t) +
               (lo(a3) <= t) + (hi(a3) <= t) + (lo(a4) <= t) + (hi(a4) <= t) + (lo(a5) <= t) + (hi(a5) <= t) +
               (lo(a6) <= t) + (hi(a6) <= t) + (lo(a7) <= t) + (hi(a7) <= t);
    }
    __device__ __forceinline__ double key_at(int i) const {  // value selects only
        double k = lo(a0);
        k = i == 1 ? hi(a0) : k;
        k = i == 2 ? lo(a1) : k;
        k = i == 3 ? hi(a1) : k;
        k = i == 4 ? lo(a2) : k;
        k = i == 5 ? hi(a2) : k;
        k = i == 6 ? lo(a3) : k;
        k = i == 7 ? hi(a3) : k;
        k = i == 8 ? lo(a4) : k;
        k = i == 9 ? hi(a4) : k;
        k = i == 10 ? lo(a5) : k;
        k = i == 11 ? hi(a5) : k;
        k = i == 12 ? lo(a6) : k;
        k = i == 13 ? hi(a6) : k;
        k = i == 14 ? lo(a7) : k;
        k = i == 15 ? hi(a7) : k;
        return k;
    }
    // Leaf: prefix sums are keys 0..4 (a0, a1, a2.lo), records are a3..a7.
    __device__ __forceinline__ int32_t leaf_count_le(double t, int32_t n) const {
        return (0 < n && lo(a0) <= t) + (1 < n && hi(a0) <= t) + (2 < n && lo(a1) <= t) + (3 < n && hi(a1) <= t) +
               (4 < n && lo(a2) <= t);
    }
    static __device__ __forceinline__ uint32_t pick5(int32_t j, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3,
                                                     uint32_t v4) {
        uint32_t r = v0;
        r = j == 1 ? v1 : r;
        r = j == 2 ? v2 : r;
        r = j == 3 ? v3 : r;
        r = j == 4 ? v4 : r;
        return r;
    }
    __device__ __forceinline__ uint4 record(int32_t j) const {
        return make_uint4(pick5(j, a3.x, a4.x, a5.x, a6.x, a7.x), pick5(j, a3.y, a4.y, a5.y, a6.y, a7.y),
                          pick5(j, a3.z, a4.z, a5.z, a6.z, a7.z), pick5(j, a3.w, a4.w, a5.w, a6.w, a7.w));
    }
};

// Alias slot e (alias_sampling.h:147-172 pick).  Returns false for a vertex
// whose weights sum to zero.
__device__ __forceinline__ bool read_alias(const DevGraph &g, int64_t e, uint32_t u1, int32_t &next, uint32_t &noff,
                                           uint32_t &ndeg) {
    if (g.arec) {
        const uint4 a = g.arec[2 * e];
        const uint4 b = g.arec[2 * e + 1];
        if (static_cast<int32_t>(a.z) < 0) return false;
        const bool take_first = real2(u1) <= __hiloint2double(static_cast<int>(a.y), static_cast<int>(a.x));
        next = static_cast<int32_t>(take_first ? a.z : a.w);
        noff = take_first ? b.x : b.z;
        ndeg = take_first ? b.y : b.w;
        return true;
    }
    const uint4 raw = *reinterpret_cast<const uint4 *>(g.alias + e);
    if (static_cast<int32_t>(raw.z) < 0) return false;
    const double prob = __hiloint2double(static_cast<int>(raw.y), static_cast<int>(raw.x));
    next = (real2(u1) <= prob) ? static_cast<int32_t>(raw.z) : static_cast<int32_t>(raw.w);
    noff = 0u;
    ndeg = kNoRange;
    return true;
}

// Update (creeper.h:298-334) after a move from `cur` (range off/d) to `next`.
__device__ __forceinline__ void finish_move(const DevGraph &g, const KindParams &P, WalkerState &s, int32_t cur,
                                            int64_t off, int32_t d, int32_t next, uint32_t noff, uint32_t ndeg,
                                            uint32_t nidx = kNoRange) {
    s.prev = cur;
    s.poff = static_cast<uint32_t>(off);
    s.pdeg = g.vrec ? static_cast<uint32_t>(d) : kNoRange;
    s.pidx = s.cidx;
    s.cur = next;
    s.coff = noff;
    s.cdeg = ndeg;
    s.cidx = nidx;
    s.len += 1;
    (void)P;
}

// One Gather + Move + Update for walker w.  Returns true when the walker moved
// (and counted a step).  On return s.st is kAlive or the walker's final status.
//
// blk holds this step's trial-0 Philox block on entry; with AHEAD, after the step's gather has been
// issued the next step's block is computed into it (free-running only); without, the caller
// computes each step's block itself.
template <int MODE, bool REPLAY, bool AHEAD = true>
__device__ __forceinline__ bool advance_mode(const DevGraph &g, const KindParams &P, const BatchIO &io, int32_t w,
                                             WalkerState &s, uint4 &blk) {
    const int32_t cur = s.cur;
    int64_t off;
    int32_t d;
    // ITS with guide records: the walker's range is its vertex's buckets, {first bucket, count};
    // no buckets at a vertex of positive degree (first bucket kNoRange) means weights summing to
    // zero, which ends the walk before a draw as the total check below does.
    const bool guide = MODE == kModeIts && g.grec != nullptr && P.its_guide;
    // MetaPath over label records: a walker that carries nothing reads its range from the current
    // vertex's label record, the line its label range comes from below.
    const bool label_rec = MODE == kModeMetaPath && g.lerec != nullptr && s.cdeg == kNoRange;
    if (guide) {
        guide_range(g, cur, s.coff, s.cdeg, off, d);
    } else if (label_rec) {
        const uint2 r = *reinterpret_cast<const uint2 *>(g.lstart + kLabelRecWords * static_cast<int64_t>(cur) + 8);
        off = static_cast<int64_t>(r.x);
        d = static_cast<int32_t>(r.y);
    } else {
        known_range(g, cur, s.coff, s.cdeg, off, d);
    }
    if (d <= 0) {
        s.st = (guide && off == static_cast<int64_t>(kNoRange)) ? kGatherTerminated : kDeadEnd;
        return false;
    }
    const uint32_t step = static_cast<uint32_t>(s.len - 1);
    const uint32_t ud = static_cast<uint32_t>(d);
    int32_t next;
    uint32_t noff = 0u, ndeg = kNoRange, nidx = kNoRange;
    double nmax = -1.0;  // rejection: the next vertex's maximum weight, when a record gave it
    uint32_t nlo = 0u, nhi = kNoRange;  // MetaPath: the next step's label range at the next vertex
    uint32_t npos_out = kNoRange;       // MetaPath: the next step's schema position
    uint4 nblk = blk;

    if (MODE == kModeUniform) {
        // uniform_sampling.h:47-63
        uint32_t u;
        if (!draw_one<REPLAY>(io, w, s, blk, u)) return false;
        next = read_neighbor(g, off + (u % ud), noff, ndeg);
        if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
    } else if (MODE == kModeAlias) {
        // alias_sampling.h:147-172: column u0 % d, then real2(u1) <= prob ? first : second.
        uint32_t u0, u1;
        if (!draw_pair<REPLAY>(io, w, s, step, 0u, blk, u0, u1)) return false;
        const bool ok = read_alias(g, off + (u0 % ud), u1, next, noff, ndeg);
        if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
        if (!ok) {
            if (REPLAY) s.k -= 2;  // the vertex's weights sum to zero: terminate without consuming draws
            s.st = kGatherTerminated;
            return false;
        }
    } else if (MODE == kModeIts && guide) {
        // inverse_transformation_sampling.h:259-286 through the guide record of the draw's bucket.
        uint32_t u;
        if (!draw_one<REPLAY>(io, w, s, blk, u)) return false;
        const int64_t e = off + static_cast<int64_t>((static_cast<uint64_t>(u) * ud) >> 32);
        const uint4 ra = g.grec[2 * e], rb = g.grec[2 * e + 1];
        if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
        if (ra.w != 0u) {
            next = guide_search(g, cur, u, noff, ndeg);
        } else {
            const bool upper = u >= ra.x;
            next = static_cast<int32_t>(upper ? ra.z : ra.y);
            noff = upper ? rb.z : rb.x;
            ndeg = upper ? rb.w : rb.y;
        }
    } else if (MODE == kModeIts) {
        // inverse_transformation_sampling.h:259-286
        uint32_t root = kNoRange;
        if (g.its_lines && d >= P.its_tree_min) root = s.cidx != kNoRange ? s.cidx : g.its_root[cur];
        if (root != kNoRange) {
            // Search structure: the first line is the tree root (its last key is the total,
            // P[d - 1]) or, for one leaf, the leaf itself.  Level k of the tree has
            // ceil(nl / 16^k) nodes, computed here rather than stored.
            const int32_t nl = its_leaves(d);
            int32_t L = 0;
            while ((static_cast<int64_t>(nl) + (int64_t(1) << (4 * L)) - 1) >> (4 * L) > 1) ++L;
            const uint4 *V = g.its_lines + static_cast<int64_t>(root) * 8;
            ItsLine line;
            line.load(V);
            const int32_t root_keys =
                L == 0 ? d : static_cast<int32_t>((static_cast<int64_t>(nl) + (int64_t(1) << (4 * (L - 1))) - 1) >>
                                                  (4 * (L - 1)));
            const double total = line.key_at(root_keys - 1);
            if (total == 0.0) {
                s.st = kGatherTerminated;
                return false;
            }
            uint32_t u;
            if (!draw_one<REPLAY>(io, w, s, blk, u)) return false;
            if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
            const double target = real2(u) * total;
            int64_t leaf = 0;
            bool beyond = false;  // target >= total: the pick clamps to the last edge
            if (L > 0) {
                int64_t node = line.count_le(target);
                int64_t base = 1;  // lines before the current level (the root's single node)
                if (node >= root_keys) {
                    beyond = true;
                    node = nl - 1;
                    for (int32_t k = L - 1; k >= 1; --k)
                        base += (static_cast<int64_t>(nl) + (int64_t(1) << (4 * k)) - 1) >> (4 * k);
                } else {
                    for (int32_t k = L - 1; k >= 1; --k) {
                        line.load(V + (base + node) * 8);
                        node = node * kItsFan + line.count_le(target);
                        base += (static_cast<int64_t>(nl) + (int64_t(1) << (4 * k)) - 1) >> (4 * k);
                    }
                }
                leaf = node;
                line.load(V + (base + leaf) * 8);  // base is now the tree's line count
            }
            const int32_t n = min(kItsLeaf, d - kItsLeaf * static_cast<int32_t>(leaf));
            int32_t j = n - 1;
            if (!beyond) j = min(line.leaf_count_le(target, n), n - 1);
            const uint4 r = line.record(j);
            next = static_cast<int32_t>(r.x);
            noff = r.y;
            ndeg = r.z;
            nidx = r.w;
        } else {
            const double *P_ = g.prefix + off;
            const double total = P_[d - 1];
            if (total == 0.0) {
                s.st = kGatherTerminated;
                return false;
            }
            uint32_t u;
            if (!draw_one<REPLAY>(io, w, s, blk, u)) return false;
            if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
            // The count is < d whenever the prefix sums are non-decreasing (non-negative
            // weights); the clamp only keeps negative-weight inputs inside the segment.
            const int32_t pos = min(count_le(P_, d, real2(u) * total), d - 1);
            next = read_neighbor_its(g, off + pos, noff, ndeg, nidx);
        }
    } else if (MODE == kModeRejection) {
        // rejection_sampling.h:334-366: redraw (u % d, real2 * max) while r > w[pos].  With
        // rejection records a trial is one read (the edge's weight), the accepted one also
        // yields the neighbour's range and maximum, and the walker carries that maximum.
        const double m = s.cmax >= 0.0 ? s.cmax : g.rs_max[cur];
        if (m == 0.0) {
            s.st = kGatherTerminated;
            return false;
        }
        uint32_t pos = 0;
        uint4 ra = make_uint4(0u, 0u, 0u, 0u), rb = ra;
        for (uint32_t trial = 0;; ++trial) {
            if (P.max_trials > 0 && trial >= static_cast<uint32_t>(P.max_trials)) {
                s.st = kTrialCap;
                return false;
            }
            uint32_t ua, ub;
            if (!draw_pair<REPLAY>(io, w, s, step, trial, blk, ua, ub)) return false;
            pos = ua % ud;
            const double r = real2(ub) * m;
            double wt;
            if (g.rrec) {
                ra = g.rrec[2 * (off + pos)];
                rb = g.rrec[2 * (off + pos) + 1];
                wt = __hiloint2double(static_cast<int>(ra.y), static_cast<int>(ra.x));
            } else {
                wt = P.unit_weights ? 1.0 : g.weight[off + pos];
            }
            if (!(r > wt)) break;
        }
        if (g.rrec) {
            next = static_cast<int32_t>(ra.z);
            noff = ra.w;
            ndeg = rb.x;
            nmax = __hiloint2double(static_cast<int>(rb.z), static_cast<int>(rb.y));
        } else {
            next = read_neighbor(g, off + pos, noff, ndeg);
        }
        if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
    } else if (MODE == kModeNode2Vec) {
        // rejection_sampling.h:368-402 with Node2Vec::weight (node2vec.cpp:24-38): a trial is
        // accepted unless real2(ub) * bound > weight(candidate), the weight being max_ on the first
        // move, p_ back to the previous vertex, 1 to one of its neighbours and q_ elsewhere (times
        // the edge weight when weighted).  The membership search runs only when its answer decides.
        int32_t c = -1;
        uint32_t croot = kNoRange;
        for (uint32_t trial = 0;; ++trial) {
            if (P.max_trials > 0 && trial >= static_cast<uint32_t>(P.max_trials)) {
                s.st = kTrialCap;
                return false;
            }
            uint32_t ua, ub;
            if (!draw_pair<REPLAY>(io, w, s, step, trial, blk, ua, ub)) return false;
            double ew = 1.0;
            if (P.n2v_weighted) c = read_neighbor_member_weighted(g, off + (ua % ud), noff, ndeg, croot, ew);
            else c = read_neighbor_member(g, off + (ua % ud), noff, ndeg, croot);
            const double y = real2(ub) * P.n2v_bound;
            bool accept;
            if (s.len == 1) accept = !(y > ew * P.n2v_max);
            else if (c == s.prev) accept = !(y > ew * P.n2v_p);
            else {
                const bool acc_nb = !(y > ew), acc_far = !(y > ew * P.n2v_q);
                accept = acc_nb;
                if (acc_nb != acc_far) {
                    int64_t poff, coff;
                    int32_t pd, cd;
                    known_range(g, s.prev, s.poff, s.pdeg, poff, pd);
                    known_range(g, c, noff, ndeg, coff, cd);
                    accept = is_neighbor_member(g, s.prev, poff, pd, s.pidx, c, coff, cd, croot) ? acc_nb : acc_far;
                }
            }
            if (accept) break;
        }
        next = c;
        nidx = croot;
        if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
    } else {  // kModeMetaPath
        // dynamic_gather (creeper.h:178-296) with MetaPath::weight (metapath.cpp:22-30)
        // then dynamic_its_move: the prefix sums are match counts, so the pick is
        // the (floor(real2(u) * K) + 1)-th neighbour whose label matches.
        // The schema position (length - 1) % |schema|, carried by the walker in cidx after its
        // first move (a modulo per step otherwise); with label records the schema's labels come as
        // nibbles (8: a label outside [0, 8), which no neighbour carries).
        const uint32_t pos = s.cidx != kNoRange ? s.cidx : static_cast<uint32_t>((s.len - 1) % P.schema_length);
        const uint32_t npos = pos + 1u == static_cast<uint32_t>(P.schema_length) ? 0u : pos + 1u;
        const int32_t target =
            g.lerec ? static_cast<int32_t>((P.schema_nibbles >> (4u * pos)) & 15u) : P.schema[pos];
        int32_t a, K;
        if (g.lerec) {
            // Matches of `target` are [lstart[target], lstart[target + 1]) in label order; after
            // the first move the record that brought the walker here carried them.
            uint32_t lo = 0u, hi = 0u;
            if (s.pdeg != kNoRange) {
                lo = s.poff;
                hi = s.pdeg;
            } else if (target < 8) {
                const uint32_t *L = g.lstart + kLabelRecWords * static_cast<int64_t>(cur);
                lo = L[target];
                hi = target < 7 ? L[target + 1] : ud;
            }
            a = static_cast<int32_t>(lo);
            K = static_cast<int32_t>(hi - lo);
        } else {
            const int32_t *ls = g.lbl_sorted + off;
            a = first_ge(ls, d, target);
            K = first_gt(ls, a, d, target) - a;
        }
        if (K == 0) {
            s.st = kGatherTerminated;
            return false;
        }
        uint32_t u;
        if (!draw_one<REPLAY>(io, w, s, blk, u)) return false;
        const int32_t t = static_cast<int32_t>(real2(u) * static_cast<double>(K));
        if (P.schema_rec) {
            // The schema's 16-byte record: the neighbour, its first edge and the next step's label
            // range there (kNoRange: the neighbour has no edges, a dead end next step).
            const uint4 r = g.msrec[off + a + t];
            next = static_cast<int32_t>(r.x);
            noff = r.y;
            ndeg = r.z == kNoRange ? 0u : 1u;  // only its sign is read: the range comes with the record
            nlo = r.z;
            nhi = r.w;
        } else if (g.lerec) {
            const uint4 *R = g.lerec + 4 * (off + a + t);
            const uint4 r = R[0];
            next = static_cast<int32_t>(r.x);
            noff = r.y;
            ndeg = r.z;
            // The next step's label range at `next`, from its label starts in the same record.
            const int32_t nt = static_cast<int32_t>((P.schema_nibbles >> (4u * npos)) & 15u);
            // Both label starts are loaded with the record (clamped indices, then value selects):
            // a load that waits on the record's degree would be a second round trip.
            const uint32_t *NL = reinterpret_cast<const uint32_t *>(R + 1);
            const uint32_t l0 = NL[nt < 8 ? nt : 0], l1 = NL[nt < 7 ? nt + 1 : 7];
            nlo = nt < 8 ? l0 : 0u;
            nhi = nt < 7 ? l1 : (nt == 7 ? ndeg : 0u);
        } else {
            next = read_neighbor(g, off + g.lbl_pos[off + a + t], noff, ndeg);
        }
        npos_out = npos;
        if (!REPLAY && AHEAD) nblk = philox_block(io, s, step + 1, 0u);
    }

    finish_move(g, P, s, cur, off, d, next, noff, ndeg, nidx);
    if (MODE == kModeRejection) s.cmax = nmax;
    if (MODE == kModeMetaPath) {  // the previous vertex's range and table root are Node2Vec's alone
        s.poff = nlo;
        s.pdeg = nhi;
        s.cidx = npos_out;
    }
    const uint4 this_blk = blk;
    blk = nblk;
    if (P.ppr) {
        uint32_t ut;
        if (!draw_stop<REPLAY>(io, w, s, this_blk, ut)) return true;
        if (real1(ut) <= P.stop_prob) s.st = kDone;  // ppr.cpp:27-29
    } else if (s.len == P.length) {
        s.st = kDone;                                 // deepwalk.cpp:23-25 and siblings
    }
    return true;
}

// Mixed batches of two walker kinds: MODE == -1 compiles every Move and reads the kind's at run
// time; pair_mode(a, b) (<= -2) compiles only Moves a and b, so a kernel for a given pair holds
// the registers of those two alone (every Move: 117 VGPRs, three wavefronts per SIMD in the queue
// form; uniform + MetaPath alone: far fewer).
constexpr int pair_mode(int a, int b) { return -2 - (a < b ? a * 8 + b : b * 8 + a); }
constexpr int pair_first(int mode) { return (-2 - mode) / 8; }
constexpr int pair_second(int mode) { return (-2 - mode) % 8; }
// Whether a kernel compiled for MODE (a Move, a pair of Moves, or -1: every Move) runs Move m.
constexpr bool runs_move(int MODE, int m) {
    return MODE == -1 || MODE == m || (MODE <= -2 && (pair_first(MODE) == m || pair_second(MODE) == m));
}

template <int MODE, bool REPLAY, bool AHEAD = true>
__device__ __forceinline__ bool advance(const DevGraph &g, const KindParams &P, const BatchIO &io, int32_t w,
                                        WalkerState &s, uint4 &blk) {
    if (MODE >= 0) return advance_mode<MODE, REPLAY, AHEAD>(g, P, io, w, s, blk);
    if (MODE <= -2) {
        constexpr int M1 = MODE <= -2 ? pair_first(MODE) : 0, M2 = MODE <= -2 ? pair_second(MODE) : 0;
        return P.mode == M1 ? advance_mode<M1, REPLAY, AHEAD>(g, P, io, w, s, blk)
                            : advance_mode<M2, REPLAY, AHEAD>(g, P, io, w, s, blk);
    }
    switch (P.mode) {
    case kModeUniform: return advance_mode<kModeUniform, REPLAY, AHEAD>(g, P, io, w, s, blk);
    case kModeAlias: return advance_mode<kModeAlias, REPLAY, AHEAD>(g, P, io, w, s, blk);
    case kModeIts: return advance_mode<kModeIts, REPLAY, AHEAD>(g, P, io, w, s, blk);
    case kModeRejection: return advance_mode<kModeRejection, REPLAY, AHEAD>(g, P, io, w, s, blk);
    case kModeNode2Vec: return advance_mode<kModeNode2Vec, REPLAY, AHEAD>(g, P, io, w, s, blk);
    default: return advance_mode<kModeMetaPath, REPLAY, AHEAD>(g, P, io, w, s, blk);
    }
}

// Trial-0 Philox block of the walker's next step (free-running; unused in replay).
template <bool REPLAY>
__device__ __forceinline__ uint4 first_block(const BatchIO &io, int32_t w, const WalkerState &s) {
    if (REPLAY) return make_uint4(0u, 0u, 0u, 0u);
    return philox_block(io, s, static_cast<uint32_t>(s.len - 1), 0u);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds per-lane counter values into io.counters with one atomic per wavefront.
// Every lane of the wavefront must call it.
__device__ __forceinline__ void flush_counters(const BatchIO &io, unsigned long long steps, unsigned long long gterm,
                                               unsigned long long dead, unsigned long long err) {
    steps = wave_sum(steps);
    gterm = wave_sum(gterm);
    dead = wave_sum(dead);
    err = wave_sum(err);
    if (io.counters != nullptr && (threadIdx.x & 63) == 0) {
        if (steps) atomicAdd(io.counters + 0, steps);
        if (gterm) atomicAdd(io.counters + 1, gterm);
        if (dead) atomicAdd(io.counters + 2, dead);
        if (err) atomicAdd(io.counters + 3, err);
    }
}

__device__ __forceinline__ void classify(int32_t st, unsigned long long &gterm, unsigned long long &dead,
                                         unsigned long long &err) {
    gterm += (st == kGatherTerminated);
    dead += (st == kDeadEnd);
    err += (st == kReplayExhausted || st == kTrialCap || st == kBadSource);
}

__device__ __forceinline__ void write_outputs(const BatchIO &io, int32_t w, int32_t len, int32_t cur, int32_t st,
                                              uint32_t k) {
    if (io.lengths) io.lengths[w] = len;
    if (io.ends) io.ends[w] = cur;
    if (io.status) io.status[w] = st;
    if (io.draws_used) io.draws_used[w] = static_cast<int32_t>(k);
}

__device__ __forceinline__ void write_outputs(const BatchIO &io, int32_t w, const WalkerState &s) {
    write_outputs(io, w, s.len, s.cur, s.st, s.k);
}

// Path row of walker w in a batch whose walkers [0, split) are one kind and [split, n) another:
// the second kind's rows (path_stride_b ints each) follow the first kind's split rows
// (path_stride ints each); with equal strides this is row w of one matrix.
__device__ __forceinline__ int32_t *path_row(const BatchIO &io, int32_t split, int32_t w) {
    const int64_t r = w < split ? static_cast<int64_t>(w) * io.path_stride
                                : static_cast<int64_t>(split) * io.path_stride +
                                      static_cast<int64_t>(w - split) * io.path_stride_b;
    return io.paths + r;
}

__device__ __forceinline__ WalkerState empty_walker() {
    return WalkerState{0, -1, 1, kDone, 0u, 0u, 0u, kNoRange, 0u, kNoRange, 0u, 0u, kNoRange, kNoRange, -1.0};
}

__device__ __forceinline__ void init_walker(const DevGraph &g, const BatchIO &io, int32_t w, WalkerState &s) {
    s = empty_walker();
    const uint64_t gid = io.walker_ids ? io.walker_ids[w] : io.gid_base + static_cast<uint64_t>(w);
    s.gid_lo = static_cast<uint32_t>(gid);
    s.gid_hi = static_cast<uint32_t>(gid >> 32);
    s.cur = io.sources[w];
    s.st = (s.cur >= 0 && s.cur < g.nv) ? kAlive : kBadSource;
}

// ---------------------------------------------------------------- fused form
//
// Path output.  Each wavefront stages T consecutive path positions of its own
// walkers in LDS and then writes them as contiguous row chunks (T * 4 bytes per
// walker row, 64 bytes for T = 16).  Only lanes of one wavefront exchange data
// through its region, so a wavefront-scope fence suffices: no workgroup barrier
// makes fast wavefronts wait for slow ones.
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));

template <int WPT>
struct WaveTile {
    int32_t *region;  // WPT * 64 rows of T + 1 ints
    int32_t lane;
    int32_t T;        // positions per row

    __device__ __forceinline__ void put(int j, int32_t p, int32_t v) { region[(j * 64 + lane) * (T + 1) + p] = v; }

    // Writes positions [t0, t0 + cnt) of every walker of this wavefront; walker
    // (j, lane) is first_walker + j * stride_j + lane.
    __device__ __forceinline__ void flush(const BatchIO &io, int32_t first_walker, int32_t stride_j, int32_t t0,
                                          int32_t cnt) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Streaming stores: path rows are written once and not read back by the walk, so
        // they should not displace the records the walk gathers from the L2.
        const int32_t Q = T >> 2;  // 16-byte pieces of a full tile row
        if ((T & 3) == 0 && Q >= 1 && cnt == T && (io.path_stride & 3) == 0 && (t0 & 3) == 0 &&
            (reinterpret_cast<uintptr_t>(io.paths) & 15) == 0) {
            // Piece i is piece q of row lw: consecutive lanes store consecutive pieces of a row,
            // and consecutive rows too when the rows are whole (path_stride == T).
            for (int32_t i = lane; i < WPT * 64 * Q; i += 64) {
                const int32_t lw = i / Q;
                const int32_t q = i - lw * Q;
                const int32_t gw = first_walker + (lw >> 6) * stride_j + (lw & 63);
                const int32_t *r = region + lw * (T + 1) + 4 * q;
                v4i32 v;
                v.x = r[0];
                v.y = r[1];
                v.z = r[2];
                v.w = r[3];
                if (gw < io.n)
                    __builtin_nontemporal_store(
                        v, reinterpret_cast<v4i32 *>(io.paths + static_cast<int64_t>(gw) * io.path_stride + t0) + q);
            }
        } else {
            for (int32_t i = lane; i < WPT * 64 * cnt; i += 64) {
                const int32_t lw = i / cnt;
                const int32_t p = i - lw * cnt;
                const int32_t gw = first_walker + (lw >> 6) * stride_j + (lw & 63);
                if (gw < io.n)
                    __builtin_nontemporal_store(region[lw * (T + 1) + p],
                                                io.paths + static_cast<int64_t>(gw) * io.path_stride + t0 + p);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// Thread per walker, the whole fixed-length walk in one launch (any Move).
// T path positions are staged per row in `tile` (dynamic LDS: blockDim x (T + 1) ints).
template <int MODE, bool REPLAY>
__device__ __forceinline__ void walk_fused_body(const DevGraph &g, const KindParams &P, const BatchIO &io, int32_t T,
                                                int32_t *tile) {
    const int32_t tid = threadIdx.x;
    const int32_t w = blockIdx.x * blockDim.x + tid;
    const int32_t wave_first = w - (tid & 63);
    WaveTile<1> out{tile + (tid >> 6) * 64 * (T + 1), tid & 63, T};
    const bool valid = w < io.n;
    WalkerState s = empty_walker();
    if (valid) init_walker(g, io, w, s);
    uint4 blk = first_block<REPLAY>(io, w, s);
    const int32_t L = P.length;
    const bool paths = io.paths != nullptr;
    if (paths) out.put(0, 0, s.cur);
    unsigned long long steps = 0;
    int32_t t0 = 0;
    for (int32_t pos = 1; pos < L; ++pos) {
        int32_t v = -1;
        if (s.st == kAlive && advance<MODE, REPLAY>(g, P, io, w, s, blk)) {
            v = s.cur;
            ++steps;
        }
        if (paths) {
            out.put(0, pos - t0, v);
            if (pos - t0 == T - 1 || pos == L - 1) {
                out.flush(io, wave_first, 0, t0, pos - t0 + 1);
                t0 = pos + 1;
            }
        }
    }
    unsigned long long gterm = 0, dead = 0, err = 0;
    if (valid) {
        if (s.st == kAlive) s.st = kDone;
        write_outputs(io, w, s);
        classify(s.st, gterm, dead, err);
    }
    flush_counters(io, steps, gterm, dead, err);
}

template <int MODE, bool REPLAY>
__global__ __launch_bounds__(kMaxFusedBlock) void walk_fused_kernel(DevGraph g, KindParams P, BatchIO io, int32_t T) {
    extern __shared__ int32_t tile[];
    walk_fused_body<MODE, REPLAY>(g, P, io, T, tile);
}

// The Node2Vec Move (membership searches) takes 99 registers, so four wavefronts per SIMD;
// asking for five spills 12 bytes per lane and runs 14% faster.  The same request costs the
// ITS Move a quarter of its rate, hence a kernel of its own.
template <bool REPLAY>
__global__ __launch_bounds__(kMaxFusedBlock) __attribute__((amdgpu_waves_per_eu(5))) void walk_fused_node2vec_kernel(
    DevGraph g, KindParams P, BatchIO io, int32_t T) {
    extern __shared__ int32_t tile[];
    walk_fused_body<kModeNode2Vec, REPLAY>(g, P, io, T, tile);
}

// Fused form, uniform, alias and static ITS Moves, WPT walkers interleaved per lane.
//
// With neighbour / alias walk records / ITS guide records a step is one dependent
// gather (the chosen entry, which also carries the next vertex's range).  Issuing the WPT walkers'
// gathers back to back keeps WPT independent requests in flight per lane, on top
// of the eight wavefronts per SIMD that occupancy provides.  Walker (t, j) of a
// block is block_base + j * blockDim + t, so every per-j access is coalesced.
template <int MODE, bool REPLAY, int WPT>
__global__ __launch_bounds__(kMaxFusedBlock) void walk_fused_multi_kernel(DevGraph g, KindParams P, BatchIO io,
                                                                          int32_t T) {
    extern __shared__ int32_t tile[];  // WPT x blockDim rows of T + 1 ints
    const int32_t tid = threadIdx.x;
    const int32_t block = static_cast<int32_t>(blockDim.x);
    const int32_t wbase = blockIdx.x * (block * WPT);
    WaveTile<WPT> out{tile + (tid >> 6) * (WPT * 64 * (T + 1)), tid & 63, T};
    const int32_t wave_first = wbase + (tid & ~63);
    const bool paths = io.paths != nullptr;
    const int32_t L = P.length;
    WalkerState s[WPT];
    uint4 blk[WPT];
    constexpr int kDraws = MODE == kModeAlias ? 2 : 1;  // move draws per step
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
        const int32_t w = wbase + j * block + tid;
        s[j] = empty_walker();
        if (w < io.n) init_walker(g, io, w, s[j]);
        if (REPLAY) blk[j] = s[j].st == kAlive ? replay_prefetch<kDraws>(io, w, s[j]) : make_uint4(0u, 0u, 0u, 0u);
        else blk[j] = first_block<REPLAY>(io, w, s[j]);
        if (paths) out.put(j, 0, s[j].cur);
    }
    unsigned long long steps = 0;
    int32_t t0 = 0;
    for (int32_t pos = 1; pos < L; ++pos) {
        int64_t off[WPT];
        int32_t d[WPT];
        uint32_t u0[WPT], u1[WPT];
        bool go[WPT];
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            go[j] = s[j].st == kAlive;
            off[j] = 0;
            d[j] = 1;
            if (!go[j]) continue;
            if (MODE == kModeIts) guide_range(g, s[j].cur, s[j].coff, s[j].cdeg, off[j], d[j]);  // d: buckets
            else known_range(g, s[j].cur, s[j].coff, s[j].cdeg, off[j], d[j]);
        }
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            const int32_t w = wbase + j * block + tid;
            const uint32_t step = static_cast<uint32_t>(s[j].len - 1);
            u0[j] = u1[j] = 0;
            if (!go[j]) continue;
            if (d[j] <= 0) {
                // ITS: no buckets at a vertex of positive degree means its weights sum to zero,
                // which ends the walk before a draw, as the Move's total check does.
                s[j].st = (MODE == kModeIts && off[j] == static_cast<int64_t>(kNoRange)) ? kGatherTerminated : kDeadEnd;
                go[j] = false;
                continue;
            }
            if (REPLAY) go[j] = replay_take_prefetched<kDraws>(io, s[j], blk[j], u0[j], u1[j]);
            else if (MODE != kModeAlias) go[j] = draw_one<REPLAY>(io, w, s[j], blk[j], u0[j]);
            else go[j] = draw_pair<REPLAY>(io, w, s[j], step, 0u, blk[j], u0[j], u1[j]);
        }
        // Issue every walker's gather, then compute next steps' Philox blocks in
        // the gathers' shadow, then consume.
        uint4 ra[WPT], rb[WPT];
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            if (!go[j]) continue;
            if (MODE == kModeIts) {
                // The draw's bucket: floor(u * m / 2^32).
                const int64_t e =
                    off[j] + static_cast<int64_t>((static_cast<uint64_t>(u0[j]) * static_cast<uint32_t>(d[j])) >> 32);
                ra[j] = g.grec[2 * e];
                rb[j] = g.grec[2 * e + 1];
                continue;
            }
            const int64_t e = off[j] + (u0[j] % static_cast<uint32_t>(d[j]));
            if (MODE == kModeUniform) {
                ra[j] = g.erec[e];
            } else {
                ra[j] = g.arec[2 * e];
                rb[j] = g.arec[2 * e + 1];
            }
        }
        // The tile holds positions [t0, pos); when full, write it out now, behind
        // the gathers just issued, so the stores overlap them instead of delaying
        // the next step's wait.
        if (paths && pos - t0 == T) {
            out.flush(io, wave_first, block, t0, T);
            t0 = pos;
        }
        if (!REPLAY) {
#pragma unroll
            for (int j = 0; j < WPT; ++j)
                if (go[j]) blk[j] = philox_block(io, s[j], static_cast<uint32_t>(s[j].len), 0u);
        } else {
#pragma unroll
            for (int j = 0; j < WPT; ++j)
                if (go[j]) blk[j] = replay_prefetch<kDraws>(io, wbase + j * block + tid, s[j]);
        }
        int32_t next[WPT];
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            next[j] = -1;
            if (!go[j]) continue;
            uint32_t noff, ndeg;
            if (MODE == kModeUniform) {
                next[j] = static_cast<int32_t>(ra[j].x);
                noff = ra[j].y;
                ndeg = ra[j].z;
            } else if (MODE == kModeIts) {
                if (ra[j].w != 0u) {
                    next[j] = guide_search(g, s[j].cur, u0[j], noff, ndeg);
                } else {
                    const bool upper = u0[j] >= ra[j].x;
                    next[j] = static_cast<int32_t>(upper ? ra[j].z : ra[j].y);
                    noff = upper ? rb[j].z : rb[j].x;
                    ndeg = upper ? rb[j].w : rb[j].y;
                }
            } else {
                if (static_cast<int32_t>(ra[j].z) < 0) {  // the vertex's weights sum to zero
                    if (REPLAY) s[j].k -= 2;
                    s[j].st = kGatherTerminated;
                    continue;
                }
                const bool take_first =
                    real2(u1[j]) <= __hiloint2double(static_cast<int>(ra[j].y), static_cast<int>(ra[j].x));
                next[j] = static_cast<int32_t>(take_first ? ra[j].z : ra[j].w);
                noff = take_first ? rb[j].x : rb[j].z;
                ndeg = take_first ? rb[j].y : rb[j].w;
            }
            finish_move(g, P, s[j], s[j].cur, off[j], d[j], next[j], noff, ndeg);
            ++steps;
            if (s[j].len == L) s[j].st = kDone;
        }
        if (paths) {
#pragma unroll
            for (int j = 0; j < WPT; ++j) out.put(j, pos - t0, next[j]);
        }
    }
    if (paths) out.flush(io, wave_first, block, t0, L - t0);
    unsigned long long gterm = 0, dead = 0, err = 0;
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
        const int32_t w = wbase + j * block + tid;
        if (w < io.n) {
            if (s[j].st == kAlive) s[j].st = kDone;
            write_outputs(io, w, s[j]);
            classify(s[j].st, gterm, dead, err);
        }
    }
    flush_counters(io, steps, gterm, dead, err);
}

// Node2Vec, fused form, with membership tables (max-weight rejection, rejection_sampling.h:368-402,
// Node2Vec::weight node2vec.cpp:24-38): a walker per lane as a state machine that issues one
// read per iteration -- the candidate record of a trial, or one half of a membership bucket -- and
// then takes the decision that read allows.  In the step-per-iteration kernel a wavefront's step
// lasts as long as its slowest lane's chain of trials and tests (5.5 reads against a lane mean of
// 1.7 on the C3 graph) while the other lanes issue nothing; here no lane waits for another.  The
// draws, comparisons and moves are the Move's: only their schedule across lanes differs.  Paths:
// each lane stages 16 positions of its own row in LDS and writes them when the chunk is complete.
constexpr int kN2vBlock = 256;
constexpr int kN2vChunk = 16;

// A lane's own path row (the Node2Vec state machine, the queue form): positions [len - staged, len) wait in
// `mine` (LDS, kN2vChunk + 1 ints per lane) and go out when kN2vChunk have gathered or the walk
// ends; at the end, positions the walk did not reach are -1, as the other forms leave them.
struct LaneRow {
    int32_t *row;   // the walker's row, or null when paths are not written
    int32_t *mine;
    int32_t staged;
    __device__ __forceinline__ void put(int32_t v, int32_t len, int32_t L) {
        if (row == nullptr) return;
        mine[staged++] = v;
        if (staged == kN2vChunk || len == L) flush(len);
    }
    // Entries before the first 16-byte boundary one by one, then 16-byte stores, then the rest:
    // a chunk is a few store requests instead of one per entry.
    __device__ __forceinline__ void flush(int32_t len) {
        int32_t *dst = row + (len - staged);
        int32_t j = 0;
        while (j < staged && (reinterpret_cast<uintptr_t>(dst + j) & 15u) != 0u) {
            dst[j] = mine[j];
            ++j;
        }
        for (; j + 4 <= staged; j += 4)
            *reinterpret_cast<uint4 *>(dst + j) =
                make_uint4(static_cast<uint32_t>(mine[j]), static_cast<uint32_t>(mine[j + 1]),
                           static_cast<uint32_t>(mine[j + 2]), static_cast<uint32_t>(mine[j + 3]));
        for (; j < staged; ++j) dst[j] = mine[j];
        staged = 0;
    }
    // Padded rows: the staged chunk filled with -1 up to its 16-position boundary (or the row's
    // end), so that it leaves as one aligned run; the positions after it were set to -1 before the
    // walk.  Returns the length to flush to.  Nothing staged: the walk ended on a chunk boundary,
    // and the chunks from there on already hold -1.
    __device__ __forceinline__ int32_t pad_chunk(int32_t len, int32_t L) {
        if (staged == 0) return len;
        const int32_t end = min(len - staged + kN2vChunk, L);
        while (len < end) {
            mine[staged++] = -1;
            ++len;
        }
        return len;
    }
    __device__ __forceinline__ void finish(int32_t len, int32_t L) {
        if (row == nullptr) return;
        flush(len);
        for (int32_t j = len; j < L; ++j) row[j] = -1;
    }
};

// WEIGHTED: Node2Vec over edge weights, a trial reading the candidate's 32-byte weighted record
// (edge weight beside the neighbour record) instead of its 16-byte neighbour record.
// BLOCK: threads per block (a launch shape the graph's tuning picks, 128 or 256).
template <bool REPLAY, bool WEIGHTED, int BLOCK>
__global__ __launch_bounds__(BLOCK) void walk_node2vec_sm_kernel(DevGraph g, KindParams P, BatchIO io) {
    __shared__ int32_t chunk[BLOCK * (kN2vChunk + 1)];
    int32_t *mine = chunk + threadIdx.x * (kN2vChunk + 1);
    const int32_t w = static_cast<int32_t>(blockIdx.x) * BLOCK + static_cast<int32_t>(threadIdx.x);
    const int32_t L = P.length;
    WalkerState s = empty_walker();
    if (w < io.n) init_walker(g, io, w, s);
    int64_t off = 0;
    int32_t d = 0;
    if (s.st == kAlive) {
        known_range(g, s.cur, s.coff, s.cdeg, off, d);
        s.cidx = g.mroot[s.cur];
    }
    LaneRow out{(io.paths != nullptr && w < io.n) ? io.paths + static_cast<int64_t>(w) * io.path_stride : nullptr, mine, 0};
    out.put(s.cur, 1, L);
    uint32_t trial = 0;
    uint4 blk = first_block<REPLAY>(io, w, s);  // free-running: the current trial's block
    uint4 rej = blk, acc = blk;                 // the next trial's and the next step's
    bool testing = false;  // the next read is a membership bucket half, not a candidate
    int32_t c = -1;        // the trial's candidate and its range and table root
    uint32_t cnoff = 0u, cndeg = 0u, croot = kNoRange;
    bool acc_nb = false, acc_far = false;  // the trial's decision if c is / is not a neighbour of prev
    uint32_t mline = 0u, mb = 0u, mmask = 0u, mleft = 0u, mhalf = 0u;  // the bucket cursor of a test
    int32_t mkey = 0;
    unsigned long long steps = 0;

    for (;;) {
        if (__ballot(s.st == kAlive) == 0) break;
        bool decided = false, accept = false;
        if (s.st == kAlive && !testing) {
            if (d <= 0) {
                s.st = kDeadEnd;
            } else if (P.max_trials > 0 && trial >= static_cast<uint32_t>(P.max_trials)) {
                s.st = kTrialCap;
            } else {
                uint32_t ua = 0u, ub = 0u;
                bool ok = true;
                if (REPLAY) {
                    ok = replay_take(io, w, s, ua) && replay_take(io, w, s, ub);
                } else {
                    ua = blk.x;
                    ub = blk.y;
                }
                if (ok) {
                    const int64_t e = off + (ua % static_cast<uint32_t>(d));
                    uint4 r, rw = make_uint4(0u, 0u, 0u, 0u);
                    if (WEIGHTED) {
                        r = g.wnrec[2 * e];
                        rw = g.wnrec[2 * e + 1];
                    } else {
                        r = g.nrec[e];
                    }
                    if (!REPLAY) {  // both possible next blocks, in the candidate read's shadow
                        rej = philox_block(io, s, static_cast<uint32_t>(s.len - 1), trial + 1u);
                        acc = philox_block(io, s, static_cast<uint32_t>(s.len), 0u);
                        pin(rej);
                        pin(acc);
                    }
                    double ew = 1.0;
                    if (WEIGHTED) {
                        ew = __hiloint2double(static_cast<int>(r.y), static_cast<int>(r.x));
                        c = static_cast<int32_t>(r.z);
                        cnoff = r.w;
                        cndeg = rw.x;
                        croot = rw.y;
                    } else {
                        c = static_cast<int32_t>(r.x);
                        cnoff = r.y;
                        cndeg = r.z;
                        croot = r.w;
                    }
                    // Accepted unless y > weight (the reference's loop condition).
                    const double y = real2(ub) * P.n2v_bound;
                    acc_nb = !(y > ew);
                    acc_far = !(y > ew * P.n2v_q);
                    if (s.len == 1) accept = !(y > ew * P.n2v_max), decided = true;
                    else if (c == s.prev) accept = !(y > ew * P.n2v_p), decided = true;
                    else if (acc_nb == acc_far) accept = acc_nb, decided = true;
                    else {
                        // is_neighbor(prev, c): the larger-degree endpoint looked up in the smaller
                        // one's table (graph.h:180-199; equal degrees: prev in c's).
                        const bool in_prev = static_cast<int32_t>(s.pdeg) < static_cast<int32_t>(cndeg);
                        const int32_t md = in_prev ? static_cast<int32_t>(s.pdeg) : static_cast<int32_t>(cndeg);
                        mline = in_prev ? s.pidx : croot;
                        mkey = in_prev ? c : s.prev;
                        if (md <= 0 || mline == kNoRange) {
                            // An empty list (a directed graph's sink) holds nothing: not a neighbour.
                            accept = acc_far;
                            decided = true;
                        } else {
                            const uint32_t bits = member_bucket_bits(md);
                            mb = member_hash(mkey, bits);
                            mmask = (1u << bits) - 1u;
                            mleft = 1u << bits;
                            mhalf = 0u;
                            testing = true;
                        }
                    }
                }
            }
        } else if (s.st == kAlive) {
            trw_dev::MemberHalf h;
            h.load(g.mlines + static_cast<uint64_t>(mline + mb) * 8 + mhalf * 4);
            bool found = h.contains(static_cast<uint32_t>(mkey)), done = found || !h.full();
            if (!done) {
                if (mhalf == 0u) {
                    mhalf = 1u;
                } else {
                    mhalf = 0u;
                    mb = (mb + 1u) & mmask;
                    done = --mleft == 0u;
                }
            }
            if (done) {
                testing = false;
                accept = found ? acc_nb : acc_far;
                decided = true;
            }
        }
        if (decided) {
            if (accept) {
                s.prev = s.cur;
                s.poff = static_cast<uint32_t>(off);
                s.pdeg = static_cast<uint32_t>(d);
                s.pidx = s.cidx;
                s.cur = c;
                s.coff = cnoff;
                s.cdeg = cndeg;
                s.cidx = croot;
                s.len += 1;
                ++steps;
                off = static_cast<int64_t>(cnoff);
                d = static_cast<int32_t>(cndeg);
                trial = 0u;
                if (s.len == L) s.st = kDone;
                out.put(c, s.len, L);
                blk = acc;
            } else {
                ++trial;
                blk = rej;
            }
        }
    }
    if (w < io.n) {
        out.finish(s.len, io.unpadded ? s.len : L);
        write_outputs(io, w, s);
    }
    unsigned long long gterm = 0, dead = 0, err = 0;
    if (w < io.n) classify(s.st, gterm, dead, err);
    flush_counters(io, steps, gterm, dead, err);
}

// ---------------------------------------------------------------- queue form
//
// Persistent: each lane walks one walker at a time and, while it walks, prepares the next one --
// its index from the wavefront's chunk of the shared queue, then its source (and global id), then
// what its first step reads before the gather (the source's vertex or guide range, MetaPath's
// label range, the rejection maximum) -- one load per stage, issued beside the step's gather.  A
// walker that ends is replaced in the same iteration by the prepared one, whose first step is then
// one gather like any other.  (Loading a new walker on demand put a chain of three dependent reads
// -- source, vertex record, first gather -- into the iteration, and a wavefront waits for its
// slowest lane: 89% of iterations carried a refill and took 4.3 times as long as the others.)

// What a kind's first step reads before its gather, for a walker starting at src (walk.hip's
// advance_mode reads the same when the walker carries nothing): .x/.y the range it carries in
// coff/cdeg (vertex record, or ITS guide range), .z/.w MetaPath's label range for schema[0]
// (poff/pdeg) or the rejection maximum's bits (cmax); kNoRange where the step reads it itself.
// (The kind's mode, ITS guide switch and first schema label come as values: a reference to one of
// two kernel-argument structures chosen at run time puts both in scratch memory.)
// Only the Moves the kernel runs (MODE) are compiled in, so the graph arrays of the others hold
// no registers.
template <int MODE>
__device__ __forceinline__ uint4 first_step_inputs(const DevGraph &g, int32_t mode, int32_t its_guide,
                                                   int32_t schema0, int32_t src) {
    uint4 r = make_uint4(0u, kNoRange, 0u, kNoRange);
    if (runs_move(MODE, kModeIts) && mode == kModeIts && g.grec != nullptr && its_guide) {
        const uint2 v = g.gvrec[src];
        r.x = v.x;
        r.y = v.y;
        return r;
    }
    if (runs_move(MODE, kModeMetaPath) && mode == kModeMetaPath && g.lerec) {
        // The source's label record: its range and the label range of schema[0], one line.
        const uint32_t *L = g.lstart + kLabelRecWords * static_cast<int64_t>(src);
        const uint2 v = *reinterpret_cast<const uint2 *>(L + 8);
        r.x = v.x;
        r.y = v.y;
        // schema[0] as a nibble (8: outside [0, 8), which no neighbour carries: an empty range);
        // label 7 ends at the degree (.y, kNoRange here).  Loads at clamped indices, then selects.
        const int32_t t = schema0;
        const uint32_t l0 = L[t < 8 ? t : 0], l1 = L[t < 7 ? t + 1 : 7];
        r.z = t < 8 ? l0 : 0u;
        r.w = t < 7 ? l1 : (t == 7 ? kNoRange : 0u);
        return r;
    }
    if (g.vrec) {
        const uint2 v = g.vrec[src];
        r.x = v.x;
        r.y = v.y;
    }
    if (runs_move(MODE, kModeRejection) && mode == kModeRejection && g.rrec) {
        const double m = g.rs_max[src];
        const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(m));
        r.z = static_cast<uint32_t>(b);
        r.w = static_cast<uint32_t>(b >> 32);
    }
    return r;
}

// The path rows of queue positions [p0, p1), filled by the wavefront that has just taken them:
// -1 in [16, L) of a padded row of length L > 16 (its walker writes its source and moves over
// them, and its first 16 positions itself, padded with -1: pad_chunk), nothing in a row of length
// 2..16 (the walker writes it whole), the source alone in a row of length 1 (PPR's, which its
// walker then leaves alone).  A walker's last chunk thus leaves as one aligned run instead of a
// few entries into a line that was filled before (a partial-line write per walker).  Consecutive
// positions are consecutive rows (position p runs walker p + rot mod n), so consecutive lanes
// store consecutive 16-byte pieces: about 1 KB per store instruction, where a separate pass over
// every row before the walk cost 12% of the PPR + MetaPath batch.  A walker's own stores come at
// least two iterations later, after waits that the counter, covering stores in issue order, makes
// wait for these too.
__device__ __forceinline__ void fill_chunk_rows(const BatchIO &io, int32_t split, int32_t La, int32_t Lb,
                                                unsigned int p0, unsigned int p1, unsigned int rot, int lane) {
    const unsigned int n = static_cast<unsigned int>(io.n);
    const auto walker = [&](unsigned int pos) {
        return static_cast<int32_t>(pos < n - rot ? pos + rot : pos - (n - rot));
    };
    // Each kind's rows are 16-byte aligned when its stride and first row are.
    const bool base16 = (reinterpret_cast<uintptr_t>(io.paths) & 15) == 0;
    const bool vec_a = base16 && (io.path_stride & 3) == 0;
    const bool vec_b = base16 && (io.path_stride_b & 3) == 0 &&
                       ((static_cast<int64_t>(split) * io.path_stride) & 3) == 0;
    // Pieces per row: L / 4 16-byte pieces (L > 1, L % 4 == 0), L entries (other L > 1), or the
    // source entry (L == 1).
    const int32_t qa = La == 1 ? 1 : io.unpadded || La <= kN2vChunk ? 0 : ((La & 3) == 0 && vec_a ? La >> 2 : La);
    const int32_t qb = Lb == 1 ? 1 : io.unpadded || Lb <= kN2vChunk ? 0 : ((Lb & 3) == 0 && vec_b ? Lb >> 2 : Lb);
    // Kinds come in runs of positions; a chunk holds both only if it crosses the position of
    // walker `split` or of walker 0.
    const unsigned int split_pos = static_cast<unsigned int>(split) >= n ? n : (static_cast<unsigned int>(split) + n - rot) % n;
    const unsigned int zero_pos = (n - rot) % n;
    const bool both = (split_pos > p0 && split_pos < p1) || (zero_pos > p0 && zero_pos < p1);
    const int32_t qmax = both ? (qa > qb ? qa : qb) : (walker(p0) >= split ? qb : qa);
    if (qmax == 0) return;  // unpadded rows of length > 1: nothing to write before the walk
    const unsigned int rows = p1 - p0;
    const auto piece = [&](unsigned int r, int32_t v) {
        const int32_t w = walker(p0 + r);
        const bool kb = w >= split;
        const int32_t L = kb ? Lb : La;
        if (v >= (kb ? qb : qa)) return;
        int32_t *row = path_row(io, split, w);
        if (L == 1) row[0] = io.sources[w];
        else if (io.unpadded) return;  // the walker writes its source and moves, nothing else
        else if ((kb ? vec_b : vec_a) && (L & 3) == 0) {
            if (v >= kN2vChunk / 4) reinterpret_cast<int4 *>(row)[v] = make_int4(-1, -1, -1, -1);
        } else if (v >= kN2vChunk) {
            row[v] = -1;
        }
    };
    if (qmax <= 64) {
        // 64 / qmax rows per pass, lane i on piece i % qmax of row i / qmax.
        const int32_t per = 64 / qmax, r_in = lane / qmax, v = lane - r_in * qmax;
        if (r_in < per)
            for (unsigned int r = static_cast<unsigned int>(r_in); r < rows; r += static_cast<unsigned int>(per))
                piece(r, v);
    } else {
        for (unsigned int r = 0; r < rows; ++r)
            for (int32_t v = lane; v < qmax; v += 64) piece(r, v);
    }
}

// A lane's path row in the QUEUE form.  Entries gather in the lane's LDS row; a full 16-position
// chunk, or a finished walk's last one, is handed over rather than stored, and goes out in the
// lane's next iteration after that iteration's loads (first-step inputs, the refill's queue, source
// and id reads) and just before its gather (drain).  Vector memory operations complete in issue
// order and one counter tracks loads and stores alike, so every wait covers the stores issued
// before it: stored at the end of an iteration, a run held up the next iteration's first wait
// (for the prepared walker's inputs) by its write latency; drained before the gather, it completes
// in the gather's shadow.
struct QueueRow {
    int32_t *row;    // the walker's row, or null when paths are not written
    int32_t *mine;   // the lane's LDS staging row (kN2vChunk + 1 ints)
    int32_t staged;
    int32_t *pdst;   // the handed-over run: destination and entries (0: none)
    int32_t pcnt;
    __device__ __forceinline__ void hand_over(int32_t len) {
        pdst = row + (len - staged);
        pcnt = staged;
        staged = 0;
    }
    __device__ __forceinline__ void put(int32_t v, int32_t len, int32_t L) {
        if (row == nullptr) return;
        mine[staged++] = v;
        if (staged == kN2vChunk || len == L) hand_over(len);
    }
    // Padded rows: the staged chunk filled with -1 to its 16-position boundary (or the row's end),
    // as LaneRow::pad_chunk.
    __device__ __forceinline__ int32_t pad_chunk(int32_t len, int32_t L) {
        if (staged == 0) return len;
        const int32_t end = min(len - staged + kN2vChunk, L);
        while (len < end) {
            mine[staged++] = -1;
            ++len;
        }
        return len;
    }
    // The handed-over run: entries before the first 16-byte boundary one by one, then 16-byte
    // stores, then the rest.
    __device__ __forceinline__ void drain() {
        if (pcnt == 0) return;
        int32_t j = 0;
        while (j < pcnt && (reinterpret_cast<uintptr_t>(pdst + j) & 15u) != 0u) {
            pdst[j] = mine[j];
            ++j;
        }
        for (; j + 4 <= pcnt; j += 4)
            *reinterpret_cast<uint4 *>(pdst + j) =
                make_uint4(static_cast<uint32_t>(mine[j]), static_cast<uint32_t>(mine[j + 1]),
                           static_cast<uint32_t>(mine[j + 2]), static_cast<uint32_t>(mine[j + 3]));
        for (; j < pcnt; ++j) pdst[j] = mine[j];
        pcnt = 0;
    }
};

// A finished walker's outputs, stored with the lane's next drain (as its last path run).
struct PendingOutputs {
    int32_t w = -1, len = 0, cur = 0, st = 0;
    uint32_t k = 0u;
    __device__ __forceinline__ void hold(int32_t walker, const WalkerState &s) {
        w = walker;
        len = s.len;
        cur = s.cur;
        st = s.st;
        k = s.k;
    }
    __device__ __forceinline__ void drain(const BatchIO &io) {
        if (w < 0) return;
        write_outputs(io, w, len, cur, st, k);
        w = -1;
    }
};

// A and B are the batch's two kinds; walkers [0, split) run A unless `swapped`, when they run B.
// A pair kernel (MODE <= -2) is launched with A the kind of the pair's first Move and B that of its
// second, so each kind's Move is known at compile time: two Move paths instead of four, and only
// the parameters those read held in registers.
template <int MODE, bool REPLAY>
__global__ __launch_bounds__(kQueueBlock) void walk_queue_kernel(DevGraph g, KindParams A, KindParams B,
                                                                  int32_t split, int32_t swapped, BatchIO io,
                                                                  unsigned int *queue, unsigned int chunk,
                                                                  unsigned int rot) {
    const int lane = threadIdx.x & 63;
    WalkerState s = empty_walker();
    uint4 blk = make_uint4(0u, 0u, 0u, 0u);
    int32_t w = -1;
    bool kind_b = false;
    unsigned long long steps = 0, gterm = 0, dead = 0, err = 0;
    const unsigned int n = static_cast<unsigned int>(io.n);
    // The wavefront takes walker indices from the shared queue `chunk` at a time and hands them to
    // its lanes from [qnext, qend): one returning atomic per chunk rather than per walker keeps the
    // single queue head off the critical path.
    unsigned int qnext = 0, qend = 0;
    bool queue_empty = false;
    // The lane's next walker: stage 0 none, 1 index held (source and id loads issued), 2 source
    // held (first-step loads issued); its inputs arrive by the next iteration.
    int32_t stage = 0, nw = -1, nsrc = 0;
    uint32_t ngid_lo = 0u, ngid_hi = 0u;
    uint4 nin = make_uint4(0u, kNoRange, 0u, kNoRange);
    // The lane's walker's path entries gather here and go out in runs (its row was set to -1 past
    // its first chunk when the wavefront took its chunk), instead of one store request per entry.
    __shared__ int32_t pchunk[kQueueBlock * (kN2vChunk + 1)];
    QueueRow out{nullptr, pchunk + threadIdx.x * (kN2vChunk + 1), 0, nullptr, 0};
    PendingOutputs done;
    bool fresh = false;  // a walker taken this iteration: its source is staged after the drain

    for (;;) {
        // 1. A lane whose walker has ended takes the prepared one.
        if (s.st != kAlive && stage == 2) {
            w = nw;
            kind_b = (w >= split) != (swapped != 0);
            const int32_t mode = kind_b ? B.mode : A.mode, length = kind_b ? B.length : A.length;
            s = empty_walker();
            s.gid_lo = ngid_lo;
            s.gid_hi = ngid_hi;
            s.cur = nsrc;
            s.st = (nsrc >= 0 && nsrc < g.nv) ? kAlive : kBadSource;
            if (s.st == kAlive) {
                s.coff = nin.x;
                s.cdeg = nin.y;
                if (runs_move(MODE, kModeMetaPath) && mode == kModeMetaPath && g.lerec) {
                    s.poff = nin.z;
                    s.pdeg = nin.w != kNoRange ? nin.w : nin.y;
                } else if (runs_move(MODE, kModeRejection) && mode == kModeRejection && g.rrec) {
                    s.cmax = __hiloint2double(static_cast<int>(nin.w), static_cast<int>(nin.z));
                }
            }
            stage = 0;
            // Position 0 already holds the source in a length-1 row (PPR's), which is then complete;
            // a longer row restates it so that its first run starts on the row's alignment.
            out.row = io.paths && length > 1 ? path_row(io, split, w) : nullptr;
            out.staged = 0;
            fresh = true;
            if (s.st != kAlive) {  // bad source: finished before the first step
                out.drain();       // the previous walker's run and outputs first (the LDS row is reused)
                done.drain(io);
                out.put(s.cur, 1, length);
                if (out.row && out.staged > 0) out.hand_over(io.unpadded ? 1 : out.pad_chunk(1, length));
                fresh = false;
                done.hold(w, s);
                classify(s.st, gterm, dead, err);
            }
        }

        // 2. Preparation, one stage per iteration; each stage's loads are issued before this
        // iteration's gather and have arrived when it has.
        if (stage == 1) {
            const bool nb = (nw >= split) != (swapped != 0);
            if (nsrc >= 0 && nsrc < g.nv)
                nin = first_step_inputs<MODE>(g, nb ? B.mode : A.mode, nb ? B.its_guide : A.its_guide,
                                        static_cast<int32_t>((nb ? B.schema_nibbles : A.schema_nibbles) & 15u), nsrc);
            stage = 2;
        }
        const bool want = stage == 0 && !queue_empty;
        const unsigned long long want_mask = __ballot(want);
        if (want_mask) {
            const unsigned int cnt = static_cast<unsigned int>(__popcll(want_mask));
            const unsigned int avail = qend - qnext;
            unsigned int fresh = n;  // first index of a newly taken chunk (n = none)
            if (avail < cnt) {
                const int leader = __ffsll(static_cast<long long>(want_mask)) - 1;
                unsigned int base = 0;
                if (lane == leader) base = atomicAdd(queue, chunk);
                base = __shfl(base, leader, 64);
                if (base < n) fresh = base;
            }
            const unsigned int rank = static_cast<unsigned int>(__popcll(want_mask & ((1ull << lane) - 1ull)));
            unsigned int idx = n;
            if (rank < avail) idx = qnext + rank;
            else if (fresh < n && fresh + (rank - avail) < min(fresh + chunk, n)) idx = fresh + (rank - avail);
            if (fresh < n) {
                const unsigned int used = cnt > avail ? cnt - avail : 0u;
                qend = min(fresh + chunk, n);
                qnext = min(fresh + used, qend);
                if (io.paths) {
                    // Row lengths of walkers [0, split) and [split, n).
                    const int32_t la = A.length > 0 ? A.length : 1, lb = B.length > 0 ? B.length : 1;
                    fill_chunk_rows(io, split, swapped ? lb : la, swapped ? la : lb, fresh, qend, rot, lane);
                }
            } else {
                qnext += min(cnt, avail);
                // No chunk was left: the queue is empty once this wavefront's own chunk is.
                if (avail < cnt) queue_empty = true;
            }
            if (want && idx < n) {
                // Queue position idx runs walker idx + rot (mod n): the kind with the longer walks
                // starts first, so the batch ends on short walks.
                nw = static_cast<int32_t>(idx < n - rot ? idx + rot : idx - (n - rot));
                nsrc = io.sources[nw];
                const uint64_t gid = io.walker_ids ? io.walker_ids[nw] : io.gid_base + static_cast<uint64_t>(nw);
                ngid_lo = static_cast<uint32_t>(gid);
                ngid_hi = static_cast<uint32_t>(gid >> 32);
                nin = make_uint4(0u, kNoRange, 0u, kNoRange);
                stage = 1;
            }
        }

        // 3. Last iteration's stores, issued after this iteration's loads; then a new walker's
        // source into its staging row (which the drain has just read).
        out.drain();
        done.drain(io);
        if (fresh) {
            out.put(s.cur, 1, kind_b ? B.length : A.length);
            fresh = false;
        }

        // 4. One step of the lane's walker.
        const bool active = s.st == kAlive;
        if (__ballot(active || stage != 0) == 0) break;  // nothing walks, nothing is prepared: drained
        if (active) {
            const int32_t length = kind_b ? B.length : A.length;
            // This step's Philox block, computed here by every walking lane alike: a lane that has
            // just taken a new walker needs no block of its own (in a wavefront every path any lane
            // takes costs every lane).
            blk = first_block<REPLAY>(io, w, s);
            bool moved;
            if constexpr (MODE <= -2) {
                moved = kind_b ? advance_mode<pair_second(MODE), REPLAY, false>(g, B, io, w, s, blk)
                               : advance_mode<pair_first(MODE), REPLAY, false>(g, A, io, w, s, blk);
            } else {
                moved = kind_b ? advance<MODE, REPLAY, false>(g, B, io, w, s, blk)
                               : advance<MODE, REPLAY, false>(g, A, io, w, s, blk);
            }
            if (moved) {
                ++steps;
                if (out.row && s.len - 1 < length) out.put(s.cur, s.len, length);
            }
            if (s.st != kAlive) {
                // The last chunk (unless put has just handed a full one over), padded to its
                // boundary in padded rows; stored with the outputs at the next drain.
                if (out.row && out.staged > 0) {
                    const int32_t e = s.len < length ? s.len : length;
                    out.hand_over(io.unpadded ? e : out.pad_chunk(e, length));
                }
                done.hold(w, s);
                classify(s.st, gterm, dead, err);
            }
        }
    }
    out.drain();
    done.drain(io);
    flush_counters(io, steps, gterm, dead, err);
}

// ----------------------------------------------------------------- step form
template <int MODE, bool REPLAY>
__global__ __launch_bounds__(kStepBlock) void walk_step_init_kernel(DevGraph g, KindParams A, KindParams B,
                                                                     int32_t split, BatchIO io, WalkerState *state) {
    const int32_t w = blockIdx.x * kStepBlock + threadIdx.x;
    if (w >= io.n) return;
    WalkerState s;
    init_walker(g, io, w, s);
    state[w] = s;
    if (io.paths) path_row(io, split, w)[0] = s.cur;
}

template <int MODE, bool REPLAY>
__global__ __launch_bounds__(kStepBlock) void walk_step_kernel(DevGraph g, KindParams A, KindParams B, int32_t split,
                                                                BatchIO io, WalkerState *state, unsigned int *alive) {
    const int32_t w = blockIdx.x * kStepBlock + threadIdx.x;
    bool still = false;
    if (w < io.n) {
        WalkerState s = state[w];
        if (s.st == kAlive) {
            const bool kind_b = w >= split;
            const int32_t plen = kind_b ? B.length : A.length;  // a value, not a reference to either kind
            uint4 blk = first_block<REPLAY>(io, w, s);
            const bool moved =
                kind_b ? advance<MODE, REPLAY>(g, B, io, w, s, blk) : advance<MODE, REPLAY>(g, A, io, w, s, blk);
            if (moved && io.paths && s.len - 1 < plen) path_row(io, split, w)[s.len - 1] = s.cur;
            still = s.st == kAlive;
            state[w] = s;
        }
    }
    // Any live walker sets the flag; a plain store, so the launch does not serialise on it.
    // Steps are counted once, from the final lengths, by the finish kernel.
    if (__ballot(still) && (threadIdx.x & 63) == 0) *alive = 1u;
}

// Appends the lanes with keep set to list_out: one returning atomic per block (the
// block's wavefront counts are combined in LDS first), so the shared counter is not
// the bottleneck.  Every thread of the block must call it.
__device__ __forceinline__ void block_append(bool keep, int32_t w, int32_t *list_out, unsigned int *count_out) {
    constexpr int kWaves = kStepBlock / 64;
    __shared__ unsigned int wave_count[kWaves];
    __shared__ unsigned int block_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long mask = __ballot(keep);
    if (lane == 0) wave_count[wave] = static_cast<unsigned int>(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int total = 0;
        for (int j = 0; j < kWaves; ++j) total += wave_count[j];
        block_base = total ? atomicAdd(count_out, total) : 0u;
    }
    __syncthreads();
    unsigned int offset = block_base;
    for (int j = 0; j < wave; ++j) offset += wave_count[j];
    if (keep) list_out[offset + __popcll(mask & ((1ull << lane) - 1ull))] = w;
    __syncthreads();  // wave_count and block_base are reused by the next call
}

// GRAPH form: each launch advances every walker on the live list by up to `burst` steps, keeping
// its state in registers between them, then stores the state of the ones still walking and appends
// them to list_out.  The first launch (first != 0) starts every walker of the batch from its source
// instead of reading a list.  A walker that ends writes its outputs there and then, so no pass over
// the batch follows the walk.  Counters rotate over three words: a launch reads count_in, appends to
// count_out and zeroes count_zero, the one the launch before it read, so a chunk of launches holds
// kernel nodes only (no memset) and can be replayed from a hipGraph.  Walks equal the other forms'.
template <int MODE, bool REPLAY>
__global__ __launch_bounds__(kStepBlock) void walk_graph_kernel(DevGraph g, KindParams A, KindParams B, int32_t split,
                                                                BatchIO io, WalkerState *state, const int32_t *list_in,
                                                                const unsigned int *count_in, int32_t *list_out,
                                                                unsigned int *count_out, unsigned int *count_zero,
                                                                int32_t first, int32_t burst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *count_zero = 0u;
    const uint32_t count = first ? static_cast<uint32_t>(io.n) : *count_in;
    const uint32_t stride = gridDim.x * kStepBlock;
    unsigned long long steps = 0, gterm = 0, dead = 0, err = 0;
    // Block-uniform trip count: every thread of the block reaches each block_append.
    for (uint32_t base = blockIdx.x * kStepBlock; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        int32_t w = 0;
        if (i < count) {
            WalkerState s;
            if (first) {
                w = static_cast<int32_t>(i);
                init_walker(g, io, w, s);
                if (io.paths && io.unpadded) path_row(io, split, w)[0] = s.cur;
            } else {
                w = list_in[i];
                s = state[w];
            }
            const bool kind_b = w >= split;
            const int32_t plen = kind_b ? B.length : A.length;  // a value, not a reference to either kind
            for (int32_t k = 0; k < burst && s.st == kAlive; ++k) {
                uint4 blk = first_block<REPLAY>(io, w, s);
                const bool moved = kind_b ? advance<MODE, REPLAY, false>(g, B, io, w, s, blk)
                                          : advance<MODE, REPLAY, false>(g, A, io, w, s, blk);
                if (moved) {
                    ++steps;
                    if (io.paths && s.len - 1 < plen) path_row(io, split, w)[s.len - 1] = s.cur;
                }
            }
            keep = s.st == kAlive;
            if (keep) {
                state[w] = s;
            } else {
                write_outputs(io, w, s);
                classify(s.st, gterm, dead, err);
            }
        }
        block_append(keep, w, list_out, count_out);
    }
    flush_counters(io, steps, gterm, dead, err);
}

// Every move adds one to a walker's length, so a batch's steps are the sum of
// (length - 1); the counters are reduced over the block before one atomic each.
__global__ __launch_bounds__(kStepBlock) void walk_step_finish_kernel(BatchIO io, const WalkerState *state) {
    constexpr int kWaves = kStepBlock / 64;
    __shared__ unsigned long long part[4][kWaves];
    const int32_t w = blockIdx.x * kStepBlock + threadIdx.x;
    unsigned long long steps = 0, gterm = 0, dead = 0, err = 0;
    if (w < io.n) {
        WalkerState s = state[w];
        if (s.st == kAlive) s.st = kDone;
        write_outputs(io, w, s);
        classify(s.st, gterm, dead, err);
        steps = static_cast<unsigned long long>(s.len > 1 ? s.len - 1 : 0);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    steps = wave_sum(steps);
    gterm = wave_sum(gterm);
    dead = wave_sum(dead);
    err = wave_sum(err);
    if (lane == 0) {
        part[0][wave] = steps;
        part[1][wave] = gterm;
        part[2][wave] = dead;
        part[3][wave] = err;
    }
    __syncthreads();
    if (io.counters != nullptr && threadIdx.x < 4) {
        unsigned long long v = 0;
        for (int j = 0; j < kWaves; ++j) v += part[threadIdx.x][j];
        if (v) atomicAdd(io.counters + threadIdx.x, v);
    }
}

// Path rows before the STEP and GRAPH forms run: the walker's source at position 0 and -1
// in [1, L), so that a walk writes only the vertices it moves to (measured: padding each row with
// 16-byte stores as its walk ends made the PPR + MetaPath mix 11% slower and MetaPath alone 30%,
// since stores share the wait counter with the gathers the walk waits on).  Rows that are 16-byte aligned
// with L a multiple of 4 are written as 16-byte pieces, a block covering 256 / (L / 4) consecutive
// rows at a time (consecutive threads, consecutive pieces); other rows an entry per thread.
__global__ void fill_rows_source_vec_kernel(int32_t *p, const int32_t *src, int64_t rows, int32_t stride, int32_t L) {
    const int32_t q = L >> 2;  // pieces per row
    const int32_t per = q <= static_cast<int32_t>(blockDim.x) ? static_cast<int32_t>(blockDim.x) / q : 1;
    const int32_t r_in = static_cast<int32_t>(threadIdx.x) / (q <= static_cast<int32_t>(blockDim.x) ? q : 1 << 30);
    const int32_t v0 = q <= static_cast<int32_t>(blockDim.x) ? static_cast<int32_t>(threadIdx.x) - r_in * q
                                                             : static_cast<int32_t>(threadIdx.x);
    if (r_in >= per) return;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * per + r_in; row < rows;
         row += static_cast<int64_t>(gridDim.x) * per) {
        int4 *r = reinterpret_cast<int4 *>(p + row * stride);
        for (int32_t v = v0; v < q; v += (q <= static_cast<int32_t>(blockDim.x) ? q : static_cast<int32_t>(blockDim.x)))
            r[v] = make_int4(v == 0 ? src[row] : -1, -1, -1, -1);
    }
}

__global__ void fill_rows_source_kernel(int32_t *p, const int32_t *src, int64_t rows, int32_t stride, int32_t L) {
    const int64_t total = rows * L;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t row = L == 1 ? i : i / L;
        const int64_t col = i - row * L;
        p[row * stride + col] = col == 0 ? src[row] : -1;
    }
}

__global__ void fill_u32_kernel(unsigned int *p, unsigned int value, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void copy_u128_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

inline unsigned grid_for(int64_t n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

// Positions per row a fused kernel with `rows` walker rows per block stages in LDS: the whole
// walk when it fits `budget` bytes, else the largest multiple of 32 that does (of 4 below 32).
int32_t fused_tile(int32_t L, int32_t rows, int32_t budget) {
    const int32_t tmax = budget / (4 * rows) - 1;
    if (L <= tmax) return L < 1 ? 1 : L;
    if (tmax >= 32) return tmax & ~31;
    return std::max(4, tmax & ~3);
}

template <int MODE, bool REPLAY>
void fused_multi_launch(const DevGraph &g, const KindParams &p, const BatchIO &io, hipStream_t stream) {
    const int32_t wpt = p.walkers_per_thread == 4 ? 4 : p.walkers_per_thread == 2 ? 2 : 1;
    const int32_t block = p.fused_block > 0 ? p.fused_block  // the graph's tuning for this Move
                          : wpt == 4        ? TRW_MULTI_BLOCK4
                          : wpt == 2        ? TRW_MULTI_BLOCK2
                                            : TRW_MULTI_BLOCK1;
    int32_t T = fused_tile(p.length, wpt * block, kTileBudget);
    if (p.fused_tile > 0 && p.fused_tile < T) T = p.fused_tile;
    const size_t lds = static_cast<size_t>(wpt) * block * (T + 1) * sizeof(int32_t);
    const dim3 grid(grid_for(io.n, wpt * block));
    switch (wpt) {
    case 4:
        hipLaunchKernelGGL((walk_fused_multi_kernel<MODE, REPLAY, 4>), grid, dim3(block), lds, stream, g, p, io, T);
        break;
    case 2:
        hipLaunchKernelGGL((walk_fused_multi_kernel<MODE, REPLAY, 2>), grid, dim3(block), lds, stream, g, p, io, T);
        break;
    default:
        hipLaunchKernelGGL((walk_fused_multi_kernel<MODE, REPLAY, 1>), grid, dim3(block), lds, stream, g, p, io, T);
        break;
    }
}

// walk_fused_kernel's launch shape per Move: threads per block and LDS bytes for path tiles.
struct FusedShape {
    int32_t block;
    int32_t budget;
};
#ifndef TRW_FUSED_SHAPE_ITS
#define TRW_FUSED_SHAPE_ITS {64, 32768}     // whole rows (+4.6% over 16-position chunks)
#endif
#ifndef TRW_FUSED_SHAPE_RS
#define TRW_FUSED_SHAPE_RS {256, 33792}     // 32-position chunks (+6%)
#endif
#ifndef TRW_FUSED_SHAPE_OTHER
#define TRW_FUSED_SHAPE_OTHER {256, 17408}  // 16-position chunks
#endif
FusedShape fused_shape(int32_t mode) {
    if (mode == kModeIts) return TRW_FUSED_SHAPE_ITS;
    if (mode == kModeRejection) return TRW_FUSED_SHAPE_RS;
    return TRW_FUSED_SHAPE_OTHER;
}

template <bool REPLAY>
hipError_t fused_dispatch(const DevGraph &g, const KindParams &p, const BatchIO &io, hipStream_t stream) {
    const bool records = p.mode == kModeUniform ? g.erec != nullptr
                         : p.mode == kModeAlias ? g.arec != nullptr
                         : p.mode == kModeIts   ? g.grec != nullptr && p.its_guide
                                                : false;
    if (records && !p.ppr) {
        if (p.mode == kModeUniform) fused_multi_launch<kModeUniform, REPLAY>(g, p, io, stream);
        else if (p.mode == kModeAlias) fused_multi_launch<kModeAlias, REPLAY>(g, p, io, stream);
        else fused_multi_launch<kModeIts, REPLAY>(g, p, io, stream);
        return hipGetLastError();
    }
    FusedShape sh = fused_shape(p.mode);
    if (p.fused_block > 0) sh.block = p.fused_block;  // the graph's tuning for this Move
    int32_t T = fused_tile(p.length, sh.block, sh.budget);
    if (p.fused_tile > 0 && p.fused_tile < T) T = p.fused_tile;
    const size_t lds = static_cast<size_t>(sh.block) * (T + 1) * sizeof(int32_t);
    const dim3 grid(grid_for(io.n, sh.block)), block(sh.block);
    switch (p.mode) {
    case kModeUniform: hipLaunchKernelGGL((walk_fused_kernel<kModeUniform, REPLAY>), grid, block, lds, stream, g, p, io, T); break;
    case kModeAlias: hipLaunchKernelGGL((walk_fused_kernel<kModeAlias, REPLAY>), grid, block, lds, stream, g, p, io, T); break;
    case kModeIts: hipLaunchKernelGGL((walk_fused_kernel<kModeIts, REPLAY>), grid, block, lds, stream, g, p, io, T); break;
    case kModeRejection: hipLaunchKernelGGL((walk_fused_kernel<kModeRejection, REPLAY>), grid, block, lds, stream, g, p, io, T); break;
    case kModeNode2Vec:
        if (g.nrec && g.mlines && g.mroot && !p.ppr && (!p.n2v_weighted || g.wnrec)) {
            if (p.n2v_block == 128) {
                if (p.n2v_weighted)
                    hipLaunchKernelGGL((walk_node2vec_sm_kernel<REPLAY, true, 128>), dim3(grid_for(io.n, 128)),
                                       dim3(128), 0, stream, g, p, io);
                else
                    hipLaunchKernelGGL((walk_node2vec_sm_kernel<REPLAY, false, 128>), dim3(grid_for(io.n, 128)),
                                       dim3(128), 0, stream, g, p, io);
            } else if (p.n2v_weighted) {
                hipLaunchKernelGGL((walk_node2vec_sm_kernel<REPLAY, true, kN2vBlock>), dim3(grid_for(io.n, kN2vBlock)),
                                   dim3(kN2vBlock), 0, stream, g, p, io);
            } else {
                hipLaunchKernelGGL((walk_node2vec_sm_kernel<REPLAY, false, kN2vBlock>), dim3(grid_for(io.n, kN2vBlock)),
                                   dim3(kN2vBlock), 0, stream, g, p, io);
            }
        } else {
            hipLaunchKernelGGL((walk_fused_node2vec_kernel<REPLAY>), grid, block, lds, stream, g, p, io, T);
        }
        break;
    default: hipLaunchKernelGGL((walk_fused_kernel<kModeMetaPath, REPLAY>), grid, block, lds, stream, g, p, io, T); break;
    }
    return hipGetLastError();
}

// Blocks of a persistent kernel that one CU holds at once.  The grid is that times the CUs: blocks
// beyond it would start only when the resident ones have drained the queue, and find it empty.
template <class K>
int blocks_per_cu(K kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kQueueBlock, 0) != hipSuccess || per_cu < 1)
        per_cu = 2;
    return per_cu;
}

// Rows [0, rows) of length L: each walker's source at position 0 (valid or not), -1 after it.
void fill_source_rows(int32_t *paths, const int32_t *src, int64_t rows, int32_t stride, int32_t L,
                      hipStream_t stream) {
    if (rows <= 0) return;
    if ((L & 3) == 0 && (stride & 3) == 0 && (reinterpret_cast<uintptr_t>(paths) & 15) == 0) {
        const int32_t q = L >> 2, per = q <= 256 ? 256 / q : 1;
        const int64_t blocks = std::min<int64_t>((rows + per - 1) / per, 256 * 64);
        hipLaunchKernelGGL(fill_rows_source_vec_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, paths,
                           src, rows, stride, L);
    } else {
        const int64_t blocks = std::min<int64_t>(grid_for(rows * L, 256), 256 * 64);
        hipLaunchKernelGGL(fill_rows_source_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, paths,
                           src, rows, stride, L);
    }
}

// Every walker's row holds its source and -1 before a STEP or GRAPH walk starts (the QUEUE form
// fills each chunk's rows as a wavefront takes it, fill_chunk_rows).
void fill_paths(const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io, hipStream_t stream) {
    if (!io.paths || io.n <= 0 || io.unpadded) return;  // unpadded: the walk writes its source itself
    const int32_t La = a.length > 0 ? a.length : 1, Lb = b.length > 0 ? b.length : 1;
    const int32_t sa = split < io.n ? split : io.n;
    fill_source_rows(io.paths, io.sources, sa, io.path_stride, La, stream);
    fill_source_rows(io.paths + static_cast<int64_t>(sa) * io.path_stride, io.sources + sa, io.n - sa, io.path_stride_b,
                     Lb, stream);
}

template <int MODE, bool REPLAY>
void queue_launch(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                  unsigned int *q, hipStream_t stream, int cus) {
    static const int per_cu = blocks_per_cu(walk_queue_kernel<MODE, REPLAY>);  // one kernel, one answer
    // The graph's tuning may keep fewer blocks per CU than fit (first kind's Move).
    const int use = a.queue_blocks_per_cu > 0 && a.queue_blocks_per_cu < per_cu ? a.queue_blocks_per_cu : per_cu;
    const int64_t resident = static_cast<int64_t>(use) * cus, needed = grid_for(io.n, kQueueBlock);
    const unsigned grid = static_cast<unsigned>(resident < needed ? resident : needed);
    // About eight chunks per wavefront of the grid, between 64 and 4096 walkers each, unless the
    // graph's tuning fixes it: a wavefront that takes its last chunk late ends the batch late.
    const int64_t waves = static_cast<int64_t>(grid) * (kQueueBlock / 64);
    int64_t chunk = a.queue_chunk > 0 ? a.queue_chunk : static_cast<int64_t>(io.n) / (waves * 8);
    // A tuned chunk is at most the batch's share per wavefront, so a small batch still spreads
    // over every wavefront of the grid.
    if (a.queue_chunk > 0) chunk = std::min<int64_t>(chunk, static_cast<int64_t>(io.n) / waves);
    chunk = chunk < 64 ? 64 : (chunk > 4096 ? 4096 : chunk);
    // Expected walk length of a kind: its length, or min(length, 1 / stop probability) for a walk
    // that stops by probability.  The kind whose walks run longer takes the queue's first positions.
    const auto expected = [](const KindParams &p) {
        const double L = p.length > 0 ? p.length : 1.0;
        return p.ppr && p.stop_prob > 0.0 ? std::min(L, 1.0 / p.stop_prob) : L;
    };
    const unsigned int rot =
        (split > 0 && split < io.n && expected(b) > expected(a)) ? static_cast<unsigned int>(split) : 0u;
    // A pair kernel takes the kind of its first Move as A (walkers [0, split) then run B).
    const bool swap = MODE <= -2 && a.mode != pair_first(MODE);
    hipLaunchKernelGGL((walk_queue_kernel<MODE, REPLAY>), dim3(grid), dim3(kQueueBlock), 0, stream, g,
                       swap ? b : a, swap ? a : b, split, static_cast<int32_t>(swap), io, q, static_cast<unsigned int>(chunk), rot);
}

template <bool REPLAY>
void queue_dispatch(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                    unsigned int *q, hipStream_t stream, int cus) {
    const bool single = (split >= io.n) || (a.mode == b.mode);
    if (!single) {
        // Two kinds: the kernel compiled for exactly their two Moves.
#define TRW_PAIR(x, y)                                                                                \
    case pair_mode(x, y):                                                                             \
        queue_launch<pair_mode(x, y), REPLAY>(g, a, b, split, io, q, stream, cus);                    \
        return;
        switch (pair_mode(a.mode, b.mode)) {
            TRW_PAIR(kModeUniform, kModeAlias)
            TRW_PAIR(kModeUniform, kModeIts)
            TRW_PAIR(kModeUniform, kModeRejection)
            TRW_PAIR(kModeUniform, kModeNode2Vec)
            TRW_PAIR(kModeUniform, kModeMetaPath)
            TRW_PAIR(kModeAlias, kModeIts)
            TRW_PAIR(kModeAlias, kModeRejection)
            TRW_PAIR(kModeAlias, kModeNode2Vec)
            TRW_PAIR(kModeAlias, kModeMetaPath)
            TRW_PAIR(kModeIts, kModeRejection)
            TRW_PAIR(kModeIts, kModeNode2Vec)
            TRW_PAIR(kModeIts, kModeMetaPath)
            TRW_PAIR(kModeRejection, kModeNode2Vec)
            TRW_PAIR(kModeRejection, kModeMetaPath)
            TRW_PAIR(kModeNode2Vec, kModeMetaPath)
        default:
            queue_launch<-1, REPLAY>(g, a, b, split, io, q, stream, cus);
            return;
        }
#undef TRW_PAIR
    }
    switch (split >= io.n ? a.mode : b.mode) {
    case kModeUniform: queue_launch<kModeUniform, REPLAY>(g, a, b, split, io, q, stream, cus); break;
    case kModeAlias: queue_launch<kModeAlias, REPLAY>(g, a, b, split, io, q, stream, cus); break;
    case kModeIts: queue_launch<kModeIts, REPLAY>(g, a, b, split, io, q, stream, cus); break;
    case kModeRejection: queue_launch<kModeRejection, REPLAY>(g, a, b, split, io, q, stream, cus); break;
    case kModeNode2Vec: queue_launch<kModeNode2Vec, REPLAY>(g, a, b, split, io, q, stream, cus); break;
    default: queue_launch<kModeMetaPath, REPLAY>(g, a, b, split, io, q, stream, cus); break;
    }
}

template <bool REPLAY>
void step_dispatch(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                   WalkerState *state, unsigned int *alive, hipStream_t stream) {
    const dim3 grid(grid_for(io.n, kStepBlock)), block(kStepBlock);
    const bool single = (split >= io.n) || (a.mode == b.mode);
    if (!single) {
        hipLaunchKernelGGL((walk_step_kernel<-1, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive);
        return;
    }
    switch (split >= io.n ? a.mode : b.mode) {
    case kModeUniform: hipLaunchKernelGGL((walk_step_kernel<kModeUniform, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive); break;
    case kModeAlias: hipLaunchKernelGGL((walk_step_kernel<kModeAlias, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive); break;
    case kModeIts: hipLaunchKernelGGL((walk_step_kernel<kModeIts, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive); break;
    case kModeRejection: hipLaunchKernelGGL((walk_step_kernel<kModeRejection, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive); break;
    case kModeNode2Vec: hipLaunchKernelGGL((walk_step_kernel<kModeNode2Vec, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive); break;
    default: hipLaunchKernelGGL((walk_step_kernel<kModeMetaPath, REPLAY>), grid, block, 0, stream, g, a, b, split, io, state, alive); break;
    }
}

}  // namespace

hipError_t launch_walk_fused(const DevGraph &g, const KindParams &p, const BatchIO &io, hipStream_t stream) {
    if (io.n <= 0) return hipSuccess;
    return io.replay ? fused_dispatch<true>(g, p, io, stream) : fused_dispatch<false>(g, p, io, stream);
}

hipError_t launch_walk_queue(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                             const BatchIO &io, unsigned int *queue_counter, hipStream_t stream,
                             bool head_zeroed) {
    if (io.n <= 0) return hipSuccess;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    if (!head_zeroed) hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, queue_counter, 0u, int64_t(1));
    if (io.replay) queue_dispatch<true>(g, a, b, split, io, queue_counter, stream, cus);
    else queue_dispatch<false>(g, a, b, split, io, queue_counter, stream, cus);
    return hipGetLastError();
}

hipError_t launch_fill_paths(const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                             hipStream_t stream) {
    if (io.n <= 0) return hipSuccess;
    fill_paths(a, b, split, io, stream);
    return hipGetLastError();
}

hipError_t launch_walk_step_init(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                                 const BatchIO &io, WalkerState *state, hipStream_t stream) {
    if (io.n <= 0) return hipSuccess;
    fill_paths(a, b, split, io, stream);
    const dim3 grid(grid_for(io.n, kStepBlock)), block(kStepBlock);
    if (io.replay) hipLaunchKernelGGL((walk_step_init_kernel<-1, true>), grid, block, 0, stream, g, a, b, split, io, state);
    else hipLaunchKernelGGL((walk_step_init_kernel<-1, false>), grid, block, 0, stream, g, a, b, split, io, state);
    return hipGetLastError();
}

hipError_t launch_walk_step(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                            const BatchIO &io, WalkerState *state, unsigned int *alive, hipStream_t stream) {
    if (io.n <= 0) return hipSuccess;
    if (io.replay) step_dispatch<true>(g, a, b, split, io, state, alive, stream);
    else step_dispatch<false>(g, a, b, split, io, state, alive, stream);
    return hipGetLastError();
}

template <int MODE, bool REPLAY>
void graph_launch(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                  WalkerState *state, const int32_t *list_in, const unsigned int *count_in, int32_t *list_out,
                  unsigned int *count_out, unsigned int *count_zero, int32_t first, int32_t burst, unsigned grid,
                  hipStream_t stream) {
    hipLaunchKernelGGL((walk_graph_kernel<MODE, REPLAY>), dim3(grid), dim3(kStepBlock), 0, stream, g, a, b, split, io,
                       state, list_in, count_in, list_out, count_out, count_zero, first, burst);
}

template <bool REPLAY>
void graph_dispatch(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                    WalkerState *state, const int32_t *list_in, const unsigned int *count_in, int32_t *list_out,
                    unsigned int *count_out, unsigned int *count_zero, int32_t first, int32_t burst, unsigned grid,
                    hipStream_t stream) {
    const bool single = (split >= io.n) || (a.mode == b.mode);
#define TRW_ARGS g, a, b, split, io, state, list_in, count_in, list_out, count_out, count_zero, first, burst, grid, stream
    if (!single) {
        graph_launch<-1, REPLAY>(TRW_ARGS);
        return;
    }
    switch (split >= io.n ? a.mode : b.mode) {
    case kModeUniform: graph_launch<kModeUniform, REPLAY>(TRW_ARGS); break;
    case kModeAlias: graph_launch<kModeAlias, REPLAY>(TRW_ARGS); break;
    case kModeIts: graph_launch<kModeIts, REPLAY>(TRW_ARGS); break;
    case kModeRejection: graph_launch<kModeRejection, REPLAY>(TRW_ARGS); break;
    case kModeNode2Vec: graph_launch<kModeNode2Vec, REPLAY>(TRW_ARGS); break;
    default: graph_launch<kModeMetaPath, REPLAY>(TRW_ARGS); break;
    }
#undef TRW_ARGS
}

hipError_t launch_walk_graph(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                             const BatchIO &io, WalkerState *state, const int32_t *list_in,
                             const unsigned int *count_in, int32_t *list_out, unsigned int *count_out,
                             unsigned int *count_zero, int32_t first, int32_t burst, hipStream_t stream) {
    if (io.n <= 0) return hipSuccess;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    // Full occupancy (8 wavefronts of 256-lane blocks per CU) or the batch, whichever is smaller.
    const int64_t needed = grid_for(io.n, kStepBlock);
    const unsigned grid = static_cast<unsigned>(needed < int64_t(cus) * 8 ? needed : int64_t(cus) * 8);
    if (io.replay)
        graph_dispatch<true>(g, a, b, split, io, state, list_in, count_in, list_out, count_out, count_zero, first,
                             burst, grid, stream);
    else
        graph_dispatch<false>(g, a, b, split, io, state, list_in, count_in, list_out, count_out, count_zero, first,
                              burst, grid, stream);
    return hipGetLastError();
}

hipError_t launch_walk_step_finish(const KindParams &, const KindParams &, int32_t, const BatchIO &io,
                                   const WalkerState *state, hipStream_t stream) {
    if (io.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(walk_step_finish_kernel, dim3(grid_for(io.n, kStepBlock)), dim3(kStepBlock), 0, stream, io,
                       state);
    return hipGetLastError();
}

hipError_t launch_copy_u128(uint4 *dst, const uint4 *src, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(grid_for(n, 256), 256 * 64);
    hipLaunchKernelGGL(copy_u128_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int *p, unsigned int value, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, value, n);
    return hipGetLastError();
}

}  // namespace trw

// Internal types shared by the HIP kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>

#include "thunderrw_device.h"

#include <cstdint>

namespace trw {

// AliasSlot, config/type_config.h:22-26.  16 bytes; read on the device as one
// 128-bit load (the table is hipMalloc'ed, so every slot is 16-byte aligned).
struct AliasSlot {
    double prob;
    int32_t first;
    int32_t second;
};
static_assert(sizeof(AliasSlot) == 16, "AliasSlot must stay 16 bytes");

// Per-walker outcome codes (same values as TRW_WALKER_* in thunderrw.h).
enum WalkerStatus : int32_t {
    kAlive = -1,
    kDone = 0,
    kGatherTerminated = 1,
    kDeadEnd = 2,
    kReplayExhausted = 3,
    kTrialCap = 4,
    kBadSource = 5,
};

// Move kernels.  One per (execution mode, sampler) pair the device provides.
enum Mode : int32_t {
    kModeUniform = 0,   // uniform_move            uniform_sampling.h:47-63
    kModeAlias = 1,     // static_alias_move       alias_sampling.h:147-172
    kModeIts = 2,       // static_its_move         inverse_transformation_sampling.h:259-286
    kModeRejection = 3, // static_rs_move          rejection_sampling.h:334-366
    kModeNode2Vec = 4,  // dynamic_max_weight_rs_move + Node2Vec::weight  rejection_sampling.h:368-402
    kModeMetaPath = 5,  // dynamic_gather + dynamic_its_move + MetaPath::weight  creeper.h:178-296
};

// Read-only CSR view in HBM (util/graph/graph.h:51-83 field meanings).
struct DevGraph {
    int32_t nv;
    int64_t ne;
    const int64_t *offset;     // nv + 1
    const int32_t *adj;        // ne, sorted per vertex
    const int32_t *degree;     // nv
    const double *weight;      // ne or null
    const int32_t *label;      // ne or null
    const AliasSlot *alias;    // ne (alias tables)
    const double *prefix;      // ne (ITS prefix sums)
    const double *rs_max;      // nv (rejection maxima)
    const int32_t *lbl_sorted; // ne: labels of each vertex's segment, sorted (stable)
    const int32_t *lbl_pos;    // ne: segment position of each entry of lbl_sorted
    const uint2 *vrec;         // nv: {first edge, degree} when num_edges < 2^32, else null
    const uint4 *erec;         // ne: neighbour records {adj[e], first edge of adj[e], degree of adj[e], 0}
    const uint4 *arec;         // 2*ne: alias walk records {prob lo, prob hi, first, second}
                               //       {first's edge, first's degree, second's edge, second's degree}
    const uint32_t *lstart;    // kLabelRecWords*nv when every label is in [0, 8), else null: v's 64-byte
                               //       label record, words k = 0..7 the neighbours of v with label < k
                               //       (the start of label k in label order), word 8 v's first edge,
                               //       word 9 its degree (a MetaPath walk's first step reads one line)
    const uint4 *lerec;        // 4*ne with lstart: 64-byte neighbour records in label order, entry j of v:
                               //       lerec[4 (off(v) + j)] = erec[off(v) + lbl_pos[off(v) + j]], then the
                               //       neighbour's 8 label starts (2 uint4), then 16 bytes of zero
    const uint4 *msrec;        // ne with lerec, built for one MetaPath schema whose labels are distinct and in
                               //       [0, 8): entry j of v's label order {neighbour u, u's first edge, the
                               //       range [lo, hi) of the schema label after this edge's at u (label
                               //       order), kNoRange twice when u has no edges}: a MetaPath step reads 16
                               //       bytes (a table a quarter of lerec's, within the TLB's reach)
    const uint4 *its_lines;    // ITS search structures (128-byte lines, 8 uint4 each), or null
    const uint32_t *its_root;  // nv: first line of v's structure, kNoRange if v has none
    const uint4 *rrec;         // 2*ne: rejection walk records {weight lo, weight hi, adj[e], first edge of adj[e]}
                               //       {degree of adj[e], its maximum weight lo, hi, 0}, or null
    const uint4 *grec;         // ITS guide records (2 uint4 per bucket), or null; see its_guide below
    const uint2 *gvrec;        // nv: {first bucket, bucket count} per vertex ({kNoRange, 0}: weights sum to zero)
    const uint4 *nrec;         // ne: neighbour records with .w = the neighbour's membership table root
    const uint4 *wnrec;        // 2*ne: weighted Node2Vec records {weight lo, weight hi, adj[e], first edge of adj[e]}
                               //       {degree of adj[e], its membership root, 0, 0}, or null
    const uint4 *mlines;       // membership tables (128-byte buckets), or null
    const uint32_t *mroot;     // nv: first line of v's membership table, kNoRange if v has none
};

// ITS search structure of one vertex, built from its prefix sums when they are
// non-decreasing: 128-byte lines, tree first.  Leaf line i holds the prefix sums of the
// vertex's edges [5i, 5i + 5) (5 doubles, +inf padded, then 8 bytes of padding) and their
// neighbour records {neighbour, its first edge, its degree, its structure's first line}
// (5 x 16 bytes).  Above the leaves, for more than one leaf, a B+-tree of 16-key
// (128-byte) nodes: level 1 holds the last prefix sum of each leaf, each higher level the
// last key of each node below, up to one root node; levels are stored root first.  A
// search reads the root, one node per level and one leaf, and that leaf also holds the
// chosen neighbour's record.  The comparisons are the `key <= target` of an upper bound,
// so for non-decreasing sums the pick is exactly the branchless binary search's.
constexpr int kItsFan = 16;  // keys per tree node
constexpr int kItsLeaf = 5;  // edges per leaf line
constexpr int kItsMaxLevels = 8;

__host__ __device__ inline int32_t its_leaves(int32_t d) { return (d + kItsLeaf - 1) / kItsLeaf; }

// Tree node counts per level (level 1 first); returns the number of levels (0 for one leaf).
__host__ __device__ inline int32_t its_levels(int32_t nl, int64_t *nodes) {
    if (nl <= 1) return 0;
    int32_t L = 0;
    int64_t keys = nl;
    do {
        nodes[L] = (keys + kItsFan - 1) / kItsFan;
        keys = nodes[L];
        ++L;
    } while (nodes[L - 1] > 1 && L < kItsMaxLevels);
    return L;
}

// ITS guide records: static ITS as one gather per step.  A vertex of degree d gets m = c * d
// buckets, c the least of 1..kGuideMaxC whose buckets all span at most two edges; bucket k
// holds the draws u with floor(u * m / 2^32) == k, so a draw's bucket is one multiply.  The
// walk's pick (the upper bound of real2(u) * total in the prefix sums) is non-decreasing in
// u, so over a bucket it runs from pick(first u) to pick(last u); when those differ by at
// most one, the bucket's 32-byte record decides every draw in it:
//   {u threshold, lower pick's neighbour, upper pick's neighbour, 0}
//   {lower's first bucket, lower's bucket count, upper's first bucket, upper's bucket count}
// with the upper pick taken iff u >= threshold (the least u of the bucket whose pick is the
// upper one, found with the Move's own expression).  A bucket spanning more edges has .w = 1
// and the walk searches the prefix sums for that step.  Built only for non-decreasing sums.
constexpr uint32_t kGuideMaxC = 4;

// Membership tables (Graph::is_neighbor's search on the device): thunderrw_device.h.
using trw_dev::kMemberEmpty;
using trw_dev::kMemberMin;
using trw_dev::kMemberSlots;
using trw_dev::member_bucket_bits;
using trw_dev::member_hash;

// Per walker kind launch parameters (app functor + g_para).
struct KindParams {
    int32_t mode;
    int32_t ppr;            // terminate by stop probability (ppr.cpp:27-29) instead of length
    int32_t length;         // g_para.length_
    int32_t max_trials;     // 0 = no cap
    int32_t walkers_per_thread;  // fused uniform/alias: 1, 2 or 4
    int32_t unit_weights;   // rejection weights are 1.0 (unit-weight tables)
    int32_t schema_length;
    double stop_prob;
    double n2v_p, n2v_q, n2v_max;  // Node2Vec p_, q_, max_ (node2vec.cpp:18-20)
    double n2v_bound;       // Node2Vec rejection bound: max_, or max_ times the largest edge weight (weighted)
    int32_t n2v_weighted;   // Node2Vec over edge weights: a candidate's weight is w_e times its factor
    int32_t schema[16];
    uint64_t schema_nibbles;  // schema[i] as nibble i when in [0, 8), else 8 (no neighbour carries it)
    int32_t schema_rec;       // MetaPath: the graph's msrec were built for this schema
    int32_t its_tree_min;   // ITS: vertices of at least this degree search their tree (when built)
    int32_t queue_chunk;    // QUEUE: walker indices a wavefront takes per atomic (0 = from the batch size)
    int32_t its_guide;      // ITS: read the guide records when the graph has them
    // Launch shapes of the Move (0: the built-in choice).
    int32_t fused_block;          // FUSED: threads per block
    int32_t fused_tile;           // FUSED: most path positions a lane stages in LDS per write-out
    int32_t queue_blocks_per_cu;  // QUEUE: resident blocks per CU of the persistent grid
    int32_t n2v_block;            // Node2Vec state machine: threads per block (128 or 256)
    int32_t graph_burst;          // GRAPH: steps per live walker per launch (0: the built-in choice)
};

// Walker batch (device pointers).
struct BatchIO {
    const int32_t *sources;
    int32_t n;
    int32_t path_stride;
    int32_t path_stride_b;       // rows of a mixed batch's walkers [split, n) (follow the first split rows)
    uint64_t gid_base;
    const uint64_t *walker_ids;  // per-walker global ids (null: gid_base + index)
    int32_t unpadded;            // path rows are not padded with -1 past the walk (TRW_PATHS_UNPADDED)
    const uint32_t *draws;
    int64_t K;
    const uint32_t *tdraws;
    int64_t KT;
    int32_t *paths;
    int32_t *lengths;
    int32_t *ends;
    int32_t *status;
    int32_t *draws_used;
    unsigned long long *counters;  // [steps, gather-terminated, dead ends, errors]
    uint32_t key0, key1;           // Philox key
    int32_t replay;
};

// Walker state held in registers (fused, queue) or HBM (step form).
struct WalkerState {
    int32_t cur;
    int32_t prev;
    int32_t len;
    int32_t st;
    uint32_t k;     // replay move-draw cursor
    uint32_t kt;    // replay termination-draw cursor
    uint32_t coff;  // current vertex: first edge and degree (cdeg == kNoRange: not known yet)
    uint32_t cdeg;
    uint32_t poff;  // previous vertex: first edge and degree (Node2Vec membership test); MetaPath: the
    uint32_t pdeg;  //   current vertex's label-order range of this step's label, from the record that
                    //   brought the walker here (pdeg == kNoRange: not known)
    uint32_t gid_lo;  // global walker id: the Philox counter's walker words
    uint32_t gid_hi;
    uint32_t cidx;    // current vertex: root of its ITS search tree (ITS) or membership table
                      // (Node2Vec); MetaPath: this step's schema position; kNoRange: not known / none
    uint32_t pidx;    // previous vertex: root of its membership table (Node2Vec)
    double cmax;      // rejection: the current vertex's maximum weight (from a rejection record), < 0: not known
};
constexpr uint32_t kNoRange = 0xffffffffu;
constexpr int64_t kLabelRecWords = 16;  // DevGraph::lstart words per vertex

// Kernel launchers (walk.hip / tables.hip).  All asynchronous on `stream`.
hipError_t launch_walk_fused(const DevGraph &g, const KindParams &p, const BatchIO &io, hipStream_t stream);
// The QUEUE form zeroes its queue head first unless the caller's own earlier work on the
// stream has (head_zeroed).
hipError_t launch_walk_queue(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                             const BatchIO &io, unsigned int *queue_counter, hipStream_t stream,
                             bool head_zeroed = false);
hipError_t launch_walk_step_init(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                                 const BatchIO &io, WalkerState *state, hipStream_t stream);
hipError_t launch_walk_step(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                            const BatchIO &io, WalkerState *state, unsigned int *alive, hipStream_t stream);
hipError_t launch_walk_step_finish(const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                                   const WalkerState *state, hipStream_t stream);
// Path rows of a STEP or GRAPH batch: source at position 0, -1 after (nothing when unpadded).
hipError_t launch_fill_paths(const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                             hipStream_t stream);
// GRAPH form: advance the live list (first: every walker of the batch, from its source) by up to
// `burst` steps per walker, append the walkers still walking to list_out, zero *count_zero.
hipError_t launch_walk_graph(const DevGraph &g, const KindParams &a, const KindParams &b, int32_t split,
                             const BatchIO &io, WalkerState *state, const int32_t *list_in,
                             const unsigned int *count_in, int32_t *list_out, unsigned int *count_out,
                             unsigned int *count_zero, int32_t first, int32_t burst, hipStream_t stream);
hipError_t launch_fill_u32(unsigned int *p, unsigned int value, int64_t n, hipStream_t stream);

// One query chunk of a serving batch: walkers [batch_offset, batch_offset + count)
// of the batch are walkers [first_walker, first_walker + count) of the query, whose
// source set is pool[pool_offset, pool_offset + n_sources).  A PPR batch cuts each
// segment into tiles of 2048 walkers; first_tile is the segment's first.
struct ServeSegment {
    int32_t batch_offset;
    int32_t count;
    int32_t pool_offset;
    int32_t n_sources;
    int32_t first_walker;
    int32_t first_tile;
    uint64_t id_base;
};
// Expands the segment table into per-walker sources and Philox ids, and zeroes
// zero[0, n_zero) (the batch's walk counters).
hipError_t launch_serve_expand(const ServeSegment *segs, int32_t n_segs, const int32_t *pool, int32_t n,
                               int32_t *sources, uint64_t *ids, uint32_t *zero, int32_t n_zero, hipStream_t stream);

// Totals of one segment (or tile) of a PPR batch.
struct PprSegmentStats {
    uint64_t steps;   // sum over the walkers of (length - 1)
    uint64_t errors;  // walkers that did not end at a vertex
};
// A slot's PPR aggregation buffers (C = chunks of 2^17 vertex ids).
struct PprScratch {
    uint32_t *tile_hist;         // [tiles][C]: walkers per chunk, then their first bucket position
    PprSegmentStats *tile_stats; // [tiles]
    int32_t *bucketed;           // [walkers]: end vertices grouped by (segment, chunk)
    int2 *pairs;                 // [walkers]: (vertex, count) runs at their bucket's positions
    uint2 *bucket;               // [segments][C]: {first position, walkers}
    uint32_t *tickets;           // [segments]: finished tiles, zero when the batch starts
    int32_t nv;
    int32_t C;
};
int32_t ppr_chunks(int32_t nv);
int32_t ppr_tiles(int32_t walkers);  // tiles of one segment
// PPR batch aggregation on the device, into host memory the device can write: each
// segment's totals (host_stats[s]), each (segment, chunk) bucket's {first position,
// runs} (host_buckets[s * C + c]), and its ascending (vertex, count) runs at
// host_pairs[first position ...].
hipError_t launch_serve_ppr_reduce(const ServeSegment *segs, int32_t n_segs, int32_t n_tiles, const int32_t *ends,
                                   const int32_t *lengths, const int32_t *status, const PprScratch &p,
                                   PprSegmentStats *host_stats, uint2 *host_buckets, int2 *host_pairs,
                                   hipStream_t stream);
hipError_t launch_copy_u128(uint4 *dst, const uint4 *src, int64_t n, hipStream_t stream);

hipError_t run_gather_probe(uint64_t table_bytes, int32_t record_bytes, int32_t chains, int32_t steps, int32_t cus,
                            hipStream_t stream, double *gathers_per_s);

// Scratch entry of the alias build: an index's normalised weight (or residual) and neighbour id.
struct AliasEntry {
    double q;
    int32_t nbr;
    int32_t pad;
};
// Table builds.  order lists the vertices by descending degree class; its first `hubs`
// entries (degree >= kWaveBuildMin) are built by a wavefront each, the rest by a lane each.
constexpr int32_t kWaveBuildMin = 1024;
hipError_t launch_alias_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, int32_t hubs,
                              AliasEntry *scratch, AliasSlot *out, unsigned int *zero_count, hipStream_t stream);
hipError_t launch_edge_records(const DevGraph &g, uint4 *erec, hipStream_t stream);
hipError_t launch_alias_records(const DevGraph &g, uint4 *arec, hipStream_t stream);
// Rejection walk records from the weights the maxima were built from (w, or 1.0 when unit).
hipError_t launch_rs_records(const DevGraph &g, const double *w, int32_t unit, uint4 *rrec, hipStream_t stream);
hipError_t launch_label_records(const DevGraph &g, uint32_t *lstart, uint4 *lerec, hipStream_t stream);
// MetaPath schema records for the schema whose next label after label l is nibble l of next_map (15:
// l is not in the schema).
hipError_t launch_schema_records(const DevGraph &g, uint32_t next_map, uint4 *msrec, hipStream_t stream);
// ITS search structures: line counts per vertex into root (then the caller turns them
// into offsets), the structures themselves, and the roots copied into the neighbour
// records' .w.
// ITS structure layout: root[v] = v's first line (kNoRange: none), start[v] likewise
// with start[nv] = all lines (nv + 1 entries).  With temp null only sizes the scan's
// temporary storage into temp_bytes.
hipError_t launch_its_tree_starts(const DevGraph &g, uint32_t *root, uint64_t *start, void *temp, size_t &temp_bytes,
                                  hipStream_t stream);
hipError_t launch_its_tree_fill(const DevGraph &g, const double *prefix, const uint64_t *start, uint64_t lines_total,
                                const uint32_t *root, uint4 *lines, hipStream_t stream);
hipError_t launch_its_tree_roots_to_records(const DevGraph &g, const uint32_t *root, uint4 *erec, hipStream_t stream);
// Membership tables: bucket counts into root, then offsets by a scan (as the ITS structures),
// then the insertion of every neighbour (fill_count: one zeroed counter per bucket, scratch),
// plus the neighbour records carrying each neighbour's root.
hipError_t launch_member_table_starts(const DevGraph &g, uint32_t *root, uint64_t *start, void *temp,
                                     size_t &temp_bytes, hipStream_t stream);
hipError_t launch_member_table_fill(const DevGraph &g, const uint64_t *start, uint64_t lines_total, uint4 *lines,
                                   uint32_t *fill_count, hipStream_t stream);
// ITS guide records (tables.hip): flag and size pass plus scan (temp == null: scan storage
// size only), then ranges and records.  start has nv + 1 entries.
hipError_t launch_its_guide_starts(const DevGraph &g, const double *prefix, uint32_t *flag, uint32_t *count,
                                   uint64_t *start, void *temp, size_t &temp_bytes, hipStream_t stream);
hipError_t launch_its_guide_fill(const DevGraph &g, const double *prefix, const uint64_t *start, uint64_t buckets,
                                 uint2 *gvrec, uint4 *grec, hipStream_t stream);
hipError_t launch_member_records(const DevGraph &g, const uint32_t *root, uint4 *nrec, hipStream_t stream);
// Weighted Node2Vec records from the neighbour records with roots (g.nrec) and the edge weights.
hipError_t launch_member_weight_records(const DevGraph &g, uint4 *wnrec, hipStream_t stream);
hipError_t launch_its_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, int32_t hubs,
                            double *out, hipStream_t stream);
hipError_t launch_rs_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, int32_t hubs,
                           double *out, hipStream_t stream);

}  // namespace trw

// Library-internal declarations shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/thunderrw.h"
#include "engine.h"

struct trw_context {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t timer[2] = {nullptr, nullptr};
    void *scratch = nullptr;  // this device's copy of trw_walk_noalloc's scratch array (api.cpp)
};

struct trw_graph {
    trw_context *ctx = nullptr;
    int32_t nv = 0;
    int64_t ne = 0;
    int32_t max_degree = 0;
    int64_t *offset = nullptr;
    int32_t *adj = nullptr;
    int32_t *degree = nullptr;
    double *weight = nullptr;
    int32_t *label = nullptr;
    int32_t *lbl_sorted = nullptr;
    int32_t *lbl_pos = nullptr;
    int32_t *order = nullptr;
    uint2 *vrec = nullptr;
    uint4 *erec = nullptr;
    uint4 *arec = nullptr;
    uint32_t *lstart = nullptr;
    uint4 *lerec = nullptr;
    uint4 *msrec = nullptr;        // MetaPath schema records (16 bytes per edge) for the schema msrec_key
    uint64_t msrec_key = 0;        // the schema's labels as nibbles, its length in bits 60-63 (0: none built)
    uint4 *its_lines = nullptr;    // ITS search structures (null: not built, or sums not non-decreasing)
    uint32_t *its_root = nullptr;  // nv: first line of each vertex's structure
    uint4 *rrec = nullptr;         // rejection walk records (2 per edge), with the rejection table
    uint4 *grec = nullptr;         // ITS guide records (2 per bucket), with the ITS table when its sums are monotone
    uint2 *gvrec = nullptr;        // nv: {first bucket, bucket count}
    uint4 *nrec = nullptr;         // neighbour records with membership roots (built by the first Node2Vec walk)
    uint4 *wnrec = nullptr;        // weighted Node2Vec records (2 per edge; built by the first weighted Node2Vec walk)
    double max_weight = 0.0;       // the largest edge weight (0 when none is positive), weighted Node2Vec's bound
    uint4 *mlines = nullptr;       // membership tables (hashed neighbour sets, 128-byte buckets)
    uint32_t *mroot = nullptr;     // nv: first line of each vertex's membership table
    bool weights_nonneg = true;    // every edge weight is finite and >= 0
    int32_t build_hubs = 0;        // vertices of degree >= kWaveBuildMin (a prefix of order)
    int64_t tune[4] = {0, 0, 0, 0};  // TRW_TUNE_* knobs 0-3
    int64_t shape[TRW_MOVES][TRW_TUNE_KNOBS - 4] = {};  // launch-shape knobs 4-8 per Move
    trw::AliasSlot *alias = nullptr;
    double *prefix = nullptr;
    double *rs_max = nullptr;
    double *rs_custom = nullptr;   // rejection weights of a TRW_WEIGHTS_CUSTOM build (E)
    int32_t alias_weights = -1, its_weights = -1, rs_weights = -1;  // TRW_WEIGHTS_* the table was built from
    unsigned int *counters = nullptr;                                 // [0] queue head [1] alive [2] zero-sum count
    trw::WalkerState *state = nullptr;
    int32_t *active = nullptr;  // three live-walker lists of state_cap entries (GRAPH form)
    struct GraphChunk *graph_chunk = nullptr;  // the GRAPH form's instantiated step chunk (api.cpp)
    int64_t state_cap = 0;
    double last_build_ms = 0.0;

    trw::DevGraph view() const {
        trw::DevGraph g;
        std::memset(&g, 0, sizeof(g));  // padding too: the GRAPH form keys its cached chunk on these bytes
        g.nv = nv;
        g.ne = ne;
        g.offset = offset;
        g.adj = adj;
        g.degree = degree;
        g.weight = weight;
        g.label = label;
        g.alias = alias;
        g.prefix = prefix;
        g.rs_max = rs_max;
        g.lbl_sorted = lbl_sorted;
        g.lbl_pos = lbl_pos;
        g.vrec = vrec;
        g.erec = erec;
        g.arec = arec;
        g.lstart = lstart;
        g.lerec = lerec;
        g.msrec = msrec;
        g.its_lines = its_lines;
        g.its_root = its_root;
        g.rrec = rrec;
        g.grec = grec;
        g.gvrec = gvrec;
        g.nrec = nrec;
        g.wnrec = wnrec;
        g.mlines = mlines;
        g.mroot = mroot;
        return g;
    }
};


namespace trw {

// Records an error message for trw_last_error() and returns code.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int set_hip_error(hipError_t e, const char *what);

// trw_walk_async with an explicit work-queue head (one per concurrent QUEUE-form
// launch; null = the graph's own), control over building missing tables, and whether
// earlier work on the stream has already zeroed that head (head_zeroed).
int walk_launch(trw_graph *g, const trw_walk_params *params, const trw_walk_io *io, hipStream_t stream,
                unsigned int *queue_head, bool allow_table_build, bool head_zeroed = false);

// Validates a walk's parameters against the graph without launching or building
// anything (the tables it reads must already be built).
int walk_check(trw_graph *g, const trw_walk_params *params);

// Builds the per-walk records a walk would build on its first launch (MetaPath schema records),
// so that a walk launched without building (the autotuner's) reads what the real walk reads.
int prepare_records(trw_graph *g, const trw_walk_params *params);

}  // namespace trw

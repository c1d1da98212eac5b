// Autotuner: picks the per-graph launch knobs by timing candidates on a sample batch.
//
// The knobs change how walkers are laid onto lanes and how a Move searches, never
// which draws a walker consumes or what it compares, so every candidate produces the
// same walks; only device time differs.  Knobs are tuned one at a time (walkers per
// lane, the ITS tree threshold, the queue chunk and blocks per CU, then the launch shapes
// of the walk's Move: Node2Vec threads per block, fused threads per block, fused path
// tile), each with the others at their current best.  The tuned knobs all start from their
// defaults; the ITS guide switch is left as the caller set it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"
#include "internal.h"

namespace {

struct Sample {
    int32_t *sources = nullptr, *paths = nullptr, *lengths = nullptr;
    uint64_t *counters = nullptr;
    int32_t n = 0, stride = 0, stride_b = 0;  // mixed sample: each kind's rows at its own length
    ~Sample() {
        (void)hipFree(sources);
        (void)hipFree(paths);
        (void)hipFree(lengths);
        (void)hipFree(counters);
    }
};

// Device time of one walk over the sample (median of three after one warm-up); with pb, the
// sample's walkers [0, split) walk as pa and the rest as pb, in one mixed batch.
int time_walk(trw_graph *g, const trw_walk_params &p, const trw_walk_params *pb, int32_t split, Sample &s,
              double *steps_per_s) {
    trw_walk_io io;
    std::memset(&io, 0, sizeof(io));
    io.sources = s.sources;
    io.n_walkers = s.n;
    io.path_stride = s.stride;
    io.path_stride_b = pb ? s.stride_b : 0;
    io.paths = s.paths;
    io.lengths = s.lengths;
    io.counters = s.counters;
    hipStream_t st = g->ctx->stream;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return trw::set_error(TRW_E_HIP, "hipEventCreate");
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return trw::set_error(TRW_E_HIP, "hipEventCreate");
    }
    std::vector<double> ms;
    uint64_t steps = 0;
    int rc = TRW_OK;
    // At least three timed runs after a warm one, and more for a short walk (a 0.1 ms batch moves by
    // a few microseconds from run to run): until 20 ms have been timed, at most 64 runs.
    double timed = 0.0;
    for (int rep = 0; rc == TRW_OK && (rep < 4 || (timed < 20.0 && rep < 64)); ++rep) {
        if (trw::launch_fill_u32(reinterpret_cast<unsigned int *>(s.counters), 0u, 8, st) != hipSuccess) {
            rc = trw::set_error(TRW_E_HIP, "autotune counter reset");
            break;
        }
        (void)hipEventRecord(a, st);
        rc = pb ? trw_walk_mix_async(g, &p, pb, split, &io, st) : trw::walk_launch(g, &p, &io, st, nullptr, false);
        (void)hipEventRecord(b, st);
        if (rc != TRW_OK) break;
        if (hipStreamSynchronize(st) != hipSuccess) {
            rc = trw::set_error(TRW_E_HIP, "autotune walk");
            break;
        }
        float t = 0.f;
        (void)hipEventElapsedTime(&t, a, b);
        if (rep > 0) {
            ms.push_back(t);
            timed += t;
        }
        if (hipMemcpy(&steps, s.counters, sizeof(steps), hipMemcpyDeviceToHost) != hipSuccess)
            rc = trw::set_error(TRW_E_HIP, "autotune counters");
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != TRW_OK) return rc;
    std::sort(ms.begin(), ms.end());
    const double med = ms[ms.size() / 2];
    *steps_per_s = med > 0 ? static_cast<double>(steps) / (med * 1e-3) : 0.0;
    return TRW_OK;
}

}  // namespace

namespace {

// pb null: one kind; else a mixed batch, walkers [0, split) of kind params, the rest of pb
// (QUEUE form; the queue kernel of a mix takes the first kind's launch shapes).
int tune(trw_graph *g, const trw_walk_params *params, const trw_walk_params *pb, int32_t split,
         int32_t sample_walkers, uint64_t seed, int32_t with_paths, trw_tune_report *report) {
    if (!g || !params || !report) return trw::set_error(TRW_E_INVALID, "null argument");
    if (params->rng != TRW_RNG_PHILOX || (pb && pb->rng != TRW_RNG_PHILOX))
        return trw::set_error(TRW_E_INVALID, "tune with the Philox stream");
    if (sample_walkers < 1 || sample_walkers > (1 << 26))
        return trw::set_error(TRW_E_INVALID, "sample_walkers must be in [1, 2^26]");
    if (pb && (split < 0 || split > sample_walkers))
        return trw::set_error(TRW_E_INVALID, "split outside [0, sample_walkers]");
    std::memset(report, 0, sizeof(*report));
    if (hipSetDevice(g->ctx->device) != hipSuccess) return trw::set_error(TRW_E_HIP, "hipSetDevice");
    int rc = trw::walk_check(g, params);
    if (rc == TRW_OK && pb) rc = trw::walk_check(g, pb);
    if (rc == TRW_OK) rc = trw::prepare_records(g, params);
    if (rc == TRW_OK && pb) rc = trw::prepare_records(g, pb);
    if (rc != TRW_OK) return rc;

    // Sources: non-isolated vertices drawn by a seeded generator (splitmix64).
    const size_t V = static_cast<size_t>(g->nv);
    std::vector<int32_t> degree(V);
    if (V && hipMemcpy(degree.data(), g->degree, sizeof(int32_t) * V, hipMemcpyDeviceToHost) != hipSuccess)
        return trw::set_error(TRW_E_HIP, "autotune degree copy");
    std::vector<int32_t> live;
    for (size_t v = 0; v < V; ++v)
        if (degree[v] > 0) live.push_back(static_cast<int32_t>(v));
    if (live.empty()) return trw::set_error(TRW_E_INVALID, "the graph has no edges");
    std::vector<int32_t> src(static_cast<size_t>(sample_walkers));
    uint64_t x = seed;
    for (auto &v : src) {
        x += 0x9e3779b97f4a7c15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        v = live[z % live.size()];
    }
    Sample s;
    s.n = sample_walkers;
    // A mixed sample writes each kind's rows at its own length (io.path_stride_b), as the mixed
    // batches trw_walk_mix_async runs with such rows; a single kind, rows of its length.
    s.stride = with_paths ? std::max(params->length, 1) : 0;
    s.stride_b = with_paths && pb ? std::max(pb->length, 1) : 0;
    const size_t N = static_cast<size_t>(s.n);
    const size_t path_ints = pb ? static_cast<size_t>(split) * static_cast<size_t>(s.stride) +
                                      (N - static_cast<size_t>(split)) * static_cast<size_t>(s.stride_b)
                                : N * static_cast<size_t>(s.stride);
    if (hipMalloc(reinterpret_cast<void **>(&s.sources), sizeof(int32_t) * N) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&s.lengths), sizeof(int32_t) * N) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&s.counters), sizeof(uint64_t) * 4) != hipSuccess ||
        (s.stride && hipMalloc(reinterpret_cast<void **>(&s.paths), sizeof(int32_t) * path_ints) != hipSuccess))
        return trw::set_error(TRW_E_NOMEM, "autotune sample buffers");
    if (hipMemcpy(s.sources, src.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice) != hipSuccess)
        return trw::set_error(TRW_E_HIP, "autotune sources");

    // Knobs that apply to this walk, with their candidates.
    const bool ppr = params->app == TRW_APP_PPR;
    const bool metapath = params->app == TRW_APP_METAPATH;
    const bool node2vec = params->app == TRW_APP_NODE2VEC && params->execution != TRW_EXEC_UNIFORM;
    int32_t form = params->form;
    if (form == TRW_FORM_AUTO) form = (ppr || metapath || pb) ? TRW_FORM_QUEUE : TRW_FORM_FUSED;
    if (pb && form != TRW_FORM_QUEUE) return trw::set_error(TRW_E_UNSUPPORTED, "mixed batches tune in the QUEUE form");
    const bool fused_records = form == TRW_FORM_FUSED && params->walkers_per_thread == 0 &&
                               (params->execution == TRW_EXEC_UNIFORM ||
                                (params->execution != TRW_EXEC_DYNAMIC &&
                                 (params->sampling == TRW_SM_ALIAS || params->sampling == TRW_SM_ITS)));
    const bool its = params->execution != TRW_EXEC_UNIFORM && params->sampling == TRW_SM_ITS && !metapath;
    // The walk's Move (the shape knobs are kept per Move).
    const int32_t move = params->execution == TRW_EXEC_UNIFORM ? TRW_MOVE_UNIFORM
                         : node2vec                            ? TRW_MOVE_NODE2VEC
                         : metapath                            ? TRW_MOVE_METAPATH
                         : params->sampling == TRW_SM_ALIAS    ? TRW_MOVE_ALIAS
                         : params->sampling == TRW_SM_ITS      ? TRW_MOVE_ITS
                                                               : TRW_MOVE_REJECTION;
    struct Knob {
        int32_t id;
        std::vector<int64_t> values;
    };
    std::vector<Knob> knobs;
    if (fused_records) knobs.push_back({TRW_TUNE_WALKERS_PER_THREAD, {1, 2, 4}});
    if (its) knobs.push_back({TRW_TUNE_ITS_TREE_MIN_DEGREE, {0, 32, 128, 512, INT32_MAX}});
    // The QUEUE form's chunk and blocks per CU act together (a smaller grid wants other chunks), so
    // every pair of them is timed rather than one knob after the other.
    const std::vector<int64_t> queue_chunks = {0, 64, 256, 1024, 4096}, queue_bpcs = {0, 1, 2, 4};
    if (form == TRW_FORM_GRAPH) knobs.push_back({TRW_TUNE_GRAPH_BURST, {0, 4, 8, 32, 64}});
    if (form == TRW_FORM_FUSED) {
        if (node2vec) knobs.push_back({TRW_TUNE_N2V_BLOCK, {256, 128}});
        knobs.push_back({TRW_TUNE_FUSED_BLOCK, {0, 64, 128, 256}});
        if (with_paths && params->length > 16) knobs.push_back({TRW_TUNE_FUSED_TILE, {0, 16, 32, 64}});
    }
    const auto get = [&](int32_t id) -> int64_t & { return id < 4 ? g->tune[id] : g->shape[move][id - 4]; };

    int64_t saved[TRW_TUNE_KNOBS];
    for (int k = 0; k < TRW_TUNE_KNOBS; ++k) saved[k] = get(k);
    const auto restore = [&] {
        for (int k = 0; k < TRW_TUNE_KNOBS; ++k) get(k) = saved[k];
    };
    for (int k = 0; k < TRW_TUNE_KNOBS; ++k)
        if (k != TRW_TUNE_ITS_GUIDE_OFF) get(k) = 0;  // the guide switch is the caller's, not tuned
    rc = time_walk(g, *params, pb, split, s, &report->default_steps_per_s);
    if (rc != TRW_OK) {
        restore();
        return rc;
    }
    report->move = move;
    double best = report->default_steps_per_s;
    if (form == TRW_FORM_QUEUE) {
        int64_t best_c = get(TRW_TUNE_QUEUE_CHUNK), best_b = get(TRW_TUNE_QUEUE_BLOCKS_PER_CU);
        for (int64_t b : queue_bpcs) {
            for (int64_t c : queue_chunks) {
                get(TRW_TUNE_QUEUE_CHUNK) = c;
                get(TRW_TUNE_QUEUE_BLOCKS_PER_CU) = b;
                double sps = 0.0;
                rc = time_walk(g, *params, pb, split, s, &sps);
                if (rc != TRW_OK) {
                    restore();
                    return rc;
                }
                if (sps > best * 1.01) {  // keep the earlier (default-first) pair unless clearly faster
                    best = sps;
                    best_c = c;
                    best_b = b;
                }
            }
        }
        get(TRW_TUNE_QUEUE_CHUNK) = best_c;
        get(TRW_TUNE_QUEUE_BLOCKS_PER_CU) = best_b;
        const int64_t chosen[2] = {best_c, best_b};
        const int32_t ids[2] = {TRW_TUNE_QUEUE_CHUNK, TRW_TUNE_QUEUE_BLOCKS_PER_CU};
        const int32_t counts[2] = {static_cast<int32_t>(queue_chunks.size()), static_cast<int32_t>(queue_bpcs.size())};
        for (int k = 0; k < 2; ++k) {
            const int i = report->knobs++;
            report->knob[i] = ids[k];
            report->chosen[i] = chosen[k];
            report->candidates[i] = counts[k];
        }
    }
    for (const Knob &kn : knobs) {
        int64_t best_v = get(kn.id);
        for (int64_t v : kn.values) {
            get(kn.id) = v;
            double sps = 0.0;
            rc = time_walk(g, *params, pb, split, s, &sps);
            if (rc != TRW_OK) {
                restore();
                return rc;
            }
            if (sps > best * 1.01) {  // keep the earlier (default-first) value unless clearly faster
                best = sps;
                best_v = v;
            }
        }
        get(kn.id) = best_v;
        const int i = report->knobs++;
        report->knob[i] = kn.id;
        report->chosen[i] = best_v;
        report->candidates[i] = static_cast<int32_t>(kn.values.size());
    }
    report->tuned_steps_per_s = best;
    return TRW_OK;
}

}  // namespace

int trw_autotune(trw_graph *g, const trw_walk_params *params, int32_t sample_walkers, uint64_t seed,
                 int32_t with_paths, trw_tune_report *report) {
    return tune(g, params, nullptr, 0, sample_walkers, seed, with_paths, report);
}

int trw_autotune_mix(trw_graph *g, const trw_walk_params *params_a, const trw_walk_params *params_b, int32_t split,
                     int32_t sample_walkers, uint64_t seed, int32_t with_paths, trw_tune_report *report) {
    if (!params_b) return trw::set_error(TRW_E_INVALID, "null argument");
    return tune(g, params_a, params_b, split, sample_walkers, seed, with_paths, report);
}

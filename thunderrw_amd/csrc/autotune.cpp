// Autotuner: picks the per-graph launch knobs by timing candidates on a sample batch.
//
// The knobs change how walkers are laid onto lanes and how a Move searches, never
// which draws a walker consumes or what it compares, so every candidate produces the
// same walks; only device time differs.  Knobs are tuned one at a time (walkers per
// lane, then the ITS tree threshold, then the queue chunk), each with the others at
// their current best.
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
    int32_t n = 0, stride = 0;
    ~Sample() {
        (void)hipFree(sources);
        (void)hipFree(paths);
        (void)hipFree(lengths);
        (void)hipFree(counters);
    }
};

// Device time of one walk over the sample (median of three after one warm-up).
int time_walk(trw_graph *g, const trw_walk_params &p, Sample &s, double *steps_per_s) {
    trw_walk_io io;
    std::memset(&io, 0, sizeof(io));
    io.sources = s.sources;
    io.n_walkers = s.n;
    io.path_stride = s.stride;
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
    for (int rep = 0; rep < 4 && rc == TRW_OK; ++rep) {
        if (trw::launch_fill_u32(reinterpret_cast<unsigned int *>(s.counters), 0u, 8, st) != hipSuccess) {
            rc = trw::set_error(TRW_E_HIP, "autotune counter reset");
            break;
        }
        (void)hipEventRecord(a, st);
        rc = trw::walk_launch(g, &p, &io, st, nullptr, false);
        (void)hipEventRecord(b, st);
        if (rc != TRW_OK) break;
        if (hipStreamSynchronize(st) != hipSuccess) {
            rc = trw::set_error(TRW_E_HIP, "autotune walk");
            break;
        }
        float t = 0.f;
        (void)hipEventElapsedTime(&t, a, b);
        if (rep > 0) ms.push_back(t);
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

int trw_autotune(trw_graph *g, const trw_walk_params *params, int32_t sample_walkers, uint64_t seed,
                 int32_t with_paths, trw_tune_report *report) {
    if (!g || !params || !report) return trw::set_error(TRW_E_INVALID, "null argument");
    if (params->rng != TRW_RNG_PHILOX) return trw::set_error(TRW_E_INVALID, "tune with the Philox stream");
    if (sample_walkers < 1 || sample_walkers > (1 << 26))
        return trw::set_error(TRW_E_INVALID, "sample_walkers must be in [1, 2^26]");
    std::memset(report, 0, sizeof(*report));
    if (hipSetDevice(g->ctx->device) != hipSuccess) return trw::set_error(TRW_E_HIP, "hipSetDevice");
    int rc = trw::walk_check(g, params);
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
    s.stride = with_paths ? std::max(params->length, 1) : 0;
    const size_t N = static_cast<size_t>(s.n);
    if (hipMalloc(reinterpret_cast<void **>(&s.sources), sizeof(int32_t) * N) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&s.lengths), sizeof(int32_t) * N) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&s.counters), sizeof(uint64_t) * 4) != hipSuccess ||
        (s.stride && hipMalloc(reinterpret_cast<void **>(&s.paths), sizeof(int32_t) * N * s.stride) != hipSuccess))
        return trw::set_error(TRW_E_NOMEM, "autotune sample buffers");
    if (hipMemcpy(s.sources, src.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice) != hipSuccess)
        return trw::set_error(TRW_E_HIP, "autotune sources");

    // Knobs that apply to this walk, with their candidates.
    const bool ppr = params->app == TRW_APP_PPR;
    const bool metapath = params->app == TRW_APP_METAPATH;
    int32_t form = params->form;
    if (form == TRW_FORM_AUTO) form = (ppr || metapath) ? TRW_FORM_QUEUE : TRW_FORM_FUSED;
    const bool fused_records = form == TRW_FORM_FUSED && params->walkers_per_thread == 0 &&
                               (params->execution == TRW_EXEC_UNIFORM ||
                                (params->execution != TRW_EXEC_DYNAMIC &&
                                 (params->sampling == TRW_SM_ALIAS || params->sampling == TRW_SM_ITS)));
    const bool its = params->execution != TRW_EXEC_UNIFORM && params->sampling == TRW_SM_ITS && !metapath;
    struct Knob {
        int32_t id;
        std::vector<int64_t> values;
    };
    std::vector<Knob> knobs;
    if (fused_records) knobs.push_back({TRW_TUNE_WALKERS_PER_THREAD, {1, 2, 4}});
    if (its) knobs.push_back({TRW_TUNE_ITS_TREE_MIN_DEGREE, {0, 32, 128, 512, INT32_MAX}});
    if (form == TRW_FORM_QUEUE) knobs.push_back({TRW_TUNE_QUEUE_CHUNK, {0, 64, 256, 1024, 4096}});

    const int64_t saved[3] = {g->tune[0], g->tune[1], g->tune[2]};
    for (int k = 0; k < 3; ++k) g->tune[k] = 0;
    rc = time_walk(g, *params, s, &report->default_steps_per_s);
    if (rc != TRW_OK) {
        std::memcpy(g->tune, saved, sizeof(saved));
        return rc;
    }
    double best = report->default_steps_per_s;
    for (const Knob &kn : knobs) {
        int64_t best_v = g->tune[kn.id];
        for (int64_t v : kn.values) {
            g->tune[kn.id] = v;
            double sps = 0.0;
            rc = time_walk(g, *params, s, &sps);
            if (rc != TRW_OK) {
                std::memcpy(g->tune, saved, sizeof(saved));
                return rc;
            }
            if (sps > best * 1.01) {  // keep the earlier (default-first) value unless clearly faster
                best = sps;
                best_v = v;
            }
        }
        g->tune[kn.id] = best_v;
        const int i = report->knobs++;
        report->knob[i] = kn.id;
        report->chosen[i] = best_v;
        report->candidates[i] = static_cast<int32_t>(kn.values.size());
    }
    report->tuned_steps_per_s = best;
    return TRW_OK;
}

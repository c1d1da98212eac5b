// C ABI of the step engine (include/thunderrw.h).
#include "../../include/thunderrw.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "engine.h"
#include "hostgraph.h"
#include "internal.h"
#include "sfmt.h"

using trw::AliasSlot;
using trw::BatchIO;
using trw::DevGraph;
using trw::KindParams;
using trw::WalkerState;

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? TRW_E_NOMEM : TRW_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define TRW_HIP(call, what)                       \
    do {                                          \
        hipError_t e_ = (call);                   \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

template <class T>
hipError_t dmalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count);
}

// Device allocations freed together when the owner goes out of scope.
struct ScratchBuffers {
    std::vector<void *> ptrs;
    ~ScratchBuffers() {
        for (void *q : ptrs) (void)hipFree(q);
    }
    template <class T>
    hipError_t get(T **p, size_t count) {
        *p = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// trw_walk_noalloc's scratch memory, a file-scope device array so that the entry point never
// allocates: the first kNoAllocControlBytes hold the QUEUE form's work-queue head (word 0) and
// the STEP form's live flag (word 1), the rest the STEP form's walker state for at most
// kNoAllocMaxStepWalkers walkers.  Every device the library runs on holds its own copy until
// the process ends.  In host code `scratch` is a stand-in, not a device address: host code
// reaches the array only through the per-device address scratch_address() resolves.
constexpr int32_t kNoAllocMaxStepWalkers = 1 << 18;
constexpr size_t kNoAllocControlBytes = 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocControlBytes + sizeof(WalkerState) * kNoAllocMaxStepWalkers;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device: its copy of `scratch`, resolved once

// The address of `scratch` on `device`, which must be the current device.
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr.size() <= static_cast<size_t>(device)) g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
    if (!g_scratch_ptr[static_cast<size_t>(device)]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[static_cast<size_t>(device)] = p;
    }
    *out = g_scratch_ptr[static_cast<size_t>(device)];
    return hipSuccess;
}

}  // namespace

struct trw_host_graph {
    trw::HostGraph g;
};

int trw::set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int trw::set_hip_error(hipError_t e, const char *what) { return hip_fail(e, what); }

// The entry points take their C linkage from the declarations in thunderrw.h.

const char *trw_last_error(void) { return g_error.c_str(); }

int trw_version(int32_t *major, int32_t *minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return TRW_OK;
}

int trw_device_count(int32_t *count) {
    if (!count) return fail(TRW_E_INVALID, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return TRW_OK;
}

int trw_context_create(int32_t device, trw_context **out) {
    if (!out) return fail(TRW_E_INVALID, "out is null");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(TRW_E_HIP, "no HIP device is visible");
    if (device < 0 || device >= n) return fail(TRW_E_INVALID, "device %d outside [0, %d)", device, n);
    TRW_HIP(hipSetDevice(device), "hipSetDevice");
    std::unique_ptr<trw_context> ctx(new trw_context);
    ctx->device = device;
    TRW_HIP(scratch_address(device, &ctx->scratch), "hipGetSymbolAddress (trw_walk_noalloc scratch)");
    TRW_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    *out = ctx.release();
    return TRW_OK;
}

int trw_context_destroy(trw_context *ctx) {
    if (!ctx) return TRW_OK;
    (void)hipSetDevice(ctx->device);
    for (hipEvent_t &e : ctx->timer)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TRW_OK;
}

static hipStream_t pick_stream(trw_context *ctx, void *stream) {
    return stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
}

int trw_device_alloc(trw_context *ctx, uint64_t bytes, void **out) {
    if (!ctx || !out) return fail(TRW_E_INVALID, "null argument");
    *out = nullptr;
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(hipMalloc(out, bytes ? bytes : 1), "hipMalloc");
    return TRW_OK;
}

int trw_device_free(trw_context *ctx, void *ptr) {
    if (!ctx) return fail(TRW_E_INVALID, "null argument");
    if (!ptr) return TRW_OK;
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
    TRW_HIP(hipFree(ptr), "hipFree");
    return TRW_OK;
}

int trw_copy_to_device(trw_context *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TRW_E_INVALID, "null argument");
    if (!bytes) return TRW_OK;
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "copy to device");
    TRW_HIP(hipStreamSynchronize(ctx->stream), "copy to device");
    return TRW_OK;
}

int trw_copy_to_host(trw_context *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TRW_E_INVALID, "null argument");
    if (!bytes) return TRW_OK;
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
    TRW_HIP(hipStreamSynchronize(ctx->stream), "copy to host");
    return TRW_OK;
}

int trw_device_zero(trw_context *ctx, void *ptr, uint64_t bytes, void *stream) {
    if (!ctx || (bytes && !ptr)) return fail(TRW_E_INVALID, "null argument");
    if (bytes % 4) return fail(TRW_E_INVALID, "byte count must be a multiple of 4");
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(trw::launch_fill_u32(static_cast<unsigned int *>(ptr), 0u, static_cast<int64_t>(bytes / 4),
                                 pick_stream(ctx, stream)), "zero fill");
    return TRW_OK;
}

int trw_device_copy(trw_context *ctx, void *dst, const void *src, uint64_t bytes, void *stream) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TRW_E_INVALID, "null argument");
    if (bytes % 16) return fail(TRW_E_INVALID, "byte count must be a multiple of 16");
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(trw::launch_copy_u128(static_cast<uint4 *>(dst), static_cast<const uint4 *>(src),
                                  static_cast<int64_t>(bytes / 16), pick_stream(ctx, stream)), "copy kernel");
    return TRW_OK;
}

int trw_probe_random_gather(trw_context *ctx, uint64_t table_bytes, int32_t record_bytes, int32_t chains,
                            int32_t steps, double *gathers_per_s) {
    if (!ctx || !gathers_per_s || steps <= 0) return fail(TRW_E_INVALID, "bad argument");
    if (record_bytes != 16 && record_bytes != 32) return fail(TRW_E_INVALID, "record_bytes must be 16 or 32");
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    int cus = 0;
    TRW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "device attribute");
    TRW_HIP(trw::run_gather_probe(table_bytes, record_bytes, chains, steps, cus, ctx->stream, gathers_per_s),
            "gather probe");
    return TRW_OK;
}

int trw_synchronize(trw_context *ctx, void *stream) {
    if (!ctx) return fail(TRW_E_INVALID, "null argument");
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(hipStreamSynchronize(pick_stream(ctx, stream)), "synchronize");
    return TRW_OK;
}

int trw_timer_begin(trw_context *ctx, void *stream) {
    if (!ctx) return fail(TRW_E_INVALID, "null argument");
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    for (hipEvent_t &e : ctx->timer)
        if (!e) TRW_HIP(hipEventCreate(&e), "hipEventCreate");
    TRW_HIP(hipEventRecord(ctx->timer[0], pick_stream(ctx, stream)), "hipEventRecord");
    return TRW_OK;
}

int trw_timer_end(trw_context *ctx, void *stream, double *elapsed_ms) {
    if (!ctx || !elapsed_ms) return fail(TRW_E_INVALID, "null argument");
    if (!ctx->timer[0]) return fail(TRW_E_STATE, "trw_timer_begin was not called");
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    TRW_HIP(hipEventRecord(ctx->timer[1], pick_stream(ctx, stream)), "hipEventRecord");
    TRW_HIP(hipEventSynchronize(ctx->timer[1]), "hipEventSynchronize");
    float ms = 0.f;
    TRW_HIP(hipEventElapsedTime(&ms, ctx->timer[0], ctx->timer[1]), "hipEventElapsedTime");
    *elapsed_ms = ms;
    return TRW_OK;
}

int trw_context_stream(trw_context *ctx, void **stream) {
    if (!ctx || !stream) return fail(TRW_E_INVALID, "null argument");
    *stream = reinterpret_cast<void *>(ctx->stream);
    return TRW_OK;
}

void release_graph_chunk(trw_graph *g);

static void graph_release(trw_graph *g) {
    dfree(g->offset);
    dfree(g->adj);
    dfree(g->degree);
    dfree(g->weight);
    dfree(g->label);
    dfree(g->lbl_sorted);
    dfree(g->lbl_pos);
    dfree(g->lstart);
    dfree(g->lerec);
    dfree(g->msrec);
    dfree(g->its_lines);
    dfree(g->its_root);
    dfree(g->grec);
    dfree(g->gvrec);
    dfree(g->mlines);
    dfree(g->nrec);
    dfree(g->wnrec);
    dfree(g->mroot);
    dfree(g->order);
    dfree(g->vrec);
    dfree(g->erec);
    dfree(g->arec);
    dfree(g->alias);
    dfree(g->prefix);
    dfree(g->rs_max);
    dfree(g->rs_custom);
    dfree(g->rrec);
    dfree(g->counters);
    dfree(g->state);
    dfree(g->active);
    release_graph_chunk(g);
}

int trw_graph_upload(trw_context *ctx, int32_t nv, int64_t ne, const int32_t *degree, const int32_t *adj,
                     const double *edge_weight, const int32_t *edge_label, trw_graph **out) {
    if (!ctx || !out || !degree || (ne > 0 && !adj)) return fail(TRW_E_INVALID, "null argument");
    *out = nullptr;
    if (nv <= 0 || ne < 0) return fail(TRW_E_INVALID, "num_vertices must be > 0 and num_edges >= 0");
    std::vector<int64_t> offset(static_cast<size_t>(nv) + 1, 0);
    int32_t maxd = 0;
    for (int32_t v = 0; v < nv; ++v) {
        if (degree[v] < 0) return fail(TRW_E_INVALID, "degree[%d] is negative", v);
        offset[v + 1] = offset[v] + degree[v];
        maxd = std::max(maxd, degree[v]);
    }
    if (offset[nv] != ne) return fail(TRW_E_INVALID, "degrees sum to %lld, num_edges is %lld",
                                      static_cast<long long>(offset[nv]), static_cast<long long>(ne));
    std::atomic<bool> bad{false};
    trw::parallel_for(ne, [&](int64_t b, int64_t e) {
        bool mine = false;
        for (int64_t i = b; i < e; ++i) mine = mine || adj[i] < 0 || adj[i] >= nv;
        if (mine) bad.store(true, std::memory_order_relaxed);
    });
    if (bad.load()) return fail(TRW_E_INVALID, "adjacency holds a vertex id outside [0, num_vertices)");

    // Vertices in descending degree order (by power-of-two class): the table
    // builds start hubs' long serial chains first.
    std::vector<int32_t> order(static_cast<size_t>(nv));
    int64_t hubs = 0;
    {
        std::vector<int64_t> cls(34, 0);
        auto klass = [](int32_t d) { return d == 0 ? 0 : 33 - __builtin_clz(static_cast<unsigned>(d)); };
        for (int32_t v = 0; v < nv; ++v) cls[klass(degree[v])]++;
        std::vector<int64_t> start(34, 0);
        int64_t acc = 0;
        for (int c = 33; c >= 0; --c) {
            start[c] = acc;
            acc += cls[c];
        }
        for (int32_t v = 0; v < nv; ++v) order[start[klass(degree[v])]++] = v;
        for (int c = klass(trw::kWaveBuildMin); c <= 33; ++c) hubs += cls[c];
    }

    std::unique_ptr<trw_graph> g(new trw_graph);
    g->ctx = ctx;
    g->nv = nv;
    g->ne = ne;
    g->max_degree = maxd;
    g->build_hubs = static_cast<int32_t>(hubs);
    TRW_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    struct Guard {
        trw_graph *g;
        ~Guard() {
            if (g) graph_release(g);
        }
    } guard{g.get()};
    const size_t V = static_cast<size_t>(nv), E = static_cast<size_t>(ne);
    TRW_HIP(dmalloc(&g->offset, V + 1), "hipMalloc offset");
    TRW_HIP(dmalloc(&g->degree, V), "hipMalloc degree");
    TRW_HIP(dmalloc(&g->order, V), "hipMalloc order");
    TRW_HIP(dmalloc(&g->adj, E + 16), "hipMalloc adj");
    TRW_HIP(dmalloc(&g->counters, 8), "hipMalloc counters");
    TRW_HIP(hipMemcpy(g->offset, offset.data(), sizeof(int64_t) * (V + 1), hipMemcpyHostToDevice), "copy offset");
    TRW_HIP(hipMemcpy(g->degree, degree, sizeof(int32_t) * V, hipMemcpyHostToDevice), "copy degree");
    TRW_HIP(hipMemcpy(g->order, order.data(), sizeof(int32_t) * V, hipMemcpyHostToDevice), "copy order");
    if (E) TRW_HIP(hipMemcpy(g->adj, adj, sizeof(int32_t) * E, hipMemcpyHostToDevice), "copy adj");
    if (ne < (int64_t(1) << 32)) {
        // Vertex records {first edge, degree}: one 8-byte gather per step instead of two offsets.
        std::vector<uint2> rec(V);
        for (size_t v = 0; v < V; ++v) rec[v] = make_uint2(static_cast<unsigned>(offset[v]), static_cast<unsigned>(degree[v]));
        TRW_HIP(dmalloc(&g->vrec, V), "hipMalloc vertex records");
        TRW_HIP(hipMemcpy(g->vrec, rec.data(), sizeof(uint2) * V, hipMemcpyHostToDevice), "copy vertex records");
        // Neighbour records {adj[e], its first edge, its degree}: a Move's read of the
        // chosen entry also yields the next vertex's range (one gather per step).
        TRW_HIP(dmalloc(&g->erec, E + 1), "hipMalloc neighbour records");
        TRW_HIP(trw::launch_edge_records(g->view(), g->erec, ctx->stream), "neighbour records");
        TRW_HIP(hipStreamSynchronize(ctx->stream), "neighbour records");
    }
    if (edge_weight) {
        for (size_t e = 0; e < E && g->weights_nonneg; ++e)
            g->weights_nonneg = edge_weight[e] >= 0.0 && edge_weight[e] < HUGE_VAL;  // NaN fails both
        for (size_t e = 0; e < E; ++e)
            if (edge_weight[e] > g->max_weight) g->max_weight = edge_weight[e];
        TRW_HIP(dmalloc(&g->weight, E + 16), "hipMalloc weight");
        if (E) TRW_HIP(hipMemcpy(g->weight, edge_weight, sizeof(double) * E, hipMemcpyHostToDevice), "copy weight");
    }
    if (edge_label) {
        std::vector<int32_t> ls(E), lp(E);
        trw::build_label_index(nv, offset.data(), edge_label, ls.data(), lp.data());
        TRW_HIP(dmalloc(&g->label, E + 16), "hipMalloc label");
        TRW_HIP(dmalloc(&g->lbl_sorted, E + 16), "hipMalloc label index");
        TRW_HIP(dmalloc(&g->lbl_pos, E + 16), "hipMalloc label index");
        if (E) {
            TRW_HIP(hipMemcpy(g->label, edge_label, sizeof(int32_t) * E, hipMemcpyHostToDevice), "copy label");
            TRW_HIP(hipMemcpy(g->lbl_sorted, ls.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice), "copy label index");
            TRW_HIP(hipMemcpy(g->lbl_pos, lp.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice), "copy label index");
        }
        // Small label alphabets (all labels in [0, 8)): per-vertex label starts and label-ordered
        // neighbour records carrying the neighbour's label starts make a MetaPath move one gather.
        bool small = g->erec != nullptr;
        for (size_t e = 0; e < E && small; ++e) small = edge_label[e] >= 0 && edge_label[e] < 8;
        if (small) {
            TRW_HIP(dmalloc(&g->lstart, static_cast<size_t>(trw::kLabelRecWords) * (V + 1)), "hipMalloc label records");
            TRW_HIP(dmalloc(&g->lerec, 4 * E + 4), "hipMalloc label records");
            TRW_HIP(trw::launch_label_records(g->view(), g->lstart, g->lerec, ctx->stream), "label records");
            TRW_HIP(hipStreamSynchronize(ctx->stream), "label records");
        }
    }
    guard.g = nullptr;
    *out = g.release();
    return TRW_OK;
}

int trw_graph_free(trw_graph *g) {
    if (!g) return TRW_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    graph_release(g);
    delete g;
    return TRW_OK;
}

int trw_graph_info(const trw_graph *g, int32_t *nv, int64_t *ne, int32_t *max_degree) {
    if (!g) return fail(TRW_E_INVALID, "graph is null");
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    if (max_degree) *max_degree = g->max_degree;
    return TRW_OK;
}

// ITS search structures over every vertex's prefix sums (128-byte lines: a 16-key B+-tree
// over leaves of 5 prefix sums and their neighbour records; see its_levels).  The tree's
// upper bound equals the binary search's only for non-decreasing sums, so weights with a
// negative or NaN entry get none; the leaves carry neighbour records, so graphs without
// them (2^32 edges or more) get none either.
static int build_its_trees(trw_graph *g, bool monotone) {
    dfree(g->its_lines);
    if (!monotone || g->nv == 0 || !g->erec) return TRW_OK;
    hipStream_t st = g->ctx->stream;
    const size_t V = static_cast<size_t>(g->nv);
    if (!g->its_root) TRW_HIP(dmalloc(&g->its_root, V), "hipMalloc ITS structure roots");
    // Line starts by a device scan of the per-vertex line counts; only the total comes back.
    ScratchBuffers sc;
    uint64_t *start = nullptr;
    uint8_t *temp = nullptr;
    size_t temp_bytes = 0;
    TRW_HIP(trw::launch_its_tree_starts(g->view(), g->its_root, nullptr, nullptr, temp_bytes, st), "ITS structure scan size");
    TRW_HIP(sc.get(&start, V + 1), "hipMalloc ITS structure starts");
    TRW_HIP(sc.get(&temp, temp_bytes), "hipMalloc ITS scan storage");
    TRW_HIP(trw::launch_its_tree_starts(g->view(), g->its_root, start, temp, temp_bytes, st), "ITS structure starts");
    uint64_t total = 0;
    TRW_HIP(hipMemcpyAsync(&total, start + V, sizeof(total), hipMemcpyDeviceToHost, st), "ITS structure size");
    TRW_HIP(hipStreamSynchronize(st), "ITS structure starts");
    if (total == 0 || total >= trw::kNoRange) return TRW_OK;  // nothing to index, or line indices overflow 32 bits
    TRW_HIP(dmalloc(&g->its_lines, static_cast<size_t>(total) * 8), "hipMalloc ITS structures");
    TRW_HIP(trw::launch_its_tree_fill(g->view(), g->prefix, start, total, g->its_root, g->its_lines, st),
            "ITS structure fill");
    TRW_HIP(trw::launch_its_tree_roots_to_records(g->view(), g->its_root, g->erec, st), "ITS structure roots");
    TRW_HIP(hipStreamSynchronize(st), "ITS structures");  // the scratch is freed on return
    return TRW_OK;
}

// ITS guide records over every vertex's prefix sums (one gather per static ITS step; the
// record layout is described with DevGraph).  Like the search structures they reproduce the
// upper bound only for non-decreasing sums, and they carry neighbours' bucket ranges in 32
// bits (a walker holds a bucket count as a positive int32), so graphs whose bucket total
// reaches 2^32 - 1 or whose degree reaches 2^29 get none.
static int build_its_guide(trw_graph *g, bool monotone) {
    dfree(g->grec);
    if (!monotone || g->nv == 0 || !g->vrec || g->max_degree >= (1 << 29)) return TRW_OK;
    hipStream_t st = g->ctx->stream;
    const size_t V = static_cast<size_t>(g->nv);
    ScratchBuffers sc;
    uint32_t *flag = nullptr, *count = nullptr;
    uint64_t *start = nullptr;
    uint8_t *temp = nullptr;
    size_t temp_bytes = 0;
    TRW_HIP(trw::launch_its_guide_starts(g->view(), g->prefix, nullptr, nullptr, nullptr, nullptr, temp_bytes, st),
            "ITS guide scan size");
    TRW_HIP(sc.get(&flag, V), "hipMalloc ITS guide flags");
    TRW_HIP(sc.get(&count, V), "hipMalloc ITS guide counts");
    TRW_HIP(sc.get(&start, V + 1), "hipMalloc ITS guide starts");
    TRW_HIP(sc.get(&temp, temp_bytes), "hipMalloc ITS guide scan storage");
    TRW_HIP(trw::launch_its_guide_starts(g->view(), g->prefix, flag, count, start, temp, temp_bytes, st),
            "ITS guide starts");
    uint64_t total = 0;
    TRW_HIP(hipMemcpyAsync(&total, start + V, sizeof(total), hipMemcpyDeviceToHost, st), "ITS guide size");
    TRW_HIP(hipStreamSynchronize(st), "ITS guide starts");
    if (total >= trw::kNoRange) return TRW_OK;
    if (!g->gvrec) TRW_HIP(dmalloc(&g->gvrec, V), "hipMalloc ITS guide ranges");
    // The records only speed the walk up (about 80 bytes per edge at U[1,5) weights): without
    // room for them the table stands and the fused walk searches it, as the other forms do.
    const hipError_t ae = dmalloc(&g->grec, 2 * static_cast<size_t>(total) + 2);
    if (ae == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        g->grec = nullptr;
        return TRW_OK;
    }
    TRW_HIP(ae, "hipMalloc ITS guide records");
    TRW_HIP(trw::launch_its_guide_fill(g->view(), g->prefix, start, total, g->gvrec, g->grec, st), "ITS guide fill");
    TRW_HIP(hipStreamSynchronize(st), "ITS guide records");  // the scratch is freed on return
    return TRW_OK;
}

// static_initialization's table build (creeper.h:102-175) from per-edge weights w (device,
// ignored when unit): `weights` is the TRW_WEIGHTS_* tag recorded for the table, `monotone`
// says the weights are all finite and >= 0 (ITS search structures are built only then).
static int build_tables(trw_graph *g, int32_t sampling, const double *w, int32_t unit, int32_t weights,
                        bool monotone) {
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    hipStream_t st = g->ctx->stream;
    const DevGraph dv = g->view();
    const size_t E = static_cast<size_t>(g->ne), V = static_cast<size_t>(g->nv);
    // The table is undefined until this build succeeds.
    (sampling == TRW_SM_ALIAS ? g->alias_weights : sampling == TRW_SM_ITS ? g->its_weights : g->rs_weights) = -1;
    hipEvent_t t0, t1;
    TRW_HIP(hipEventCreate(&t0), "hipEventCreate");
    TRW_HIP(hipEventCreate(&t1), "hipEventCreate");
    struct Ev {
        hipEvent_t a, b;
        ~Ev() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{t0, t1};
    if (sampling == TRW_SM_ALIAS) {
        if (!g->alias) TRW_HIP(dmalloc(&g->alias, E + 1), "hipMalloc alias table");
        trw::AliasEntry *scratch = nullptr;
        TRW_HIP(dmalloc(&scratch, E + 1), "hipMalloc alias scratch");
        {
            const hipError_t ze = trw::launch_fill_u32(g->counters + 2, 0u, 1, st);
            if (ze != hipSuccess) {
                dfree(scratch);
                return fail(TRW_E_HIP, "clear the zero-sum count: %s", hipGetErrorString(ze));
            }
        }
        (void)hipEventRecord(t0, st);
        hipError_t e = trw::launch_alias_build(dv, w, unit, g->order, g->build_hubs, scratch, g->alias,
                                              g->counters + 2, st);
        (void)hipEventRecord(t1, st);
        hipError_t e2 = hipStreamSynchronize(st);
        dfree(scratch);
        if (e != hipSuccess) return hip_fail(e, "alias build launch");
        if (e2 != hipSuccess) return hip_fail(e2, "alias build");
        g->alias_weights = weights;
        if (g->vrec) {
            // Alias walk records: the slot plus both candidates' ranges (32 bytes per slot).
            if (!g->arec) TRW_HIP(dmalloc(&g->arec, 2 * E + 2), "hipMalloc alias walk records");
            TRW_HIP(trw::launch_alias_records(g->view(), g->arec, st), "alias walk records");
            TRW_HIP(hipStreamSynchronize(st), "alias walk records");
        }
    } else if (sampling == TRW_SM_ITS) {
        // Padded so that a search may read the whole aligned 16-entry chunk holding any entry.
        if (!g->prefix) TRW_HIP(dmalloc(&g->prefix, E + 2 * trw::kItsFan), "hipMalloc prefix sums");
        (void)hipEventRecord(t0, st);
        hipError_t e = trw::launch_its_build(dv, w, unit, g->order, g->build_hubs, g->prefix, st);
        if (e != hipSuccess) return hip_fail(e, "ITS build launch");
        int rc = build_its_trees(g, unit || monotone);
        if (rc == TRW_OK) rc = build_its_guide(g, unit || monotone);
        (void)hipEventRecord(t1, st);
        if (rc != TRW_OK) return rc;
        TRW_HIP(hipStreamSynchronize(st), "ITS build");
        g->its_weights = weights;
    } else {
        if (!g->rs_max) TRW_HIP(dmalloc(&g->rs_max, V), "hipMalloc rejection maxima");
        (void)hipEventRecord(t0, st);
        hipError_t e = trw::launch_rs_build(dv, w, unit, g->order, g->build_hubs, g->rs_max, st);
        (void)hipEventRecord(t1, st);
        if (e != hipSuccess) return hip_fail(e, "rejection build launch");
        TRW_HIP(hipStreamSynchronize(st), "rejection build");
        if (g->vrec) {
            // Rejection walk records: weight, neighbour range and neighbour maximum (32 bytes per edge).
            if (!g->rrec) TRW_HIP(dmalloc(&g->rrec, 2 * E + 2), "hipMalloc rejection walk records");
            TRW_HIP(trw::launch_rs_records(g->view(), w, unit, g->rrec, st), "rejection walk records");
            TRW_HIP(hipStreamSynchronize(st), "rejection walk records");
        }
        g->rs_weights = weights;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess) g->last_build_ms = ms;
    return TRW_OK;
}

// Membership tables (hash sets in 128-byte buckets) over every adjacency of degree > kMemberMin,
// and neighbour records carrying each neighbour's root, for Node2Vec's is_neighbor tests (built
// by the first Node2Vec walk; a graph without neighbour records, 2^32 edges or more, keeps the
// binary search).
static int build_member_tables(trw_graph *g) {
    if (g->nrec || g->nv == 0 || !g->erec) return TRW_OK;
    hipStream_t st = g->ctx->stream;
    const size_t V = static_cast<size_t>(g->nv);
    ScratchBuffers sc;
    uint64_t *start = nullptr;
    uint8_t *temp = nullptr;
    size_t temp_bytes = 0;
    uint32_t *root = nullptr;
    uint4 *lines = nullptr, *nrec = nullptr;
    struct Owned {  // freed unless handed to the graph
        uint32_t *&r;
        uint4 *&l, *&n;
        ~Owned() {
            dfree(r);
            dfree(l);
            dfree(n);
        }
    } owned{root, lines, nrec};
    TRW_HIP(dmalloc(&root, V), "hipMalloc membership roots");
    TRW_HIP(trw::launch_member_table_starts(g->view(), root, nullptr, nullptr, temp_bytes, st), "membership scan size");
    TRW_HIP(sc.get(&start, V + 1), "hipMalloc membership starts");
    TRW_HIP(sc.get(&temp, temp_bytes), "hipMalloc membership scan storage");
    TRW_HIP(trw::launch_member_table_starts(g->view(), root, start, temp, temp_bytes, st), "membership starts");
    uint64_t total = 0;
    TRW_HIP(hipMemcpyAsync(&total, start + V, sizeof(total), hipMemcpyDeviceToHost, st), "membership size");
    TRW_HIP(hipStreamSynchronize(st), "membership starts");
    if (total >= trw::kNoRange) return TRW_OK;  // line indices would overflow 32 bits: binary search
    if (total) TRW_HIP(dmalloc(&lines, static_cast<size_t>(total) * 8), "hipMalloc membership tables");
    TRW_HIP(dmalloc(&nrec, static_cast<size_t>(g->ne) + 1), "hipMalloc Node2Vec neighbour records");
    uint32_t *fill_count = nullptr;
    TRW_HIP(sc.get(&fill_count, static_cast<size_t>(total)), "hipMalloc membership bucket counters");
    TRW_HIP(trw::launch_member_table_fill(g->view(), start, total, lines, fill_count, st), "membership table fill");
    TRW_HIP(trw::launch_member_records(g->view(), root, nrec, st), "Node2Vec neighbour records");
    TRW_HIP(hipStreamSynchronize(st), "membership tables");
    g->mroot = root;
    g->mlines = lines;
    g->nrec = nrec;
    root = nullptr;
    lines = nullptr;
    nrec = nullptr;
    return TRW_OK;
}

// Weighted Node2Vec records (two per edge) from the neighbour records with roots and the edge
// weights: built by the first weighted Node2Vec walk once the membership tables exist.
static int build_member_weight_records(trw_graph *g) {
    if (g->wnrec || !g->nrec || !g->weight || g->ne == 0) return TRW_OK;
    uint4 *w = nullptr;
    TRW_HIP(dmalloc(&w, 2 * static_cast<size_t>(g->ne) + 2), "hipMalloc weighted Node2Vec records");
    hipError_t e = trw::launch_member_weight_records(g->view(), w, g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) {
        dfree(w);
        return hip_fail(e, "weighted Node2Vec records");
    }
    g->wnrec = w;
    return TRW_OK;
}

int trw_build_membership(trw_graph *g) {
    if (!g) return fail(TRW_E_INVALID, "graph is null");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    return build_member_tables(g);
}

int trw_build_tables(trw_graph *g, int32_t sampling, int32_t weights) {
    if (!g) return fail(TRW_E_INVALID, "graph is null");
    if (weights != TRW_WEIGHTS_EDGE && weights != TRW_WEIGHTS_UNIT) return fail(TRW_E_INVALID, "unknown weights %d", weights);
    if (weights == TRW_WEIGHTS_EDGE && !g->weight) return fail(TRW_E_STATE, "graph was uploaded without edge weights");
    if (sampling != TRW_SM_ALIAS && sampling != TRW_SM_ITS && sampling != TRW_SM_REJECTION)
        return fail(TRW_E_INVALID, "static tables exist for ITS, alias and rejection sampling only");
    return build_tables(g, sampling, g->weight, weights == TRW_WEIGHTS_UNIT, weights, g->weights_nonneg);
}

int trw_build_tables_from(trw_graph *g, int32_t sampling, const double *device_weights) {
    if (!g) return fail(TRW_E_INVALID, "graph is null");
    if (sampling != TRW_SM_ALIAS && sampling != TRW_SM_ITS && sampling != TRW_SM_REJECTION)
        return fail(TRW_E_INVALID, "static tables exist for ITS, alias and rejection sampling only");
    if (!device_weights && g->ne > 0) return fail(TRW_E_INVALID, "device_weights is null");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    const double *w = device_weights;
    if (sampling == TRW_SM_REJECTION) {
        // static_rs_move reads the weights beside the maxima (rejection_sampling.h:334-366): keep a copy.
        if (!g->rs_custom) TRW_HIP(dmalloc(&g->rs_custom, static_cast<size_t>(g->ne) + 1), "hipMalloc rejection weights");
        if (g->ne > 0)
            TRW_HIP(hipMemcpyAsync(g->rs_custom, device_weights, sizeof(double) * static_cast<size_t>(g->ne),
                                   hipMemcpyDeviceToDevice, g->ctx->stream),
                    "copy rejection weights");
        w = g->rs_custom;
    }
    // Custom weights get no ITS search structures (their walks binary-search the prefix sums).
    return build_tables(g, sampling, w, 0, TRW_WEIGHTS_CUSTOM, false);
}

int trw_graph_device_view(const trw_graph *g, trw_device_view *out) {
    if (!g || !out) return fail(TRW_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->num_vertices = g->nv;
    out->max_degree = g->max_degree;
    out->num_edges = g->ne;
    out->offset = g->offset;
    out->degree = g->degree;
    out->adj = g->adj;
    out->edge_weight = g->weight;
    out->edge_label = g->label;
    out->alias_table = g->alias_weights >= 0 ? reinterpret_cast<const trw_alias_slot *>(g->alias) : nullptr;
    out->prefix_sum = g->its_weights >= 0 ? g->prefix : nullptr;
    out->rejection_max = g->rs_weights >= 0 ? g->rs_max : nullptr;
    out->rejection_weight = g->rs_weights == TRW_WEIGHTS_EDGE     ? g->weight
                            : g->rs_weights == TRW_WEIGHTS_CUSTOM ? g->rs_custom
                                                                  : nullptr;
    out->alias_weights = g->alias_weights;
    out->its_weights = g->its_weights;
    out->rejection_weights = g->rs_weights;
    out->membership_lines = g->mlines;
    out->membership_root = g->mroot;
    return TRW_OK;
}

// ---- table files (graph.cpp:18-30, 184-212) ----

namespace {

const char *table_file(int32_t sampling) {
    return sampling == TRW_SM_ALIAS ? "b_edge_weight_alias_table.bin"
           : sampling == TRW_SM_ITS ? "b_edge_weight_prefix_sum.bin"
                                    : "b_edge_weight_max.bin";
}

int write_device_array(const std::string &path, const void *dev, size_t bytes) {
    std::vector<char> host(bytes);
    if (bytes) TRW_HIP(hipMemcpy(host.data(), dev, bytes, hipMemcpyDeviceToHost), "copy table to host");
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(TRW_E_IO, "cannot create %s", path.c_str());
    const size_t w = bytes ? std::fwrite(host.data(), 1, bytes, f) : 0;
    const int c = std::fclose(f);
    if (w != bytes || c != 0) return fail(TRW_E_IO, "write failed: %s", path.c_str());
    return TRW_OK;
}

}  // namespace

int trw_graph_store_tables(const trw_graph *g, const char *dir) {
    if (!g || !dir) return fail(TRW_E_INVALID, "null argument");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    TRW_HIP(hipStreamSynchronize(g->ctx->stream), "synchronize");
    const std::string d(dir);
    const size_t E = static_cast<size_t>(g->ne), V = static_cast<size_t>(g->nv);
    int written = 0, rc = TRW_OK;
    if (g->alias && g->alias_weights >= 0 && rc == TRW_OK) {
        rc = write_device_array(d + "/" + table_file(TRW_SM_ALIAS), g->alias, sizeof(AliasSlot) * E);
        ++written;
    }
    if (g->prefix && g->its_weights >= 0 && rc == TRW_OK) {
        rc = write_device_array(d + "/" + table_file(TRW_SM_ITS), g->prefix, sizeof(double) * E);
        ++written;
    }
    if (g->rs_max && g->rs_weights >= 0 && rc == TRW_OK) {
        rc = write_device_array(d + "/" + table_file(TRW_SM_REJECTION), g->rs_max, sizeof(double) * V);
        ++written;
    }
    if (rc != TRW_OK) return rc;
    if (!written) return fail(TRW_E_STATE, "no table has been built");
    return TRW_OK;
}

int trw_graph_load_tables(trw_graph *g, const char *dir, int32_t sampling, int32_t weights) {
    if (!g || !dir) return fail(TRW_E_INVALID, "null argument");
    if (sampling != TRW_SM_ALIAS && sampling != TRW_SM_ITS && sampling != TRW_SM_REJECTION)
        return fail(TRW_E_INVALID, "table files exist for ITS, alias and rejection sampling only");
    if (weights != TRW_WEIGHTS_EDGE && weights != TRW_WEIGHTS_UNIT) return fail(TRW_E_INVALID, "unknown weights %d", weights);
    const size_t E = static_cast<size_t>(g->ne), V = static_cast<size_t>(g->nv);
    const size_t bytes = sampling == TRW_SM_ALIAS ? sizeof(AliasSlot) * E
                         : sampling == TRW_SM_ITS ? sizeof(double) * E
                                                  : sizeof(double) * V;
    const std::string path = std::string(dir) + "/" + table_file(sampling);
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(TRW_E_IO, "cannot open %s", path.c_str());
    std::vector<char> host(bytes + 1);
    const size_t got = std::fread(host.data(), 1, bytes + 1, f);
    std::fclose(f);
    if (got != bytes) return fail(TRW_E_IO, "%s holds %zu bytes, the graph needs %zu", path.c_str(), got, bytes);
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    hipStream_t st = g->ctx->stream;
    TRW_HIP(hipStreamSynchronize(st), "synchronize");
    if (sampling == TRW_SM_ALIAS) {
        if (!g->alias) TRW_HIP(dmalloc(&g->alias, E + 1), "hipMalloc alias table");
        if (E) TRW_HIP(hipMemcpy(g->alias, host.data(), bytes, hipMemcpyHostToDevice), "copy alias table");
        g->alias_weights = weights;
        if (g->vrec) {
            if (!g->arec) TRW_HIP(dmalloc(&g->arec, 2 * E + 2), "hipMalloc alias walk records");
            TRW_HIP(trw::launch_alias_records(g->view(), g->arec, st), "alias walk records");
        }
    } else if (sampling == TRW_SM_ITS) {
        if (!g->prefix) TRW_HIP(dmalloc(&g->prefix, E + 2 * trw::kItsFan), "hipMalloc prefix sums");
        if (E) TRW_HIP(hipMemcpy(g->prefix, host.data(), bytes, hipMemcpyHostToDevice), "copy prefix sums");
        // Search trees need non-decreasing sums within every vertex (so a file written from
        // weights with a negative entry gets none, as a device build would).
        const double *P = reinterpret_cast<const double *>(host.data());
        std::vector<int32_t> degree(V);
        if (V) TRW_HIP(hipMemcpy(degree.data(), g->degree, sizeof(int32_t) * V, hipMemcpyDeviceToHost), "copy degree");
        bool monotone = true, nonneg = true;
        size_t e = 0;
        for (size_t v = 0; v < V && monotone; ++v) {
            if (degree[v] > 0) nonneg = nonneg && P[e] >= 0.0;
            for (int32_t j = 1; j < degree[v] && monotone; ++j) monotone = P[e + j - 1] <= P[e + j];
            e += static_cast<size_t>(degree[v]);
        }
        g->its_weights = weights;
        int rc = build_its_trees(g, monotone);
        // Guide records also need sums >= 0: they end a walk at a vertex whose sums total zero
        // without a draw, which for sums running below zero would not be the Move's pick.
        if (rc == TRW_OK) rc = build_its_guide(g, monotone && nonneg);
        if (rc != TRW_OK) return rc;
    } else {
        if (!g->rs_max) TRW_HIP(dmalloc(&g->rs_max, V), "hipMalloc rejection maxima");
        if (V) TRW_HIP(hipMemcpy(g->rs_max, host.data(), bytes, hipMemcpyHostToDevice), "copy rejection maxima");
        if (g->vrec && (weights == TRW_WEIGHTS_UNIT || g->weight)) {
            if (!g->rrec) TRW_HIP(dmalloc(&g->rrec, 2 * E + 2), "hipMalloc rejection walk records");
            TRW_HIP(trw::launch_rs_records(g->view(), g->weight, weights == TRW_WEIGHTS_UNIT, g->rrec, st),
                    "rejection walk records");
        } else {
            dfree(g->rrec);
        }
        g->rs_weights = weights;
    }
    TRW_HIP(hipStreamSynchronize(st), "load tables");
    return TRW_OK;
}

namespace {

int check_shape(int32_t knob, int64_t value) {
    switch (knob) {
    case TRW_TUNE_FUSED_BLOCK:
        if (value != 0 && value != 64 && value != 128 && value != 256)
            return fail(TRW_E_INVALID, "fused threads per block: 0, 64, 128 or 256");
        return TRW_OK;
    case TRW_TUNE_FUSED_TILE:
        if (value != 0 && (value < 4 || value > 4096)) return fail(TRW_E_INVALID, "fused path tile: 0 or 4..4096");
        return TRW_OK;
    case TRW_TUNE_QUEUE_BLOCKS_PER_CU:
        if (value < 0 || value > 16) return fail(TRW_E_INVALID, "queue blocks per CU outside [0, 16]");
        return TRW_OK;
    case TRW_TUNE_N2V_BLOCK:
        if (value != 0 && value != 128 && value != 256)
            return fail(TRW_E_INVALID, "Node2Vec threads per block: 0, 128 or 256");
        return TRW_OK;
    case TRW_TUNE_GRAPH_BURST:
        if (value < 0 || value > 256) return fail(TRW_E_INVALID, "GRAPH-form burst outside [0, 256]");
        return TRW_OK;
    default:
        return fail(TRW_E_INVALID, "unknown tuning knob %d", knob);
    }
}

}  // namespace

int trw_graph_set_tuning(trw_graph *g, int32_t knob, int64_t value) {
    if (!g) return fail(TRW_E_INVALID, "graph is null");
    switch (knob) {
    case TRW_TUNE_WALKERS_PER_THREAD:
        if (value != 0 && value != 1 && value != 2 && value != 4) return fail(TRW_E_INVALID, "walkers per lane: 0, 1, 2 or 4");
        break;
    case TRW_TUNE_ITS_TREE_MIN_DEGREE:
        if (value < 0 || value > INT32_MAX) return fail(TRW_E_INVALID, "degree threshold outside [0, 2^31)");
        break;
    case TRW_TUNE_QUEUE_CHUNK:
        if (value < 0 || value > 4096) return fail(TRW_E_INVALID, "queue chunk outside [0, 4096]");
        break;
    case TRW_TUNE_ITS_GUIDE_OFF:
        if (value != 0 && value != 1) return fail(TRW_E_INVALID, "ITS guide switch: 0 or 1");
        break;
    default: {
        const int rc = check_shape(knob, value);
        if (rc != TRW_OK) return rc;
        for (int m = 0; m < TRW_MOVES; ++m) g->shape[m][knob - 4] = value;
        return TRW_OK;
    }
    }
    g->tune[knob] = value;
    return TRW_OK;
}

int trw_graph_get_tuning(const trw_graph *g, int32_t knob, int64_t *value) {
    if (!g || !value) return fail(TRW_E_INVALID, "null argument");
    if (knob < 0 || knob >= TRW_TUNE_KNOBS) return fail(TRW_E_INVALID, "unknown tuning knob %d", knob);
    *value = knob < 4 ? g->tune[knob] : g->shape[TRW_MOVE_UNIFORM][knob - 4];
    return TRW_OK;
}

int trw_graph_set_move_tuning(trw_graph *g, int32_t move, int32_t knob, int64_t value) {
    if (!g) return fail(TRW_E_INVALID, "graph is null");
    if (move < 0 || move >= TRW_MOVES) return fail(TRW_E_INVALID, "unknown Move %d", move);
    if (knob < 4 || knob >= TRW_TUNE_KNOBS) return fail(TRW_E_INVALID, "knob %d is not a launch shape (4..8)", knob);
    const int rc = check_shape(knob, value);
    if (rc != TRW_OK) return rc;
    g->shape[move][knob - 4] = value;
    return TRW_OK;
}

int trw_graph_get_move_tuning(const trw_graph *g, int32_t move, int32_t knob, int64_t *value) {
    if (!g || !value) return fail(TRW_E_INVALID, "null argument");
    if (move < 0 || move >= TRW_MOVES) return fail(TRW_E_INVALID, "unknown Move %d", move);
    if (knob < 4 || knob >= TRW_TUNE_KNOBS) return fail(TRW_E_INVALID, "knob %d is not a launch shape (4..8)", knob);
    *value = g->shape[move][knob - 4];
    return TRW_OK;
}

int trw_last_build_ms(const trw_graph *g, double *ms) {
    if (!g || !ms) return fail(TRW_E_INVALID, "null argument");
    *ms = g->last_build_ms;
    return TRW_OK;
}

template <class T>
static int copy_table(const trw_graph *g, const T *src, int32_t built, size_t count, T *host, const char *what) {
    if (!g || !host) return fail(TRW_E_INVALID, "null argument");
    if (!src || built < 0) return fail(TRW_E_STATE, "%s has not been built", what);
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    TRW_HIP(hipStreamSynchronize(g->ctx->stream), "synchronize");
    if (count) TRW_HIP(hipMemcpy(host, src, sizeof(T) * count, hipMemcpyDeviceToHost), what);
    return TRW_OK;
}

int trw_copy_alias_table(const trw_graph *g, trw_alias_slot *host_out) {
    static_assert(sizeof(trw_alias_slot) == sizeof(AliasSlot), "layout");
    return copy_table(g, g ? g->alias : nullptr, g ? g->alias_weights : -1, g ? static_cast<size_t>(g->ne) : 0,
                      reinterpret_cast<AliasSlot *>(host_out), "alias table");
}

int trw_copy_prefix_sum(const trw_graph *g, double *host_out) {
    return copy_table(g, g ? g->prefix : nullptr, g ? g->its_weights : -1, g ? static_cast<size_t>(g->ne) : 0,
                      host_out, "prefix-sum table");
}

int trw_copy_rejection_max(const trw_graph *g, double *host_out) {
    return copy_table(g, g ? g->rs_max : nullptr, g ? g->rs_weights : -1, g ? static_cast<size_t>(g->nv) : 0,
                      host_out, "rejection maxima");
}

// ---------------------------------------------------------------- walks

int trw_walk_supported(const trw_walk_params *p) {
    if (!p) return fail(TRW_E_INVALID, "params is null");
    if (p->app < TRW_APP_DEEPWALK || p->app > TRW_APP_GENERIC) return fail(TRW_E_INVALID, "unknown app %d", p->app);
    if (p->execution < TRW_EXEC_UNIFORM || p->execution > TRW_EXEC_DYNAMIC)
        return fail(TRW_E_INVALID, "unknown execution mode %d", p->execution);
    if (p->sampling < TRW_SM_UNIFORM || p->sampling > TRW_SM_MAX_WEIGHT_REJECTION)
        return fail(TRW_E_INVALID, "unknown sampling method %d", p->sampling);
    if (p->execution == TRW_EXEC_UNIFORM) return TRW_OK;  // uniform_move for every app
    if (p->app == TRW_APP_NODE2VEC) {
        if (p->execution != TRW_EXEC_DYNAMIC || p->sampling != TRW_SM_MAX_WEIGHT_REJECTION)
            return fail(TRW_E_UNSUPPORTED, "Node2Vec runs in Dynamic mode with max-weight rejection sampling");
        return TRW_OK;
    }
    if (p->app == TRW_APP_METAPATH) {
        if (p->execution != TRW_EXEC_DYNAMIC || p->sampling != TRW_SM_ITS)
            return fail(TRW_E_UNSUPPORTED, "MetaPath runs in Dynamic mode with inverse-transform sampling");
        return TRW_OK;
    }
    if (p->sampling == TRW_SM_UNIFORM || p->sampling == TRW_SM_MAX_WEIGHT_REJECTION)
        return fail(TRW_E_UNSUPPORTED, "sampling method %d is not defined for this app in mode %d", p->sampling,
                    p->execution);
    return TRW_OK;
}

// The weights (TRW_WEIGHTS_EDGE / UNIT) the static table a walk of DeepWalk, GenericWalk or PPR
// samples from is built from: the app's F::weight unless params.weights overrides it.
static int32_t table_weights(const trw_walk_params *p) {
    return p->weights == TRW_WALK_WEIGHTS_EDGE   ? TRW_WEIGHTS_EDGE
           : p->weights == TRW_WALK_WEIGHTS_UNIT ? TRW_WEIGHTS_UNIT
           : p->app == TRW_APP_PPR               ? TRW_WEIGHTS_UNIT
                                                 : TRW_WEIGHTS_EDGE;
}

// Whether a walk samples from a graph-wide static table (alias, ITS or rejection), and which.
static bool reads_static_table(const trw_walk_params *p) {
    return p->execution != TRW_EXEC_UNIFORM && p->app != TRW_APP_NODE2VEC && p->app != TRW_APP_METAPATH &&
           (p->sampling == TRW_SM_ALIAS || p->sampling == TRW_SM_ITS || p->sampling == TRW_SM_REJECTION);
}

// Maps an app + g_para to a device mode, building the static tables the mode
// reads when they are missing.
// A MetaPath schema whose labels are distinct and all in [0, 8): its labels as nibbles and its
// length in bits 60-63 (never 0); 0 for any other schema.
static uint64_t schema_key(const trw_walk_params *p) {
    if (p->schema_length <= 0 || p->schema_length > 8) return 0;
    uint32_t seen = 0;
    uint64_t key = static_cast<uint64_t>(p->schema_length) << 60;
    for (int i = 0; i < p->schema_length; ++i) {
        const int32_t l = p->schema[i];
        if (l < 0 || l >= 8 || (seen >> l & 1u)) return 0;
        seen |= 1u << l;
        key |= static_cast<uint64_t>(l) << (4 * i);
    }
    return key;
}

static int build_schema_records(trw_graph *g, const trw_walk_params *p, uint64_t key) {
    uint32_t next_map = 0xffffffffu;  // nibble l: the label after l in the schema (15: l not in it)
    for (int i = 0; i < p->schema_length; ++i) {
        const uint32_t l = static_cast<uint32_t>(p->schema[i]);
        const uint32_t nl = static_cast<uint32_t>(p->schema[(i + 1) % p->schema_length]);
        next_map = (next_map & ~(15u << (4 * l))) | (nl << (4 * l));
    }
    if (!g->msrec) TRW_HIP(dmalloc(&g->msrec, static_cast<size_t>(g->ne) + 1), "hipMalloc MetaPath schema records");
    g->msrec_key = 0;
    TRW_HIP(trw::launch_schema_records(g->view(), next_map, g->msrec, g->ctx->stream), "MetaPath schema records");
    TRW_HIP(hipStreamSynchronize(g->ctx->stream), "MetaPath schema records");
    g->msrec_key = key;
    return TRW_OK;
}

int trw::prepare_records(trw_graph *g, const trw_walk_params *p) {
    if (p->app == TRW_APP_METAPATH && p->execution != TRW_EXEC_UNIFORM && g->lerec) {
        const uint64_t key = schema_key(p);
        if (key != 0 && g->msrec_key != key) return build_schema_records(g, p, key);
    }
    return TRW_OK;
}

static int resolve_move(trw_graph *g, const trw_walk_params *p, KindParams &k, bool allow_build) {
    std::memset(&k, 0, sizeof(k));
    if (p->app < TRW_APP_DEEPWALK || p->app > TRW_APP_GENERIC) return fail(TRW_E_INVALID, "unknown app %d", p->app);
    if (p->execution < TRW_EXEC_UNIFORM || p->execution > TRW_EXEC_DYNAMIC)
        return fail(TRW_E_INVALID, "unknown execution mode %d", p->execution);
    if (p->sampling < TRW_SM_UNIFORM || p->sampling > TRW_SM_MAX_WEIGHT_REJECTION)
        return fail(TRW_E_INVALID, "unknown sampling method %d", p->sampling);
    const bool ppr = p->app == TRW_APP_PPR;
    if (ppr) {
        if (p->length < 1) return fail(TRW_E_INVALID, "PPR needs length >= 1");
        if (!(p->stop_probability > 0.0 && p->stop_probability <= 1.0))
            return fail(TRW_E_INVALID, "stop_probability must lie in (0, 1]");
    } else if (p->length < 2) {
        return fail(TRW_E_INVALID, "length must be >= 2 (a walk of length l makes l-1 moves)");
    }
    if (p->max_trials < 0) return fail(TRW_E_INVALID, "max_trials must be >= 0");
    if (p->walkers_per_thread != 0 && p->walkers_per_thread != 1 && p->walkers_per_thread != 2 &&
        p->walkers_per_thread != 4)
        return fail(TRW_E_INVALID, "walkers_per_thread must be 0, 1, 2 or 4");
    if (p->weights < TRW_WALK_WEIGHTS_APP || p->weights > TRW_WALK_WEIGHTS_UNIT)
        return fail(TRW_E_INVALID, "unknown weights selector %d", p->weights);
    k.walkers_per_thread = p->walkers_per_thread ? p->walkers_per_thread : static_cast<int32_t>(g->tune[TRW_TUNE_WALKERS_PER_THREAD]);
    k.its_tree_min = static_cast<int32_t>(g->tune[TRW_TUNE_ITS_TREE_MIN_DEGREE]);
    k.queue_chunk = static_cast<int32_t>(g->tune[TRW_TUNE_QUEUE_CHUNK]);
    k.its_guide = g->tune[TRW_TUNE_ITS_GUIDE_OFF] == 0;
    k.ppr = ppr;
    k.length = p->length;
    k.max_trials = p->max_trials;
    k.stop_prob = p->stop_probability;

    {
        const int rc = trw_walk_supported(p);
        if (rc != TRW_OK) return rc;
    }
    if (p->execution == TRW_EXEC_UNIFORM) {
        k.mode = trw::kModeUniform;
        return TRW_OK;
    }
    if (p->app == TRW_APP_NODE2VEC) {
        if (!(p->p > 0.0) || !(p->q > 0.0)) return fail(TRW_E_INVALID, "Node2Vec p and q must be > 0");
        // weights EDGE: Node2Vec over edge weights -- a candidate's weight is w_e times Node2Vec's
        // factor, the bound the largest edge weight times max_; APP and UNIT: the reference's functor.
        const bool weighted = p->weights == TRW_WALK_WEIGHTS_EDGE;
        if (weighted && !g->weight) return fail(TRW_E_STATE, "graph was uploaded without edge weights");
        k.mode = trw::kModeNode2Vec;
        if (allow_build) {
            int rc = build_member_tables(g);
            if (rc == TRW_OK && weighted) rc = build_member_weight_records(g);
            if (rc != TRW_OK) return rc;
        }
        k.n2v_p = p->p;  // Node2Vec::p_ (node2vec.cpp:18, after the inversion at :69)
        k.n2v_q = p->q;  // Node2Vec::q_ (:76)
        k.n2v_max = std::max(1.0, std::max(k.n2v_q, k.n2v_p));  // node2vec.cpp:19
        k.n2v_weighted = weighted;
        k.n2v_bound = weighted ? g->max_weight * k.n2v_max : k.n2v_max;
        return TRW_OK;
    }
    if (p->app == TRW_APP_METAPATH) {
        if (!g->label) return fail(TRW_E_STATE, "MetaPath needs edge labels");
        if (p->schema_length <= 0 || p->schema_length > 16) return fail(TRW_E_INVALID, "schema length must be 1..16");
        k.mode = trw::kModeMetaPath;
        k.schema_length = p->schema_length;
        for (int i = 0; i < p->schema_length; ++i) {
            k.schema[i] = p->schema[i];
            const uint64_t nib = (p->schema[i] >= 0 && p->schema[i] < 8) ? static_cast<uint64_t>(p->schema[i]) : 8u;
            k.schema_nibbles |= nib << (4 * i);
        }
        // Distinct labels, all in [0, 8): 16-byte schema records (built once per schema and kept
        // with the graph; another schema replaces them).
        const uint64_t key = schema_key(p);
        if (key != 0 && g->lerec) {
            if (g->msrec_key != key && allow_build) {
                int rc = build_schema_records(g, p, key);
                if (rc != TRW_OK) return rc;
            }
            k.schema_rec = g->msrec_key == key;
        }
        return TRW_OK;
    }
    // DeepWalk, GenericWalk (edge weights) and PPR (unit weights): walker-independent
    // weights, so the Dynamic gather rebuilds exactly the Static tables every step.
    const int32_t wsrc = table_weights(p);
    if (wsrc == TRW_WEIGHTS_EDGE && !g->weight) return fail(TRW_E_STATE, "graph was uploaded without edge weights");
    int32_t *built = p->sampling == TRW_SM_ALIAS ? &g->alias_weights
                     : p->sampling == TRW_SM_ITS ? &g->its_weights
                                                 : &g->rs_weights;
    if (*built != wsrc) {
        if (!allow_build) return fail(TRW_E_STATE, "the tables this walk reads have not been built (trw_build_tables)");
        int rc = trw_build_tables(g, p->sampling, wsrc);
        if (rc != TRW_OK) return rc;
    }
    k.mode = p->sampling == TRW_SM_ALIAS ? trw::kModeAlias : p->sampling == TRW_SM_ITS ? trw::kModeIts : trw::kModeRejection;
    k.unit_weights = wsrc == TRW_WEIGHTS_UNIT;
    return TRW_OK;
}

static_assert(trw::kModeUniform == TRW_MOVE_UNIFORM && trw::kModeAlias == TRW_MOVE_ALIAS &&
                  trw::kModeIts == TRW_MOVE_ITS && trw::kModeRejection == TRW_MOVE_REJECTION &&
                  trw::kModeNode2Vec == TRW_MOVE_NODE2VEC && trw::kModeMetaPath == TRW_MOVE_METAPATH,
              "TRW_MOVE_* index the graph's per-Move shapes by the walk's Move");

// The walk's kind, then the launch shapes the graph keeps for its Move.
static int resolve_kind(trw_graph *g, const trw_walk_params *p, KindParams &k, bool allow_build = true) {
    const int rc = resolve_move(g, p, k, allow_build);
    if (rc != TRW_OK) return rc;
    if (k.mode >= 0 && k.mode < TRW_MOVES) {
        const int64_t *sh = g->shape[k.mode];
        k.fused_block = static_cast<int32_t>(sh[TRW_TUNE_FUSED_BLOCK - 4]);
        k.fused_tile = static_cast<int32_t>(sh[TRW_TUNE_FUSED_TILE - 4]);
        k.queue_blocks_per_cu = static_cast<int32_t>(sh[TRW_TUNE_QUEUE_BLOCKS_PER_CU - 4]);
        k.n2v_block = static_cast<int32_t>(sh[TRW_TUNE_N2V_BLOCK - 4]);
        k.graph_burst = static_cast<int32_t>(sh[TRW_TUNE_GRAPH_BURST - 4]);
    }
    return TRW_OK;
}

static int make_io(const trw_graph *g, const trw_walk_params *p, const trw_walk_io *io, const KindParams &a,
                   const KindParams &b, BatchIO &out) {
    if (!io) return fail(TRW_E_INVALID, "io is null");
    if (io->n_walkers < 0) return fail(TRW_E_INVALID, "n_walkers must be >= 0");
    if (io->n_walkers > 0 && !io->sources) return fail(TRW_E_INVALID, "sources is null");
    if (io->n_walkers > (1 << 30)) return fail(TRW_E_INVALID, "at most 2^30 walkers per call");
    const int32_t La = std::max(a.length, 1), Lb = std::max(b.length, 1);
    if (io->path_stride_b < 0) return fail(TRW_E_INVALID, "path_stride_b must be >= 0");
    const int32_t stride_b = io->path_stride_b > 0 ? io->path_stride_b : io->path_stride;
    if (io->paths && io->path_stride_b == 0 && io->path_stride < std::max(La, Lb))
        return fail(TRW_E_INVALID, "path_stride %d < length %d", io->path_stride, std::max(La, Lb));
    if (io->paths && io->path_stride < La) return fail(TRW_E_INVALID, "path_stride %d < length %d", io->path_stride, La);
    if (io->paths && stride_b < Lb) return fail(TRW_E_INVALID, "path_stride_b %d < length %d", stride_b, Lb);
    if (p->rng != TRW_RNG_REPLAY && p->rng != TRW_RNG_PHILOX) return fail(TRW_E_INVALID, "unknown rng %d", p->rng);
    const bool replay = p->rng == TRW_RNG_REPLAY;
    if (replay) {
        if (io->draws_per_walker < 0 || io->term_draws_per_walker < 0) return fail(TRW_E_INVALID, "negative draw count");
        if (io->draws_per_walker > 0 && !io->draws) return fail(TRW_E_INVALID, "draws is null");
        if ((a.ppr || b.ppr) && io->term_draws_per_walker > 0 && !io->term_draws)
            return fail(TRW_E_INVALID, "term_draws is null");
        if (io->draws_per_walker > 0xffffffffLL || io->term_draws_per_walker > 0xffffffffLL)
            return fail(TRW_E_INVALID, "draws_per_walker must fit in 32 bits");
    }
    std::memset(&out, 0, sizeof(out));
    out.sources = io->sources;
    out.n = io->n_walkers;
    out.path_stride = io->path_stride;
    out.path_stride_b = stride_b;
    out.gid_base = static_cast<uint64_t>(io->walker_id_base);
    out.walker_ids = io->walker_ids;
    if (io->path_flags != TRW_PATHS_PADDED && io->path_flags != TRW_PATHS_UNPADDED)
        return fail(TRW_E_INVALID, "unknown path_flags %d", io->path_flags);
    out.unpadded = io->path_flags == TRW_PATHS_UNPADDED;
    out.draws = replay ? io->draws : nullptr;
    out.K = replay ? io->draws_per_walker : 0;
    out.tdraws = replay ? io->term_draws : nullptr;
    out.KT = replay ? io->term_draws_per_walker : 0;
    out.paths = io->paths;
    out.lengths = io->lengths;
    out.ends = io->ends;
    out.status = io->status;
    out.draws_used = io->draws_used;
    out.counters = reinterpret_cast<unsigned long long *>(io->counters);
    out.key0 = static_cast<uint32_t>(p->seed);
    out.key1 = static_cast<uint32_t>(p->seed >> 32);
    out.replay = replay;
    (void)g;
    return TRW_OK;
}

static int ensure_state(trw_graph *g, int64_t n) {
    if (g->state_cap >= n) return TRW_OK;
    dfree(g->state);
    dfree(g->active);
    g->state_cap = 0;
    TRW_HIP(dmalloc(&g->state, static_cast<size_t>(n)), "hipMalloc walker state");
    TRW_HIP(dmalloc(&g->active, 3 * static_cast<size_t>(n)), "hipMalloc live-walker lists");
    g->state_cap = n;
    return TRW_OK;
}

// Plain per-step form over walker state in `state` (io.n entries) with a live flag at `alive_flag`.
static int run_step_form_on(trw_graph *g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                            hipStream_t st, WalkerState *state, unsigned int *alive_flag) {
    if (io.n == 0) return TRW_OK;
    const DevGraph dv = g->view();
    TRW_HIP(trw::launch_walk_step_init(dv, a, b, split, io, state, st), "step-form init");
    unsigned int alive = 1;
    const int kChunk = 8;
    while (alive) {
        for (int i = 0; i < kChunk; ++i) {
            TRW_HIP(trw::launch_fill_u32(alive_flag, 0u, 1, st), "step-form reset");
            TRW_HIP(trw::launch_walk_step(dv, a, b, split, io, state, alive_flag, st), "step-form step");
        }
        TRW_HIP(hipMemcpyAsync(&alive, alive_flag, sizeof(alive), hipMemcpyDeviceToHost, st), "read alive");
        TRW_HIP(hipStreamSynchronize(st), "step-form synchronize");
    }
    TRW_HIP(trw::launch_walk_step_finish(a, b, split, io, state, st), "step-form finish");
    return TRW_OK;
}

static int run_step_form(trw_graph *g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                         hipStream_t st) {
    if (io.n == 0) return TRW_OK;
    int rc = ensure_state(g, io.n);
    if (rc != TRW_OK) return rc;
    return run_step_form_on(g, a, b, split, io, st, g->state, g->counters + 1);
}

// The GRAPH form's chunk: six launches captured once and replayed while walkers remain.  Kept per
// graph with the bytes of every argument the launches were captured with; a call with other
// arguments (other kinds, buffers or batch) captures a new one.
struct GraphChunk {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
    ~GraphChunk() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

void release_graph_chunk(trw_graph *g) {
    delete g->graph_chunk;
    g->graph_chunk = nullptr;
}

namespace {
// Steps per live walker per launch unless the first kind's Move is tuned otherwise.  Measured on the
// PPR + MetaPath mix (2^23 walkers, walks of 7 - 8 positions on average): 4 steps 4.2 ms per batch,
// 8 steps 3.1 ms, 16 steps 2.6 ms, 32 steps 2.8 ms, 64 steps 3.4 ms: each launch reads and writes every
// live walker's state, so fewer launches save that traffic until lanes idle on walks that end early.
constexpr int kGraphBurst = 16;
constexpr int kGraphChunk = 6;   // launches per captured chunk (a multiple of the three rotating lists)

template <class T>
void append_bytes(std::vector<unsigned char> &k, const T &v) {
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}
}  // namespace

// GRAPH form: live walkers on three rotating lists (launch k reads list k % 3, appends the walkers
// still walking to list (k + 1) % 3 and zeroes count (k + 2) % 3, the one launch k - 1 read).  The
// first launch starts every walker from its source and writes list 1; a chunk of six launches
// (k = 1..6) ends with the live list again in list 1, so it is captured once, kept, and replayed
// until the list is empty.  Each launch takes a walker up to `burst` steps (the first kind's tuned
// choice, else kGraphBurst).
static int run_graph_form(trw_graph *g, const KindParams &a, const KindParams &b, int32_t split, const BatchIO &io,
                          hipStream_t st) {
    if (io.n == 0) return TRW_OK;
    int rc = ensure_state(g, io.n);
    if (rc != TRW_OK) return rc;
    const int32_t burst = a.graph_burst > 0 ? a.graph_burst : kGraphBurst;
    const DevGraph dv = g->view();
    int32_t *list[3] = {g->active, g->active + g->state_cap, g->active + 2 * g->state_cap};
    unsigned int *count[3] = {g->counters + 4, g->counters + 5, g->counters + 6};
    TRW_HIP(trw::launch_fill_paths(a, b, split, io, st), "graph-form path rows");
    TRW_HIP(trw::launch_fill_u32(count[1], 0u, 1, st), "graph-form reset");
    TRW_HIP(trw::launch_walk_graph(dv, a, b, split, io, g->state, nullptr, nullptr, list[1], count[1], count[2], 1,
                                   burst, st),
            "graph-form first launch");
    std::vector<unsigned char> key;
    append_bytes(key, dv);
    append_bytes(key, a);
    append_bytes(key, b);
    append_bytes(key, split);
    append_bytes(key, io);
    append_bytes(key, g->state);
    append_bytes(key, g->active);
    if (!g->graph_chunk || g->graph_chunk->key != key) {
        release_graph_chunk(g);
        std::unique_ptr<GraphChunk> c(new GraphChunk);
        TRW_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
        hipError_t e = hipSuccess;
        for (int k = 1; k <= kGraphChunk && e == hipSuccess; ++k)
            e = trw::launch_walk_graph(dv, a, b, split, io, g->state, list[k % 3], count[k % 3], list[(k + 1) % 3],
                                       count[(k + 1) % 3], count[(k + 2) % 3], 0, burst, st);
        const hipError_t ec = hipStreamEndCapture(st, &c->graph);
        TRW_HIP(e, "capturing the step chunk");
        TRW_HIP(ec, "end capture");
        TRW_HIP(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0), "instantiate the step chunk");
        c->key = std::move(key);
        g->graph_chunk = c.release();
    }
    unsigned int live = 0;
    TRW_HIP(hipMemcpyAsync(&live, count[1], sizeof(live), hipMemcpyDeviceToHost, st), "read live count");
    TRW_HIP(hipStreamSynchronize(st), "graph-form synchronize");
    while (live) {
        TRW_HIP(hipGraphLaunch(g->graph_chunk->exec, st), "replay the step chunk");
        TRW_HIP(hipMemcpyAsync(&live, count[1], sizeof(live), hipMemcpyDeviceToHost, st), "read live count");
        TRW_HIP(hipStreamSynchronize(st), "graph-form synchronize");
    }
    return TRW_OK;
}

int trw::walk_launch(trw_graph *g, const trw_walk_params *params, const trw_walk_io *io, hipStream_t stream,
                     unsigned int *queue_head, bool allow_table_build, bool head_zeroed) {
    if (!g || !params) return fail(TRW_E_INVALID, "null argument");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    KindParams k;
    int rc = resolve_kind(g, params, k, allow_table_build);
    if (rc != TRW_OK) return rc;
    BatchIO bio;
    rc = make_io(g, params, io, k, k, bio);
    if (rc != TRW_OK) return rc;
    hipStream_t st = stream ? stream : g->ctx->stream;
    int32_t form = params->form;
    if (form == TRW_FORM_AUTO) form = (k.ppr || k.mode == trw::kModeMetaPath) ? TRW_FORM_QUEUE : TRW_FORM_FUSED;
    const DevGraph dv = g->view();
    switch (form) {
    case TRW_FORM_FUSED:
        if (k.ppr) return fail(TRW_E_UNSUPPORTED, "the fused form runs fixed-length walks; use QUEUE or STEP for PPR");
        TRW_HIP(trw::launch_walk_fused(dv, k, bio, st), "fused walk launch");
        return TRW_OK;
    case TRW_FORM_QUEUE:
        TRW_HIP(trw::launch_walk_queue(dv, k, k, bio.n, bio, queue_head ? queue_head : g->counters, st,
                                       head_zeroed && queue_head != nullptr),
                "queue walk launch");
        return TRW_OK;
    case TRW_FORM_STEP:
        return run_step_form(g, k, k, bio.n, bio, st);
    case TRW_FORM_GRAPH:
        return run_graph_form(g, k, k, bio.n, bio, st);
    default:
        return fail(TRW_E_INVALID, "unknown form %d", params->form);
    }
}

int trw::walk_check(trw_graph *g, const trw_walk_params *params) {
    if (!g || !params) return fail(TRW_E_INVALID, "null argument");
    KindParams k;
    int rc = resolve_kind(g, params, k, false);
    if (rc != TRW_OK) return rc;
    if (params->rng != TRW_RNG_REPLAY && params->rng != TRW_RNG_PHILOX)
        return fail(TRW_E_INVALID, "unknown rng %d", params->rng);
    if (params->form < TRW_FORM_AUTO || params->form > TRW_FORM_GRAPH)
        return fail(TRW_E_INVALID, "unknown form %d", params->form);
    if (params->form == TRW_FORM_FUSED && k.ppr)
        return fail(TRW_E_UNSUPPORTED, "the fused form runs fixed-length walks; use QUEUE or STEP for PPR");
    return TRW_OK;
}

int trw_walk_async(trw_graph *g, const trw_walk_params *params, const trw_walk_io *io, void *stream) {
    return trw::walk_launch(g, params, io, reinterpret_cast<hipStream_t>(stream), nullptr, true);
}

// trw_walk_async without allocation: tables are never built here, and the QUEUE form's queue
// head and the STEP form's live flag and walker state live in this device's `scratch` array
// (through the address resolved at context creation, never the array's name).  The kernels are
// trw_walk_async's, so the walks are identical.  Calls that overlap on one device share that
// array and race.
int trw_walk_noalloc(trw_graph *g, const trw_walk_params *params, const trw_walk_io *io, void *stream) {
    if (!g || !params || !io) return fail(TRW_E_INVALID, "null argument");
    if (!g->ctx->scratch) return fail(TRW_E_STATE, "the context holds no scratch address for its device");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    KindParams k;
    int rc = resolve_kind(g, params, k, false);
    if (rc != TRW_OK) return rc;
    BatchIO bio;
    rc = make_io(g, params, io, k, k, bio);
    if (rc != TRW_OK) return rc;
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : g->ctx->stream;
    unsigned int *control = static_cast<unsigned int *>(g->ctx->scratch);
    WalkerState *state = reinterpret_cast<WalkerState *>(static_cast<unsigned char *>(g->ctx->scratch) +
                                                         kNoAllocControlBytes);
    int32_t form = params->form;
    if (form == TRW_FORM_AUTO) form = (k.ppr || k.mode == trw::kModeMetaPath) ? TRW_FORM_QUEUE : TRW_FORM_FUSED;
    switch (form) {
    case TRW_FORM_FUSED:
        if (k.ppr) return fail(TRW_E_UNSUPPORTED, "the fused form runs fixed-length walks; use QUEUE or STEP for PPR");
        TRW_HIP(trw::launch_walk_fused(g->view(), k, bio, st), "fused walk launch");
        return TRW_OK;
    case TRW_FORM_QUEUE:
        TRW_HIP(trw::launch_walk_queue(g->view(), k, k, bio.n, bio, control, st), "queue walk launch");
        return TRW_OK;
    case TRW_FORM_STEP:
        if (bio.n > kNoAllocMaxStepWalkers)
            return fail(TRW_E_INVALID,
                        "trw_walk_noalloc's STEP form takes at most %d walkers (its walker state is a fixed device "
                        "array); got %d",
                        kNoAllocMaxStepWalkers, bio.n);
        return run_step_form_on(g, k, k, bio.n, bio, st, state, control + 1);
    case TRW_FORM_GRAPH:
        return fail(TRW_E_UNSUPPORTED, "trw_walk_noalloc runs the AUTO, FUSED, QUEUE and STEP forms (instantiating "
                                       "the GRAPH form's hipGraph allocates)");
    default:
        return fail(TRW_E_INVALID, "unknown form %d", params->form);
    }
}

int trw_walk_mix_async(trw_graph *g, const trw_walk_params *pa, const trw_walk_params *pb, int32_t split,
                       const trw_walk_io *io, void *stream) {
    if (!g || !pa || !pb || !io) return fail(TRW_E_INVALID, "null argument");
    if (pa->rng != pb->rng || pa->seed != pb->seed) return fail(TRW_E_INVALID, "both kinds must share rng and seed");
    if (split < 0 || split > io->n_walkers) return fail(TRW_E_INVALID, "split outside [0, n_walkers]");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    // A graph holds one table per sampler: two kinds sampling the same static table from different
    // weights (e.g. PPR's unit-weight alias and DeepWalk's edge-weight alias) cannot share a batch.
    if (reads_static_table(pa) && reads_static_table(pb) && pa->sampling == pb->sampling &&
        table_weights(pa) != table_weights(pb))
        return fail(TRW_E_UNSUPPORTED, "both kinds sample the static table of sampler %d, from different weights; "
                                       "a graph holds one table per sampler", pa->sampling);
    KindParams a, b;
    int rc = resolve_kind(g, pa, a);
    if (rc != TRW_OK) return rc;
    rc = resolve_kind(g, pb, b);
    if (rc != TRW_OK) return rc;
    // Two MetaPath kinds with different schemas: the graph keeps the records of the one resolved
    // last, and the other reads the general label-ordered records.
    if (a.schema_rec) a.schema_rec = g->msrec_key == schema_key(pa);
    if (b.schema_rec) b.schema_rec = g->msrec_key == schema_key(pb);
    BatchIO bio;
    rc = make_io(g, pa, io, a, b, bio);
    if (rc != TRW_OK) return rc;
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : g->ctx->stream;
    const int32_t form = pa->form == TRW_FORM_AUTO ? TRW_FORM_QUEUE : pa->form;
    if (form == TRW_FORM_QUEUE) {
        TRW_HIP(trw::launch_walk_queue(g->view(), a, b, split, bio, g->counters, st), "queue walk launch");
        return TRW_OK;
    }
    if (form == TRW_FORM_STEP) return run_step_form(g, a, b, split, bio, st);
    if (form == TRW_FORM_GRAPH) return run_graph_form(g, a, b, split, bio, st);
    return fail(TRW_E_UNSUPPORTED, "mixed batches run in the QUEUE, STEP or GRAPH form");
}

int trw_walk_host(trw_graph *g, const trw_walk_params *p, const int32_t *sources, int32_t n, const uint32_t *draws,
                  int64_t K, const uint32_t *tdraws, int64_t KT, int32_t *paths, int32_t *lengths, int32_t *ends,
                  int32_t *status, uint64_t *steps) {
    if (!g || !p || (n > 0 && !sources)) return fail(TRW_E_INVALID, "null argument");
    if (n < 0) return fail(TRW_E_INVALID, "n_walkers must be >= 0");
    TRW_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    const size_t N = static_cast<size_t>(n);
    const int32_t L = std::max(p->length, 1);
    const bool replay = p->rng == TRW_RNG_REPLAY;
    ScratchBuffers buf;
    int32_t *d_src = nullptr, *d_paths = nullptr, *d_len = nullptr, *d_end = nullptr, *d_st = nullptr;
    uint32_t *d_draws = nullptr, *d_tdraws = nullptr;
    uint64_t *d_cnt = nullptr;
    TRW_HIP(buf.get(&d_src, N), "hipMalloc");
    TRW_HIP(buf.get(&d_cnt, 4), "hipMalloc");
    if (paths) TRW_HIP(buf.get(&d_paths, N * static_cast<size_t>(L)), "hipMalloc");
    if (lengths) TRW_HIP(buf.get(&d_len, N), "hipMalloc");
    if (ends) TRW_HIP(buf.get(&d_end, N), "hipMalloc");
    if (status) TRW_HIP(buf.get(&d_st, N), "hipMalloc");
    hipStream_t st = g->ctx->stream;
    if (N) TRW_HIP(hipMemcpyAsync(d_src, sources, sizeof(int32_t) * N, hipMemcpyHostToDevice, st), "copy sources");
    if (replay && K > 0 && draws) {
        TRW_HIP(buf.get(&d_draws, N * static_cast<size_t>(K)), "hipMalloc draws");
        TRW_HIP(hipMemcpyAsync(d_draws, draws, sizeof(uint32_t) * N * static_cast<size_t>(K), hipMemcpyHostToDevice, st),
                "copy draws");
    }
    if (replay && KT > 0 && tdraws) {
        TRW_HIP(buf.get(&d_tdraws, N * static_cast<size_t>(KT)), "hipMalloc draws");
        TRW_HIP(hipMemcpyAsync(d_tdraws, tdraws, sizeof(uint32_t) * N * static_cast<size_t>(KT), hipMemcpyHostToDevice,
                               st), "copy draws");
    }
    TRW_HIP(hipMemsetAsync(d_cnt, 0, sizeof(uint64_t) * 4, st), "clear counters");
    trw_walk_io io;
    std::memset(&io, 0, sizeof(io));
    io.sources = d_src;
    io.n_walkers = n;
    io.path_stride = L;
    io.draws = d_draws;
    io.draws_per_walker = K;
    io.term_draws = d_tdraws;
    io.term_draws_per_walker = KT;
    io.paths = d_paths;
    io.lengths = d_len;
    io.ends = d_end;
    io.status = d_st;
    io.counters = d_cnt;
    int rc = trw_walk_async(g, p, &io, st);
    if (rc != TRW_OK) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    uint64_t cnt[4] = {0, 0, 0, 0};
    TRW_HIP(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, st), "copy counters");
    if (paths) TRW_HIP(hipMemcpyAsync(paths, d_paths, sizeof(int32_t) * N * L, hipMemcpyDeviceToHost, st), "copy paths");
    if (lengths) TRW_HIP(hipMemcpyAsync(lengths, d_len, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st), "copy lengths");
    if (ends) TRW_HIP(hipMemcpyAsync(ends, d_end, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st), "copy ends");
    if (status) TRW_HIP(hipMemcpyAsync(status, d_st, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st), "copy status");
    TRW_HIP(hipStreamSynchronize(st), "walk");
    if (steps) *steps = cnt[0];
    return TRW_OK;
}

// ---------------------------------------------------------------- replay stream

int trw_replay_fill(uint32_t seed, int64_t first_walker, int32_t n, int64_t K, uint32_t *out) {
    if (n < 0 || K < 0 || first_walker < 0 || (n > 0 && K > 0 && !out)) return fail(TRW_E_INVALID, "bad argument");
    if (n == 0 || K == 0) return TRW_OK;
    trw::Sfmt19937 gen(seed);
    gen.skip(static_cast<uint64_t>(first_walker) * static_cast<uint64_t>(K));
    const int32_t kTileW = 64;
    std::vector<uint32_t> chunk(static_cast<size_t>(kTileW) * static_cast<size_t>(K));
    for (int32_t w0 = 0; w0 < n; w0 += kTileW) {
        const int32_t tw = std::min(kTileW, n - w0);
        gen.fill(chunk.data(), static_cast<int64_t>(tw) * K);
        for (int64_t k = 0; k < K; ++k) {
            uint32_t *row = out + k * n + w0;
            for (int32_t i = 0; i < tw; ++i) row[i] = chunk[static_cast<size_t>(i) * K + k];
        }
    }
    return TRW_OK;
}

// ---------------------------------------------------------------- walker sources

int trw_sources_all(const int32_t *degree, int32_t nv, int32_t *out, int32_t *count) {
    if (!degree || !out || !count || nv < 0) return fail(TRW_E_INVALID, "bad argument");
    std::vector<int32_t> src;
    src.reserve(static_cast<size_t>(nv));
    for (int32_t v = 0; v < nv; ++v)
        if (degree[v] != 0) src.push_back(v);
    std::mt19937 rng(0);  // Utility::shuffle with debug_mode (utility.h:96-100)
    std::shuffle(src.begin(), src.end(), rng);
    std::copy(src.begin(), src.end(), out);
    *count = static_cast<int32_t>(src.size());
    return TRW_OK;
}

int trw_sources_single(const int32_t *degree, int32_t nv, int32_t n, int32_t *out) {
    if (!degree || (!out && n > 0) || nv <= 0 || n < 0) return fail(TRW_E_INVALID, "bad argument");
    int32_t u = 0, d = degree[0];
    for (int32_t v = 1; v < nv; ++v)
        if (degree[v] > d) {
            d = degree[v];
            u = v;
        }
    for (int32_t i = 0; i < n; ++i) out[i] = u;
    return TRW_OK;
}

int trw_sources_random(const int32_t *degree, int32_t nv, int32_t n, uint32_t seed, int32_t *out) {
    if (!degree || (!out && n > 0) || nv <= 0 || n < 0) return fail(TRW_E_INVALID, "bad argument");
    bool any = false;
    for (int32_t v = 0; v < nv && !any; ++v) any = degree[v] != 0;
    if (!any && n > 0) return fail(TRW_E_INVALID, "every vertex is isolated");
    trw::Sfmt19937 gen(seed);
    for (int32_t i = 0; i < n;) {
        const int32_t u = static_cast<int32_t>(gen.next() % static_cast<uint32_t>(nv));
        if (degree[u] != 0) out[i++] = u;
    }
    return TRW_OK;
}

// ---------------------------------------------------------------- host graphs

int trw_host_graph_chung_lu(int32_t nv, int64_t m, double gamma, double dmin, double dmax, uint64_t seed,
                            int32_t with_weights, int32_t n_labels, trw_host_graph **out) {
    if (!out) return fail(TRW_E_INVALID, "out is null");
    *out = nullptr;
    std::unique_ptr<trw_host_graph> h(new trw_host_graph);
    std::string err;
    if (!trw::chung_lu(nv, m, gamma, dmin, dmax, seed, with_weights != 0, n_labels, h->g, err))
        return fail(TRW_E_INVALID, "%s", err.c_str());
    *out = h.release();
    return TRW_OK;
}

int trw_host_graph_load(const char *dir, int32_t with_weights, int32_t with_labels, trw_host_graph **out) {
    if (!dir || !out) return fail(TRW_E_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<trw_host_graph> h(new trw_host_graph);
    std::string err;
    if (!trw::load_reference_csr(dir, with_weights != 0, with_labels != 0, h->g, err))
        return fail(TRW_E_IO, "%s", err.c_str());
    *out = h.release();
    return TRW_OK;
}

int trw_host_graph_load_edge_list(const char *dir, int32_t skip, trw_host_graph **out) {
    if (!dir || !out) return fail(TRW_E_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<trw_host_graph> h(new trw_host_graph);
    std::string err;
    if (!trw::load_edge_list(dir, static_cast<char>(skip), h->g, err)) return fail(TRW_E_IO, "%s", err.c_str());
    *out = h.release();
    return TRW_OK;
}

int trw_host_graph_store(const trw_host_graph *h, const char *dir) {
    if (!h || !dir) return fail(TRW_E_INVALID, "null argument");
    std::string err;
    if (!trw::store_reference_csr(h->g, dir, err)) return fail(TRW_E_IO, "%s", err.c_str());
    return TRW_OK;
}

int trw_host_graph_arrays(const trw_host_graph *h, int32_t *nv, int64_t *ne, const int32_t **degree,
                          const int32_t **adj, const double **w, const int32_t **label) {
    if (!h) return fail(TRW_E_INVALID, "null argument");
    if (nv) *nv = h->g.nv;
    if (ne) *ne = h->g.ne;
    if (degree) *degree = h->g.degree.data();
    if (adj) *adj = h->g.adj.data();
    if (w) *w = h->g.weight.empty() ? nullptr : h->g.weight.data();
    if (label) *label = h->g.label.empty() ? nullptr : h->g.label.data();
    return TRW_OK;
}

int trw_host_graph_free(trw_host_graph *h) {
    delete h;
    return TRW_OK;
}

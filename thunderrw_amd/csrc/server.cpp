// Query server: continuous batching of walk queries over one resident graph.
//
// The reference has no serving path; its compute<F> (random_walk/creeper.h:689-831)
// runs one blocking batch.  Here a dispatcher thread owns the device side:
//   * submit() appends a query (copied) to a FIFO of pending queries;
//   * whenever a slot is free, the dispatcher packs pending queries that share the
//     head query's walk parameters, in arrival order, into one batch (up to the
//     slot's walker, path-byte, query and source-pool capacities; a query that does
//     not fit whole is split and its remainder stays at the head);
//   * a batch crosses PCIe as a segment table plus the queries' source sets, is
//     expanded on the device into per-walker start vertices and Philox ids, walks
//     on the slot's own stream, and its per-walker outputs come back into the
//     slot's pinned staging;
//   * a PPR batch is reduced on the device to each segment's ascending end vertices
//     with walker counts (per-segment bitmaps over the vertex ids, no sort), laid
//     end to end; the device writes the segments' step, error and run totals into
//     the slot's pinned staging, and the completion thread copies back exactly the
//     runs by DMA; a walk batch copies back its per-walker outputs;
//   * the dispatcher polls the slots' completion events; a finished slot is handed
//     to a completion thread, which hands each segment's results to its query
//     (merging the runs of a query split over several batches) and frees the slot.
// Batches on different slots overlap on the device.  Each walker's random stream is
// keyed by (seed, walker_id_base + index in its query), so a result does not depend
// on how queries were batched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "internal.h"

using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr int32_t kPoolInts = 1 << 22;  // source-set pool of one batch (16 MiB)
// A slot's control block in u32 words, zeroed by each batch's expand launch: the walk's
// counters in words [0, 16) (64 bytes), the QUEUE form's work-queue head at kCtlHead (its own
// 64 bytes), then one PPR ticket per segment from kCtlTickets.
constexpr int32_t kCtlHead = 16;
constexpr int32_t kCtlTickets = 32;
// Bound on (segments of a PPR batch) x (chunks of 2^17 vertex ids): the per-bucket
// arrays of a slot (8 MiB each at most).
constexpr int64_t kPprBuckets = int64_t(1) << 20;

// One PPR run: an end vertex and its walker count (the device's pair layout).
struct Run {
    int32_t vertex;
    uint32_t count;
};
static_assert(sizeof(Run) == sizeof(int2), "runs are copied from the device's (vertex, count) pairs");

struct Query {
    uint64_t id = 0;
    trw_walk_params params{};
    int32_t result = TRW_RESULT_PPR;
    int32_t path_len = 1;
    int32_t n_walkers = 0;
    uint64_t id_base = 0;
    std::vector<int32_t> sources;
    // Owned by the dispatcher.
    int32_t next_walker = 0;
    // Completion threads write disjoint walker ranges; the one that completes the
    // query builds its counts.
    std::atomic<int32_t> done_walkers{0};
    std::atomic<uint64_t> steps{0}, errors{0};
    std::vector<int32_t> ends, lengths, status, paths;  // TRW_RESULT_PATHS queries
    std::mutex merge_mu;                                // PPR runs of a query split over batches
    std::vector<Run> ppr;                               // ascending end vertices with walker counts
    Clock::time_point device_done{};                    // latest batch of its walkers seen done (merge_mu)
    // Guarded by the server mutex.
    Clock::time_point arrival, dispatched, finished;
    bool was_dispatched = false;
    int state = TRW_QUERY_PENDING;
    int error_code = TRW_OK;
    std::string error;
};

bool same_walk(const Query &a, const Query &b) {
    const trw_walk_params &x = a.params, &y = b.params;
    if (a.result != b.result) return false;
    if (x.app != y.app || x.execution != y.execution || x.sampling != y.sampling || x.length != y.length ||
        x.rng != y.rng || x.seed != y.seed || x.form != y.form || x.max_trials != y.max_trials ||
        x.walkers_per_thread != y.walkers_per_thread || x.weights != y.weights ||
        x.schema_length != y.schema_length)
        return false;
    if (std::memcmp(&x.stop_probability, &y.stop_probability, sizeof(double)) != 0 ||
        std::memcmp(&x.p, &y.p, sizeof(double)) != 0 || std::memcmp(&x.q, &y.q, sizeof(double)) != 0)
        return false;
    return std::equal(x.schema, x.schema + std::max(0, std::min(x.schema_length, 16)), y.schema);
}

// Adds the ascending runs b[0, nb) into the ascending runs a.
void merge_runs(std::vector<Run> &a, const Run *b, size_t nb) {
    std::vector<Run> m;
    m.reserve(a.size() + nb);
    size_t i = 0, j = 0;
    while (i < a.size() || j < nb) {
        if (j == nb || (i < a.size() && a[i].vertex < b[j].vertex)) {
            m.push_back(a[i++]);
        } else if (i == a.size() || b[j].vertex < a[i].vertex) {
            m.push_back(b[j++]);
        } else {
            m.push_back({a[i].vertex, a[i].count + b[j].count});
            ++i;
            ++j;
        }
    }
    a.swap(m);
}

struct Segment {
    std::shared_ptr<Query> q;
    int32_t first = 0;
    int32_t count = 0;
    int32_t batch_offset = 0;
};

enum class SlotState { Free, InFlight, Draining };

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // A batch's segment table and then its source pool, one copy to the device.
    unsigned char *d_stage = nullptr, *h_stage = nullptr;
    int32_t *d_sources = nullptr;
    uint64_t *d_ids = nullptr;
    int32_t *d_ends = nullptr, *d_lengths = nullptr, *d_status = nullptr, *d_paths = nullptr;
    // Control block zeroed by the expand launch (walk counters, kCtlHead, kCtlTickets).
    uint32_t *d_ctl = nullptr;
    int32_t *h_ends = nullptr, *h_lengths = nullptr, *h_status = nullptr, *h_paths = nullptr;
    // PPR aggregation: device scratch, and the pinned memory the device writes each
    // segment's totals, each bucket's {first position, runs} and the runs into.
    trw::PprScratch agg{};
    trw::PprSegmentStats *h_stats = nullptr;
    uint2 *h_buckets = nullptr;
    int2 *h_pairs = nullptr;
    std::vector<Segment> segs;
    int32_t n = 0;
    int32_t path_stride = 0;
    int32_t pool_used = 0;
    bool ppr = false;  // the batch's queries want end-vertex counts
    SlotState state = SlotState::Free;
    Clock::time_point done_at{};  // when the dispatcher saw the batch done on the device
};

class WorkerPool {
  public:
    explicit WorkerPool(int n) {
        for (int i = 0; i < n; ++i) threads_.emplace_back([this] { run(); });
    }
    ~WorkerPool() { stop(); }
    void push(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            tasks_.push_back(std::move(f));
        }
        cv_.notify_one();
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stopping_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_)
            if (t.joinable()) t.join();
        threads_.clear();
    }

  private:
    void run() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stopping_ || !tasks_.empty(); });
                if (tasks_.empty()) return;
                f = std::move(tasks_.front());
                tasks_.pop_front();
            }
            f();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> tasks_;
    std::vector<std::thread> threads_;
    bool stopping_ = false;
};

}  // namespace

struct trw_server {
    trw_graph *g = nullptr;
    int32_t device = 0;
    int32_t max_walkers = 0;
    int64_t path_ints = 0;
    int32_t max_pending = 0;
    int32_t max_segments = 0;
    int32_t ppr_segments = 0;  // segments of a PPR batch
    int32_t ppr_chunks = 0;    // chunks of 2^17 vertex ids the PPR aggregation buckets by
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv;       // dispatcher: new work, a slot freed, stop
    std::condition_variable cv_done;  // waiters: a query finished
    std::deque<std::shared_ptr<Query>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Query>> queries;
    uint64_t next_id = 1;
    bool stopping = false;
    trw_server_stats stats{};
    uint64_t batch_walkers = 0, batch_queries = 0;

    std::thread dispatcher;
    std::unique_ptr<WorkerPool> workers;

    int submit(const trw_query *q, Clock::time_point arrival, uint64_t *id_out);
    void dispatch_loop();
    bool form_batch(Slot &sl, std::unique_lock<std::mutex> &lk);
    void launch(Slot &sl);
    void drain(Slot *sl, hipError_t err, std::string msg);
    void finalize(const std::shared_ptr<Query> &q);
    int warm_up();
    void free_slots();
};

// ---------------------------------------------------------------- dispatcher

bool trw_server::form_batch(Slot &sl, std::unique_lock<std::mutex> &) {
    // Called with mu held and a non-empty pending queue.
    const std::shared_ptr<Query> head = pending.front();
    const int32_t L = head->result == TRW_RESULT_PATHS ? head->path_len : 0;
    const int32_t seg_cap = head->result == TRW_RESULT_PPR ? ppr_segments : max_segments;
    int64_t room_paths = L > 0 ? path_ints / L : INT64_MAX;
    int32_t n = 0, pool = 0;
    sl.segs.clear();
    const auto now = Clock::now();
    for (auto it = pending.begin(); it != pending.end();) {
        const std::shared_ptr<Query> &q = *it;
        if (q != head && !same_walk(*q, *head)) {
            ++it;
            continue;
        }
        if (static_cast<int32_t>(sl.segs.size()) == seg_cap) break;
        if (pool + static_cast<int32_t>(q->sources.size()) > kPoolInts) break;
        const int64_t room = std::min<int64_t>(max_walkers - n, room_paths - n);
        if (room <= 0) break;
        const int32_t take = static_cast<int32_t>(std::min<int64_t>(q->n_walkers - q->next_walker, room));
        Segment s;
        s.q = q;
        s.first = q->next_walker;
        s.count = take;
        s.batch_offset = n;
        sl.segs.push_back(s);
        n += take;
        pool += static_cast<int32_t>(q->sources.size());
        q->next_walker += take;
        if (!q->was_dispatched) {
            q->was_dispatched = true;
            q->dispatched = now;
            q->state = TRW_QUERY_RUNNING;
        }
        if (q->next_walker == q->n_walkers) it = pending.erase(it);
        else break;  // the slot is full; the rest of this query waits at the head
    }
    sl.n = n;
    sl.path_stride = L;
    sl.pool_used = pool;
    sl.ppr = head->result == TRW_RESULT_PPR;
    ++stats.batches;
    batch_walkers += static_cast<uint64_t>(n);
    batch_queries += sl.segs.size();
    return !sl.segs.empty();
}

void trw_server::launch(Slot &sl) {
    // Runs without mu: the slot is InFlight and owned by the dispatcher until its
    // completion event fires.
    const size_t seg_bytes = sizeof(trw::ServeSegment) * sl.segs.size();  // a multiple of 16
    trw::ServeSegment *h_segs = reinterpret_cast<trw::ServeSegment *>(sl.h_stage);
    int32_t *h_pool = reinterpret_cast<int32_t *>(sl.h_stage + seg_bytes);
    const trw::ServeSegment *d_segs = reinterpret_cast<const trw::ServeSegment *>(sl.d_stage);
    const int32_t *d_pool = reinterpret_cast<const int32_t *>(sl.d_stage + seg_bytes);
    int32_t pool = 0, tiles = 0;
    for (size_t i = 0; i < sl.segs.size(); ++i) {
        const Segment &s = sl.segs[i];
        trw::ServeSegment d;
        d.batch_offset = s.batch_offset;
        d.count = s.count;
        d.pool_offset = pool;
        d.n_sources = static_cast<int32_t>(s.q->sources.size());
        d.first_walker = s.first;
        d.first_tile = tiles;
        if (sl.ppr) tiles += trw::ppr_tiles(s.count);
        d.id_base = s.q->id_base;
        h_segs[i] = d;
        std::memcpy(h_pool + pool, s.q->sources.data(), sizeof(int32_t) * s.q->sources.size());
        pool += d.n_sources;
    }
    const int32_t nseg = static_cast<int32_t>(sl.segs.size());
    const trw_walk_params params = sl.segs.front().q->params;
    const size_t N = static_cast<size_t>(sl.n);
    hipStream_t st = sl.stream;
    std::string msg;
    hipError_t e = hipMemcpyAsync(sl.d_stage, sl.h_stage, seg_bytes + sizeof(int32_t) * static_cast<size_t>(pool),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = trw::launch_serve_expand(d_segs, nseg, d_pool, sl.n, sl.d_sources, sl.d_ids, sl.d_ctl, kCtlTickets + nseg,
                                     st);
    if (e == hipSuccess) {
        trw_walk_io io;
        std::memset(&io, 0, sizeof(io));
        io.sources = sl.d_sources;
        io.n_walkers = sl.n;
        io.path_stride = sl.path_stride;
        io.walker_ids = sl.d_ids;
        io.paths = sl.path_stride > 0 ? sl.d_paths : nullptr;
        io.lengths = sl.d_lengths;
        io.ends = sl.d_ends;
        io.status = sl.d_status;
        io.counters = reinterpret_cast<uint64_t *>(sl.d_ctl);
        const int rc = trw::walk_launch(g, &params, &io, st, sl.d_ctl + kCtlHead, false, true);
        if (rc != TRW_OK) {
            e = hipErrorLaunchFailure;
            msg = trw_last_error();
        }
    }
    if (sl.ppr) {
        // The device writes the totals, bucket extents and runs straight into the pinned staging.
        if (e == hipSuccess)
            e = trw::launch_serve_ppr_reduce(d_segs, nseg, tiles, sl.d_ends, sl.d_lengths, sl.d_status, sl.agg,
                                             sl.h_stats, sl.h_buckets, sl.h_pairs, st);
    } else {
        if (e == hipSuccess) e = hipMemcpyAsync(sl.h_ends, sl.d_ends, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sl.h_lengths, sl.d_lengths, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sl.h_status, sl.d_status, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && sl.path_stride > 0)
            e = hipMemcpyAsync(sl.h_paths, sl.d_paths, sizeof(int32_t) * N * static_cast<size_t>(sl.path_stride),
                               hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipEventRecord(sl.done, st);
    if (e != hipSuccess) {
        // Let whatever was queued finish before the staging buffers are reused.
        (void)hipStreamSynchronize(st);
        if (msg.empty()) msg = std::string("serving batch launch: ") + hipGetErrorString(e);
        std::lock_guard<std::mutex> lk(mu);
        sl.state = SlotState::Draining;
        sl.done_at = Clock::now();
        Slot *p = &sl;
        workers->push([this, p, e, msg] { drain(p, e, msg); });
    }
}

void trw_server::dispatch_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progressed = false;
        bool in_flight = false;
        for (Slot &sl : slots) {
            if (sl.state != SlotState::InFlight) continue;
            const hipError_t e = hipEventQuery(sl.done);
            if (e == hipErrorNotReady) {
                in_flight = true;
                continue;
            }
            sl.state = SlotState::Draining;
            sl.done_at = Clock::now();
            Slot *p = &sl;
            std::string msg = e == hipSuccess ? std::string() : std::string("serving batch: ") + hipGetErrorString(e);
            workers->push([this, p, e, msg] { drain(p, e, msg); });
            progressed = true;
        }
        while (!pending.empty()) {
            Slot *free = nullptr;
            for (Slot &sl : slots)
                if (sl.state == SlotState::Free) {
                    free = &sl;
                    break;
                }
            if (!free) break;
            if (!form_batch(*free, lk)) break;
            free->state = SlotState::InFlight;
            in_flight = true;
            progressed = true;
            lk.unlock();
            launch(*free);
            lk.lock();
        }
        if (stopping && pending.empty()) {
            bool idle = true;
            for (const Slot &sl : slots) idle = idle && sl.state == SlotState::Free;
            if (idle) return;
        }
        if (progressed) continue;
        if (in_flight) {
            // Poll the completion events; yielding keeps the latency of a finished
            // batch to a few microseconds without a HIP callback thread.
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            cv.wait(lk);
        }
    }
}

// ---------------------------------------------------------------- completion

void trw_server::drain(Slot *sl, hipError_t err, std::string msg) {
    bool ok = err == hipSuccess && msg.empty();
    const size_t C = static_cast<size_t>(ppr_chunks);
    if (ok && sl->ppr) {
        // A segment's buckets lie inside its walker positions, one after another.
        for (size_t j = 0; j < sl->segs.size() && ok; ++j) {
            const uint64_t lo = static_cast<uint64_t>(sl->segs[j].batch_offset);
            const uint64_t hi = lo + static_cast<uint64_t>(sl->segs[j].count);
            uint64_t at = lo;
            for (size_t c = 0; c < C; ++c) {
                const uint2 b = sl->h_buckets[j * C + c];
                if (b.x < at || static_cast<uint64_t>(b.x) + b.y > hi) {
                    ok = false;
                    msg = "serving batch: inconsistent PPR buckets";
                    break;
                }
                at = static_cast<uint64_t>(b.x) + b.y;
            }
        }
    }
    for (size_t j = 0; j < sl->segs.size(); ++j) {
        const Segment &s = sl->segs[j];
        Query &q = *s.q;
        if (ok && sl->ppr) {
            // Segment j's runs, ascending: its buckets' runs in chunk order.
            const trw::PprSegmentStats &st = sl->h_stats[j];
            const uint2 *bk = sl->h_buckets + j * C;
            size_t nr = 0;
            for (size_t c = 0; c < C; ++c) nr += bk[c].y;
            // Appended bucket by bucket into reserved storage: a sized vector would first zero
            // every run (half a megabyte for a query of 10^5 walkers) only to overwrite it.
            std::vector<Run> runs;
            runs.reserve(nr);
            for (size_t c = 0; c < C; ++c) {
                const Run *from = reinterpret_cast<const Run *>(sl->h_pairs + bk[c].x);
                runs.insert(runs.end(), from, from + bk[c].y);
            }
            {
                std::lock_guard<std::mutex> lk(q.merge_mu);
                if (q.ppr.empty()) q.ppr.swap(runs);
                else merge_runs(q.ppr, runs.data(), nr);
                q.device_done = std::max(q.device_done, sl->done_at);
            }
            q.steps += st.steps;
            q.errors += st.errors;
        } else if (ok) {
            const size_t o = static_cast<size_t>(s.batch_offset), f = static_cast<size_t>(s.first),
                         c = static_cast<size_t>(s.count);
            std::memcpy(q.ends.data() + f, sl->h_ends + o, sizeof(int32_t) * c);
            std::memcpy(q.lengths.data() + f, sl->h_lengths + o, sizeof(int32_t) * c);
            std::memcpy(q.status.data() + f, sl->h_status + o, sizeof(int32_t) * c);
            if (sl->path_stride > 0) {
                const size_t L = static_cast<size_t>(sl->path_stride);
                std::memcpy(q.paths.data() + f * L, sl->h_paths + o * L, sizeof(int32_t) * c * L);
            }
            uint64_t steps = 0, errors = 0;
            for (size_t i = 0; i < c; ++i) {
                steps += static_cast<uint64_t>(std::max(sl->h_lengths[o + i] - 1, 0));
                errors += sl->h_status[o + i] >= TRW_WALKER_REPLAY_EXHAUSTED;
            }
            q.steps += steps;
            q.errors += errors;
            std::lock_guard<std::mutex> lk(q.merge_mu);
            q.device_done = std::max(q.device_done, sl->done_at);
        } else {
            std::lock_guard<std::mutex> lk(mu);
            if (q.error_code == TRW_OK) {
                q.error_code = TRW_E_HIP;
                q.error = msg;
            }
        }
        if (q.done_walkers.fetch_add(s.count) + s.count == q.n_walkers) finalize(s.q);
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        sl->segs.clear();
        sl->state = SlotState::Free;
    }
    cv.notify_one();
}

void trw_server::finalize(const std::shared_ptr<Query> &qp) {
    Query &q = *qp;
    bool failed;
    {
        std::lock_guard<std::mutex> lk(mu);
        failed = q.error_code != TRW_OK;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        q.finished = Clock::now();
        q.state = failed ? TRW_QUERY_FAILED : TRW_QUERY_DONE;
        ++stats.queries_done;
        stats.walkers += static_cast<uint64_t>(q.n_walkers);
        stats.steps += q.steps.load();
    }
    cv_done.notify_all();
}

// ---------------------------------------------------------------- submit

int trw_server::submit(const trw_query *q, Clock::time_point arrival, uint64_t *id_out) {
    if (!q) return trw::set_error(TRW_E_INVALID, "query is null");
    if (q->n_walkers < 0) return trw::set_error(TRW_E_INVALID, "n_walkers must be >= 0");
    if (q->n_walkers > 0 && (q->n_sources <= 0 || !q->sources))
        return trw::set_error(TRW_E_INVALID, "a query with walkers needs at least one source");
    if (q->n_sources > kPoolInts)
        return trw::set_error(TRW_E_INVALID, "at most %d sources per query", kPoolInts);
    if (q->result != TRW_RESULT_PPR && q->result != TRW_RESULT_PATHS)
        return trw::set_error(TRW_E_INVALID, "unknown result kind %d", q->result);
    if (q->params.rng != TRW_RNG_PHILOX)
        return trw::set_error(TRW_E_UNSUPPORTED, "served queries draw from the Philox stream (TRW_RNG_PHILOX)");
    // Slots run side by side on their own streams; the STEP and GRAPH forms keep walker state in
    // graph-wide scratch, so only the forms whose state stays in registers are served.
    if (q->params.form != TRW_FORM_AUTO && q->params.form != TRW_FORM_FUSED && q->params.form != TRW_FORM_QUEUE)
        return trw::set_error(TRW_E_UNSUPPORTED, "the server runs the AUTO, FUSED or QUEUE form (got %d)",
                              q->params.form);
    int rc = trw::walk_check(g, &q->params);
    if (rc != TRW_OK) return rc;
    const int32_t L = std::max(q->params.length, 1);
    if (q->result == TRW_RESULT_PATHS && L > path_ints)
        return trw::set_error(TRW_E_INVALID, "one path row (%d ints) exceeds the batch path capacity", L);
    for (int32_t i = 0; i < q->n_sources; ++i)
        if (q->sources[i] < 0 || q->sources[i] >= g->nv)
            return trw::set_error(TRW_E_INVALID, "source %d outside [0, %d)", q->sources[i], g->nv);

    auto qp = std::make_shared<Query>();
    qp->params = q->params;
    qp->result = q->result;
    qp->path_len = L;
    qp->n_walkers = q->n_walkers;
    qp->id_base = q->walker_id_base;
    qp->sources.assign(q->sources, q->sources + std::max(q->n_sources, 0));
    const size_t N = static_cast<size_t>(q->n_walkers);
    if (q->result == TRW_RESULT_PATHS) {
        qp->ends.resize(N);
        qp->lengths.resize(N);
        qp->status.resize(N);
        qp->paths.resize(N * static_cast<size_t>(L));
    }
    qp->arrival = arrival;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (stopping) return trw::set_error(TRW_E_STATE, "the server is shutting down");
        if (static_cast<int64_t>(pending.size()) >= max_pending) {
            ++stats.rejected;
            return trw::set_error(TRW_E_BUSY, "%d queries already wait for admission", max_pending);
        }
        qp->id = next_id++;
        queries[qp->id] = qp;
        if (N == 0) {
            qp->state = TRW_QUERY_DONE;
            qp->dispatched = qp->finished = Clock::now();
            ++stats.queries_done;
        } else {
            pending.push_back(qp);
        }
        if (id_out) *id_out = qp->id;
    }
    cv.notify_one();
    cv_done.notify_all();
    return TRW_OK;
}

void trw_server::free_slots() {
    for (Slot &sl : slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        (void)hipFree(sl.d_stage);
        (void)hipFree(sl.d_sources);
        (void)hipFree(sl.d_ids);
        (void)hipFree(sl.d_ends);
        (void)hipFree(sl.d_lengths);
        (void)hipFree(sl.d_status);
        (void)hipFree(sl.d_paths);
        (void)hipFree(sl.d_ctl);
        (void)hipHostFree(sl.h_stage);
        (void)hipHostFree(sl.h_ends);
        (void)hipHostFree(sl.h_lengths);
        (void)hipHostFree(sl.h_status);
        (void)hipHostFree(sl.h_paths);
        (void)hipFree(sl.agg.tile_hist);
        (void)hipFree(sl.agg.tile_stats);
        (void)hipFree(sl.agg.bucketed);
        (void)hipFree(sl.agg.pairs);
        (void)hipFree(sl.agg.bucket);
        (void)hipHostFree(sl.h_stats);
        (void)hipHostFree(sl.h_buckets);
        (void)hipHostFree(sl.h_pairs);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        sl = Slot();
    }
}

namespace {

template <class T>
hipError_t dev(T **p, size_t count) {
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * std::max<size_t>(count, 1));
}

template <class T>
hipError_t pinned(T **p, size_t count) {
    return hipHostMalloc(reinterpret_cast<void **>(p), sizeof(T) * std::max<size_t>(count, 1), hipHostMallocDefault);
}

std::shared_ptr<Query> find_query(trw_server *s, uint64_t id) {
    auto it = s->queries.find(id);
    return it == s->queries.end() ? nullptr : it->second;
}

void fill_info(const Query &q, trw_query_info *info) {
    std::memset(info, 0, sizeof(*info));
    info->state = q.state;
    info->n_walkers = q.n_walkers;
    const bool finished = q.state == TRW_QUERY_DONE || q.state == TRW_QUERY_FAILED;
    info->steps = finished ? q.steps.load() : 0;
    info->errors = finished ? q.errors.load() : 0;
    info->ppr_entries = finished ? static_cast<int64_t>(q.ppr.size()) : 0;
    info->queue_ms = q.was_dispatched || finished ? ms_between(q.arrival, q.dispatched) : -1.0;
    info->latency_ms = finished ? ms_between(q.arrival, q.finished) : -1.0;
    // Written by the completion threads before the one that finished the query took the server
    // mutex to mark it done; read here under that mutex.
    info->device_ms = finished && q.was_dispatched ? ms_between(q.dispatched, q.device_done) : -1.0;
}

}  // namespace

// ---------------------------------------------------------------- C ABI

// The first time many full-size batches are in flight together, the runtime grows what it keeps
// for them, once (measured on the 2^24-vertex serving graph: one stall of 7 - 12 ms inside the
// latency of the batches then in flight, the first time an open-loop run reached 25,000 or 50,000
// queries/s, and none after a burst of full-size batches on every slot).  Before the first query
// every slot runs two full-size PPR batches at once (uniform moves, which need no table, from the
// highest-degree vertex); then the statistics start from zero.
int trw_server::warm_up() {
    if (g->nv <= 0) return TRW_OK;  // no vertex to start from
    // Sources: the graph's vertices in degree order (most of them non-isolated), as many as a
    // batch's queries bring at most, so that batches carry many segments and a full source pool.
    const int32_t per_query = std::min<int32_t>(std::min<int32_t>(kPoolInts / std::max(max_segments, 1), 4096), g->nv);
    std::vector<int32_t> src(static_cast<size_t>(std::max(per_query, 1)), 0);
    if (g->order && per_query > 0 &&
        hipMemcpy(src.data(), g->order, sizeof(int32_t) * src.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return trw::set_error(TRW_E_HIP, "reading the graph's degree order");
    const int32_t queries_per_batch = std::max(1, std::min(max_segments, 64));
    trw_query wq;
    std::memset(&wq, 0, sizeof(wq));
    wq.sources = src.data();
    wq.n_sources = static_cast<int32_t>(src.size());
    wq.n_walkers = std::max(1, max_walkers / queries_per_batch);
    wq.result = TRW_RESULT_PPR;
    wq.params.app = TRW_APP_PPR;
    wq.params.execution = TRW_EXEC_UNIFORM;
    wq.params.sampling = TRW_SM_UNIFORM;
    wq.params.length = 1;
    wq.params.stop_probability = 0.2;
    wq.params.rng = TRW_RNG_PHILOX;
    wq.params.form = TRW_FORM_QUEUE;
    // Submitted as fast as admission takes them (a full pending list: the oldest is waited for).
    const size_t total = slots.size() * 2 * static_cast<size_t>(queries_per_batch);
    std::deque<uint64_t> open;
    int rc = TRW_OK;
    const auto retire = [&] {
        trw_query_info info;
        const int w = trw_server_wait(this, open.front(), -1.0, &info);
        if (rc == TRW_OK && w != TRW_OK) rc = w;
        (void)trw_server_release(this, open.front());
        open.pop_front();
    };
    for (size_t i = 0; i < total && rc == TRW_OK;) {
        uint64_t id = 0;
        const int r = submit(&wq, Clock::now(), &id);
        if (r == TRW_OK) {
            open.push_back(id);
            ++i;
        } else if (r == TRW_E_BUSY && !open.empty()) {
            retire();
        } else {
            rc = r;
        }
    }
    while (!open.empty()) retire();
    std::lock_guard<std::mutex> lk(mu);
    stats = trw_server_stats{};
    batch_walkers = batch_queries = 0;
    return rc;
}

int trw_server_create(trw_graph *g, const trw_server_config *config, trw_server **out) {
    if (!g || !out) return trw::set_error(TRW_E_INVALID, "null argument");
    *out = nullptr;
    trw_server_config c;
    std::memset(&c, 0, sizeof(c));
    if (config) c = *config;
    if (c.streams == 0) c.streams = 4;
    if (c.max_batch_walkers == 0) c.max_batch_walkers = 1 << 20;
    if (c.path_bytes_per_batch == 0) c.path_bytes_per_batch = 64ll << 20;
    if (c.max_pending_queries == 0) c.max_pending_queries = 1 << 16;
    if (c.completion_threads == 0) c.completion_threads = 4;
    if (c.max_queries_per_batch == 0) c.max_queries_per_batch = 1024;
    if (c.streams < 1 || c.streams > 16) return trw::set_error(TRW_E_INVALID, "streams must be in [1, 16]");
    if (c.max_batch_walkers < 1 || c.max_batch_walkers > (1 << 28))
        return trw::set_error(TRW_E_INVALID, "max_batch_walkers must be in [1, 2^28]");
    if (c.path_bytes_per_batch < 4) return trw::set_error(TRW_E_INVALID, "path_bytes_per_batch must be >= 4");
    if (c.max_pending_queries < 1 || c.completion_threads < 1 || c.completion_threads > 64 ||
        c.max_queries_per_batch < 1 || c.max_queries_per_batch > (1 << 20))
        return trw::set_error(TRW_E_INVALID, "bad queue, thread or per-batch query limit");

    auto s = std::make_unique<trw_server>();
    s->g = g;
    s->device = g->ctx->device;
    s->max_walkers = c.max_batch_walkers;
    s->path_ints = c.path_bytes_per_batch / 4;
    s->max_pending = c.max_pending_queries;
    s->max_segments = c.max_queries_per_batch;
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return trw::set_hip_error(e, "hipSetDevice");
    s->slots.resize(static_cast<size_t>(c.streams));
    const size_t W = static_cast<size_t>(c.max_batch_walkers), P = static_cast<size_t>(s->path_ints),
                 S = static_cast<size_t>(c.max_queries_per_batch);
    const size_t stage_bytes = sizeof(trw::ServeSegment) * S + sizeof(int32_t) * static_cast<size_t>(kPoolInts);
    // PPR batches: buckets of (segment, chunk of 2^17 vertex ids), kPprBuckets at most.
    s->ppr_chunks = trw::ppr_chunks(g->nv);
    s->ppr_segments = static_cast<int32_t>(std::max<int64_t>(
        1, std::min<int64_t>(static_cast<int64_t>(S), std::min<int64_t>(kPprBuckets / s->ppr_chunks, 65535))));
    const size_t PC = static_cast<size_t>(s->ppr_chunks), PS = static_cast<size_t>(s->ppr_segments),
                 PT = static_cast<size_t>(trw::ppr_tiles(c.max_batch_walkers)) + PS;
    for (Slot &sl : s->slots) {
        // Default stream priority: every stream of the library runs at the same one.
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = dev(&sl.d_stage, stage_bytes);
        if (e == hipSuccess) e = dev(&sl.d_sources, W);
        if (e == hipSuccess) e = dev(&sl.d_ids, W);
        if (e == hipSuccess) e = dev(&sl.d_ends, W);
        if (e == hipSuccess) e = dev(&sl.d_lengths, W);
        if (e == hipSuccess) e = dev(&sl.d_status, W);
        if (e == hipSuccess) e = dev(&sl.d_paths, P);
        if (e == hipSuccess) e = dev(&sl.d_ctl, static_cast<size_t>(kCtlTickets) + S);
        if (e == hipSuccess) e = pinned(&sl.h_stage, stage_bytes);
        if (e == hipSuccess) e = pinned(&sl.h_ends, W);
        if (e == hipSuccess) e = pinned(&sl.h_lengths, W);
        if (e == hipSuccess) e = pinned(&sl.h_status, W);
        if (e == hipSuccess) e = pinned(&sl.h_paths, P);
        trw::PprScratch &x = sl.agg;
        x.nv = g->nv;
        x.C = s->ppr_chunks;
        if (e == hipSuccess) e = dev(&x.tile_hist, PT * PC);
        if (e == hipSuccess) e = dev(&x.tile_stats, PT);
        if (e == hipSuccess) e = dev(&x.bucketed, W);
        if (e == hipSuccess) e = dev(&x.pairs, W);
        if (e == hipSuccess) e = dev(&x.bucket, PS * PC);
        x.tickets = sl.d_ctl ? sl.d_ctl + kCtlTickets : nullptr;
        if (e == hipSuccess) e = pinned(&sl.h_stats, PS);
        if (e == hipSuccess) e = pinned(&sl.h_buckets, PS * PC);
        if (e == hipSuccess) e = pinned(&sl.h_pairs, W);
    }
    if (e != hipSuccess) {
        s->free_slots();
        return trw::set_hip_error(e, "allocating the server's slots");
    }
    // Every slot's stream does once what a batch does -- copies both ways between its pinned staging
    // and the device, a kernel, its event -- before the first query: the runtime sets a stream up
    // on its first use, which otherwise lands inside the latency of the first batches on that slot.
    for (Slot &sl : s->slots) {
        if (e == hipSuccess) e = hipMemcpyAsync(sl.d_stage, sl.h_stage, sizeof(trw::ServeSegment), hipMemcpyHostToDevice, sl.stream);
        if (e == hipSuccess) e = trw::launch_fill_u32(sl.d_ctl, 0u, kCtlTickets, sl.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(sl.h_ends, sl.d_ends, sizeof(int32_t), hipMemcpyDeviceToHost, sl.stream);
        if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    }
    for (Slot &sl : s->slots)
        if (e == hipSuccess) e = hipEventSynchronize(sl.done);
    if (e != hipSuccess) {
        s->free_slots();
        return trw::set_hip_error(e, "setting up the server's streams");
    }
    s->workers = std::make_unique<WorkerPool>(c.completion_threads);
    trw_server *raw = s.get();
    s->dispatcher = std::thread([raw] { raw->dispatch_loop(); });
    const int rc = raw->warm_up();
    if (rc != TRW_OK) {
        const std::string msg = trw_last_error();
        trw_server_destroy(s.release());
        return trw::set_error(rc, "server warm-up: %s", msg.c_str());
    }
    *out = s.release();
    return TRW_OK;
}

int trw_server_destroy(trw_server *s) {
    if (!s) return TRW_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    s->workers->stop();
    (void)hipSetDevice(s->device);
    s->free_slots();
    delete s;
    return TRW_OK;
}

int trw_server_submit(trw_server *s, const trw_query *q, uint64_t *query_id) {
    if (!s) return trw::set_error(TRW_E_INVALID, "server is null");
    return s->submit(q, Clock::now(), query_id);
}

int trw_server_wait(trw_server *s, uint64_t query_id, double timeout_ms, trw_query_info *info) {
    if (!s) return trw::set_error(TRW_E_INVALID, "server is null");
    std::unique_lock<std::mutex> lk(s->mu);
    std::shared_ptr<Query> q = find_query(s, query_id);
    if (!q) return trw::set_error(TRW_E_INVALID, "unknown query %llu", static_cast<unsigned long long>(query_id));
    auto finished = [&] { return q->state == TRW_QUERY_DONE || q->state == TRW_QUERY_FAILED; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, finished);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
        if (info) fill_info(*q, info);
        return trw::set_error(TRW_E_TIMEOUT, "query %llu not finished", static_cast<unsigned long long>(query_id));
    }
    if (info) fill_info(*q, info);
    if (q->state == TRW_QUERY_FAILED) return trw::set_error(q->error_code, "%s", q->error.c_str());
    return TRW_OK;
}

int trw_server_query_info(trw_server *s, uint64_t query_id, trw_query_info *info) {
    if (!s || !info) return trw::set_error(TRW_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::shared_ptr<Query> q = find_query(s, query_id);
    if (!q) return trw::set_error(TRW_E_INVALID, "unknown query %llu", static_cast<unsigned long long>(query_id));
    fill_info(*q, info);
    return TRW_OK;
}

int trw_server_ppr_result(trw_server *s, uint64_t query_id, int32_t *vertices, uint32_t *counts, int64_t capacity) {
    if (!s) return trw::set_error(TRW_E_INVALID, "server is null");
    std::lock_guard<std::mutex> lk(s->mu);
    std::shared_ptr<Query> q = find_query(s, query_id);
    if (!q) return trw::set_error(TRW_E_INVALID, "unknown query %llu", static_cast<unsigned long long>(query_id));
    if (q->state != TRW_QUERY_DONE) return trw::set_error(TRW_E_STATE, "query is not done");
    if (q->result != TRW_RESULT_PPR) return trw::set_error(TRW_E_STATE, "query was not a PPR query");
    const int64_t n = static_cast<int64_t>(q->ppr.size());
    if (capacity < n) return trw::set_error(TRW_E_INVALID, "capacity %lld < %lld entries", (long long)capacity, (long long)n);
    if (n && (!vertices || !counts)) return trw::set_error(TRW_E_INVALID, "null output");
    for (int64_t i = 0; i < n; ++i) {
        vertices[i] = q->ppr[static_cast<size_t>(i)].vertex;
        counts[i] = q->ppr[static_cast<size_t>(i)].count;
    }
    return TRW_OK;
}

int trw_server_walk_result(trw_server *s, uint64_t query_id, int32_t *paths, int32_t *lengths, int32_t *ends,
                           int32_t *status) {
    if (!s) return trw::set_error(TRW_E_INVALID, "server is null");
    std::lock_guard<std::mutex> lk(s->mu);
    std::shared_ptr<Query> q = find_query(s, query_id);
    if (!q) return trw::set_error(TRW_E_INVALID, "unknown query %llu", static_cast<unsigned long long>(query_id));
    if (q->state != TRW_QUERY_DONE) return trw::set_error(TRW_E_STATE, "query is not done");
    if (q->result != TRW_RESULT_PATHS)
        return trw::set_error(TRW_E_STATE, "a PPR query keeps its end-vertex counts only (trw_server_ppr_result)");
    if (paths) std::copy(q->paths.begin(), q->paths.end(), paths);
    if (lengths) std::copy(q->lengths.begin(), q->lengths.end(), lengths);
    if (ends) std::copy(q->ends.begin(), q->ends.end(), ends);
    if (status) std::copy(q->status.begin(), q->status.end(), status);
    return TRW_OK;
}

int trw_server_release(trw_server *s, uint64_t query_id) {
    if (!s) return trw::set_error(TRW_E_INVALID, "server is null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->queries.find(query_id);
    if (it == s->queries.end())
        return trw::set_error(TRW_E_INVALID, "unknown query %llu", static_cast<unsigned long long>(query_id));
    if (it->second->state != TRW_QUERY_DONE && it->second->state != TRW_QUERY_FAILED)
        return trw::set_error(TRW_E_STATE, "query is still pending or running");
    s->queries.erase(it);
    return TRW_OK;
}

int trw_server_stats_get(trw_server *s, trw_server_stats *out) {
    if (!s || !out) return trw::set_error(TRW_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->mean_batch_walkers = s->stats.batches ? double(s->batch_walkers) / double(s->stats.batches) : 0.0;
    out->mean_batch_queries = s->stats.batches ? double(s->batch_queries) / double(s->stats.batches) : 0.0;
    return TRW_OK;
}

int trw_server_run_open_loop(trw_server *s, const trw_query *queries, int32_t n, double qps, uint64_t seed,
                             uint64_t *ids, double *latency_ms, trw_open_loop_report *report) {
    if (!s || (n > 0 && !queries) || !report) return trw::set_error(TRW_E_INVALID, "null argument");
    if (n < 0 || !(qps > 0.0)) return trw::set_error(TRW_E_INVALID, "need n >= 0 and qps > 0");
    std::memset(report, 0, sizeof(*report));
    trw_server_stats before;
    trw_server_stats_get(s, &before);
    uint64_t batch_queries_before;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        batch_queries_before = s->batch_queries;
    }
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(qps);
    std::vector<Clock::time_point> when(static_cast<size_t>(n));
    const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(2);
    double t = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        if (i) t += gap(rng);
        when[static_cast<size_t>(i)] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t));
    }
    std::vector<uint64_t> id(static_cast<size_t>(n), 0);
    int first_error = TRW_OK;
    std::string first_message;
    for (int32_t i = 0; i < n; ++i) {
        const Clock::time_point at = when[static_cast<size_t>(i)];
        std::this_thread::sleep_until(at - std::chrono::microseconds(100));
        while (Clock::now() < at) {
        }
        const int rc = s->submit(&queries[i], at, &id[static_cast<size_t>(i)]);
        if (rc == TRW_E_BUSY) {
            ++report->rejected;
        } else if (rc != TRW_OK && first_error == TRW_OK) {
            first_error = rc;
            first_message = trw_last_error();
        }
    }
    std::vector<double> lat;
    lat.reserve(static_cast<size_t>(n));
    Clock::time_point last = t0;
    uint64_t steps = 0;
    for (int32_t i = 0; i < n; ++i) {
        const uint64_t qid = id[static_cast<size_t>(i)];
        if (ids) ids[i] = qid;
        if (latency_ms) latency_ms[i] = -1.0;
        if (!qid) continue;
        trw_query_info info;
        const int rc = trw_server_wait(s, qid, -1.0, &info);
        if (rc != TRW_OK && first_error == TRW_OK) {
            first_error = rc;
            first_message = trw_last_error();
        }
        std::lock_guard<std::mutex> lk(s->mu);
        std::shared_ptr<Query> q = find_query(s, qid);
        if (!q) continue;
        last = std::max(last, q->finished);
        steps += q->steps.load();
        lat.push_back(info.latency_ms);
        if (latency_ms) latency_ms[i] = info.latency_ms;
    }
    trw_server_stats after;
    trw_server_stats_get(s, &after);
    report->wall_ms = ms_between(t0, last);
    report->offered_qps = qps;
    report->achieved_qps = report->wall_ms > 0 ? 1e3 * double(lat.size()) / report->wall_ms : 0.0;
    report->steps = steps;
    report->steps_per_s = report->wall_ms > 0 ? 1e3 * double(steps) / report->wall_ms : 0.0;
    report->batches = after.batches - before.batches;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        report->mean_batch_queries =
            report->batches ? double(s->batch_queries - batch_queries_before) / double(report->batches) : 0.0;
    }
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            const size_t k = static_cast<size_t>(std::min<double>(lat.size() - 1, std::ceil(p * lat.size()) - 1));
            return lat[k];
        };
        report->p50_ms = pct(0.50);
        report->p90_ms = pct(0.90);
        report->p99_ms = pct(0.99);
        report->p999_ms = pct(0.999);
        report->max_ms = lat.back();
        double sum = 0;
        for (double x : lat) sum += x;
        report->mean_ms = sum / double(lat.size());
    }
    if (first_error != TRW_OK) return trw::set_error(first_error, "%s", first_message.c_str());
    return TRW_OK;
}

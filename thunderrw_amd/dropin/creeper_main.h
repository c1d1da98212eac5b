// Drop-in app surface of the reference's step engine, backed by libthunderrw.so.
//
// The reference's apps (random_walk/apps/{deepwalk,ppr,node2vec,metapath,genericwalk}.cpp)
// include "creeper_main.h" and use: Graph and its load flags, WalkerMeta,
// HyperParameter g_para, the ExecutionMode / SamplingMethod / SourceType enums,
// InputParser, Utility::split, the walker generators, the SFMT API their functors
// seed, and compute<F>(Graph&, std::vector<WalkerMeta>&, F).  This header provides
// all of them with the same names, signatures and meanings, so those app sources
// compile unchanged against it (hipcc -x hip --offload-arch=gfx950 -std=c++17
// -I thunderrw_amd/dropin -I include app.cpp -L thunderrw_amd -lthunderrw) and run their
// walks on the GPU.
//
// compute<F> recognises the reference's built-in functors by type name and members
// (PPR: tp_; MetaPath: l_, schema_length_, schema_; Node2Vec: l_, p_, q_; DeepWalk and
// GenericWalk: l_) and runs the library's Move kernel for them where the library has one for
// the (execution mode, sampling method) pair (trw_walk_supported).  Every other functor, and
// those pairs the library lacks (e.g. Node2Vec with dynamic ITS, MetaPath with dynamic alias
// or rejection, DeepWalk with max-weight rejection), runs in the functor engine
// (functor_walk.h): kernels instantiated for F in the app's translation unit call its
// weight(), update() and max_weight() on the device.  That needs hipcc; built with g++ the
// header serves the library's Moves only.
//
// Environment (no counterpart in the reference):
//   TRW_DEVICE      HIP device (default 0)
//   TRW_RNG         philox (default) or replay: bit-exact SFMT stream, walker i reads
//                   words [i*K, (i+1)*K) of init_gen_rand(TRW_SEED), PPR stop draws of
//                   init_gen_rand(TRW_TSEED)
//   TRW_SEED, TRW_TSEED, TRW_REPLAY_K, TRW_REPLAY_KT
//   TRW_DUMP_DIR    write paths.bin (walkers x length int32, LOG_SEQUENCE layout),
//                   lengths.bin, ends.bin, status.bin there after the walk
//   TRW_MAX_MOVES   functor engine: moves per walker before it ends as TRW_WALKER_STEP_CAP
//                   (default max(65536, 4 x length); the reference has no cap)
//   TRW_MAX_TRIALS  functor engine: rejection trials per move (default 65536)
#ifndef THUNDERRW_DROPIN_CREEPER_MAIN_H
#define THUNDERRW_DROPIN_CREEPER_MAIN_H

#include <algorithm>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <numeric>
#include <queue>
#include <random>
#include <sstream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "SFMT.h"
#include "thunderrw.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>

#include "thunderrw_device.h"
#define TRW_DROPIN_HD __host__ __device__
#else
#define TRW_DROPIN_HD
#endif

typedef int intT;
typedef unsigned int uintT;

// walker.h:9-15
struct WalkerMeta {
    int id_;
    int source_;
    int current_;
    int length_;
    intT *seq_;
};

// types.h:41-77
enum ExecutionMode { Uniform = 0, Static = 1, Dynamic = 2 };
enum SamplingMethod {
    UniformSampling = 0,
    InverseTransformationSampling = 1,
    AliasSampling = 2,
    RejectionSampling = 3,
    MaxWeightRejectionSampling = 4
};
enum SourceType { All = 0, Single = 1, Multiple = 2 };

struct HyperParameter {
    ExecutionMode execution_;
    SamplingMethod sample_;
    int length_;
    std::pair<intT *, int> meta_path_;
};

extern HyperParameter g_para;
extern int g_num_threads;

namespace trw_dropin {

inline void log_line(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline void log_line(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "[thunderrw] ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    va_end(ap);
}

[[noreturn]] inline void die(const char *what) {
    std::fprintf(stderr, "[thunderrw] %s: %s\n", what, trw_last_error());
    std::exit(-1);
}

inline long env_long(const char *name, long dflt) {
    const char *v = std::getenv(name);
    return (v && *v) ? std::strtol(v, nullptr, 10) : dflt;
}

inline trw_context *context() {
    static trw_context *ctx = nullptr;
    if (!ctx && trw_context_create(static_cast<int32_t>(env_long("TRW_DEVICE", 0)), &ctx) != TRW_OK)
        die("trw_context_create");
    return ctx;
}

}  // namespace trw_dropin

// util/command_parser.h:12-39
class InputParser {
public:
    InputParser(int &argc, char **argv) {
        for (int i = 1; i < argc; ++i) tokens_.emplace_back(argv[i]);
    }
    std::string get_cmd_option(const std::string &option) const {
        auto itr = std::find(tokens_.begin(), tokens_.end(), option);
        if (itr != tokens_.end() && ++itr != tokens_.end()) return *itr;
        return "";
    }
    bool check_cmd_option_exists(const std::string &option) const {
        return std::find(tokens_.begin(), tokens_.end(), option) != tokens_.end();
    }
    std::string get_cmd() { return std::accumulate(tokens_.begin(), tokens_.end(), std::string(" ")); }

private:
    std::vector<std::string> tokens_;
};

// util/utility.h: the helpers the apps call.
class Utility {
public:
    static void split(const std::string &s, char delimiter, std::vector<std::string> &tokens) {
        std::istringstream iss(s);
        std::string item;
        while (std::getline(iss, item, delimiter)) tokens.emplace_back(item);
    }
};

// util/graph/graph.h: the CSR view and load flags the apps and functors use.  The
// arrays live in a library-owned host graph; the device copy is made on first compute().
class Graph {
public:
    bool is_directed_ = false;
    bool is_vertex_labeled_ = false;
    bool is_vertex_weighted_ = false;
    bool is_edge_labeled_ = false;
    bool is_edge_weighted_ = false;
    bool is_edge_weight_prefix_summed_ = false;
    bool is_edge_weight_alias_generated_ = false;
    bool is_edge_weight_rejection_generated_ = false;
    bool is_offset_pair_ = false;

    intT num_vertices_ = 0;
    int64_t num_edges_ = 0;
    intT max_degree_ = 0;
    intT node_with_max_degree_ = 0;

    int64_t *offset_ = nullptr;
    intT *adj_ = nullptr;
    intT *degree_ = nullptr;
    int *edge_label_ = nullptr;
    double *edge_weight_ = nullptr;

    trw_graph *device_ = nullptr;

    // Set in the device image only: the hashed neighbour sets is_neighbor reads there.
    const void *member_lines_ = nullptr;
    const uint32_t *member_root_ = nullptr;

    Graph() = default;
    Graph(const Graph &) = delete;
    Graph &operator=(const Graph &) = delete;
    ~Graph() {
        if (image_) trw_device_free(trw_dropin::context(), image_);
        if (device_) trw_graph_free(device_);
        if (host_) trw_host_graph_free(host_);
    }

    // Graph::load_csr (util/graph/graph.cpp:45-101).
    void load_csr(const std::string &graph_dir) {
        auto t0 = std::chrono::high_resolution_clock::now();
        if (trw_host_graph_load(graph_dir.c_str(), is_edge_weighted_, is_edge_labeled_, &host_) != TRW_OK)
            trw_dropin::die("load_csr");
        const int32_t *deg = nullptr, *adj = nullptr, *lab = nullptr;
        const double *w = nullptr;
        trw_host_graph_arrays(host_, &num_vertices_, &num_edges_, &deg, &adj, &w, &lab);
        degree_ = const_cast<intT *>(deg);
        adj_ = const_cast<intT *>(adj);
        edge_weight_ = const_cast<double *>(w);
        edge_label_ = const_cast<int *>(lab);
        offsets_.assign(static_cast<size_t>(num_vertices_) + 1, 0);
        for (intT v = 0; v < num_vertices_; ++v) {
            offsets_[v + 1] = offsets_[v] + degree_[v];
            if (degree_[v] > max_degree_) {
                max_degree_ = degree_[v];
                node_with_max_degree_ = v;
            }
        }
        offset_ = offsets_.data();
        auto t1 = std::chrono::high_resolution_clock::now();
        trw_dropin::log_line("Load graph time: %.3lf seconds",
                             std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count() / 1000.0);
    }

    void print_metadata() {
        trw_dropin::log_line("|V| = %d, |E| = %lld, d = %.2lf, node with max degree = %d, max degree = %d",
                             num_vertices_, static_cast<long long>(num_edges_), avg_degree(), node_with_max_degree_,
                             max_degree_);
    }

    // The accessors functors call; under hipcc they also run on the device, where the
    // functor engine hands functors a copy of this object whose arrays are the graph's in HBM.
    TRW_DROPIN_HD intT num_vertices() const { return num_vertices_; }
    TRW_DROPIN_HD int64_t num_edges() const { return num_edges_; }
    TRW_DROPIN_HD double avg_degree() const { return num_edges_ / static_cast<double>(num_vertices_); }
    TRW_DROPIN_HD intT max_degree() const { return max_degree_; }
    TRW_DROPIN_HD intT node_with_max_degree() const { return node_with_max_degree_; }
    TRW_DROPIN_HD intT degree(intT u) const { return degree_[u]; }
    TRW_DROPIN_HD std::pair<intT *, intT> neighbors(intT u) const { return {adj_ + offset_[u], degree(u)}; }
    TRW_DROPIN_HD std::pair<int64_t, int64_t> edges(intT u) const { return {offset_[u], offset_[u + 1]}; }
    TRW_DROPIN_HD double edge_weight(int64_t eid) const { return edge_weight_[eid]; }
    TRW_DROPIN_HD int edge_label(int64_t eid) const { return edge_label_[eid]; }

    // graph.h:180-199: the larger-degree endpoint searched in the smaller one's list.
    // On the device, a vertex with a membership table answers from it (one 128-byte read for
    // the same answer as the search).
    TRW_DROPIN_HD bool is_neighbor(intT u, intT v) const {
        if (degree(u) < degree(v)) {
            const intT t = u;
            u = v;
            v = t;
        }
#if defined(__HIPCC__)
        if (member_root_) {
            const uint32_t r = member_root_[v];
            if (r != trw_dev::kNoRoot) return trw_dev::member_contains(static_cast<const uint4 *>(member_lines_), r, degree(v), u);
        }
#endif
        auto nbrs = neighbors(v);
        int b = 0, e = nbrs.second - 1;
        while (b <= e) {
            int m = b + ((e - b) >> 1);
            if (nbrs.first[m] == u) return true;
            if (nbrs.first[m] > u) e = m - 1;
            else b = m + 1;
        }
        return false;
    }

    trw_graph *device() {
        if (!device_ && trw_graph_upload(trw_dropin::context(), num_vertices_, num_edges_, degree_, adj_,
                                         edge_weight_, edge_label_, &device_) != TRW_OK)
            trw_dropin::die("trw_graph_upload");
        return device_;
    }

    // A device copy of this object whose array fields (offset_, adj_, degree_, edge_weight_,
    // edge_label_) are the graph's arrays in HBM, with the membership tables is_neighbor reads
    // there (built here); made once, freed with the graph.
    const void *device_image() {
        if (image_) return image_;
        if (trw_build_membership(device()) != TRW_OK) trw_dropin::die("trw_build_membership");
        trw_device_view v;
        if (trw_graph_device_view(device(), &v) != TRW_OK) trw_dropin::die("trw_graph_device_view");
        std::vector<unsigned char> bytes(sizeof(Graph));
        std::memcpy(bytes.data(), static_cast<const void *>(this), sizeof(Graph));
        auto put = [&](const void *field, const void *value) {
            const size_t at = static_cast<size_t>(static_cast<const char *>(field) - reinterpret_cast<const char *>(this));
            std::memcpy(bytes.data() + at, &value, sizeof(value));
        };
        put(&offset_, v.offset);
        put(&adj_, v.adj);
        put(&degree_, v.degree);
        put(&edge_weight_, v.edge_weight);
        put(&edge_label_, v.edge_label);
        put(&member_lines_, v.membership_lines);
        put(&member_root_, v.membership_root);
        put(&device_, nullptr);
        put(&host_, nullptr);
        if (trw_device_alloc(trw_dropin::context(), sizeof(Graph), &image_) != TRW_OK ||
            trw_copy_to_device(trw_dropin::context(), image_, bytes.data(), sizeof(Graph)) != TRW_OK)
            trw_dropin::die("graph device image");
        return image_;
    }

private:
    trw_host_graph *host_ = nullptr;
    std::vector<int64_t> offsets_;
    void *image_ = nullptr;
};

// walker_generator.cpp:9-25
inline void generate_all_source(Graph &graph, std::vector<WalkerMeta> &walkers) {
    std::vector<int32_t> src(static_cast<size_t>(graph.num_vertices()));
    int32_t n = 0;
    if (trw_sources_all(graph.degree_, graph.num_vertices(), src.data(), &n) != TRW_OK)
        trw_dropin::die("generate_all_source");
    walkers.reserve(2048);
    for (intT i = 0; i < n; ++i) walkers.push_back({i, src[i], src[i], 1, nullptr});
}

// walker_generator.cpp:27-40
inline void generate_random_source(Graph &graph, std::vector<WalkerMeta> &walkers, int target_num_vertex, int seed) {
    const int have = static_cast<int>(walkers.size());
    const int need = target_num_vertex - have;
    if (need <= 0) return;
    std::vector<int32_t> src(static_cast<size_t>(need));
    if (trw_sources_random(graph.degree_, graph.num_vertices(), need, static_cast<uint32_t>(seed), src.data()) != TRW_OK)
        trw_dropin::die("generate_random_source");
    walkers.reserve(2048);
    for (int i = 0; i < need; ++i) walkers.push_back({static_cast<int>(walkers.size()), src[i], src[i], 1, nullptr});
}

// walker_generator.cpp:42-63
inline void generate_single_source(Graph &graph, std::vector<WalkerMeta> &walkers, int target_num_vertex, int) {
    intT u = 0;
    intT d = graph.degree(u);
    for (intT v = 1; v < graph.num_vertices(); ++v)
        if (graph.degree(v) > d) {
            d = graph.degree(v);
            u = v;
        }
    std::cout << u << ' ' << graph.degree(u) << std::endl;
    walkers.reserve(2048);
    for (int i = 0; i < target_num_vertex; ++i) walkers.push_back({i, u, u, 1, nullptr});
}

namespace trw_dropin {

#define TRW_DROPIN_MEMBER_TRAIT(name)                                                    \
    template <class F, class = void>                                                   \
    struct has_##name : std::false_type {};                                            \
    template <class F>                                                                 \
    struct has_##name<F, decltype(void(std::declval<F &>().name))> : std::true_type {};

TRW_DROPIN_MEMBER_TRAIT(tp_)
TRW_DROPIN_MEMBER_TRAIT(l_)
TRW_DROPIN_MEMBER_TRAIT(p_)
TRW_DROPIN_MEMBER_TRAIT(q_)
TRW_DROPIN_MEMBER_TRAIT(schema_)
TRW_DROPIN_MEMBER_TRAIT(schema_length_)
TRW_DROPIN_MEMBER_TRAIT(sfmt_)
#undef TRW_DROPIN_MEMBER_TRAIT

// Whether F's type is named `want`: the compiler's function signature spells "F = <type>", and
// the type's last component -- after any namespace, named or anonymous, or enclosing class --
// must equal `want` (an alias names the type it aliases).
template <class F>
bool functor_named(const char *want) {
    const char *p = std::strstr(__PRETTY_FUNCTION__, "F = ");
    if (!p) return false;
    p += 4;
    const char *end = p;
    while (*end && *end != ']' && *end != ';' && *end != ',') ++end;
    const char *last = p;
    for (const char *q = p; q + 1 < end; ++q)
        if (q[0] == ':' && q[1] == ':') last = q + 2;
    const size_t n = std::strlen(want);
    return static_cast<size_t>(end - last) == n && std::strncmp(last, want, n) == 0;
}

// The reference's built-in functors, recognised by name and members, as library apps with
// their constructor arguments.  Returns false for any other functor.
template <class F>
bool describe(const F &f, trw_walk_params &p) {
    if constexpr (has_tp_<F>::value) {
        if (!functor_named<F>("PPR")) return false;
        p.app = TRW_APP_PPR;  // apps/ppr.cpp:10-34
        p.stop_probability = f.tp_;
        return true;
    } else if constexpr (has_schema_<F>::value && has_schema_length_<F>::value && has_l_<F>::value) {
        if (!functor_named<F>("MetaPath")) return false;
        p.app = TRW_APP_METAPATH;  // apps/metapath.cpp:9-39
        if (f.schema_length_ < 1 || f.schema_length_ > 16) {
            std::fprintf(stderr, "[thunderrw] MetaPath schema length must be 1..16\n");
            std::exit(-1);
        }
        p.schema_length = f.schema_length_;
        for (int i = 0; i < f.schema_length_; ++i) p.schema[i] = f.schema_[i];
        return true;
    } else if constexpr (has_p_<F>::value && has_q_<F>::value && has_l_<F>::value) {
        if (!functor_named<F>("Node2Vec")) return false;
        p.app = TRW_APP_NODE2VEC;  // apps/node2vec.cpp:9-47
        p.p = f.p_;
        p.q = f.q_;
        return true;
    } else if constexpr (has_l_<F>::value) {
        if (!functor_named<F>("DeepWalk") && !functor_named<F>("GenericWalk")) return false;
        p.app = TRW_APP_DEEPWALK;  // apps/deepwalk.cpp:9-30, apps/genericwalk.cpp:9-30
        return true;
    } else {
#if !defined(__HIPCC__)
        static_assert(has_l_<F>::value,
                      "compute<F>: a functor other than the reference's runs its weight()/update() on the device; "
                      "compile the app with hipcc");
#endif
        (void)f;
        (void)p;
        return false;
    }
}

inline int64_t replay_budget(const trw_walk_params &p) {
    const int64_t moves = p.length > 1 ? p.length - 1 : 1;
    if (p.app == TRW_APP_PPR) return 4096;
    if (p.execution != TRW_EXEC_UNIFORM &&
        (p.sampling == TRW_SM_REJECTION || p.sampling == TRW_SM_MAX_WEIGHT_REJECTION))
        return 32 * moves;
    if (p.execution != TRW_EXEC_UNIFORM && p.sampling == TRW_SM_ALIAS) return 2 * moves;
    return moves;
}

template <class T>
void dump(const std::string &dir, const char *name, const std::vector<T> &v) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(sizeof(T) * v.size()));
}

// Device buffers of one compute() call, freed together.
struct Buffers {
    trw_context *ctx;
    std::vector<void *> held;
    explicit Buffers(trw_context *c) : ctx(c) {}
    ~Buffers() {
        for (void *q : held) trw_device_free(ctx, q);
    }
    void *get(uint64_t bytes) {
        void *q = nullptr;
        if (trw_device_alloc(ctx, bytes ? bytes : 4, &q) != TRW_OK) die("trw_device_alloc");
        held.push_back(q);
        return q;
    }
};

}  // namespace trw_dropin

#if defined(__HIPCC__)
#include "functor_walk.h"

namespace trw_dropin {

// A functor walk ready to launch: the device copy of the functor and the kernel arguments.
template <class F>
struct FunctorWalk {
    F *functor = nullptr;
    trw_functor::Args args;
    hipStream_t stream = nullptr;
};

// The functor engine (functor_walk.h): F's methods run on the device.  io holds the walker
// batch (paths non-null, with readable guard rows before and after).  Everything before the
// walk itself -- the functor's device copy and, in Static mode, static_gather and the tables --
// happens here, so the walk's timing covers the walk alone (as creeper.h:700-803 times it).
template <class F>
FunctorWalk<F> prepare_functor(Graph &graph, F &f, const trw_walk_params &p, const trw_walk_io &io, Buffers &bufs) {
    trw_context *ctx = context();
    trw_graph *g = graph.device();
    void *sv = nullptr;
    trw_context_stream(ctx, &sv);
    const hipStream_t st = static_cast<hipStream_t>(sv);
    trw_device_view v;
    if (trw_graph_device_view(g, &v) != TRW_OK) die("trw_graph_device_view");
    const size_t nv = static_cast<size_t>(graph.num_vertices_), ne = static_cast<size_t>(graph.num_edges_);

    // The functor's bytes for the device: its Graph pointer (and pointers into the graph's
    // arrays) now address the device copies.
    std::vector<trw_functor::Relocation> rel = {
        {&graph, sizeof(Graph), graph.device_image()},
        {graph.offset_, 8 * (nv + 1), v.offset},
        {graph.degree_, 4 * nv, v.degree},
        {graph.adj_, 4 * ne, v.adj},
        {graph.edge_weight_, 8 * ne, v.edge_weight},
        {graph.edge_label_, 4 * ne, v.edge_label},
    };
    if constexpr (has_schema_<F>::value && has_schema_length_<F>::value) {
        // MetaPath's schema (apps/metapath.cpp:13-14) lives in host memory.
        const size_t bytes = 4 * static_cast<size_t>(f.schema_length_ > 0 ? f.schema_length_ : 0);
        void *d = bufs.get(bytes);
        if (bytes) trw_copy_to_device(ctx, d, f.schema_, bytes);
        rel.push_back({f.schema_, bytes, d});
    }
    std::vector<unsigned char> fb(sizeof(F));
    std::memcpy(fb.data(), static_cast<const void *>(&f), sizeof(F));
    const int64_t bad = trw_functor::relocate(fb.data(), sizeof(F), rel.data(), static_cast<int>(rel.size()));
    if (bad >= 0) {
        // Possibly a host pointer member (a device read through it would fault), possibly an
        // integer and padding that happen to look like one: warn, the walk still runs.
        log_line("warning: the functor's word at byte %lld holds a host address; the device can reach graph "
                 "data only through the functor's Graph pointer",
                 static_cast<long long>(bad));
    }
    F *df = static_cast<F *>(bufs.get(sizeof(F)));
    trw_copy_to_device(ctx, df, fb.data(), sizeof(F));

    if (p.execution == TRW_EXEC_STATIC) {
        // static_initialization (creeper.h:102-175): F::weight per edge, then the tables.
        double *w = static_cast<double *>(bufs.get(8 * ne));
        hipError_t e = trw_functor::launch_weights<F, WalkerMeta>(df, v, w, st);
        if (e != hipSuccess) {
            std::fprintf(stderr, "[thunderrw] static_gather: %s\n", hipGetErrorString(e));
            std::exit(-1);
        }
        if (trw_build_tables_from(g, p.sampling, w) != TRW_OK) die("static_initialization");
        double ms = 0;
        trw_last_build_ms(g, &ms);
        log_line("Preprocessing time: %.6lf seconds", ms / 1e3);
        if (trw_graph_device_view(g, &v) != TRW_OK) die("trw_graph_device_view");
    }

    trw_functor::Args a;
    std::memset(&a, 0, sizeof(a));
    a.offset = v.offset;
    a.adj = v.adj;
    a.alias = v.alias_table;
    a.prefix = v.prefix_sum;
    a.rs_max = v.rejection_max;
    a.rs_weight = v.rejection_weight;
    a.nv = v.num_vertices;
    a.execution = p.execution;
    a.sampling = p.sampling;
    a.length = p.length;
    a.max_trials = p.max_trials > 0 ? p.max_trials : static_cast<int32_t>(env_long("TRW_MAX_TRIALS", 1 << 16));
    a.max_moves = static_cast<int32_t>(env_long("TRW_MAX_MOVES", std::max<long>(1 << 16, 4L * p.length)));
    a.sources = io.sources;
    a.n = io.n_walkers;
    a.path_stride = io.path_stride;
    a.paths = io.paths;
    a.lengths = io.lengths;
    a.ends = io.ends;
    a.status = io.status;
    a.draws_used = io.draws_used;
    a.counters = reinterpret_cast<unsigned long long *>(io.counters);
    a.draws = p.rng == TRW_RNG_REPLAY ? io.draws : nullptr;
    a.K = p.rng == TRW_RNG_REPLAY ? io.draws_per_walker : 0;
    a.seed = p.seed;
    a.id_base = static_cast<uint64_t>(io.walker_id_base);
    if (p.rng == TRW_RNG_REPLAY && !a.draws) {
        // No draws: a zero-length replay stream (every walker that needs one is exhausted).
        a.draws = static_cast<const uint32_t *>(bufs.get(4));
    }
    if (p.execution == TRW_EXEC_DYNAMIC && p.sampling != TRW_SM_MAX_WEIGHT_REJECTION) {
        const int32_t maxd = std::max<int32_t>(v.max_degree, 1);
        const int64_t waves = trw_functor::gather_waves(a.n, maxd, p.sampling);
        const uint64_t rows = static_cast<uint64_t>(waves) * static_cast<uint64_t>(maxd);
        a.scratch_stride = maxd;
        a.scratch_w = static_cast<double *>(bufs.get(8 * rows));
        if (p.sampling == TRW_SM_ALIAS) {
            a.scratch_s = static_cast<trw_functor::Entry *>(bufs.get(sizeof(trw_functor::Entry) * rows));
            a.scratch_g = static_cast<trw_functor::Entry *>(bufs.get(sizeof(trw_functor::Entry) * rows));
        }
    }
    if (trw_synchronize(ctx, nullptr) != TRW_OK) die("functor walk set-up");
    FunctorWalk<F> fw;
    fw.functor = df;
    fw.args = a;
    fw.stream = st;
    return fw;
}

template <class F>
void launch_functor(const FunctorWalk<F> &fw) {
    const hipError_t e = trw_functor::launch_walk<F, WalkerMeta>(fw.functor, fw.args, fw.stream);
    if (e != hipSuccess) {
        std::fprintf(stderr, "[thunderrw] functor walk launch: %s\n", hipGetErrorString(e));
        std::exit(-1);
    }
}

}  // namespace trw_dropin
#endif  // __HIPCC__

// compute<F> (random_walk/creeper.h:689-831) on the GPU: the library's Move kernels for the
// reference's five functors where the library has that (mode, sampler) pair, the functor
// engine for everything else (built with hipcc).
template <class F>
void compute(Graph &graph, std::vector<WalkerMeta> &walkers, F f) {
    static const char *kExec[] = {"Uniform", "Static", "Dynamic"};
    static const char *kSample[] = {"Uniform", "ITS", "Alias", "RejectionSampling", "MaxRejectionSampling"};
    trw_walk_params p;
    std::memset(&p, 0, sizeof(p));
    const bool builtin = trw_dropin::describe(f, p);
    p.execution = g_para.execution_;
    p.sampling = g_para.sample_;
    p.length = g_para.length_;
    const char *rng = std::getenv("TRW_RNG");
    const bool replay = rng && std::strcmp(rng, "replay") == 0;
    p.rng = replay ? TRW_RNG_REPLAY : TRW_RNG_PHILOX;
    p.seed = static_cast<uint64_t>(trw_dropin::env_long("TRW_SEED", replay ? 1234 : 0));
    p.form = TRW_FORM_AUTO;
    const bool library_move = builtin && trw_walk_supported(&p) == TRW_OK;
    trw_dropin::log_line("Configuration: %s, %s, GPU %ld, length %d, %s", kExec[g_para.execution_ % 3],
                         kSample[g_para.sample_ % 5], trw_dropin::env_long("TRW_DEVICE", 0), g_para.length_,
                         library_move ? "library Move" : "functor engine");
#if !defined(__HIPCC__)
    if (!library_move) {
        std::fprintf(stderr, "[thunderrw] this functor / mode / sampler runs in the functor engine: build the app "
                             "with hipcc\n");
        std::exit(-1);
    }
#endif
    if (!builtin) p.app = TRW_APP_GENERIC;

    trw_graph *g = graph.device();
    if (library_move && g_para.execution_ == Static) {  // static_initialization (creeper.h:102-175)
        const int32_t w = p.app == TRW_APP_PPR ? TRW_WEIGHTS_UNIT : TRW_WEIGHTS_EDGE;
        if (trw_build_tables(g, g_para.sample_, w) != TRW_OK) trw_dropin::die("static_initialization");
        double ms = 0;
        trw_last_build_ms(g, &ms);
        trw_dropin::log_line("Preprocessing time: %.6lf seconds", ms / 1e3);
    }

    const int32_t n = static_cast<int32_t>(walkers.size());
    const int32_t L = std::max(g_para.length_, 1);
    std::vector<int32_t> src(static_cast<size_t>(n)), len(static_cast<size_t>(n)), end(static_cast<size_t>(n)),
        st(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) src[i] = walkers[i].source_;
    const char *dump_dir = std::getenv("TRW_DUMP_DIR");
    std::vector<int32_t> paths(dump_dir ? static_cast<size_t>(n) * L : 0);

    trw_context *ctx = trw_dropin::context();
    trw_dropin::Buffers bufs(ctx);
    trw_walk_io io;
    std::memset(&io, 0, sizeof(io));
    io.n_walkers = n;
    io.path_stride = L;
    io.sources = static_cast<const int32_t *>(bufs.get(4ull * n));
    trw_copy_to_device(ctx, const_cast<int32_t *>(io.sources), src.data(), 4ull * n);
    io.lengths = static_cast<int32_t *>(bufs.get(4ull * n));
    io.ends = static_cast<int32_t *>(bufs.get(4ull * n));
    io.status = static_cast<int32_t *>(bufs.get(4ull * n));
    io.counters = static_cast<uint64_t *>(bufs.get(32));
    trw_device_zero(ctx, io.counters, 32, nullptr);
    if (!library_move) {
        // The functor engine always keeps paths: functors read WalkerMeta::seq_ (node2vec.cpp:29).
        // Guard rows on both sides keep a functor's seq_[-1] or seq_[length] readable.
        const uint64_t guard = 64;
        int32_t *base = static_cast<int32_t *>(bufs.get(4ull * (static_cast<uint64_t>(n) * L + 2 * guard)));
        void *sv = nullptr;
        trw_context_stream(ctx, &sv);
#if defined(__HIPCC__)
        (void)hipMemsetAsync(base, 0xff, 4ull * (static_cast<uint64_t>(n) * L + 2 * guard), static_cast<hipStream_t>(sv));
#endif
        io.paths = base + guard;
    } else if (dump_dir) {
        io.paths = static_cast<int32_t *>(bufs.get(4ull * n * L));
    }
    if (replay) {
        const int64_t K = trw_dropin::env_long("TRW_REPLAY_K", trw_dropin::replay_budget(p));
        std::vector<uint32_t> d(static_cast<size_t>(K) * n);
        trw_replay_fill(static_cast<uint32_t>(p.seed), 0, n, K, d.data());
        io.draws = static_cast<const uint32_t *>(bufs.get(4ull * d.size()));
        trw_copy_to_device(ctx, const_cast<uint32_t *>(io.draws), d.data(), 4ull * d.size());
        io.draws_per_walker = K;
        if (p.app == TRW_APP_PPR) {
            const int64_t KT = trw_dropin::env_long("TRW_REPLAY_KT", 4096);
            std::vector<uint32_t> t(static_cast<size_t>(KT) * n);
            trw_replay_fill(static_cast<uint32_t>(trw_dropin::env_long("TRW_TSEED", 4321)), 0, n, KT, t.data());
            io.term_draws = static_cast<const uint32_t *>(bufs.get(4ull * t.size()));
            trw_copy_to_device(ctx, const_cast<uint32_t *>(io.term_draws), t.data(), 4ull * t.size());
            io.term_draws_per_walker = KT;
        }
    }

    trw_dropin::log_line("Compute...");
    if (library_move) {
        trw_timer_begin(ctx, nullptr);
        if (trw_walk_async(g, &p, &io, nullptr) != TRW_OK) trw_dropin::die("compute");
    } else {
#if defined(__HIPCC__)
        if constexpr (trw_dropin::has_sfmt_<F>::value) {
            // PPR's update draws from the functor's own host SFMT stream (apps/ppr.cpp:27-29).
            std::fprintf(stderr, "[thunderrw] a functor with a host SFMT stream runs only in the library's Moves "
                                 "(sampling methods 0-3)\n");
            std::exit(-1);
        } else {
            const auto fw = trw_dropin::prepare_functor(graph, f, p, io, bufs);
            trw_timer_begin(ctx, nullptr);
            trw_dropin::launch_functor(fw);
        }
#endif
    }
    double ms = 0;
    trw_timer_end(ctx, nullptr, &ms);
    uint64_t counters[4] = {0, 0, 0, 0};
    trw_copy_to_host(ctx, counters, io.counters, sizeof(counters));
    trw_copy_to_host(ctx, len.data(), io.lengths, 4ull * n);
    trw_copy_to_host(ctx, end.data(), io.ends, 4ull * n);
    trw_copy_to_host(ctx, st.data(), io.status, 4ull * n);
    if (dump_dir) trw_copy_to_host(ctx, paths.data(), io.paths, 4ull * n * L);

    for (int32_t i = 0; i < n; ++i) {
        walkers[i].length_ = len[i];
        walkers[i].current_ = end[i];
    }
    const double secs = ms / 1e3;
    trw_dropin::log_line("Walking time (seconds): %.6lf", secs);
    trw_dropin::log_line("Num of steps: %llu", static_cast<unsigned long long>(counters[0]));
    trw_dropin::log_line("Num of terminated walkers: %llu", static_cast<unsigned long long>(counters[1]));
    trw_dropin::log_line("Throughput (steps per second): %.2lf", secs > 0 ? counters[0] / secs : 0.0);
    if (counters[3]) trw_dropin::log_line("Walkers ended by an error: %llu", static_cast<unsigned long long>(counters[3]));
    if (dump_dir) {
        trw_dropin::dump(dump_dir, "paths.bin", paths);
        trw_dropin::dump(dump_dir, "lengths.bin", len);
        trw_dropin::dump(dump_dir, "ends.bin", end);
        trw_dropin::dump(dump_dir, "status.bin", st);
    }
}

#if defined(__HIPCC__) && !defined(TRW_DROPIN_NO_FORCE_HD)
#pragma clang force_cuda_host_device begin
#endif
void execute(Graph &graph, InputParser &cmd_parser);
#if defined(__HIPCC__) && !defined(TRW_DROPIN_NO_FORCE_HD)
#pragma clang force_cuda_host_device end
#endif

int g_num_threads = 1;

// creeper_main.h:14-72 (the reference driver), loading the graph for the GPU engine.
int main(int argc, char *argv[]) {
    InputParser cmd_parser(argc, argv);
    Graph graph;
    graph.is_vertex_labeled_ = cmd_parser.check_cmd_option_exists("-vl");
    graph.is_edge_labeled_ = cmd_parser.check_cmd_option_exists("-el");
    graph.is_vertex_weighted_ = cmd_parser.check_cmd_option_exists("-vw");
    graph.is_edge_weighted_ = cmd_parser.check_cmd_option_exists("-ew");
    graph.is_edge_weight_prefix_summed_ = cmd_parser.check_cmd_option_exists("-ep");
    graph.is_edge_weight_alias_generated_ = cmd_parser.check_cmd_option_exists("-al");
    graph.is_edge_weight_rejection_generated_ = cmd_parser.check_cmd_option_exists("-rj");
    g_num_threads = 1;
    if (!cmd_parser.get_cmd_option("-n").empty()) g_num_threads = std::stoi(cmd_parser.get_cmd_option("-n"));
    graph.is_offset_pair_ = true;
    graph.load_csr(cmd_parser.get_cmd_option("-f"));
    graph.print_metadata();
    execute(graph, cmd_parser);
    trw_dropin::log_line("Done.");
    return 0;
}

// Under hipcc, everything the app defines after this header -- its functor above all -- is
// compiled for the host and the device, so that the functor engine can call weight(),
// update() and max_weight() in its kernels.  In the device pass execute() is renamed to an
// unused internal function, so its host-only body (argument parsing, I/O) is never emitted
// for the GPU while the kernels compute<F> instantiates are.  Define TRW_DROPIN_NO_FORCE_HD
// to mark the functor __host__ __device__ yourself instead.
#if defined(__HIPCC__) && !defined(TRW_DROPIN_NO_FORCE_HD)
#if defined(__HIP_DEVICE_COMPILE__)
#define execute(...) __attribute__((unused)) static trw_dropin_execute_host_only(__VA_ARGS__)
#endif
#pragma clang force_cuda_host_device begin
#endif

#endif  // THUNDERRW_DROPIN_CREEPER_MAIN_H

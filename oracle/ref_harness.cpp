// Golden-vector harness around the reference's own code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's sources where they lie (util/graph/graph.cpp,
// util/log/log.cpp, random_walk/walker_generator.cpp, SFMT.c) and with one
// reference app translation unit (random_walk/apps/<app>.cpp, whose main() is
// renamed away; or tests/dropin_apps/customwalk.cpp, an app written against the
// reference's interface), into oracle/_ref/harness_<app>.  Nothing in the product uses
// it; tests/golden/make_golden.py runs it in the build container to write the
// fixtures under tests/golden/.
//
// What it runs (all reference code, unmodified):
//   Graph::load_csr                          util/graph/graph.cpp:45-101
//   static_initialization                    random_walk/creeper.h:102-175
//   uniform_move / static_*_move / dynamic_*_move / dynamic_max_weight_rs_move
//   dynamic_gather, update                   random_walk/creeper.h:178-334
//   the app functor (weight/update/max_weight) of the app TU it is built with
// with exactly one live ring slot, so every walker consumes its random draws
// in its own sequential order.  Walker i's move draws are words
// [i*K, (i+1)*K) of the SFMT-19937 stream initialised by init_gen_rand(seed);
// for PPR its termination draws are words [i*KT, (i+1)*KT) of the stream
// initialised by init_gen_rand(tseed), installed as the functor's own sfmt_.

#define main trw_reference_app_main_unused
#if defined(REF_APP_deepwalk)
#include "apps/deepwalk.cpp"
#elif defined(REF_APP_ppr)
#include "apps/ppr.cpp"
#elif defined(REF_APP_node2vec)
#include "apps/node2vec.cpp"
#elif defined(REF_APP_metapath)
#include "apps/metapath.cpp"
#elif defined(REF_APP_genericwalk)
#include "apps/genericwalk.cpp"
#elif defined(REF_APP_customwalk)
#include "customwalk.cpp"  // tests/dropin_apps: a functor without a library Move
#else
#error "define REF_APP_<app>"
#endif
#undef main

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace {

template <class T>
void dump(const std::string &path, const T *data, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(data), sizeof(T) * n);
}

template <class T>
std::vector<T> slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    size_t bytes = static_cast<size_t>(f.tellg());
    std::vector<T> v(bytes / sizeof(T));
    f.seekg(0);
    f.read(reinterpret_cast<char *>(v.data()), bytes);
    return v;
}

// Number of words generator `after` has advanced past `before` (bounded by limit).
int64_t words_consumed(const sfmt_t &before, const sfmt_t &after, int64_t limit) {
    sfmt_t probe = before;
    for (int64_t k = 0; k <= limit; ++k) {
        if (probe.idx == after.idx && std::memcmp(probe.state, after.state, sizeof(probe.state)) == 0) return k;
        sfmt_genrand_uint32(&probe);
    }
    return -1;
}

double opt_d(InputParser &c, const char *k, double dflt) {
    auto s = c.get_cmd_option(k);
    return s.empty() ? dflt : std::stod(s);
}
long opt_l(InputParser &c, const char *k, long dflt) {
    auto s = c.get_cmd_option(k);
    return s.empty() ? dflt : std::stol(s);
}

}  // namespace

int main(int argc, char **argv) {
    InputParser cmd(argc, argv);
    Graph graph;
    graph.is_edge_weighted_ = cmd.check_cmd_option_exists("-ew");
    graph.is_edge_labeled_ = cmd.check_cmd_option_exists("-el");
    graph.is_offset_pair_ = true;
    graph.load_csr(cmd.get_cmd_option("-f"));
    const std::string out = cmd.get_cmd_option("-o");
    const std::string mode = cmd.get_cmd_option("-mode");

    const int L = static_cast<int>(opt_l(cmd, "-l", 80));
    g_para.length_ = L;
    g_para.execution_ = static_cast<ExecutionMode>(opt_l(cmd, "-em", g_para.execution_));
    g_para.sample_ = static_cast<SamplingMethod>(opt_l(cmd, "-sm", g_para.sample_));
    g_num_threads = 1;

#if defined(REF_APP_deepwalk)
    DeepWalk f(L, &graph);
#elif defined(REF_APP_genericwalk)
    GenericWalk f(L, &graph);
#elif defined(REF_APP_customwalk)
    BiasedWalk f(L, static_cast<int>(opt_l(cmd, "-kind", 0)), 1.0 / opt_d(cmd, "-p", 2.0), 1.0 / opt_d(cmd, "-q", 0.5),
                 &graph);
#elif defined(REF_APP_ppr)
    PPR f(opt_d(cmd, "-sp", 0.2));
#elif defined(REF_APP_node2vec)
    // node2vec.cpp:64-76 inverts p and q before constructing the functor.
    Node2Vec f(L, 1.0 / opt_d(cmd, "-p", 2.0), 1.0 / opt_d(cmd, "-q", 0.5), &graph);
#elif defined(REF_APP_metapath)
    std::vector<int> schema;
    {
        std::vector<std::string> tokens;
        Utility::split(cmd.get_cmd_option("-s"), ',', tokens);
        for (auto &t : tokens) schema.push_back(std::stoi(t));
    }
    MetaPath f(L, static_cast<int>(schema.size()), schema.data(), &graph);
#endif

    if (mode == "tables") {
        // static_initialization with each sampler in turn (creeper.h:102-175).
        const SamplingMethod keep = g_para.sample_;
        g_para.sample_ = AliasSampling;
        static_initialization(&graph, f);
        dump(out + "/alias.bin", graph.edge_weight_alias_table_, graph.num_edges());
        g_para.sample_ = InverseTransformationSampling;
        static_initialization(&graph, f);
        dump(out + "/prefix.bin", graph.edge_weight_prefix_sum_, graph.num_edges());
        g_para.sample_ = RejectionSampling;
        static_initialization(&graph, f);
        dump(out + "/rs_weight.bin", graph.edge_weight_rejection_, graph.num_edges());
        dump(out + "/rs_max.bin", graph.edge_weight_rejection_max_, graph.num_vertices());
        g_para.sample_ = keep;
        return 0;
    }

    if (mode == "sources") {
        // walker_generator.cpp:9-63: every non-isolated vertex shuffled (generate_all_source), -wn
        // random sources of seed -wseed (generate_random_source) and -wn copies of the max-degree vertex
        // (generate_single_source), the WalkerMeta::source_ of each walker in order.
        const int wn = static_cast<int>(opt_l(cmd, "-wn", 1000));
        const int wseed = static_cast<int>(opt_l(cmd, "-wseed", 0));
        const char *names[3] = {"all", "random", "single"};
        for (int k = 0; k < 3; ++k) {
            std::vector<WalkerMeta> walkers;
            if (k == 0) generate_all_source(graph, walkers);
            else if (k == 1) generate_random_source(graph, walkers, wn, wseed);
            else generate_single_source(graph, walkers, wn, wseed);
            std::vector<int32_t> s;
            for (const WalkerMeta &w : walkers) s.push_back(w.source_);
            dump(out + "/sources_" + names[k] + ".bin", s.data(), s.size());
        }
        return 0;
    }

    // mode == "walk"
    if (g_para.execution_ == Static) static_initialization(&graph, f);

    const std::vector<int32_t> sources = slurp<int32_t>(cmd.get_cmd_option("-src"));
    const int64_t n = static_cast<int64_t>(sources.size());
    const int64_t K = opt_l(cmd, "-K", 1024);
    const int64_t KT = opt_l(cmd, "-KT", 1024);
    const int path_len = L > 0 ? L : 1;

    std::vector<int32_t> paths(static_cast<size_t>(n * path_len), -1), lengths(n), ends(n), status(n), used(n);
    uint64_t total_steps = 0;

    sfmt_t gen, tgen;
    sfmt_init_gen_rand(&gen, static_cast<uint32_t>(opt_l(cmd, "-seed", 1234)));
    sfmt_init_gen_rand(&tgen, static_cast<uint32_t>(opt_l(cmd, "-tseed", 4321)));

    const int64_t maxd = graph.max_degree() > 0 ? graph.max_degree() : 1;
    std::vector<double> wbuf(static_cast<size_t>(maxd));
    std::vector<AliasSlot> abuf(static_cast<size_t>(maxd));
    std::queue<intT> large, small;

    for (int64_t i = 0; i < n; ++i) {
        BufferSlot ring[RING_SIZE];
        for (auto &s : ring) s.empty_ = true;
        BufferSlot &slot = ring[0];
        int32_t *seq = paths.data() + i * path_len;
        WalkerMeta w = {static_cast<int>(i), sources[i], sources[i], 1, seq};
        slot.empty_ = false;
        slot.local_id_ = 0;
        slot.w_ = w;
        slot.seq_ = seq;
        slot.seq_[0] = w.source_;
        slot.weight_ = wbuf.data();
        slot.alias_ = abuf.data();

        sfmt_t s = gen;
        const sfmt_t s0 = gen;
#if defined(REF_APP_ppr)
        f.sfmt_ = tgen;
#endif
        int completed = 0, terminated = 0, next = 1;
        uint64_t steps = 0;
        while (!slot.empty_) {
            if (g_para.execution_ == Uniform) {
                uniform_move(&graph, ring, &s, g_para.length_);
            } else if (g_para.execution_ == Static) {
                if (g_para.sample_ == InverseTransformationSampling) static_its_move(&graph, ring, &s, g_para.length_);
                else if (g_para.sample_ == AliasSampling) static_alias_move(&graph, ring, &s, g_para.length_);
                else if (g_para.sample_ == RejectionSampling) static_rs_move(&graph, ring, &s, g_para.length_);
            } else if (g_para.sample_ != MaxWeightRejectionSampling) {
                dynamic_gather(&graph, ring, f, g_para, completed, terminated, large, small);
                if (g_para.sample_ == InverseTransformationSampling) dynamic_its_move(&graph, ring, &s, g_para.length_);
                else if (g_para.sample_ == AliasSampling) dynamic_alias_move(&graph, ring, &s, g_para.length_);
                else if (g_para.sample_ == RejectionSampling) dynamic_rs_move(&graph, ring, &s, g_para.length_);
            } else {
                if (!slot.empty_) slot.dr_ = f.max_weight(slot.w_);
                dynamic_max_weight_rs_move(&graph, ring, &s, g_para.length_, f);
            }
            update(&graph, ring, completed, next, 1, &w, f, steps);
        }
        lengths[i] = slot.w_.length_;
        ends[i] = slot.w_.current_;
        status[i] = terminated ? 1 : 0;
        used[i] = static_cast<int32_t>(words_consumed(s0, s, K));
        total_steps += steps;

        for (int64_t k = 0; k < K; ++k) sfmt_genrand_uint32(&gen);
#if defined(REF_APP_ppr)
        for (int64_t k = 0; k < KT; ++k) sfmt_genrand_uint32(&tgen);
#endif
    }

    dump(out + "/paths.bin", paths.data(), paths.size());
    dump(out + "/lengths.bin", lengths.data(), lengths.size());
    dump(out + "/ends.bin", ends.data(), ends.size());
    dump(out + "/status.bin", status.data(), status.size());
    dump(out + "/draws_used.bin", used.data(), used.size());
    dump(out + "/steps.bin", &total_steps, 1);
    return 0;
}

// An app whose functor the library has no Move kernel for (test infrastructure).
//
// Written against the reference's app interface only -- creeper_main.h, Graph, WalkerMeta,
// g_para, generate_all_source, compute<F> -- so the same file compiles unchanged
//   * against the reference (oracle/Makefile: oracle/_ref/harness_customwalk, whose walks are
//     the golden fixtures tests/golden/functor_*.npz), and
//   * against thunderrw_amd/dropin (thunderrw_amd/dropin/bin/customwalk), where compute<F>
//     runs BiasedWalk's weight()/update()/max_weight() on the GPU in the functor engine.
//
//   customwalk -f <graph dir> -ew [-l 20] [-kind k] [-em e] [-sm s] [-p 2] [-q 0.5] [-wn n]
//
// -wn n walks n walkers from random sources (generate_random_source, seed 0) instead of one per
// non-isolated vertex.
//
// kind 0: walker-independent, edge weight over (1 + neighbour degree) (Static and Dynamic).
// kind 1: second order: the edge weight, x0.25 back to the previous vertex, x1 to a common
//         neighbour, x2 elsewhere (reads WalkerMeta::seq_ and Graph::is_neighbor).
// kind 2: the edge weight where (neighbour + length) % 4 == 0, else 0 (walkers end by the
//         gather); update() also ends a walk at a vertex of degree 1.
// kind 3: DeepWalk's weight and update under another name (apps/deepwalk.cpp:19-25).
// kind 4: Node2Vec's weight and update under another name, p_ = 1/p, q_ = 1/q
//         (apps/node2vec.cpp:24-42, 64-76).
// kind 5: DeepWalk's weight with an update() that never ends a walk (the reference would run
//         forever; the functor engine stops it at its move cap).
// kind 6: every weight 0 (max-weight rejection never accepts; the engine's trial cap ends it).
// kind 7: Node2Vec over edge weights: the edge weight times Node2Vec's factor (x max(1, p_, q_) on
//         the first move, x p_ back to the previous vertex, x1 to a common neighbour, x q_
//         elsewhere), bounded by the largest edge weight times max(1, p_, q_).
#include "creeper_main.h"

HyperParameter g_para = {ExecutionMode::Dynamic, SamplingMethod::InverseTransformationSampling, 0, {nullptr, 0}};

struct BiasedWalk {
    int l_;
    int kind_;
    double p_;
    double q_;
    double max_;
    double n2v_;  // max(1, p_, q_): Node2Vec's bound (apps/node2vec.cpp:19)
    double mw_;   // kind 7: the largest edge weight
    Graph* g_;

    BiasedWalk() {}

    BiasedWalk(int l, int kind, double p, double q, Graph* g) : l_(l), kind_(kind), p_(p), q_(q), g_(g) {
        n2v_ = q > p ? (q > 1.0 ? q : 1.0) : (p > 1.0 ? p : 1.0);
        mw_ = 0.0;
        if (kind == 7)
            for (int64_t e = 0; e < g->num_edges(); ++e)
                if (g->edge_weight_[e] > mw_) mw_ = g->edge_weight_[e];
        max_ = kind == 0 ? 2.5 : kind == 1 ? 10.0 : kind == 4 ? n2v_ : kind == 7 ? mw_ * n2v_ : 5.0;
    }

    BiasedWalk(const BiasedWalk& b)
        : l_(b.l_), kind_(b.kind_), p_(b.p_), q_(b.q_), max_(b.max_), n2v_(b.n2v_), mw_(b.mw_), g_(b.g_) {}

    inline double weight(WalkerMeta& w, intT begin, intT end, int64_t eid) {
        const double ew = g_->edge_weight_[eid];
        if (kind_ == 0) return ew / (1.0 + g_->degree(end));
        if (kind_ == 1) {
            if (w.length_ == 1) return ew;
            const int prev = w.seq_[w.length_ - 2];
            if (end == prev) return 0.25 * ew;
            return g_->is_neighbor(prev, end) ? ew : 2.0 * ew;
        }
        if (kind_ == 2) return ((end + w.length_) % 4 == 0) ? ew : 0.0;
        if (kind_ == 3 || kind_ == 5) return ew;
        if (kind_ == 6) return 0.0;
        if (kind_ == 7) {
            if (w.length_ == 1) return ew * n2v_;
            const int prev = w.seq_[w.length_ - 2];
            if (end == prev) return ew * p_;
            return g_->is_neighbor(prev, end) ? ew : ew * q_;
        }
        if (w.length_ == 1) return max_;
        const int prev = w.seq_[w.length_ - 2];
        if (end == prev) return p_;
        return g_->is_neighbor(prev, end) ? 1.0 : q_;
    }

    inline bool update(WalkerMeta& w, intT begin, intT end, int64_t eid) {
        if (kind_ == 5) return false;
        return w.length_ == l_ || (kind_ == 2 && g_->degree(end) == 1);
    }

    inline double max_weight(WalkerMeta& w) {
        return max_;
    }
};

void execute(Graph &graph, InputParser &cmd_parser) {
    int l = 20;
    if (!cmd_parser.get_cmd_option("-l").empty()) l = std::stoi(cmd_parser.get_cmd_option("-l"));
    int kind = 0;
    if (!cmd_parser.get_cmd_option("-kind").empty()) kind = std::stoi(cmd_parser.get_cmd_option("-kind"));
    double p = 2.0, q = 0.5;
    if (!cmd_parser.get_cmd_option("-p").empty()) p = std::stod(cmd_parser.get_cmd_option("-p"));
    if (!cmd_parser.get_cmd_option("-q").empty()) q = std::stod(cmd_parser.get_cmd_option("-q"));
    if (!cmd_parser.get_cmd_option("-em").empty())
        g_para.execution_ = static_cast<ExecutionMode>(std::stoi(cmd_parser.get_cmd_option("-em")));
    if (!cmd_parser.get_cmd_option("-sm").empty())
        g_para.sample_ = static_cast<SamplingMethod>(std::stoi(cmd_parser.get_cmd_option("-sm")));
    g_para.length_ = l;

    std::vector<WalkerMeta> walkers;
    if (!cmd_parser.get_cmd_option("-wn").empty())
        generate_random_source(graph, walkers, std::stoi(cmd_parser.get_cmd_option("-wn")), 0);
    else
        generate_all_source(graph, walkers);
    compute(graph, walkers, BiasedWalk(l, kind, 1.0 / p, 1.0 / q, &graph));
}

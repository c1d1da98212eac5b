// Host-side graph data: synthetic generator, the reference's CSR files, walker sources.
#pragma once

#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace trw {

struct HostGraph {
    int32_t nv = 0;
    int64_t ne = 0;
    std::vector<int32_t> degree;
    std::vector<int32_t> adj;
    std::vector<double> weight;  // empty when unweighted
    std::vector<int32_t> label;  // empty when unlabelled
};

// Number of worker threads for host-side preprocessing (OMP_NUM_THREADS if set,
// else the hardware concurrency, at most 64).
int host_threads();

// Runs fn(begin, end) over [0, n) split into contiguous chunks across threads.
template <class Fn>
void parallel_for(int64_t n, Fn fn) {
    int t = host_threads();
    if (n < 4096 || t <= 1) {
        fn(int64_t(0), n);
        return;
    }
    if (t > n / 1024) t = static_cast<int>(n / 1024);
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(t));
    for (int i = 0; i < t; ++i) {
        const int64_t b = n * i / t, e = n * (i + 1) / t;
        pool.emplace_back([=, &fn] { fn(b, e); });
    }
    for (auto &th : pool) th.join();
}

bool chung_lu(int32_t nv, int64_t undirected_edges, double gamma, double dmin, double dmax, uint64_t seed,
              bool weights, int32_t n_labels, HostGraph &out, std::string &err);
bool load_reference_csr(const std::string &dir, bool weights, bool labels, HostGraph &out, std::string &err);
bool store_reference_csr(const HostGraph &g, const std::string &dir, std::string &err);
bool load_edge_list(const std::string &dir, char skip, HostGraph &out, std::string &err);

// Stable per-vertex ordering of segment positions by label (the MetaPath label index).
void build_label_index(int32_t nv, const int64_t *offset, const int32_t *label, int32_t *sorted_labels,
                       int32_t *positions);

}  // namespace trw

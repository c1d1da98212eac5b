#include "hostgraph.h"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <numeric>
#include <sys/stat.h>

namespace trw {

int host_threads() {
    const char *e = std::getenv("OMP_NUM_THREADS");
    int t = e ? std::atoi(e) : 0;
    if (t <= 0) t = static_cast<int>(std::thread::hardware_concurrency());
    if (t <= 0) t = 1;
    return t > 64 ? 64 : t;
}

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Counter-based hash stream: value i of stream s under seed.
inline uint64_t hash_at(uint64_t seed, uint64_t s, uint64_t i) {
    return splitmix64(splitmix64(seed ^ (s * 0xD1B54A32D192ED03ull)) ^ i);
}

inline double unit_interval(uint64_t h) { return static_cast<double>(h >> 11) * 0x1.0p-53; }

void parallel_sort(std::vector<uint64_t> &v) {
    const size_t n = v.size();
    int t = host_threads();
    if (n < (size_t(1) << 20) || t <= 1) {
        std::sort(v.begin(), v.end());
        return;
    }
    int chunks = 1;
    while (chunks * 2 <= t) chunks *= 2;
    std::vector<size_t> b(static_cast<size_t>(chunks) + 1);
    for (int i = 0; i <= chunks; ++i) b[static_cast<size_t>(i)] = n * static_cast<size_t>(i) / static_cast<size_t>(chunks);
    {
        std::vector<std::thread> pool;
        for (int i = 0; i < chunks; ++i)
            pool.emplace_back([&, i] { std::sort(v.begin() + b[i], v.begin() + b[i + 1]); });
        for (auto &th : pool) th.join();
    }
    for (int width = 1; width < chunks; width *= 2) {
        std::vector<std::thread> pool;
        for (int i = 0; i + width < chunks; i += 2 * width) {
            const size_t lo = b[i], mid = b[i + width], hi = b[std::min(i + 2 * width, chunks)];
            pool.emplace_back([&, lo, mid, hi] { std::inplace_merge(v.begin() + lo, v.begin() + mid, v.begin() + hi); });
        }
        for (auto &th : pool) th.join();
    }
}

// Vose alias table over vertex weights, for O(1) endpoint draws.
struct EndpointSampler {
    std::vector<double> prob;
    std::vector<int32_t> alias;

    explicit EndpointSampler(const std::vector<double> &theta) {
        const size_t n = theta.size();
        prob.resize(n);
        alias.resize(n);
        const double total = std::accumulate(theta.begin(), theta.end(), 0.0);
        std::vector<int32_t> small, large;
        std::vector<double> q(n);
        for (size_t i = 0; i < n; ++i) {
            q[i] = theta[i] * static_cast<double>(n) / total;
            (q[i] < 1.0 ? small : large).push_back(static_cast<int32_t>(i));
        }
        while (!small.empty() && !large.empty()) {
            const int32_t s = small.back(), l = large.back();
            small.pop_back();
            prob[s] = q[s];
            alias[s] = l;
            q[l] = (q[l] + q[s]) - 1.0;
            if (q[l] < 1.0) {
                large.pop_back();
                small.push_back(l);
            }
        }
        for (int32_t s : small) prob[s] = 1.0, alias[s] = s;
        for (int32_t l : large) prob[l] = 1.0, alias[l] = l;
    }

    int32_t draw(uint64_t h) const {
        const uint64_t col = ((h & 0xffffffffull) * static_cast<uint64_t>(prob.size())) >> 32;
        const double x = static_cast<double>(h >> 32) * 0x1.0p-32;
        return x < prob[col] ? static_cast<int32_t>(col) : alias[col];
    }
};

bool read_exact(std::ifstream &f, void *dst, size_t bytes) {
    f.read(static_cast<char *>(dst), static_cast<std::streamsize>(bytes));
    return static_cast<size_t>(f.gcount()) == bytes;
}

template <class T>
bool read_array_file(const std::string &path, std::vector<T> &out, int64_t count, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    out.resize(static_cast<size_t>(count));
    if (!read_exact(f, out.data(), sizeof(T) * static_cast<size_t>(count))) {
        err = path + " is shorter than the degree file announces";
        return false;
    }
    return true;
}

template <class T>
bool write_array_file(const std::string &path, const T *data, int64_t count, std::string &err) {
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot create " + path;
        return false;
    }
    f.write(reinterpret_cast<const char *>(data), static_cast<std::streamsize>(sizeof(T) * static_cast<size_t>(count)));
    if (!f) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

}  // namespace

bool chung_lu(int32_t nv, int64_t target, double gamma, double dmin, double dmax, uint64_t seed, bool weights,
              int32_t n_labels, HostGraph &out, std::string &err) {
    if (nv < 2 || target < 1 || gamma <= 1.0 || dmin <= 0.0 || dmax < dmin) {
        err = "chung_lu: invalid parameters";
        return false;
    }
    if (static_cast<double>(target) > 0.45 * static_cast<double>(nv) * static_cast<double>(nv - 1)) {
        err = "chung_lu: too many edges for the vertex count";
        return false;
    }
    std::vector<double> theta(static_cast<size_t>(nv));
    const double expo = -1.0 / (gamma - 1.0);
    parallel_for(nv, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            theta[i] = std::min(dmax, dmin * std::pow(1.0 - unit_interval(hash_at(seed, 1, i)), expo));
    });
    const EndpointSampler sampler(theta);
    theta.clear();
    theta.shrink_to_fit();

    const uint64_t kNone = ~0ull;
    std::vector<uint64_t> keys;
    uint64_t drawn = 0;
    const int64_t enough = static_cast<int64_t>(0.98 * static_cast<double>(target));
    for (int round = 0; static_cast<int64_t>(keys.size()) < enough; ++round) {
        if (round >= 64) {
            err = "chung_lu: could not reach the edge target (graph too dense for its degree cap)";
            return false;
        }
        const int64_t missing = target - static_cast<int64_t>(keys.size());
        const int64_t m = round == 0 ? target : missing + missing / 8 + 1024;
        const size_t base = keys.size();
        keys.resize(base + static_cast<size_t>(m));
        parallel_for(m, [&](int64_t b, int64_t e) {
            for (int64_t j = b; j < e; ++j) {
                const uint64_t idx = drawn + static_cast<uint64_t>(j);
                const uint64_t u = static_cast<uint64_t>(sampler.draw(hash_at(seed, 2, idx)));
                const uint64_t v = static_cast<uint64_t>(sampler.draw(hash_at(seed, 3, idx)));
                keys[base + static_cast<size_t>(j)] = (u == v) ? kNone : ((std::min(u, v) << 32) | std::max(u, v));
            }
        });
        drawn += static_cast<uint64_t>(m);
        parallel_sort(keys);
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        if (!keys.empty() && keys.back() == kNone) keys.pop_back();
    }

    out.nv = nv;
    out.ne = 2 * static_cast<int64_t>(keys.size());
    out.degree.assign(static_cast<size_t>(nv), 0);
    for (uint64_t k : keys) {
        out.degree[k >> 32] += 1;
        out.degree[k & 0xffffffffull] += 1;
    }
    std::vector<int64_t> fillpos(static_cast<size_t>(nv) + 1, 0);
    for (int32_t v = 0; v < nv; ++v) fillpos[v + 1] = fillpos[v] + out.degree[v];
    out.adj.resize(static_cast<size_t>(out.ne));
    // Keys are sorted by (lo, hi): every vertex receives its lower neighbours (as hi)
    // before its higher ones (as lo), each group ascending, so rows come out sorted.
    for (uint64_t k : keys) {
        const int32_t lo = static_cast<int32_t>(k >> 32), hi = static_cast<int32_t>(k & 0xffffffffull);
        out.adj[fillpos[lo]++] = hi;
        out.adj[fillpos[hi]++] = lo;
    }
    keys.clear();
    keys.shrink_to_fit();

    if (weights || n_labels > 0) {
        if (weights) out.weight.resize(static_cast<size_t>(out.ne));
        if (n_labels > 0) out.label.resize(static_cast<size_t>(out.ne));
        std::vector<int64_t> off(static_cast<size_t>(nv) + 1, 0);
        for (int32_t v = 0; v < nv; ++v) off[v + 1] = off[v] + out.degree[v];
        parallel_for(nv, [&](int64_t b, int64_t e) {
            for (int64_t u = b; u < e; ++u)
                for (int64_t j = off[u]; j < off[u + 1]; ++j) {
                    const uint64_t x = static_cast<uint64_t>(out.adj[j]);
                    const uint64_t key = (std::min<uint64_t>(u, x) << 32) | std::max<uint64_t>(u, x);
                    if (weights) out.weight[j] = 1.0 + 4.0 * unit_interval(hash_at(seed, 4, key));
                    if (n_labels > 0)
                        out.label[j] = static_cast<int32_t>(hash_at(seed, 5, key) % static_cast<uint64_t>(n_labels));
                }
        });
    }
    return true;
}

// Graph::load_csr_degree_file / load_csr_adj_file / edge weight and label
// loaders (util/graph/graph.cpp:233-370), with the size checks the reference omits.
bool load_reference_csr(const std::string &dir, bool weights, bool labels, HostGraph &out, std::string &err) {
    std::ifstream f(dir + "/b_degree.bin", std::ios::binary);
    if (!f) {
        err = "cannot open " + dir + "/b_degree.bin";
        return false;
    }
    uint32_t elem = 0;
    int32_t nv = 0;
    int64_t ne = 0;
    if (!read_exact(f, &elem, 4) || !read_exact(f, &nv, 4) || !read_exact(f, &ne, 8)) {
        err = "b_degree.bin: short header";
        return false;
    }
    if (elem != 4 || nv < 0 || ne < 0) {
        err = "b_degree.bin: unsupported element size or negative sizes";
        return false;
    }
    out.nv = nv;
    out.ne = ne;
    out.degree.resize(static_cast<size_t>(nv));
    if (!read_exact(f, out.degree.data(), 4 * static_cast<size_t>(nv))) {
        err = "b_degree.bin: fewer degrees than |V|";
        return false;
    }
    int64_t sum = 0;
    for (int32_t d : out.degree) {
        if (d < 0) {
            err = "b_degree.bin: negative degree";
            return false;
        }
        sum += d;
    }
    if (sum != ne) {
        err = "b_degree.bin: degrees do not sum to |E|";
        return false;
    }
    if (!read_array_file(dir + "/b_adj.bin", out.adj, ne, err)) return false;
    out.weight.clear();
    out.label.clear();
    if (weights && !read_array_file(dir + "/b_edge_weight.bin", out.weight, ne, err)) return false;
    if (labels && !read_array_file(dir + "/b_edge_label.bin", out.label, ne, err)) return false;
    return true;
}

// Graph::store_csr (util/graph/graph.cpp:123-212).
bool store_reference_csr(const HostGraph &g, const std::string &dir, std::string &err) {
    if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) {
        err = "cannot create directory " + dir;
        return false;
    }
    {
        std::ofstream f(dir + "/b_degree.bin", std::ios::binary);
        if (!f) {
            err = "cannot create " + dir + "/b_degree.bin";
            return false;
        }
        const uint32_t elem = 4;
        f.write(reinterpret_cast<const char *>(&elem), 4);
        f.write(reinterpret_cast<const char *>(&g.nv), 4);
        f.write(reinterpret_cast<const char *>(&g.ne), 8);
        f.write(reinterpret_cast<const char *>(g.degree.data()), static_cast<std::streamsize>(4 * g.degree.size()));
        if (!f) {
            err = "write failed: b_degree.bin";
            return false;
        }
    }
    if (!write_array_file(dir + "/b_adj.bin", g.adj.data(), g.ne, err)) return false;
    if (!g.weight.empty() && !write_array_file(dir + "/b_edge_weight.bin", g.weight.data(), g.ne, err)) return false;
    if (!g.label.empty() && !write_array_file(dir + "/b_edge_label.bin", g.label.data(), g.ne, err)) return false;
    return true;
}

// Graph::load_edge_list(dir, skip) (util/graph/graph.cpp:428-519): text lines "src dst"
// of <dir>/b_edge_list.bin, lines starting with `skip` ignored; each pair ordered
// (min, max), pairs sorted, duplicates and self-loops dropped, both directions added.
// The vertex count is the largest id seen on any line + 1.  The adjacency of v comes
// out sorted: its smaller neighbours (edges whose first end is smaller) precede its
// larger ones, each in ascending order, exactly as the reference's append order gives.
namespace {

// `ss >> src >> dst` for int32: skips whitespace, optional sign, digits; false on failure.
bool parse_int32(const char *&p, const char *end, int32_t &out) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    bool neg = false;
    if (p < end && (*p == '+' || *p == '-')) neg = *p++ == '-';
    if (p >= end || *p < '0' || *p > '9') return false;
    int64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        v = v * 10 + (*p++ - '0');
        if (v > int64_t(INT32_MAX) + 1) return false;
    }
    if (neg) v = -v;
    if (v > INT32_MAX || v < INT32_MIN) return false;
    out = static_cast<int32_t>(v);
    return true;
}

}  // namespace

bool load_edge_list(const std::string &dir, char skip, HostGraph &out, std::string &err) {
    const std::string path = dir + "/b_edge_list.bin";
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<uint64_t> keys;
    int32_t max_id = 0;
    uint64_t line_no = 0;
    const char *p = text.data(), *end = text.data() + text.size();
    while (p < end) {
        const char *eol = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
        if (!eol) eol = end;
        ++line_no;
        // std::string::operator[] on an empty line gives '\0', which is never the skip
        // character, so (as in the reference) an empty line must parse as an edge.
        const char first = p < eol ? *p : '\0';
        if (first != skip) {
            const char *q = p;
            int32_t a, b;
            if (!parse_int32(q, eol, a) || !parse_int32(q, eol, b)) {
                err = "cannot convert line " + std::to_string(line_no) + " to edge";
                return false;
            }
            if (a < 0 || b < 0) {
                err = "negative vertex id on line " + std::to_string(line_no);
                return false;
            }
            if (a == INT32_MAX || b == INT32_MAX) {  // the vertex count (largest id + 1) must fit int32
                err = "vertex id 2147483647 on line " + std::to_string(line_no) + " leaves no int32 vertex count";
                return false;
            }
            if (a > b) std::swap(a, b);
            keys.push_back((static_cast<uint64_t>(a) << 32) | static_cast<uint32_t>(b));
            max_id = std::max(max_id, b);
        }
        p = eol < end ? eol + 1 : end;
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int32_t nv = max_id + 1;
    std::vector<int64_t> off(static_cast<size_t>(nv) + 1, 0);
    std::vector<int32_t> deg(static_cast<size_t>(nv), 0);
    int64_t ne = 0;
    for (uint64_t k : keys) {
        const int32_t a = static_cast<int32_t>(k >> 32), b = static_cast<int32_t>(k & 0xffffffffu);
        if (a == b) continue;
        ++deg[a];
        ++deg[b];
        ne += 2;
    }
    for (int32_t v = 0; v < nv; ++v) off[v + 1] = off[v] + deg[v];
    std::vector<int32_t> adj(static_cast<size_t>(ne));
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (uint64_t k : keys) {
        const int32_t a = static_cast<int32_t>(k >> 32), b = static_cast<int32_t>(k & 0xffffffffu);
        if (a == b) continue;
        adj[fill[a]++] = b;
        adj[fill[b]++] = a;
    }
    out = HostGraph();
    out.nv = nv;
    out.ne = ne;
    out.degree.swap(deg);
    out.adj.swap(adj);
    return true;
}

void build_label_index(int32_t nv, const int64_t *offset, const int32_t *label, int32_t *sorted_labels,
                       int32_t *positions) {
    parallel_for(nv, [&](int64_t b, int64_t e) {
        std::vector<int32_t> pos;
        for (int64_t v = b; v < e; ++v) {
            const int64_t off = offset[v];
            const int32_t d = static_cast<int32_t>(offset[v + 1] - off);
            pos.resize(static_cast<size_t>(d));
            for (int32_t i = 0; i < d; ++i) pos[i] = i;
            std::stable_sort(pos.begin(), pos.end(),
                             [&](int32_t x, int32_t y) { return label[off + x] < label[off + y]; });
            for (int32_t i = 0; i < d; ++i) {
                positions[off + i] = pos[i];
                sorted_labels[off + i] = label[off + pos[i]];
            }
        }
    });
}

}  // namespace trw

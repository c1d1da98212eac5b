// Static table builds on the device: static_initialization (creeper.h:102-175).
//
// Every table is byte-identical to the reference's, so each vertex's FP64 work
// runs in the reference's order on one lane: a left-to-right sum, IEEE division,
// serial prefix sums.  Vertices are visited in descending degree order so the
// long serial chains of hub vertices start first.
#include "engine.h"

namespace trw {
namespace {

constexpr int kBuildBlock = 128;

__device__ __forceinline__ double weight_at(const double *w, int32_t unit, int64_t e) { return unit ? 1.0 : w[e]; }

// alias_initialization (alias_sampling.h:12-54) without queues.
//
// The reference keeps two FIFO queues.  `large` is only ever popped at its
// front, so it is the list of initial large indices in index order, consumed by
// a cursor li.  `small` receives the initial small indices (index order) and then,
// at its back, each large index at the moment it drops below 1 -- which happens
// in `large` order.  So the small queue is always initial_smalls ++ larges[0..li)
// and a second cursor si walks that concatenation.  lists[] holds the small
// indices ascending from the segment start and the large indices from the
// segment end backwards; q[] holds the normalised weights, and a large index's
// residual is written back to q[] when it moves to the small side.
__global__ __launch_bounds__(kBuildBlock) void alias_build_kernel(DevGraph g, const double *w, int32_t unit,
                                                                   const int32_t *order, double *q, int32_t *lists,
                                                                   AliasSlot *out, unsigned int *zero_count) {
    const int32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= g.nv) return;
    const int32_t v = order ? order[i] : i;
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    if (d == 0) return;

    double sum = 0.0;  // std::accumulate(src, src + length, 0.0)
    for (int32_t j = 0; j < d; ++j) sum = sum + weight_at(w, unit, off + j);
    const double avg = sum / d;
    AliasSlot *t = out + off;
    if (avg == 0.0) {
        for (int32_t j = 0; j < d; ++j) t[j] = AliasSlot{0.0, -1, -1};
        atomicAdd(zero_count, 1u);
        return;
    }

    double *qv = q + off;
    int32_t *lst = lists + off;
    const int32_t *nb = g.adj + off;
    int32_t ns = 0, nl = 0;
    for (int32_t j = 0; j < d; ++j) {
        const double x = weight_at(w, unit, off + j) / avg;
        qv[j] = x;
        if (x < 1) lst[ns++] = j;
        else lst[d - 1 - nl++] = j;
    }

    int32_t k = 0, si = 0, li = 0;
    int32_t l = nl > 0 ? lst[d - 1] : 0;
    double wl = nl > 0 ? qv[l] : 0.0;
    while (si < ns + li && li < nl) {
        const int32_t s = si < ns ? lst[si] : lst[d - 1 - (si - ns)];
        const double ws = qv[s];
        t[k++] = AliasSlot{ws, nb[s], nb[l]};
        ++si;
        wl = wl + ws - 1;
        if (wl < 1) {
            qv[l] = wl;
            ++li;
            if (li < nl) {
                l = lst[d - 1 - li];
                wl = qv[l];
            }
        }
    }
    for (; si < ns + li; ++si) {
        const int32_t s = si < ns ? lst[si] : lst[d - 1 - (si - ns)];
        t[k++] = AliasSlot{1.0, nb[s], nb[s]};
    }
    for (; li < nl; ++li) {
        const int32_t x = lst[d - 1 - li];
        t[k++] = AliasSlot{1.0, nb[x], nb[x]};
    }
}

// its_initialization (inverse_transformation_sampling.h:19-24).
__global__ __launch_bounds__(kBuildBlock) void its_build_kernel(DevGraph g, const double *w, int32_t unit,
                                                                 const int32_t *order, double *out) {
    const int32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= g.nv) return;
    const int32_t v = order ? order[i] : i;
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    if (d == 0) return;
    double acc = weight_at(w, unit, off);
    out[off] = acc;
    for (int32_t j = 1; j < d; ++j) {
        acc = acc + weight_at(w, unit, off + j);
        out[off + j] = acc;
    }
}

// rs_initialization (rejection_sampling.h:12-14): *std::max_element.
__global__ __launch_bounds__(kBuildBlock) void rs_build_kernel(DevGraph g, const double *w, int32_t unit,
                                                                const int32_t *order, double *out) {
    const int32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= g.nv) return;
    const int32_t v = order ? order[i] : i;
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    if (d == 0) {
        out[v] = 0.0;
        return;
    }
    double m = weight_at(w, unit, off);
    for (int32_t j = 1; j < d; ++j) {
        const double x = weight_at(w, unit, off + j);
        if (m < x) m = x;
    }
    out[v] = m;
}

// Neighbour records: a Move that lands on adj[e] also learns where that
// vertex's adjacency starts and how long it is, so the next step needs no
// separate vertex gather.
__global__ void edge_records_kernel(const int32_t *adj, const uint2 *vrec, int64_t ne, uint4 *erec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t v = adj[e];
        const uint2 r = vrec[v];
        erec[e] = make_uint4(static_cast<uint32_t>(v), r.x, r.y, 0u);
    }
}

// Label starts: for each vertex the number of neighbours with label < k, k = 0..7,
// by binary search of its label-sorted segment.
__global__ void label_start_kernel(const int64_t *offset, const int32_t *lbl_sorted, int32_t nv, uint32_t *lstart) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int64_t off = offset[v];
    const int32_t d = static_cast<int32_t>(offset[v + 1] - off);
    const int32_t *ls = lbl_sorted + off;
    for (int32_t k = 0; k < 8; ++k) {
        int32_t lo = 0, hi = d;
        while (lo < hi) {
            const int32_t m = (lo + hi) >> 1;
            if (ls[m] < k) lo = m + 1;
            else hi = m;
        }
        lstart[8 * static_cast<int64_t>(v) + k] = static_cast<uint32_t>(lo);
    }
}

// Neighbour records in label order: entry j of v's label-sorted segment is the
// neighbour record of segment position lbl_pos[off + j].
__global__ void label_records_kernel(const int64_t *offset, int32_t nv, const int32_t *lbl_pos, const uint4 *erec,
                                     int64_t ne, uint4 *lerec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        // The vertex whose segment holds e: the last v with offset[v] <= e.
        int32_t lo = 0, hi = nv;
        while (hi - lo > 1) {
            const int32_t m = (lo + hi) >> 1;
            if (offset[m] <= e) lo = m;
            else hi = m;
        }
        lerec[e] = erec[offset[lo] + lbl_pos[e]];
    }
}

// ITS search trees.  One thread per vertex; hubs write a few thousand keys once.
__global__ void its_tree_sizes_kernel(const int64_t *offset, int32_t nv, uint32_t *count) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int64_t off = offset[v];
    int64_t nodes[kItsMaxLevels];
    const int32_t L = its_levels(its_leaf_blocks(off, static_cast<int32_t>(offset[v + 1] - off)), nodes);
    int64_t total = 0;
    for (int32_t k = 0; k < L; ++k) total += nodes[k];
    count[v] = static_cast<uint32_t>(total);
}

__global__ void its_tree_fill_kernel(const int64_t *offset, int32_t nv, const double *P, const uint32_t *root,
                                     double *idx) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv || root[v] == kNoRange) return;
    const int64_t off = offset[v];
    const int32_t d = static_cast<int32_t>(offset[v + 1] - off);
    const int32_t nb = its_leaf_blocks(off, d);
    int64_t nodes[kItsMaxLevels];
    const int32_t L = its_levels(nb, nodes);
    // Level k (0 = level 1) starts after the levels above it.
    int64_t base[kItsMaxLevels];
    int64_t acc = 0;
    for (int32_t k = L - 1; k >= 0; --k) {
        base[k] = acc;
        acc += nodes[k];
    }
    double *tree = idx + static_cast<int64_t>(root[v]) * kItsFan;
    const int64_t first_end = ((off >> 4) + 1) << 4;
    // Level 1: the last prefix value of every leaf block.
    double *l1 = tree + base[0] * kItsFan;
    for (int64_t i = 0; i < nodes[0] * kItsFan; ++i) {
        double key = __longlong_as_double(0x7ff0000000000000LL);  // +inf padding
        if (i < nb) {
            const int64_t end = (i == 0) ? first_end : first_end + kItsFan * i;
            key = P[(end < off + d ? end : off + d) - 1];
        }
        l1[i] = key;
    }
    // Higher levels: the last key of every node below.
    int64_t keys_below = nb;
    for (int32_t k = 1; k < L; ++k) {
        const double *below = tree + base[k - 1] * kItsFan;
        double *lk = tree + base[k] * kItsFan;
        for (int64_t i = 0; i < nodes[k] * kItsFan; ++i) {
            double key = __longlong_as_double(0x7ff0000000000000LL);
            if (i < nodes[k - 1]) {
                const int64_t last = (i + 1) * kItsFan - 1;
                key = below[last < keys_below ? last : keys_below - 1];
            }
            lk[i] = key;
        }
        keys_below = nodes[k - 1];
    }
}

__global__ void its_tree_roots_kernel(const int32_t *adj, int64_t ne, const uint32_t *root, uint4 *erec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        erec[e].w = root[adj[e]];
}

// Alias walk records: the reference AliasSlot plus both candidates' ranges.
__global__ void alias_records_kernel(const AliasSlot *alias, const uint2 *vrec, int64_t ne, uint4 *arec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const AliasSlot a = alias[e];
        const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(a.prob));
        arec[2 * e] = make_uint4(static_cast<uint32_t>(bits), static_cast<uint32_t>(bits >> 32),
                                 static_cast<uint32_t>(a.first), static_cast<uint32_t>(a.second));
        if (a.first < 0) {
            arec[2 * e + 1] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            const uint2 f = vrec[a.first], s = vrec[a.second];
            arec[2 * e + 1] = make_uint4(f.x, f.y, s.x, s.y);
        }
    }
}

inline unsigned grid_for(int64_t n) { return static_cast<unsigned>((n + kBuildBlock - 1) / kBuildBlock); }

}  // namespace

hipError_t launch_alias_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, double *qbuf,
                              int32_t *lists, AliasSlot *out, unsigned int *zero_count, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(alias_build_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g, w, unit, order,
                       qbuf, lists, out, zero_count);
    return hipGetLastError();
}

hipError_t launch_edge_records(const DevGraph &g, uint4 *erec, hipStream_t stream) {
    if (g.ne <= 0) return hipSuccess;
    hipLaunchKernelGGL(edge_records_kernel, dim3(8192), dim3(256), 0, stream, g.adj, g.vrec, g.ne, erec);
    return hipGetLastError();
}

hipError_t launch_alias_records(const DevGraph &g, uint4 *arec, hipStream_t stream) {
    if (g.ne <= 0) return hipSuccess;
    hipLaunchKernelGGL(alias_records_kernel, dim3(8192), dim3(256), 0, stream, g.alias, g.vrec, g.ne, arec);
    return hipGetLastError();
}

hipError_t launch_label_records(const DevGraph &g, uint32_t *lstart, uint4 *lerec, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(label_start_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.lbl_sorted,
                       g.nv, lstart);
    if (g.ne > 0)
        hipLaunchKernelGGL(label_records_kernel, dim3(8192), dim3(256), 0, stream, g.offset, g.nv, g.lbl_pos, g.erec,
                           g.ne, lerec);
    return hipGetLastError();
}

hipError_t launch_its_tree_sizes(const DevGraph &g, uint32_t *root, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(its_tree_sizes_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.nv, root);
    return hipGetLastError();
}

hipError_t launch_its_tree_fill(const DevGraph &g, const double *prefix, const uint32_t *root, double *idx,
                                hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(its_tree_fill_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.nv, prefix,
                       root, idx);
    return hipGetLastError();
}

hipError_t launch_its_tree_roots_to_records(const DevGraph &g, const uint32_t *root, uint4 *erec, hipStream_t stream) {
    if (g.ne <= 0 || !erec) return hipSuccess;
    hipLaunchKernelGGL(its_tree_roots_kernel, dim3(8192), dim3(256), 0, stream, g.adj, g.ne, root, erec);
    return hipGetLastError();
}

hipError_t launch_its_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, double *out,
                            hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(its_build_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g, w, unit, order, out);
    return hipGetLastError();
}

hipError_t launch_rs_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, double *out,
                           hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(rs_build_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g, w, unit, order, out);
    return hipGetLastError();
}

}  // namespace trw

// Static table builds on the device: static_initialization (creeper.h:102-175).
//
// Every table is byte-identical to the reference's, so each vertex's FP64 work
// runs in the reference's order on one lane: a left-to-right sum, IEEE division,
// serial prefix sums.  Vertices are visited in descending degree order so the
// long serial chains of hub vertices start first.
#include "engine.h"
#include "thunderrw_device.h"

#include <rocprim/device/device_scan.hpp>

#include <algorithm>

namespace trw {
namespace {

constexpr int kBuildBlock = 128;
constexpr int kAliasChunk = 16;      // alias build: entries loaded per batch of the serial loops
constexpr int kAliasLargeChunk = 8;  // alias build: large entries held ahead

__device__ __forceinline__ double weight_at(const double *w, int32_t unit, int64_t e) { return unit ? 1.0 : w[e]; }

// Waits for a conditional refill's loads where they are issued.  Vector memory
// operations complete in issue order and the counter covers stores too, so a wait
// placed where the refilled and the untouched paths merge (the compiler's choice)
// would also wait for every store since, on every pass of the pairing loop.
__device__ __forceinline__ void refill_wait() { __builtin_amdgcn_s_waitcnt(0); }

// alias_initialization (alias_sampling.h:12-54) without queues.
//
// The reference keeps two FIFO queues.  `large` is only ever popped at its
// front, so it is the list of initial large indices in index order, consumed by
// a cursor li.  `small` receives the initial small indices (index order) and then,
// at its back, each large index at the moment it drops below 1 -- which happens
// in `large` order.  So the small queue is always initial_smalls ++ larges[0..li)
// and a second cursor si walks that concatenation.  The scratch entries hold each
// index's normalised weight and neighbour id: smalls ascending from the segment
// start, larges from the segment end backwards, and a large entry's weight is
// replaced by its residual when it moves to the small side.  Every load address in
// the pairing loop then depends on the cursors only, never on a loaded value.
__device__ __forceinline__ void alias_build_vertex(const DevGraph &g, const double *w, int32_t unit, int32_t v,
                                                   AliasEntry *scratch, AliasSlot *out, unsigned int *zero_count) {
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    if (d == 0) return;

    double sum = 0.0;  // std::accumulate(src, src + length, 0.0): left to right
    if (unit) {
        for (int32_t j = 0; j < d; ++j) sum = sum + 1.0;
    } else {
        const double *wv = w + off;
        int32_t j = 0;
        for (; j + kAliasChunk <= d; j += kAliasChunk) {  // a chunk of loads in flight, added in order
            double x[kAliasChunk];
#pragma unroll
            for (int c = 0; c < kAliasChunk; ++c) x[c] = wv[j + c];
#pragma unroll
            for (int c = 0; c < kAliasChunk; ++c) sum = sum + x[c];
        }
        for (; j < d; ++j) sum = sum + wv[j];
    }
    const double avg = sum / d;
    AliasSlot *t = out + off;
    if (avg == 0.0) {
        for (int32_t j = 0; j < d; ++j) t[j] = AliasSlot{0.0, -1, -1};
        atomicAdd(zero_count, 1u);
        return;
    }

    AliasEntry *E = scratch + off;
    const int32_t *nb = g.adj + off;
    int32_t ns = 0, nl = 0;
    for (int32_t j0 = 0; j0 < d; j0 += kAliasChunk) {  // loads a chunk ahead, classifies in order
        double wc[kAliasChunk];
        int32_t nc[kAliasChunk];
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c) {  // unconditional (clamped) loads issue back to back
            const int32_t j = min(j0 + c, d - 1);
            wc[c] = weight_at(w, unit, off + j);
            nc[c] = nb[j];
        }
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c) {
            if (j0 + c >= d) break;
            const double x = wc[c] / avg;
            const AliasEntry en{x, nc[c], 0};
            if (x < 1) E[ns++] = en;
            else E[d - 1 - nl++] = en;
        }
    }

    int32_t k = 0, si = 0, li = 0;
    // Larges are consumed in order from the segment end; a chunk of their entries is
    // held ahead (their weights change only once they are consumed).
    AliasEntry lbuf[kAliasLargeChunk];
    int32_t lbase = 0;  // large index of lbuf[0]
#pragma unroll
    for (int c = 0; c < kAliasLargeChunk; ++c) lbuf[c] = E[max(d - 1 - c, 0)];
    int32_t lpos = d - 1;
    double wl = nl > 0 ? lbuf[0].q : 0.0;
    int32_t lnb = nl > 0 ? lbuf[0].nbr : 0;
    double drop_q = 0.0;  // residual and neighbour of the large that dropped last
    int32_t drop_nbr = 0;
    while (si < ns + li && li < nl) {
        if (si >= ns && si - ns == li - 1) {
            // The only pending small is the large that just dropped: once the initial
            // smalls are used up this is the common case, and its entry is in registers.
            t[k++] = AliasSlot{drop_q, drop_nbr, lnb};
            ++si;
            wl = wl + drop_q - 1;
            if (wl < 1) {
                E[lpos].q = wl;
                drop_q = wl;
                drop_nbr = lnb;
                ++li;
                if (li < nl) {
                    if (li - lbase == kAliasLargeChunk) {
                        lbase = li;
#pragma unroll
                        for (int c2 = 0; c2 < kAliasLargeChunk; ++c2) lbuf[c2] = E[max(d - 1 - (lbase + c2), 0)];
                        refill_wait();
                    }
                    lpos = d - 1 - li;
                    AliasEntry le = lbuf[0];
#pragma unroll
                    for (int c2 = 1; c2 < kAliasLargeChunk; ++c2)
                        if (li - lbase == c2) le = lbuf[c2];
                    wl = le.q;
                    lnb = le.nbr;
                }
            }
            continue;
        }
        // The next small entries that are already final (initial smalls, and larges that
        // have dropped with their residual written), a chunk at a time so that their
        // loads overlap; the pairing itself stays strictly sequential.
        const int32_t avail = min(kAliasChunk, ns + li - si);
        AliasEntry buf[kAliasChunk];
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c) {  // clamped to the segment; entries past avail are unused
            const int32_t p = si + c < ns ? si + c : d - 1 - (si + c - ns);
            buf[c] = E[max(0, min(p, d - 1))];
        }
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c) {
            if (c >= avail || li >= nl) break;
            const AliasEntry se = buf[c];
            t[k++] = AliasSlot{se.q, se.nbr, lnb};
            ++si;
            wl = wl + se.q - 1;
            if (wl < 1) {
                E[lpos].q = wl;
                drop_q = wl;
                drop_nbr = lnb;
                ++li;
                if (li < nl) {
                    if (li - lbase == kAliasLargeChunk) {
                        lbase = li;
#pragma unroll
                        for (int c2 = 0; c2 < kAliasLargeChunk; ++c2) lbuf[c2] = E[max(d - 1 - (lbase + c2), 0)];
                        refill_wait();
                    }
                    lpos = d - 1 - li;
                    AliasEntry le = lbuf[0];
#pragma unroll
                    for (int c2 = 1; c2 < kAliasLargeChunk; ++c2)
                        if (li - lbase == c2) le = lbuf[c2];
                    wl = le.q;
                    lnb = le.nbr;
                }
            }
        }
    }
    // Leftovers take probability 1 (alias_sampling.h:43-53); their loads are independent.
    for (int32_t s0 = si; s0 < ns + li; s0 += kAliasChunk) {
        int32_t x[kAliasChunk];
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c) {
            const int32_t p = s0 + c < ns ? s0 + c : d - 1 - (s0 + c - ns);
            x[c] = E[max(0, min(p, d - 1))].nbr;
        }
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c)
            if (s0 + c < ns + li) t[k + (s0 + c - si)] = AliasSlot{1.0, x[c], x[c]};
    }
    k += ns + li - si;
    for (int32_t l0 = li; l0 < nl; l0 += kAliasChunk) {
        int32_t x[kAliasChunk];
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c) x[c] = E[max(d - 1 - (l0 + c), 0)].nbr;
#pragma unroll
        for (int c = 0; c < kAliasChunk; ++c)
            if (l0 + c < nl) t[k + (l0 + c - li)] = AliasSlot{1.0, x[c], x[c]};
    }
}

// its_initialization (inverse_transformation_sampling.h:19-24).
__device__ __forceinline__ void its_build_vertex(const DevGraph &g, const double *w, int32_t unit, int32_t v,
                                                 double *out) {
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
__device__ __forceinline__ void rs_build_vertex(const DevGraph &g, const double *w, int32_t unit, int32_t v,
                                                double *out) {
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


// ---------------------------------------------------------------- hub builds
//
// A vertex of degree >= kWaveBuildMin gets a whole wavefront.  The reference's
// arithmetic stays sequential (a left-to-right FP64 sum, the prefix chain, Vose's
// pairing), so the tables are the same bytes; what the lanes share is the memory
// traffic: they load 64 consecutive entries at a time, the sequential chain reads
// them from registers (readlane), and results are written back 64 at a time.

__device__ __forceinline__ double readlane_f64(double x, int32_t lane) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

__device__ __forceinline__ int32_t readlane_i32(int32_t x, int32_t lane) { return __builtin_amdgcn_readlane(x, lane); }

// Hub loops take their entries 256 at a time: four 64-entry groups are loaded before
// the sequential chain reads the first, so one memory latency covers 256 entries.
constexpr int kWaveGroups = 4;

// std::accumulate(w, w + d, 0.0): the same additions in the same order.
__device__ __forceinline__ double wave_ordered_sum(const double *w, int32_t unit, int64_t off, int32_t d, int lane) {
    double sum = 0.0;
    for (int32_t base = 0; base < d; base += 64 * kWaveGroups) {
        double x[kWaveGroups];
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t j = base + 64 * h + lane;
            x[h] = j < d ? weight_at(w, unit, off + j) : 0.0;
        }
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t n = min(64, d - (base + 64 * h));
            for (int32_t c = 0; c < n; ++c) sum = sum + readlane_f64(x[h], c);
        }
    }
    return sum;
}

__device__ __forceinline__ void alias_build_hub(const DevGraph &g, const double *w, int32_t unit, int32_t v,
                                                AliasEntry *scratch, AliasSlot *out, unsigned int *zero_count) {
    const int lane = static_cast<int>(threadIdx.x);
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    const double avg = wave_ordered_sum(w, unit, off, d, lane) / d;
    AliasSlot *t = out + off;
    if (avg == 0.0) {
        for (int32_t j = lane; j < d; j += 64) t[j] = AliasSlot{0.0, -1, -1};
        if (lane == 0) atomicAdd(zero_count, 1u);
        return;
    }

    // Classification in index order: a ballot ranks each lane's entry among the smalls
    // (or larges) of its 64-entry group.
    AliasEntry *E = scratch + off;
    const int32_t *nb = g.adj + off;
    const unsigned long long below = (1ull << lane) - 1ull;
    int32_t ns = 0, nl = 0;
    for (int32_t base = 0; base < d; base += 64 * kWaveGroups) {
        double wx[kWaveGroups];
        int32_t nx[kWaveGroups];
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t j = base + 64 * h + lane;
            wx[h] = j < d ? weight_at(w, unit, off + j) : 0.0;
            nx[h] = j < d ? nb[j] : 0;
        }
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const bool valid = base + 64 * h + lane < d;
            const double x = valid ? wx[h] / avg : 0.0;
            const bool small = valid && x < 1;
            const bool large = valid && !(x < 1);
            const unsigned long long ms = __ballot(small), ml = __ballot(large);
            if (small) E[ns + __popcll(ms & below)] = AliasEntry{x, nx[h], 0};
            if (large) E[d - 1 - (nl + __popcll(ml & below))] = AliasEntry{x, nx[h], 0};
            ns += __popcll(ms);
            nl += __popcll(ml);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // entries written by one lane, read by another

    // Pairing, exactly the lane kernel's sequence; slot k of the table is filled by the
    // k-th small consumed, so k == si throughout.  Small-queue position p is entry p of
    // the segment for p < ns and the residual of large p - ns after that.
    //   * Larges are held 64 at a time (window lw: my_l); a lane keeps its large's
    //     residual once it drops (my_res), and the 64 residuals are stored when the
    //     window moves on.  (Loading the next window ahead would rotate registers
    //     through every drop, and the rotation waits for the load.)
    //   * Small-queue entries are consumed from a window of up to 64 final entries
    //     (initial smalls, and residuals already stored): my_sq / my_snb, each lane also
    //     keeping the neighbour of the large its entry pairs with (my_lnb), so the
    //     window's slots are written together.  The next window loads while this one is
    //     consumed (my_n), as far as its entries are final.  A residual of the current
    //     large window is read from my_res.
    //   * The inner loop over a window issues memory operations only when the large
    //     window moves on.  Small-window switches wait for loads issued a window earlier
    //     and stores older than that; every small window load follows such a wait, so
    //     it sees every residual stored before it.
    const auto small_pos = [&](int32_t p) { return max(0, min(p < ns ? p : d - 1 - (p - ns), d - 1)); };
    int32_t si = 0, li = 0;
    int32_t lw = 0;
    AliasEntry my_l = E[max(d - 1 - lane, 0)];
    int32_t sw = 0, swn = 0, nw = 0, nwn = min(64, ns);
    AliasEntry my_n = E[small_pos(lane)];
    refill_wait();
    double wl = nl > 0 ? readlane_f64(my_l.q, 0) : 0.0;
    int32_t lnb = nl > 0 ? readlane_i32(my_l.nbr, 0) : 0;
    double my_res = 0.0;
    double my_sq = 0.0;
    int32_t my_snb = 0, my_lnb = 0;
    // The large in use drops below 1: keep its residual, take the next large (slots from
    // small-window position c_next on pair with it).  True when no large is left.
    const auto drop = [&](int32_t c_next) {
        if (lane == li - lw) my_res = wl;
        ++li;
        if (li == nl) return true;
        if (li - lw == 64) {
            E[d - 1 - (lw + lane)].q = my_res;
            lw = li;
            my_l = E[max(d - 1 - (lw + lane), 0)];
            refill_wait();
        }
        wl = readlane_f64(my_l.q, li - lw);
        lnb = readlane_i32(my_l.nbr, li - lw);
        if (lane >= c_next) my_lnb = lnb;
        return false;
    };
    bool done = !(ns > 0 && nl > 0);
    while (!done) {
        const int32_t jr = si - ns;
        if (jr >= lw) {  // the next small is a residual of the current large window
            const double ws = readlane_f64(my_res, jr - lw);
            const int32_t snb = readlane_i32(my_l.nbr, jr - lw);
            if (lane == 0) t[si] = AliasSlot{ws, snb, lnb};
            ++si;
            wl = wl + ws - 1;
            if (wl < 1 && drop(64)) break;
            done = !(si < ns + li);
            continue;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // residuals stored by other lanes
        refill_wait();
        if (lane < swn) t[sw + lane] = AliasSlot{my_sq, my_snb, my_lnb};
        if (nw == si && nwn > 0) {
            my_sq = my_n.q;
            my_snb = my_n.nbr;
            sw = nw;
            swn = nwn;
        } else {
            sw = si;
            swn = min(64, ns + lw - si);
            const AliasEntry e = E[small_pos(sw + lane)];
            refill_wait();
            my_sq = e.q;
            my_snb = e.nbr;
        }
        nw = sw + swn;
        nwn = max(0, min(64, ns + lw - nw));
        if (nwn > 0) my_n = E[small_pos(nw + lane)];
        my_lnb = lnb;
        int32_t c = 0;
        do {
            const double ws = readlane_f64(my_sq, c);
            ++c;
            wl = wl + ws - 1;
            if (wl < 1 && drop(c)) {
                done = true;
                break;
            }
        } while (c < swn);
        si = sw + c;
        if (!done) done = !(si < ns + li);
    }
    if (lane < min(si - sw, swn)) t[sw + lane] = AliasSlot{my_sq, my_snb, my_lnb};
    int32_t k = si;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // Leftovers take probability 1 (alias_sampling.h:43-53), in queue order.
    for (int32_t b = si; b < ns + li; b += 64) {
        const int32_t s2 = b + lane;
        if (s2 < ns + li) {
            const int32_t x = E[s2 < ns ? s2 : d - 1 - (s2 - ns)].nbr;
            t[k + (s2 - si)] = AliasSlot{1.0, x, x};
        }
    }
    k += ns + li - si;
    for (int32_t b = li; b < nl; b += 64) {
        const int32_t l2 = b + lane;
        if (l2 < nl) {
            const int32_t x = E[d - 1 - l2].nbr;
            t[k + (l2 - li)] = AliasSlot{1.0, x, x};
        }
    }
}

// its_initialization for hubs: the prefix chain runs on uniform values, each lane keeps
// the running sum at its own index of the 64-entry group.
__device__ __forceinline__ void its_build_hub(const DevGraph &g, const double *w, int32_t unit, int32_t v,
                                              double *out) {
    const int lane = static_cast<int>(threadIdx.x);
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    double acc = 0.0;
    for (int32_t base = 0; base < d; base += 64 * kWaveGroups) {
        double x[kWaveGroups];
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t j = base + 64 * h + lane;
            x[h] = j < d ? weight_at(w, unit, off + j) : 0.0;
        }
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t g0 = base + 64 * h;
            const int32_t n = min(64, d - g0);
            double mine = 0.0;
            for (int32_t c = 0; c < n; ++c) {
                const double y = readlane_f64(x[h], c);
                acc = (g0 + c == 0) ? y : acc + y;  // the first entry is copied, not added to 0.0
                if (lane == c) mine = acc;
            }
            if (g0 + lane < d) out[off + g0 + lane] = mine;
        }
    }
}

// rs_initialization for hubs: the same comparisons in the same order.
__device__ __forceinline__ void rs_build_hub(const DevGraph &g, const double *w, int32_t unit, int32_t v,
                                             double *out) {
    const int lane = static_cast<int>(threadIdx.x);
    const int64_t off = g.offset[v];
    const int32_t d = static_cast<int32_t>(g.offset[v + 1] - off);
    double m = 0.0;
    for (int32_t base = 0; base < d; base += 64 * kWaveGroups) {
        double x[kWaveGroups];
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t j = base + 64 * h + lane;
            x[h] = j < d ? weight_at(w, unit, off + j) : 0.0;
        }
#pragma unroll
        for (int h = 0; h < kWaveGroups; ++h) {
            const int32_t g0 = base + 64 * h;
            const int32_t n = min(64, d - g0);
            for (int32_t c = 0; c < n; ++c) {
                const double y = readlane_f64(x[h], c);
                if (g0 + c == 0) m = y;
                else if (m < y) m = y;
            }
        }
    }
    if (lane == 0) out[v] = m;
}

// One launch builds every vertex: blocks [0, hubs) take one hub each with the whole
// wavefront (hubs lead the degree-class order, largest first); later blocks take 64
// vertices, one per lane.  The hubs' long sequential chains run beside the lane builds,
// at raised wave priority so the SIMD issues their next instruction first.
__device__ __forceinline__ int32_t lane_vertex(const int32_t *order, int32_t hubs) {
    const int32_t i = hubs + (static_cast<int32_t>(blockIdx.x) - hubs) * 64 + static_cast<int32_t>(threadIdx.x);
    return order ? order[i] : i;
}

__device__ __forceinline__ bool lane_in_range(int32_t hubs, int32_t nv) {
    return hubs + (static_cast<int32_t>(blockIdx.x) - hubs) * 64 + static_cast<int32_t>(threadIdx.x) < nv;
}

__global__ __launch_bounds__(64) void alias_build_kernel(DevGraph g, const double *w, int32_t unit,
                                                         const int32_t *order, int32_t hubs, AliasEntry *scratch,
                                                         AliasSlot *out, unsigned int *zero_count) {
    if (static_cast<int32_t>(blockIdx.x) < hubs) {
        __builtin_amdgcn_s_setprio(3);
        alias_build_hub(g, w, unit, order[blockIdx.x], scratch, out, zero_count);
    } else if (lane_in_range(hubs, g.nv))
        alias_build_vertex(g, w, unit, lane_vertex(order, hubs), scratch, out, zero_count);
}

__global__ __launch_bounds__(64) void its_build_kernel(DevGraph g, const double *w, int32_t unit,
                                                       const int32_t *order, int32_t hubs, double *out) {
    if (static_cast<int32_t>(blockIdx.x) < hubs) {
        __builtin_amdgcn_s_setprio(3);
        its_build_hub(g, w, unit, order[blockIdx.x], out);
    } else if (lane_in_range(hubs, g.nv)) its_build_vertex(g, w, unit, lane_vertex(order, hubs), out);
}

__global__ __launch_bounds__(64) void rs_build_kernel(DevGraph g, const double *w, int32_t unit,
                                                      const int32_t *order, int32_t hubs, double *out) {
    if (static_cast<int32_t>(blockIdx.x) < hubs) {
        __builtin_amdgcn_s_setprio(3);
        rs_build_hub(g, w, unit, order[blockIdx.x], out);
    } else if (lane_in_range(hubs, g.nv)) rs_build_vertex(g, w, unit, lane_vertex(order, hubs), out);
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

// Label records: for each vertex the number of neighbours with label < k, k = 0..7, by binary
// search of its label-sorted segment, then its first edge and degree (num_edges < 2^32), padded
// to 64 bytes: the first MetaPath step from v reads its range and label range from one line.
__global__ void label_start_kernel(const int64_t *offset, const int32_t *lbl_sorted, int32_t nv, uint32_t *lstart) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int64_t off = offset[v];
    const int32_t d = static_cast<int32_t>(offset[v + 1] - off);
    const int32_t *ls = lbl_sorted + off;
    uint32_t k8[8];
    for (int32_t k = 0; k < 8; ++k) {
        int32_t lo = 0, hi = d;
        while (lo < hi) {
            const int32_t m = (lo + hi) >> 1;
            if (ls[m] < k) lo = m + 1;
            else hi = m;
        }
        k8[k] = static_cast<uint32_t>(lo);
    }
    uint4 *rec = reinterpret_cast<uint4 *>(lstart + kLabelRecWords * static_cast<int64_t>(v));
    rec[0] = make_uint4(k8[0], k8[1], k8[2], k8[3]);
    rec[1] = make_uint4(k8[4], k8[5], k8[6], k8[7]);
    rec[2] = make_uint4(static_cast<uint32_t>(off), static_cast<uint32_t>(d), 0u, 0u);
    rec[3] = make_uint4(0u, 0u, 0u, 0u);
}

// Neighbour records in label order, 64 bytes each: entry j of v's label-sorted segment is the
// neighbour record of segment position lbl_pos[off + j], followed by that neighbour's label
// starts, so the move that picks it also has the next step's label range.
__global__ void label_records_kernel(const int64_t *offset, int32_t nv, const int32_t *lbl_pos, const uint4 *erec,
                                     const uint32_t *lstart, int64_t ne, uint4 *lerec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        // The vertex whose segment holds e: the last v with offset[v] <= e.
        int32_t lo = 0, hi = nv;
        while (hi - lo > 1) {
            const int32_t m = (lo + hi) >> 1;
            if (offset[m] <= e) lo = m;
            else hi = m;
        }
        const uint4 r = erec[offset[lo] + lbl_pos[e]];
        const uint4 *ls = reinterpret_cast<const uint4 *>(lstart + kLabelRecWords * static_cast<int64_t>(r.x));
        lerec[4 * e] = r;
        lerec[4 * e + 1] = ls[0];
        lerec[4 * e + 2] = ls[1];
        lerec[4 * e + 3] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// MetaPath schema records: entry e of the label order (lerec's), {neighbour u, u's first edge, the
// label range at u of the label that follows this edge's in the schema, in u's label order}.  A
// schema with distinct labels fixes a walker's position from the label of the edge it moved along,
// so the next step's label is known per edge.  Edges whose label is not in the schema are never
// taken; their records are zero.  A neighbour without edges gets kNoRange twice (a dead end).
__global__ void schema_records_kernel(const uint4 *lerec, const int32_t *lbl_sorted, const uint32_t *lstart,
                                      int64_t ne, uint32_t next_map, uint4 *msrec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t l = lbl_sorted[e];
        const uint32_t nl = (l >= 0 && l < 8) ? (next_map >> (4 * l)) & 15u : 15u;
        uint4 r = make_uint4(0u, 0u, 0u, 0u);
        if (nl < 8u) {
            const uint32_t u = lerec[4 * e].x;
            const uint32_t *L = lstart + kLabelRecWords * static_cast<int64_t>(u);
            const uint32_t first = L[8], deg = L[9];
            r.x = u;
            r.y = first;
            r.z = deg == 0u ? kNoRange : L[nl];
            r.w = deg == 0u ? kNoRange : (nl < 7u ? L[nl + 1] : deg);
        }
        msrec[e] = r;
    }
}

// ITS search structures.  Sizes: one thread per vertex writes its line count.
__global__ void its_tree_sizes_kernel(const int64_t *offset, int32_t nv, uint32_t *count, uint64_t *start) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v == 0) start[0] = 0;
    if (v >= nv) return;
    const int32_t d = static_cast<int32_t>(offset[v + 1] - offset[v]);
    const int32_t nl = its_leaves(d);
    int64_t nodes[kItsMaxLevels];
    const int32_t L = its_levels(nl, nodes);
    int64_t total = nl;
    for (int32_t k = 0; k < L; ++k) total += nodes[k];
    count[v] = static_cast<uint32_t>(total);
}

// start[v] is v's first line (start[nv] = all lines); a vertex with no lines has no root.
__global__ void its_tree_starts_to_roots_kernel(const uint64_t *start, int32_t nv, uint32_t *root) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    root[v] = start[v + 1] > start[v] ? static_cast<uint32_t>(start[v]) : kNoRange;
}

// One thread per 16-byte piece of every line, so a hub's thousands of lines spread
// over the whole grid.  v's lines are its tree levels, root level first, then its
// leaves.  Key j of tree level k (0 = the level over the leaves) is the last prefix
// sum of leaf min((j + 1) * 16^k, nl) - 1 when j is below the level's real key count
// (nl at level 0, the node count of level k - 1 above that), +inf otherwise.  A leaf
// holds 5 prefix sums (+inf past the last edge), 8 bytes of zero, then the 5 edges'
// neighbour records with each neighbour's root.
__global__ void its_tree_fill_kernel(const int64_t *offset, const uint64_t *start, int32_t nv, const double *P,
                                     const uint4 *erec, const uint32_t *root, uint4 *lines, uint64_t pieces) {
    const double inf = __longlong_as_double(0x7ff0000000000000LL);
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < pieces;
         t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t l = t >> 3;
        const int32_t q = static_cast<int32_t>(t & 7);
        // The vertex that owns line l: the last v with start[v] <= l (start[0] = 0 <= l < start[nv]).
        int32_t lo = 0, hi = nv;
        while (hi - lo > 1) {
            const int32_t m = (lo + hi) >> 1;
            if (start[m] <= l) lo = m;
            else hi = m;
        }
        const int64_t off = offset[lo];
        const int32_t d = static_cast<int32_t>(offset[lo + 1] - off);
        const int32_t nl = its_leaves(d);
        int64_t nodes[kItsMaxLevels];
        const int32_t L = its_levels(nl, nodes);
        int64_t tree = 0;
        for (int32_t k = 0; k < L; ++k) tree += nodes[k];
        const int64_t r = static_cast<int64_t>(l - start[lo]);
        uint4 out;
        if (r < tree) {
            int32_t k = L - 1;
            int64_t b = 0;
            while (r >= b + nodes[k]) b += nodes[k--];
            const int64_t real = k == 0 ? nl : nodes[k - 1];
            const int64_t j0 = (r - b) * kItsFan + 2 * q;
            double key[2];
            for (int32_t h = 0; h < 2; ++h) {
                const int64_t j = j0 + h;
                const int64_t leaf = (j + 1) << (4 * k);
                const int64_t last = (leaf < nl ? leaf : nl) * kItsLeaf;
                key[h] = j < real ? P[off + (last < d ? last : d) - 1] : inf;
            }
            out = make_uint4(static_cast<uint32_t>(__double_as_longlong(key[0])),
                             static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(key[0])) >> 32),
                             static_cast<uint32_t>(__double_as_longlong(key[1])),
                             static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(key[1])) >> 32));
        } else if (q < 3) {
            const int64_t j = (r - tree) * kItsLeaf + 2 * q;
            const double a = j < d ? P[off + j] : inf;
            const double c = q == 2 ? 0.0 : j + 1 < d ? P[off + j + 1] : inf;
            out = make_uint4(static_cast<uint32_t>(__double_as_longlong(a)),
                             static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(a)) >> 32),
                             static_cast<uint32_t>(__double_as_longlong(c)),
                             static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(c)) >> 32));
        } else {
            const int64_t j = (r - tree) * kItsLeaf + (q - 3);
            if (j < d) {
                const uint4 e = erec[off + j];
                out = make_uint4(e.x, e.y, e.z, root[e.x]);
            } else {
                out = make_uint4(0u, 0u, 0u, kNoRange);
            }
        }
        lines[t] = out;
    }
}

__global__ void its_tree_roots_kernel(const int32_t *adj, int64_t ne, const uint32_t *root, uint4 *erec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        erec[e].w = root[adj[e]];
}

// Membership tables.  Sizes: one thread per vertex writes its bucket count (none for degree <=
// kMemberMin, i.e. a vertex without neighbours).
__global__ void member_table_sizes_kernel(const int64_t *offset, int32_t nv, uint32_t *count, uint64_t *start) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v == 0) start[0] = 0;
    if (v >= nv) return;
    const int32_t d = static_cast<int32_t>(offset[v + 1] - offset[v]);
    count[v] = d > kMemberMin ? 1u << member_bucket_bits(d) : 0u;
}

// One thread per edge e of a vertex with a table: its neighbour takes the next free slot of its
// bucket, or of the next bucket with room (fill_count[b] counts the claims on bucket b, so slots
// [0, min(count, 32)) end up holding keys; the slots were set to kMemberEmpty beforehand).  The
// order of the keys inside a bucket depends on the order of the claims; which keys a bucket holds
// does not matter to a test, which looks at all of them.
__global__ void member_table_fill_kernel(const int64_t *offset, const uint64_t *start, int32_t nv, const int32_t *adj,
                                        int64_t ne, uint32_t *slots, uint32_t *fill_count) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int32_t lo = 0, hi = nv;  // the vertex v with offset[v] <= e < offset[v + 1]
        while (hi - lo > 1) {
            const int32_t m = (lo + hi) >> 1;
            if (offset[m] <= e) lo = m;
            else hi = m;
        }
        const int32_t d = static_cast<int32_t>(offset[lo + 1] - offset[lo]);
        if (d <= kMemberMin) continue;
        const uint32_t bits = member_bucket_bits(d), mask = (1u << bits) - 1u;
        const int32_t u = adj[e];
        const uint64_t base = start[lo];
        // At most d <= 16 nb keys in 32 nb slots: a free slot exists, the loop ends.
        for (uint32_t b = member_hash(u, bits);; b = (b + 1u) & mask) {
            const uint32_t slot = atomicAdd(fill_count + base + b, 1u);
            if (slot < static_cast<uint32_t>(kMemberSlots)) {
                slots[(base + b) * kMemberSlots + slot] = static_cast<uint32_t>(u);
                break;
            }
        }
    }
}

// Neighbour records for Node2Vec: the edge records with each neighbour's membership root.
__global__ void member_records_kernel(const uint4 *erec, int64_t ne, const uint32_t *root, uint4 *nrec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint4 r = erec[e];
        nrec[e] = make_uint4(r.x, r.y, r.z, root[r.x]);
    }
}

// Weighted Node2Vec records: a candidate's edge weight beside its neighbour record with root, so
// a trial of the weighted Move is one 32-byte read.
__global__ void member_weight_records_kernel(const uint4 *nrec, const double *w, int64_t ne, uint4 *wnrec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint4 r = nrec[e];
        const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(w[e]));
        wnrec[2 * e] = make_uint4(static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32), r.x, r.y);
        wnrec[2 * e + 1] = make_uint4(r.z, r.w, 0u, 0u);
    }
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

// Rejection walk records: an edge's weight beside its neighbour's range and maximum weight, so
// a rejection trial is one read and the accepted one also carries the next step's range.
__global__ void rs_records_kernel(const int32_t *adj, const uint2 *vrec, const double *rs_max, const double *w,
                                  int32_t unit, int64_t ne, uint4 *rrec) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t nb = adj[e];
        const unsigned long long wb = static_cast<unsigned long long>(__double_as_longlong(unit ? 1.0 : w[e]));
        const unsigned long long mb = static_cast<unsigned long long>(__double_as_longlong(rs_max[nb]));
        const uint2 r = vrec[nb];
        rrec[2 * e] = make_uint4(static_cast<uint32_t>(wb), static_cast<uint32_t>(wb >> 32), static_cast<uint32_t>(nb),
                                 r.x);
        rrec[2 * e + 1] = make_uint4(r.y, static_cast<uint32_t>(mb), static_cast<uint32_t>(mb >> 32), 0u);
    }
}

inline unsigned grid_for(int64_t n) { return static_cast<unsigned>((n + kBuildBlock - 1) / kBuildBlock); }

}  // namespace

inline unsigned build_grid(int32_t nv, int32_t hubs) {
    return static_cast<unsigned>(hubs + (static_cast<int64_t>(nv) - hubs + 63) / 64);
}

hipError_t launch_alias_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, int32_t hubs,
                              AliasEntry *scratch, AliasSlot *out, unsigned int *zero_count, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(alias_build_kernel, dim3(build_grid(g.nv, hubs)), dim3(64), 0, stream, g, w, unit, order, hubs,
                       scratch, out, zero_count);
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

hipError_t launch_rs_records(const DevGraph &g, const double *w, int32_t unit, uint4 *rrec, hipStream_t stream) {
    if (g.ne <= 0) return hipSuccess;
    hipLaunchKernelGGL(rs_records_kernel, dim3(8192), dim3(256), 0, stream, g.adj, g.vrec, g.rs_max, w, unit, g.ne,
                       rrec);
    return hipGetLastError();
}

hipError_t launch_schema_records(const DevGraph &g, uint32_t next_map, uint4 *msrec, hipStream_t stream) {
    if (g.ne > 0)
        hipLaunchKernelGGL(schema_records_kernel, dim3(8192), dim3(256), 0, stream, g.lerec, g.lbl_sorted, g.lstart,
                           g.ne, next_map, msrec);
    return hipGetLastError();
}

hipError_t launch_label_records(const DevGraph &g, uint32_t *lstart, uint4 *lerec, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(label_start_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.lbl_sorted,
                       g.nv, lstart);
    if (g.ne > 0)
        hipLaunchKernelGGL(label_records_kernel, dim3(8192), dim3(256), 0, stream, g.offset, g.nv, g.lbl_pos, g.erec,
                           lstart, g.ne, lerec);
    return hipGetLastError();
}

hipError_t launch_its_tree_starts(const DevGraph &g, uint32_t *root, uint64_t *start, void *temp, size_t &temp_bytes,
                                  hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    const size_t n = static_cast<size_t>(g.nv);
    if (!temp)
        return rocprim::inclusive_scan<rocprim::default_config, const uint32_t *, uint64_t *, rocprim::plus<uint64_t>,
                                       uint64_t>(nullptr, temp_bytes, root, start + 1, n, rocprim::plus<uint64_t>(),
                                                 stream);
    hipLaunchKernelGGL(its_tree_sizes_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.nv, root,
                       start);
    hipError_t e = rocprim::inclusive_scan<rocprim::default_config, const uint32_t *, uint64_t *,
                                           rocprim::plus<uint64_t>, uint64_t>(temp, temp_bytes, root, start + 1, n,
                                                                              rocprim::plus<uint64_t>(), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(its_tree_starts_to_roots_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, start, g.nv,
                       root);
    return hipGetLastError();
}

hipError_t launch_its_tree_fill(const DevGraph &g, const double *prefix, const uint64_t *start, uint64_t lines_total,
                                const uint32_t *root, uint4 *lines, hipStream_t stream) {
    if (g.nv <= 0 || lines_total == 0) return hipSuccess;
    const uint64_t pieces = lines_total * 8;
    const uint64_t blocks = std::min<uint64_t>((pieces + 255) / 256, uint64_t(1) << 20);
    hipLaunchKernelGGL(its_tree_fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, g.offset, start,
                       g.nv, prefix, g.erec, root, lines, pieces);
    return hipGetLastError();
}

hipError_t launch_its_tree_roots_to_records(const DevGraph &g, const uint32_t *root, uint4 *erec, hipStream_t stream) {
    if (g.ne <= 0 || !erec) return hipSuccess;
    hipLaunchKernelGGL(its_tree_roots_kernel, dim3(8192), dim3(256), 0, stream, g.adj, g.ne, root, erec);
    return hipGetLastError();
}

hipError_t launch_member_table_starts(const DevGraph &g, uint32_t *root, uint64_t *start, void *temp,
                                     size_t &temp_bytes, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    const size_t n = static_cast<size_t>(g.nv);
    if (!temp)
        return rocprim::inclusive_scan<rocprim::default_config, const uint32_t *, uint64_t *, rocprim::plus<uint64_t>,
                                       uint64_t>(nullptr, temp_bytes, root, start + 1, n, rocprim::plus<uint64_t>(),
                                                 stream);
    hipLaunchKernelGGL(member_table_sizes_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.nv,
                       root, start);
    hipError_t e = rocprim::inclusive_scan<rocprim::default_config, const uint32_t *, uint64_t *,
                                           rocprim::plus<uint64_t>, uint64_t>(temp, temp_bytes, root, start + 1, n,
                                                                              rocprim::plus<uint64_t>(), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(its_tree_starts_to_roots_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, start, g.nv,
                       root);
    return hipGetLastError();
}

hipError_t launch_member_table_fill(const DevGraph &g, const uint64_t *start, uint64_t lines_total, uint4 *lines,
                                   uint32_t *fill_count, hipStream_t stream) {
    if (g.nv <= 0 || lines_total == 0) return hipSuccess;
    uint32_t *slots = reinterpret_cast<uint32_t *>(lines);
    hipError_t e = launch_fill_u32(slots, kMemberEmpty, static_cast<int64_t>(lines_total) * kMemberSlots, stream);
    if (e != hipSuccess) return e;
    e = launch_fill_u32(fill_count, 0u, static_cast<int64_t>(lines_total), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(member_table_fill_kernel, dim3(8192), dim3(256), 0, stream, g.offset, start, g.nv, g.adj, g.ne,
                       slots, fill_count);
    return hipGetLastError();
}

hipError_t launch_member_records(const DevGraph &g, const uint32_t *root, uint4 *nrec, hipStream_t stream) {
    if (g.ne <= 0 || !g.erec) return hipSuccess;
    hipLaunchKernelGGL(member_records_kernel, dim3(8192), dim3(256), 0, stream, g.erec, g.ne, root, nrec);
    return hipGetLastError();
}

hipError_t launch_member_weight_records(const DevGraph &g, uint4 *wnrec, hipStream_t stream) {
    if (g.ne <= 0 || !g.nrec || !g.weight) return hipSuccess;
    hipLaunchKernelGGL(member_weight_records_kernel, dim3(8192), dim3(256), 0, stream, g.nrec, g.weight, g.ne, wnrec);
    return hipGetLastError();
}

hipError_t launch_its_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, int32_t hubs,
                            double *out, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(its_build_kernel, dim3(build_grid(g.nv, hubs)), dim3(64), 0, stream, g, w, unit, order, hubs,
                       out);
    return hipGetLastError();
}

hipError_t launch_rs_build(const DevGraph &g, const double *w, int32_t unit, const int32_t *order, int32_t hubs,
                           double *out, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(rs_build_kernel, dim3(build_grid(g.nv, hubs)), dim3(64), 0, stream, g, w, unit, order, hubs,
                       out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- ITS guide records
namespace {

// The walk's ITS pick for draw u (inverse_transformation_sampling.h:259-286): the number
// of prefix sums <= real2(u) * total, clamped to the last edge.  Same expression as the
// Move's, so the records below reproduce its picks exactly.
__device__ __forceinline__ int32_t guide_pick(const double *P, int32_t d, double total, uint32_t u) {
    const int32_t c = trw_dev::upper_bound_branchless(P, d, trw_dev::real2(u) * total);
    return c < d - 1 ? c : d - 1;
}

// First draw of bucket k of m: the least u with floor(u * m / 2^32) >= k (2^32 for k = m).
__device__ __forceinline__ uint64_t guide_first_u(uint64_t k, uint64_t m) { return ((k << 32) + m - 1) / m; }

// Pass c (1..kGuideMaxC), one thread per edge: at c buckets per edge, edge e's share is the
// buckets [c * k, c * k + c) of its vertex (k = e's position); flag bit c - 1 of the vertex is
// set when one of them reaches more than two edges.  Pass c > 1 runs only for vertices whose
// pass c - 1 set its bit.
__global__ void its_guide_flag_kernel(const int64_t *offset, int32_t nv, int64_t ne, const double *P, int32_t c,
                                      uint32_t *flag) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int32_t lo = 0, hi = nv;  // the vertex v with offset[v] <= e < offset[v + 1]
        while (hi - lo > 1) {
            const int32_t m = (lo + hi) >> 1;
            if (offset[m] <= e) lo = m;
            else hi = m;
        }
        if (c > 1 && !(flag[lo] & (1u << (c - 2)))) continue;
        const int64_t off = offset[lo];
        const int32_t d = static_cast<int32_t>(offset[lo + 1] - off);
        const double total = P[off + d - 1];
        if (!(total > 0.0)) continue;
        const uint64_t m = static_cast<uint64_t>(c) * static_cast<uint64_t>(d);
        for (uint64_t k = static_cast<uint64_t>(c) * static_cast<uint64_t>(e - off), kend = k + c; k < kend; ++k) {
            const int32_t a = guide_pick(P + off, d, total, static_cast<uint32_t>(guide_first_u(k, m)));
            const int32_t b = guide_pick(P + off, d, total, static_cast<uint32_t>(guide_first_u(k + 1, m) - 1));
            if (b - a > 1) {
                atomicOr(flag + lo, 1u << (c - 1));
                break;
            }
        }
    }
}

// Bucket counts: c * d for the least c whose buckets all reach at most two edges (kGuideMaxC
// when none does), none for a vertex of degree 0 or whose weights sum to zero.
__global__ void its_guide_sizes_kernel(const int64_t *offset, int32_t nv, const double *P, const uint32_t *flag,
                                       uint32_t *count, uint64_t *start) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v == 0) start[0] = 0;
    if (v >= nv) return;
    const int64_t off = offset[v];
    const int32_t d = static_cast<int32_t>(offset[v + 1] - off);
    uint32_t c = 1;
    while (c < kGuideMaxC && (flag[v] & (1u << (c - 1)))) ++c;
    count[v] = (d > 0 && P[off + d - 1] > 0.0) ? static_cast<uint32_t>(d) * c : 0u;
}

// {first bucket, bucket count} per vertex; a vertex of positive degree whose weights sum to
// zero is {kNoRange, 0} (the walk ends there without a draw, as the Move does).
// Written without a short-circuit: hipcc (ROCm 7.2, -O3) compiled the `m == 0 && degree > 0`
// form of this selection so that a degree-0 vertex also got kNoRange (the -1 was moved into
// the m == 0 branch over the register holding start[v]).  Both loads here are unconditional
// and the selection is one v_cndmask.
__global__ void its_guide_ranges_kernel(const int64_t *offset, int32_t nv, const uint64_t *start, uint2 *gvrec) {
    const int32_t v = static_cast<int32_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const uint64_t s0 = start[v], s1 = start[v + 1];
    const int64_t d = offset[v + 1] - offset[v];
    const uint32_t m = static_cast<uint32_t>(s1 - s0);
    const uint32_t zero_total = static_cast<uint32_t>(m == 0u) & static_cast<uint32_t>(d > 0);
    gvrec[v] = make_uint2(zero_total ? kNoRange : static_cast<uint32_t>(s0), m);
}

// One thread per bucket.  Bucket k of vertex v (m buckets) holds the draws u in
// [guide_first_u(k), guide_first_u(k + 1)); the picks over it run from a = pick(first) to
// b = pick(last).  When b <= a + 1 the record decides the step: the least u of the bucket
// whose pick is b (found by bisection on P[a] <= real2(u) * total, which is pick(u) > a),
// both candidates and their bucket ranges.  Otherwise the record says so (.w = 1) and the
// walk searches the prefix sums.
__global__ void its_guide_fill_kernel(const int64_t *offset, int32_t nv, const int32_t *adj, const double *P,
                                      const uint64_t *start, const uint2 *gvrec, uint4 *grec, uint64_t buckets) {
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < buckets;
         t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        int32_t lo = 0, hi = nv;  // the last v with start[v] <= t: the vertex whose buckets hold t
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (start[mid] <= t) lo = mid;
            else hi = mid;
        }
        const int32_t v = lo;
        const int64_t off = offset[v];
        const int32_t d = static_cast<int32_t>(offset[v + 1] - off);
        const uint64_t m = start[v + 1] - start[v], k = t - start[v];
        const double *Pv = P + off;
        const double total = Pv[d - 1];
        const uint32_t u_lo = static_cast<uint32_t>(guide_first_u(k, m));
        const uint32_t u_hi = static_cast<uint32_t>(guide_first_u(k + 1, m) - 1);
        const int32_t a = guide_pick(Pv, d, total, u_lo);
        const int32_t b = guide_pick(Pv, d, total, u_hi);
        if (b - a > 1) {
            grec[2 * t] = make_uint4(0u, 0u, 0u, 1u);
            grec[2 * t + 1] = make_uint4(0u, 0u, 0u, 0u);
            continue;
        }
        uint32_t thr = 0u;
        if (b > a) {
            const double pa = Pv[a];
            uint32_t l = u_lo, h = u_hi;  // pick(l) = a, pick(h) = b
            while (h - l > 1) {
                const uint32_t mid = l + ((h - l) >> 1);
                if (pa <= trw_dev::real2(mid) * total) h = mid;
                else l = mid;
            }
            thr = h;
        }
        const int32_t na = adj[off + a], nb = adj[off + b];
        const uint2 ra = gvrec[na], rb = gvrec[nb];
        grec[2 * t] = make_uint4(thr, static_cast<uint32_t>(na), static_cast<uint32_t>(nb), 0u);
        grec[2 * t + 1] = make_uint4(ra.x, ra.y, rb.x, rb.y);
    }
}

}  // namespace

hipError_t launch_its_guide_starts(const DevGraph &g, const double *prefix, uint32_t *flag, uint32_t *count,
                                   uint64_t *start, void *temp, size_t &temp_bytes, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    const size_t n = static_cast<size_t>(g.nv);
    if (!temp)
        return rocprim::inclusive_scan<rocprim::default_config, const uint32_t *, uint64_t *, rocprim::plus<uint64_t>,
                                       uint64_t>(nullptr, temp_bytes, count, start + 1, n, rocprim::plus<uint64_t>(),
                                                 stream);
    launch_fill_u32(flag, 0u, g.nv, stream);
    for (int32_t c = 1; c < static_cast<int32_t>(kGuideMaxC) && g.ne > 0; ++c)
        hipLaunchKernelGGL(its_guide_flag_kernel, dim3(8192), dim3(256), 0, stream, g.offset, g.nv, g.ne, prefix, c,
                           flag);
    hipLaunchKernelGGL(its_guide_sizes_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.nv,
                       prefix, flag, count, start);
    return rocprim::inclusive_scan<rocprim::default_config, const uint32_t *, uint64_t *, rocprim::plus<uint64_t>,
                                   uint64_t>(temp, temp_bytes, count, start + 1, n, rocprim::plus<uint64_t>(), stream);
}

hipError_t launch_its_guide_fill(const DevGraph &g, const double *prefix, const uint64_t *start, uint64_t buckets,
                                 uint2 *gvrec, uint4 *grec, hipStream_t stream) {
    if (g.nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(its_guide_ranges_kernel, dim3(grid_for(g.nv)), dim3(kBuildBlock), 0, stream, g.offset, g.nv,
                       start, gvrec);
    if (buckets == 0) return hipGetLastError();
    const uint64_t blocks = std::min<uint64_t>((buckets + 255) / 256, uint64_t(1) << 20);
    hipLaunchKernelGGL(its_guide_fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, g.offset,
                       g.nv, g.adj, prefix, start, gvrec, grec, buckets);
    return hipGetLastError();
}

}  // namespace trw

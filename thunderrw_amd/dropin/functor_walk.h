// Functor engine: compute<F> for app functors that have no built-in Move, run on the GPU.
//
// The reference's engine calls the app functor's weight / update / max_weight for every
// walker step (random_walk/creeper.h:24-334, 543-666).  The library's own kernels compile the
// five reference functors' arithmetic in; any other functor -- or a reference functor in an
// (execution mode, sampler) pair the library has no Move for -- runs here, in kernels
// instantiated for F in the application's translation unit (hipcc), which call F's methods
// on the device.  Included by creeper_main.h under hipcc; WalkerMeta is the app's type.
//
//   Uniform (em 0)            uniform_move (uniform_sampling.h:47-63), then F::update
//   Static (em 1)             static_gather (creeper.h:24-99) evaluates F::weight per edge
//                             (weights_kernel), the library builds the tables from those
//                             weights (trw_build_tables_from), then static_{its,alias,rs}_move
//   Dynamic ITS/alias/RS      dynamic_gather (creeper.h:178-296): F::weight for all d
//   (em 2, sm 1/2/3)          neighbours, serial prefix sums / alias_initialization / maximum,
//                             then dynamic_{its,alias,rs}_move -- a wavefront per walker
//                             (gather_kernel): lanes evaluate 64 weights at a time, the serial
//                             FP64 chains read them across lanes in index order
//   Dynamic max-weight RS     dynamic_max_weight_rs_move (rejection_sampling.h:368-402):
//   (em 2, sm 4)              F::max_weight per step, F::weight per trial
//
// Random draws are the library's (thunderrw_device.h): replay words [k][walker] consumed in
// the reference's per-walker order, or Philox keyed by (seed, walker id, step, trial), so a
// functor equal to a built-in app walks exactly as the library's kernel does.
//
// Every lane holds its own bitwise copy of the functor (the reference copies it per thread,
// creeper.h:552).  weight() is evaluated by several lanes at once for one walker, so it must
// not depend on earlier weight() calls; update() is called once per move.
#ifndef THUNDERRW_DROPIN_FUNCTOR_WALK_H
#define THUNDERRW_DROPIN_FUNCTOR_WALK_H

#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <unistd.h>

#include <cstdint>
#include <cstring>

#include "thunderrw.h"
#include "thunderrw_device.h"

namespace trw_functor {

enum : int32_t {
    kAlive = -1,
    kDone = TRW_WALKER_DONE,
    kGatherTerminated = TRW_WALKER_GATHER_TERMINATED,
    kDeadEnd = TRW_WALKER_DEAD_END,
    kReplayExhausted = TRW_WALKER_REPLAY_EXHAUSTED,
    kTrialCap = TRW_WALKER_TRIAL_CAP,
    kBadSource = TRW_WALKER_BAD_SOURCE,
    kStepCap = TRW_WALKER_STEP_CAP,
};

// Scratch entry of the per-step alias build: an index's normalised weight (or residual)
// and its neighbour.
struct Entry {
    double q;
    int32_t nbr;
    int32_t pad;
};

struct Args {
    // graph (device)
    const int64_t *offset;
    const int32_t *adj;
    const trw_alias_slot *alias;
    const double *prefix;
    const double *rs_max;
    const double *rs_weight;  // null: unit weights
    int32_t nv;
    // g_para and limits
    int32_t execution, sampling, length;
    int32_t max_trials;  // rejection trials per move before kTrialCap
    int32_t max_moves;   // moves per walker before kStepCap (update() never ended the walk)
    // walkers
    const int32_t *sources;
    int32_t n;
    int32_t path_stride;
    int32_t *paths;  // n x path_stride: WalkerMeta::seq_ rows
    int32_t *lengths, *ends, *status, *draws_used;
    unsigned long long *counters;  // steps, gather-terminated, dead ends, errors
    const uint32_t *draws;         // replay: [k][n]
    int64_t K;
    uint64_t seed;  // Philox key
    uint64_t id_base;
    // gather_kernel: per wavefront, scratch_stride (>= max degree) entries of each
    double *scratch_w;
    Entry *scratch_s, *scratch_g;
    int64_t scratch_stride;
};

// A walker's draws: replay words in order, or Philox words of (step, trial) blocks.
template <bool REPLAY>
struct Draws {
    const uint32_t *draws;  // replay [k][n]
    int64_t n, K;
    int32_t w;
    uint64_t seed, gid;
    uint32_t k;

    __device__ Draws(const Args &a, int32_t walker)
        : draws(a.draws), n(a.n), K(a.K), w(walker), seed(a.seed), gid(a.id_base + static_cast<uint64_t>(walker)),
          k(0u) {}
    __device__ bool take(uint32_t &u) {
        if (static_cast<int64_t>(k) >= K) return false;
        u = draws[static_cast<int64_t>(k) * n + w];
        k += 1;
        return true;
    }
    // One move draw (uniform, ITS): word x of the step's trial-0 block.
    __device__ bool one(uint32_t step, uint32_t &u) {
        if (REPLAY) return take(u);
        u = trw_dev::walk_block(seed, gid, step, 0u).x;
        return true;
    }
    // Two move draws of one trial (alias: trial 0; rejection: every trial), in order.
    __device__ bool pair(uint32_t step, uint32_t trial, uint32_t &x, uint32_t &y) {
        if (REPLAY) return take(x) && take(y);
        const uint4 b = trw_dev::walk_block(seed, gid, step, trial);
        x = b.x;
        y = b.y;
        return true;
    }
};

// Stores of one wavefront made visible to its own later loads of other lanes' addresses.
__device__ __forceinline__ void wave_publish() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ void flush_counters(const Args &a, unsigned long long steps, unsigned long long gterm,
                                               unsigned long long dead, unsigned long long errs) {
    if (!a.counters) return;
    steps = wave_sum(steps);
    gterm = wave_sum(gterm);
    dead = wave_sum(dead);
    errs = wave_sum(errs);
    if ((threadIdx.x & 63) == 0) {
        if (steps) atomicAdd(a.counters + 0, steps);
        if (gterm) atomicAdd(a.counters + 1, gterm);
        if (dead) atomicAdd(a.counters + 2, dead);
        if (errs) atomicAdd(a.counters + 3, errs);
    }
}

__device__ __forceinline__ void tally(int32_t st, unsigned long long &gterm, unsigned long long &dead,
                                      unsigned long long &errs) {
    gterm += st == kGatherTerminated;
    dead += st == kDeadEnd;
    errs += st >= kReplayExhausted;
}

template <class WM>
__device__ __forceinline__ void write_outputs(const Args &a, int32_t w, const WM &wm, int32_t st, uint32_t k) {
    if (a.lengths) a.lengths[w] = wm.length_;
    if (a.ends) a.ends[w] = wm.current_;
    if (a.status) a.status[w] = st;
    if (a.draws_used) a.draws_used[w] = static_cast<int32_t>(k);
}

// One walker from its source to the end of its walk, in one lane: Uniform, Static, and
// Dynamic max-weight rejection (no per-step gather).  Returns the walker's status.
template <class F, class WM, bool REPLAY>
__device__ int32_t walk_lane(F &f, const Args &a, int32_t w, WM &wm, Draws<REPLAY> &dr, unsigned long long &steps) {
    const int32_t src = a.sources[w];
    wm.id_ = w;
    wm.source_ = src;
    wm.current_ = src;
    wm.length_ = 1;
    wm.seq_ = a.paths + static_cast<int64_t>(w) * a.path_stride;
    if (src < 0 || src >= a.nv) return kBadSource;
    wm.seq_[0] = src;
    for (int32_t moves = 0;; ++moves) {
        if (moves >= a.max_moves) return kStepCap;
        const int32_t cur = wm.current_;
        const int64_t off = a.offset[cur];
        const int32_t d = static_cast<int32_t>(a.offset[cur + 1] - off);
        if (d <= 0) return kDeadEnd;
        const uint32_t step = static_cast<uint32_t>(wm.length_ - 1);
        const uint32_t ud = static_cast<uint32_t>(d);
        int32_t next;
        if (a.execution == TRW_EXEC_UNIFORM) {
            uint32_t u;  // uniform_sampling.h:47-63
            if (!dr.one(step, u)) return kReplayExhausted;
            next = a.adj[off + (u % ud)];
        } else if (a.execution == TRW_EXEC_STATIC && a.sampling == TRW_SM_ALIAS) {
            uint32_t u0, u1;  // alias_sampling.h:147-172
            if (!dr.pair(step, 0u, u0, u1)) return kReplayExhausted;
            const trw_alias_slot t = a.alias[off + (u0 % ud)];
            if (t.first < 0) {  // the vertex's weights sum to zero (its rows are {0, -1, -1})
                if (REPLAY) dr.k -= 2;
                return kGatherTerminated;
            }
            next = trw_dev::real2(u1) <= t.alias_value ? t.first : t.second;
        } else if (a.execution == TRW_EXEC_STATIC && a.sampling == TRW_SM_ITS) {
            const double *P = a.prefix + off;  // inverse_transformation_sampling.h:259-286
            const double total = P[d - 1];
            if (total == 0.0) return kGatherTerminated;
            uint32_t u;
            if (!dr.one(step, u)) return kReplayExhausted;
            const int32_t pos = min(trw_dev::upper_bound_branchless(P, d, trw_dev::real2(u) * total), d - 1);
            next = a.adj[off + pos];
        } else if (a.execution == TRW_EXEC_STATIC) {
            const double m = a.rs_max[cur];  // rejection_sampling.h:334-366
            if (m == 0.0) return kGatherTerminated;
            uint32_t pos;
            for (uint32_t trial = 0;; ++trial) {
                if (trial >= static_cast<uint32_t>(a.max_trials)) return kTrialCap;
                uint32_t ua, ub;
                if (!dr.pair(step, trial, ua, ub)) return kReplayExhausted;
                pos = ua % ud;
                const double r = trw_dev::real2(ub) * m;
                if (!(r > (a.rs_weight ? a.rs_weight[off + pos] : 1.0))) break;
            }
            next = a.adj[off + pos];
        } else {
            // Dynamic max-weight rejection (rejection_sampling.h:368-402; creeper.h:648-654).
            const double m = f.max_weight(wm);
            int32_t c;
            for (uint32_t trial = 0;; ++trial) {
                if (trial >= static_cast<uint32_t>(a.max_trials)) return kTrialCap;
                uint32_t ua, ub;
                if (!dr.pair(step, trial, ua, ub)) return kReplayExhausted;
                const uint32_t p0 = ua % ud;
                c = a.adj[off + p0];
                const double sel = f.weight(wm, cur, c, off + p0);
                const double r = trw_dev::real2(ub) * m;
                if (!(r > sel)) break;
            }
            next = c;
        }
        // The move's bookkeeping (e.g. alias_sampling.h:163-169), then update (creeper.h:298-334).
        wm.current_ = next;
        if (wm.length_ < a.length) wm.seq_[wm.length_] = next;
        wm.length_ += 1;
        steps += 1;
        if (f.update(wm, cur, wm.current_, 0)) return kDone;
    }
}

template <class F, class WM, bool REPLAY>
__global__ void __launch_bounds__(256) lane_kernel(const F *fsrc, Args a) {
    alignas(F) unsigned char fbuf[sizeof(F)];
    __builtin_memcpy(fbuf, fsrc, sizeof(F));
    F &f = *reinterpret_cast<F *>(fbuf);
    unsigned long long steps = 0, gterm = 0, dead = 0, errs = 0;
    const int32_t stride = static_cast<int32_t>(gridDim.x * blockDim.x);
    for (int32_t w = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); w < a.n; w += stride) {
        WM wm;
        Draws<REPLAY> dr(a, w);
        const int32_t st = walk_lane<F, WM, REPLAY>(f, a, w, wm, dr, steps);
        tally(st, gterm, dead, errs);
        write_outputs(a, w, wm, st, dr.k);
    }
    flush_counters(a, steps, gterm, dead, errs);
}

constexpr int kGatherBatch = 4;

// dynamic_gather's weights for one move, and the serial FP64 chain its sampler folds them with
// (ITS: the in-place prefix sums w[x] += w[x - 1], creeper.h:263-265, stored to W; alias:
// std::accumulate from 0.0, alias_sampling.h:16, weights stored to W; rejection:
// std::max_element, creeper.h:284, weights stored to W).  acc ends as the total / sum / maximum.
//
// The chain runs in every lane on the same values: its operands come from the wavefront's LDS
// row as broadcast reads and its prefix sums go back there, so each element costs one vector
// ALU instruction (the add, or the compare-select of the maximum); reading them across lanes
// with readlane cost about fifteen.  lds_in / lds_out are this wavefront's 64-entry rows.
template <int SM>
__device__ __forceinline__ void chain_step(double &acc, double x, double *__restrict__ lds_out, int32_t j) {
    if (SM == TRW_SM_ITS) {
        acc = acc + x;
        lds_out[j] = acc;
    } else if (SM == TRW_SM_ALIAS) {
        acc = acc + x;
    } else {
        acc = acc < x ? x : acc;
    }
}

template <int SM, class F, class WM>
__device__ __forceinline__ void gather_chain(F &f, const Args &a, WM &wm, int32_t cur, int64_t off, int32_t d,
                                             int lane, double *W, double *__restrict__ lds_in,
                                             double *__restrict__ lds_out, double &acc) {
    for (int32_t base = 0; base < d; base += 64 * kGatherBatch) {
        double wv[kGatherBatch];
#pragma unroll
        for (int q = 0; q < kGatherBatch; ++q) {
            const int32_t i = base + 64 * q + lane;
            wv[q] = 0.0;
            if (i < d) wv[q] = f.weight(wm, cur, a.adj[off + i], off + i);
        }
#pragma unroll
        for (int q = 0; q < kGatherBatch; ++q) {
            const int32_t cb = base + 64 * q;
            if (cb >= d) break;
            const int32_t cnt = min(64, d - cb);
            lds_in[lane] = wv[q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            int32_t j = 0;
            if (cb == 0) {  // the chain starts at the first weight
                acc = lds_in[0];
                if (SM == TRW_SM_ITS) lds_out[0] = acc;
                if (SM == TRW_SM_ALIAS) acc = 0.0 + acc;
                j = 1;
            }
            // Eight operands read ahead of their adds, so the reads' latency overlaps.
            for (; j + 8 <= cnt; j += 8) {
                double x[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) x[t] = lds_in[j + t];
#pragma unroll
                for (int t = 0; t < 8; ++t) chain_step<SM>(acc, x[t], lds_out, j + t);
            }
            for (; j < cnt; ++j) chain_step<SM>(acc, lds_in[j], lds_out, j);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (cb + lane < d) W[cb + lane] = SM == TRW_SM_ITS ? lds_out[lane] : wv[q];
            __builtin_amdgcn_wave_barrier();  // the next chunk's writes of lds_in follow every read
        }
    }
}

// Dynamic ITS / alias / rejection: one walker per wavefront.  Everything but the weight
// evaluation is wave-uniform: every lane runs the same serial chains on the same values, so
// no result needs broadcasting.  W, S and G are this wavefront's scratch rows.
template <class F, class WM, bool REPLAY>
__device__ int32_t walk_wave(F &f, const Args &a, int32_t w, WM &wm, Draws<REPLAY> &dr, unsigned long long &steps,
                             double *W, Entry *S, Entry *G, double *lds_in, double *lds_out) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int32_t src = a.sources[w];
    wm.id_ = w;
    wm.source_ = src;
    wm.current_ = src;
    wm.length_ = 1;
    wm.seq_ = a.paths + static_cast<int64_t>(w) * a.path_stride;
    if (src < 0 || src >= a.nv) return kBadSource;
    if (lane == 0) wm.seq_[0] = src;
    for (int32_t moves = 0;; ++moves) {
        if (moves >= a.max_moves) return kStepCap;
        const int32_t cur = wm.current_;
        const int64_t off = a.offset[cur];
        const int32_t d = static_cast<int32_t>(a.offset[cur + 1] - off);
        if (d <= 0) return kDeadEnd;
        const uint32_t step = static_cast<uint32_t>(wm.length_ - 1);
        const uint32_t ud = static_cast<uint32_t>(d);

        // Gather (creeper.h:178-247): weight i by lane i % 64, kGatherBatch chunks of 64 at a time
        // (their weight reads overlap); the serial chain of each sampler then reads each chunk's
        // weights across lanes in index order.
        double acc = 0.0;  // ITS: prefix sum; alias: accumulate from 0.0; RS: max_element
        if (a.sampling == TRW_SM_ITS)
            gather_chain<TRW_SM_ITS>(f, a, wm, cur, off, d, lane, W, lds_in, lds_out, acc);
        else if (a.sampling == TRW_SM_ALIAS)
            gather_chain<TRW_SM_ALIAS>(f, a, wm, cur, off, d, lane, W, lds_in, lds_out, acc);
        else
            gather_chain<TRW_SM_REJECTION>(f, a, wm, cur, off, d, lane, W, lds_in, lds_out, acc);
        wave_publish();

        int32_t next;
        if (a.sampling == TRW_SM_ITS) {
            // creeper.h:259-272 then dynamic_its_move (inverse_transformation_sampling.h:288-313).
            if (acc == 0.0) return kGatherTerminated;
            uint32_t u;
            if (!dr.one(step, u)) return kReplayExhausted;
            const int32_t pos = min(trw_dev::upper_bound_branchless(W, d, trw_dev::real2(u) * acc), d - 1);
            next = a.adj[off + pos];
        } else if (a.sampling == TRW_SM_ALIAS) {
            // alias_initialization (alias_sampling.h:12-54) then dynamic_alias_move (:174-199).  The
            // move reads one slot, column u0 % d, so the build stops once that slot is emitted.
            const double avg = acc / d;
            if (avg == 0.0) return kGatherTerminated;
            uint32_t u0, u1;
            if (!dr.pair(step, 0u, u0, u1)) return kReplayExhausted;
            const int64_t col = u0 % ud;
            // The FIFO queues: initial smalls (w/avg < 1) and larges, each in index order.
            int64_t ns = 0, ng = 0;
            for (int32_t base = 0; base < d; base += 64) {
                const int32_t i = base + lane;
                const bool in = i < d;
                const double q = in ? W[i] / avg : 0.0;
                const int32_t nb = in ? a.adj[off + i] : 0;
                const bool small = in && q < 1.0;
                const bool large = in && !(q < 1.0);
                const unsigned long long ms = __ballot(small), ml = __ballot(large);
                const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
                if (small) S[ns + __popcll(ms & below)] = Entry{q, nb, 0};
                if (large) G[ng + __popcll(ml & below)] = Entry{q, nb, 0};
                ns += __popcll(ms);
                ng += __popcll(ml);
            }
            wave_publish();
            // Pairing: the front large's residual stays in a register; a large that drops below 1
            // is appended to the small queue with its residual.
            double prob = 1.0;
            int32_t first = -1, second = -1;
            int64_t idx = 0, sh = 0, st = ns, gh = 0;
            double lres = 0.0;
            int32_t lnbr = -1;
            if (ng > 0) {
                const Entry e = G[0];
                lres = e.q;
                lnbr = e.nbr;
            }
            bool found = false;
            while (sh < st && gh < ng) {
                const Entry e = S[sh];
                if (idx == col) {
                    prob = e.q;
                    first = e.nbr;
                    second = lnbr;
                    found = true;
                    break;
                }
                ++idx;
                ++sh;
                lres = (lres + e.q) - 1;
                if (lres < 1) {
                    S[st++] = Entry{lres, lnbr, 0};
                    ++gh;
                    if (gh < ng) {
                        const Entry g2 = G[gh];
                        lres = g2.q;
                        lnbr = g2.nbr;
                    }
                }
            }
            if (!found) {
                // Leftover smalls, then leftover larges: {1.0, nbr, nbr}.
                const int64_t r = col - idx;
                first = second = (r < st - sh) ? S[sh + r].nbr : G[gh + (r - (st - sh))].nbr;
                prob = 1.0;
            }
            next = trw_dev::real2(u1) <= prob ? first : second;
        } else {
            // creeper.h:282-292 then dynamic_rs_move (rejection_sampling.h:404-436).
            if (acc == 0) return kGatherTerminated;
            uint32_t p0;
            for (uint32_t trial = 0;; ++trial) {
                if (trial >= static_cast<uint32_t>(a.max_trials)) return kTrialCap;
                uint32_t ua, ub;
                if (!dr.pair(step, trial, ua, ub)) return kReplayExhausted;
                p0 = ua % ud;
                const double r = trw_dev::real2(ub) * acc;
                if (!(r > W[p0])) break;
            }
            next = a.adj[off + p0];
        }
        wm.current_ = next;
        if (wm.length_ < a.length && lane == 0) wm.seq_[wm.length_] = next;
        wm.length_ += 1;
        steps += 1;
        if (f.update(wm, cur, wm.current_, 0)) return kDone;
    }
}

template <class F, class WM, bool REPLAY>
__global__ void __launch_bounds__(256) gather_kernel(const F *fsrc, Args a) {
    alignas(F) unsigned char fbuf[sizeof(F)];
    __builtin_memcpy(fbuf, fsrc, sizeof(F));
    F &f = *reinterpret_cast<F *>(fbuf);
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    double *W = a.scratch_w + wave * a.scratch_stride;
    Entry *S = a.scratch_s ? a.scratch_s + wave * a.scratch_stride : nullptr;  // alias only
    Entry *G = a.scratch_g ? a.scratch_g + wave * a.scratch_stride : nullptr;
    __shared__ double chain_in[256], chain_out[256];  // per wavefront: 64 chain operands, 64 prefix sums
    double *lds_in = chain_in + (threadIdx.x & ~63u);
    double *lds_out = chain_out + (threadIdx.x & ~63u);
    unsigned long long steps = 0, gterm = 0, dead = 0, errs = 0;
    const bool lead = (threadIdx.x & 63) == 0;
    for (int64_t w = wave; w < a.n; w += nwaves) {
        WM wm;
        Draws<REPLAY> dr(a, static_cast<int32_t>(w));
        unsigned long long s = 0;
        const int32_t st = walk_wave<F, WM, REPLAY>(f, a, static_cast<int32_t>(w), wm, dr, s, W, S, G, lds_in, lds_out);
        if (lead) {
            steps += s;
            tally(st, gterm, dead, errs);
            write_outputs(a, static_cast<int32_t>(w), wm, st, dr.k);
        }
    }
    flush_counters(a, steps, gterm, dead, errs);
}

// static_gather (creeper.h:24-99) for every edge: w[e] = F::weight(temp, v, adj[e], e) with v
// the vertex whose segment holds e.  The reference passes an uninitialised WalkerMeta; here it
// is {0, v, v, 1, seq_ -> {v, v} with a readable entry on each side}.
template <class F, class WM>
__global__ void __launch_bounds__(256) weights_kernel(const F *fsrc, const int64_t *offset, const int32_t *adj,
                                                      int32_t nv, int64_t ne, double *out) {
    alignas(F) unsigned char fbuf[sizeof(F)];
    __builtin_memcpy(fbuf, fsrc, sizeof(F));
    F &f = *reinterpret_cast<F *>(fbuf);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < ne; e += stride) {
        int32_t lo = 0, hi = nv;  // the last v with offset[v] <= e (offset[nv] = ne > e)
        while (hi - lo > 1) {
            const int32_t mid = lo + ((hi - lo) >> 1);
            if (offset[mid] <= e) lo = mid;
            else hi = mid;
        }
        int32_t seq[4] = {lo, lo, lo, lo};
        WM temp;
        temp.id_ = 0;
        temp.source_ = lo;
        temp.current_ = lo;
        temp.length_ = 1;
        temp.seq_ = seq + 2;
        out[e] = f.weight(temp, lo, adj[e], e);
    }
}

// ---- host side ----

// A functor word to rewrite in the device copy: a host address in [host, host + bytes)
// becomes the same offset from dev.
struct Relocation {
    const void *host;
    size_t bytes;
    const void *dev;
};

// True when x is the address of host memory the device cannot read: mapped in this process
// and not an allocation HIP knows (device, pinned or managed memory).
inline bool host_only_address(uint64_t x) {
    if (x < 65536 || x >= (1ull << 47)) return false;
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, reinterpret_cast<void *>(x)) == hipSuccess && at.type != hipMemoryTypeUnregistered)
        return false;
    (void)hipGetLastError();
    const long ps = sysconf(_SC_PAGESIZE);
    void *page = reinterpret_cast<void *>(x & ~static_cast<uint64_t>(ps - 1));
    return msync(page, static_cast<size_t>(ps), MS_ASYNC) == 0;
}

// The functor's bytes for the device: every 8-byte word that addresses one of the relocated
// host ranges is rewritten.  Returns the byte offset of the first word left addressing other
// host memory (a device dereference of it would fault), or -1.
inline int64_t relocate(unsigned char *bytes, size_t size, const Relocation *rel, int nrel) {
    for (size_t o = 0; o + 8 <= size; o += 8) {
        uint64_t x;
        std::memcpy(&x, bytes + o, 8);
        bool done = false;
        for (int r = 0; r < nrel && !done; ++r) {
            const uint64_t h = reinterpret_cast<uint64_t>(rel[r].host);
            const uint64_t span = rel[r].bytes ? rel[r].bytes : 1;
            if (rel[r].host && x >= h && x < h + span) {
                const uint64_t y = reinterpret_cast<uint64_t>(rel[r].dev) + (x - h);
                std::memcpy(bytes + o, &y, 8);
                done = true;
            }
        }
        if (!done && host_only_address(x)) return static_cast<int64_t>(o);
    }
    return -1;
}

template <class F, class WM>
hipError_t launch_weights(const F *dev_f, const trw_device_view &v, double *out, hipStream_t st) {
    if (v.num_edges == 0) return hipSuccess;
    const int64_t blocks = (v.num_edges + 255) / 256;
    const unsigned grid = static_cast<unsigned>(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL((weights_kernel<F, WM>), dim3(grid), dim3(256), 0, st, dev_f, v.offset, v.adj, v.num_vertices,
                       v.num_edges, out);
    return hipGetLastError();
}

// Scratch bytes per entry of a wavefront's row: the weights (prefix sums), and for alias the
// two queues as well.
inline int64_t gather_entry_bytes(int32_t sampling) {
    return sampling == TRW_SM_ALIAS ? 8 + 2 * static_cast<int64_t>(sizeof(Entry)) : 8;
}

// Wavefronts gather_kernel runs at most (one walker each at a time): up to 32 per CU, fewer when
// their scratch rows (max degree entries each) would pass 4 GiB.  A Dynamic move is a serial
// FP64 chain over the vertex's neighbours, so throughput grows with the walkers in flight.
inline int64_t gather_waves(int32_t n, int32_t max_degree, int32_t sampling) {
    const int64_t per_wave = static_cast<int64_t>(max_degree > 0 ? max_degree : 1) * gather_entry_bytes(sampling);
    int64_t waves = 8192;  // 256 CUs x 32 wavefronts
    while (waves > 64 && waves * per_wave > (int64_t(4) << 30)) waves >>= 1;
    if (waves > n) waves = n > 0 ? n : 1;
    return (waves + 3) & ~int64_t(3);  // whole 256-thread blocks
}

template <class F, class WM>
hipError_t launch_walk(const F *dev_f, const Args &a, hipStream_t st) {
    if (a.n == 0) return hipSuccess;
    const bool replay = a.draws != nullptr;
    if (a.execution == TRW_EXEC_DYNAMIC && a.sampling != TRW_SM_MAX_WEIGHT_REJECTION) {
        const int64_t waves = gather_waves(a.n, static_cast<int32_t>(a.scratch_stride), a.sampling);
        const dim3 grid(static_cast<unsigned>(waves / 4)), block(256);
        if (replay) hipLaunchKernelGGL((gather_kernel<F, WM, true>), grid, block, 0, st, dev_f, a);
        else hipLaunchKernelGGL((gather_kernel<F, WM, false>), grid, block, 0, st, dev_f, a);
    } else {
        int64_t blocks = (a.n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        if (replay) hipLaunchKernelGGL((lane_kernel<F, WM, true>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                                       st, dev_f, a);
        else hipLaunchKernelGGL((lane_kernel<F, WM, false>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                                dev_f, a);
    }
    return hipGetLastError();
}

}  // namespace trw_functor

#endif  // THUNDERRW_DROPIN_FUNCTOR_WALK_H

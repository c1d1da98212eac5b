/*
 * CPU oracle for the ThunderRW Gather-Move-Update step engine.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared against.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product never does.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -shared -fPIC (oracle/Makefile).
 * Contraction stays off so every expression rounds exactly as written, which is
 * what the reference's own build does for these expressions (no multiply-add
 * pairs exist in them).
 *
 * Pins (what checks each function; the tests under tests/ run them):
 *   SFMT              known-answer vectors of dependency/SFMT-src-1.5.1/SFMT.19937.out.txt
 *   real1/real2       closed form (exact power-of-two / folded constant)
 *   Philox-4x32-10    published known-answer vectors (Salmon et al., SC'11)
 *   alias build       golden tables written by the reference's alias_initialization
 *                     + reconstruction of w/sum(w) to 1e-12
 *   ITS build         numpy add.accumulate (sequential) + reference golden tables
 *   RS max            numpy max + reference golden tables
 *   walks             golden paths written by the reference's own move/update
 *                     functions driven walker-sequentially on the same stream;
 *                     brute-force closed forms on tiny graphs
 *   PPR power         closed form on a 2-vertex graph and on a cycle;
 *                     agreement with long Monte-Carlo runs of the walk oracle
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* SFMT-19937.  Parameters: SFMT-params19937.h:5-17.  Recursion:       */
/* SFMT-common.h:142-158, gen_rand_all SFMT.c:85-101 (generic path),   */
/* init_gen_rand SFMT.c:349-362, init_by_array SFMT.c:370-432,         */
/* period certification SFMT.c:170-200.                                */
/* ------------------------------------------------------------------ */

#define SFMT_WORDS 624          /* 32-bit words of state */
#define SFMT_LANES 156          /* 128-bit lanes of state */
#define SFMT_POS1 122
#define SFMT_SL1 18
#define SFMT_SL2_BYTES 1
#define SFMT_SR1 11
#define SFMT_SR2_BYTES 1

static const uint32_t k_sfmt_mask[4] = {0xdfffffefU, 0xddfecb7fU, 0xbffaffffU, 0xbffffff6U};
static const uint32_t k_sfmt_parity[4] = {0x00000001U, 0x00000000U, 0x00000000U, 0x13c9e684U};

/* 128-bit value held as four little-endian 32-bit lanes, shifted by whole bytes. */
static void lanes_shl_bytes(uint32_t out[4], const uint32_t in[4], int nbytes) {
    uint64_t hi = ((uint64_t)in[3] << 32) | in[2];
    uint64_t lo = ((uint64_t)in[1] << 32) | in[0];
    int bits = nbytes * 8;
    uint64_t nhi = (hi << bits) | (lo >> (64 - bits));
    uint64_t nlo = lo << bits;
    out[0] = (uint32_t)nlo;
    out[1] = (uint32_t)(nlo >> 32);
    out[2] = (uint32_t)nhi;
    out[3] = (uint32_t)(nhi >> 32);
}

static void lanes_shr_bytes(uint32_t out[4], const uint32_t in[4], int nbytes) {
    uint64_t hi = ((uint64_t)in[3] << 32) | in[2];
    uint64_t lo = ((uint64_t)in[1] << 32) | in[0];
    int bits = nbytes * 8;
    uint64_t nlo = (lo >> bits) | (hi << (64 - bits));
    uint64_t nhi = hi >> bits;
    out[0] = (uint32_t)nlo;
    out[1] = (uint32_t)(nlo >> 32);
    out[2] = (uint32_t)nhi;
    out[3] = (uint32_t)(nhi >> 32);
}

/* r = a ^ (a <<128 SL2) ^ ((b >>32 SR1) & MSK) ^ (c >>128 SR2) ^ (d <<32 SL1) */
static void sfmt_mix(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t *c, const uint32_t *d) {
    uint32_t x[4], y[4], t[4];
    lanes_shl_bytes(x, a, SFMT_SL2_BYTES);
    lanes_shr_bytes(y, c, SFMT_SR2_BYTES);
    for (int i = 0; i < 4; ++i)
        t[i] = a[i] ^ x[i] ^ ((b[i] >> SFMT_SR1) & k_sfmt_mask[i]) ^ y[i] ^ (d[i] << SFMT_SL1);
    memcpy(r, t, sizeof(t));
}

static void sfmt_next_block(orc_sfmt *s) {
    uint32_t *w = s->st;
    const uint32_t *r1 = w + 4 * (SFMT_LANES - 2);
    const uint32_t *r2 = w + 4 * (SFMT_LANES - 1);
    int i;
    for (i = 0; i < SFMT_LANES - SFMT_POS1; ++i) {
        sfmt_mix(w + 4 * i, w + 4 * i, w + 4 * (i + SFMT_POS1), r1, r2);
        r1 = r2;
        r2 = w + 4 * i;
    }
    for (; i < SFMT_LANES; ++i) {
        sfmt_mix(w + 4 * i, w + 4 * i, w + 4 * (i + SFMT_POS1 - SFMT_LANES), r1, r2);
        r1 = r2;
        r2 = w + 4 * i;
    }
}

static void sfmt_certify_period(orc_sfmt *s) {
    uint32_t inner = 0;
    for (int i = 0; i < 4; ++i) inner ^= s->st[i] & k_sfmt_parity[i];
    for (int sh = 16; sh > 0; sh >>= 1) inner ^= inner >> sh;
    if (inner & 1U) return;
    for (int i = 0; i < 4; ++i) {
        uint32_t bit = 1U;
        for (int j = 0; j < 32; ++j, bit <<= 1) {
            if (bit & k_sfmt_parity[i]) {
                s->st[i] ^= bit;
                return;
            }
        }
    }
}

void orc_sfmt_init_gen_rand(orc_sfmt *s, uint32_t seed) {
    s->st[0] = seed;
    for (uint32_t i = 1; i < SFMT_WORDS; ++i)
        s->st[i] = 1812433253U * (s->st[i - 1] ^ (s->st[i - 1] >> 30)) + i;
    s->idx = SFMT_WORDS;
    sfmt_certify_period(s);
}

static uint32_t sfmt_scramble_a(uint32_t x) { return (x ^ (x >> 27)) * 1664525U; }
static uint32_t sfmt_scramble_b(uint32_t x) { return (x ^ (x >> 27)) * 1566083941U; }

void orc_sfmt_init_by_array(orc_sfmt *s, const uint32_t *key, int32_t key_length) {
    const int n = SFMT_WORDS;
    const int lag = 11;                  /* n >= 623 */
    const int mid = (n - lag) / 2;
    uint32_t *w = s->st;
    for (int i = 0; i < n; ++i) w[i] = 0x8b8b8b8bU;
    int count = (key_length + 1 > n) ? key_length + 1 : n;
    uint32_t r = sfmt_scramble_a(w[0] ^ w[mid] ^ w[n - 1]);
    w[mid] += r;
    r += (uint32_t)key_length;
    w[mid + lag] += r;
    w[0] = r;
    count--;
    int i = 1, j = 0;
    for (; j < count && j < key_length; ++j) {
        r = sfmt_scramble_a(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += key[j] + (uint32_t)i;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (; j < count; ++j) {
        r = sfmt_scramble_a(w[i] ^ w[(i + mid) % n] ^ w[(i + n - 1) % n]);
        w[(i + mid) % n] += r;
        r += (uint32_t)i;
        w[(i + mid + lag) % n] += r;
        w[i] = r;
        i = (i + 1) % n;
    }
    for (j = 0; j < n; ++j) {
        r = sfmt_scramble_b(w[i] + w[(i + mid) % n] + w[(i + n - 1) % n]);
        w[(i + mid) % n] ^= r;
        r -= (uint32_t)i;
        w[(i + mid + lag) % n] ^= r;
        w[i] = r;
        i = (i + 1) % n;
    }
    s->idx = n;
    sfmt_certify_period(s);
}

uint32_t orc_sfmt_genrand_uint32(orc_sfmt *s) {
    if (s->idx >= SFMT_WORDS) {
        sfmt_next_block(s);
        s->idx = 0;
    }
    return s->st[s->idx++];
}

void orc_sfmt_fill(orc_sfmt *s, uint32_t *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_sfmt_genrand_uint32(s);
}

void orc_sfmt_stream(uint32_t seed, uint64_t skip, int64_t n, uint32_t *out) {
    orc_sfmt s;
    orc_sfmt_init_gen_rand(&s, seed);
    for (uint64_t i = 0; i < skip; ++i) (void)orc_sfmt_genrand_uint32(&s);
    orc_sfmt_fill(&s, out, n);
}

/* SFMT.h:193-228 */
double orc_to_real1(uint32_t v) { return v * (1.0 / 4294967295.0); }
double orc_to_real2(uint32_t v) { return v * (1.0 / 4294967296.0); }

/* ------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                 */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9U;
            k1 += 0xBB67AE85U;
        }
        uint64_t p0 = (uint64_t)0xD2511F53U * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57U * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------ */
/* Table builds                                                        */
/* ------------------------------------------------------------------ */

/* alias_sampling.h:12-54: Vose's method with FIFO queues, FP64.  Returns 0
 * (and writes nothing) when the average weight is 0. */
int32_t orc_alias_initialization(double *src, const int32_t *nbrs, orc_alias_slot *table, int32_t length) {
    double sum = 0.0;                         /* std::accumulate, left to right */
    for (int32_t j = 0; j < length; ++j) sum = sum + src[j];
    double avg = sum / length;
    if (avg == 0.0) return 0;
    if (length <= 0) return 1;

    /* Two FIFO queues as ring-free arrays: every index is pushed to `small`
     * at most twice in total (once initially or once after leaving `large`),
     * so `length` slots each suffice. */
    int32_t *small = (int32_t *)malloc(sizeof(int32_t) * (size_t)length);
    int32_t *large = (int32_t *)malloc(sizeof(int32_t) * (size_t)length);
    int32_t sh = 0, st = 0, lh = 0, lt = 0;
    for (int32_t j = 0; j < length; ++j) {
        src[j] /= avg;
        if (src[j] < 1) small[st++] = j;
        else large[lt++] = j;
    }
    int32_t k = 0;
    while (sh < st && lh < lt) {
        int32_t s = small[sh], l = large[lh];
        table[k].prob = src[s];
        table[k].first = nbrs[s];
        table[k].second = nbrs[l];
        ++k;
        ++sh;
        src[l] = src[l] + src[s] - 1;
        if (src[l] < 1) {
            ++lh;
            small[st++] = l;
        }
    }
    while (sh < st) {
        int32_t s = small[sh++];
        table[k].prob = 1.0;
        table[k].first = nbrs[s];
        table[k].second = nbrs[s];
        ++k;
    }
    while (lh < lt) {
        int32_t l = large[lh++];
        table[k].prob = 1.0;
        table[k].first = nbrs[l];
        table[k].second = nbrs[l];
        ++k;
    }
    free(small);
    free(large);
    return 1;
}

/* creeper.h:107-130 with the weights already gathered into w.  Vertices
 * whose weights sum to zero get {0.0, -1, -1} entries (the reference leaves
 * them uninitialised).  Returns the number of such vertices. */
int32_t orc_alias_build(const orc_graph *g, const double *w, orc_alias_slot *out) {
    int32_t zero = 0;
    int64_t maxd = 1;
    for (int32_t v = 0; v < g->nv; ++v) {
        int64_t d = g->offset[v + 1] - g->offset[v];
        if (d > maxd) maxd = d;
    }
    double *buf = (double *)malloc(sizeof(double) * (size_t)maxd);
    for (int32_t v = 0; v < g->nv; ++v) {
        int64_t off = g->offset[v];
        int32_t d = (int32_t)(g->offset[v + 1] - off);
        if (d == 0) continue;
        memcpy(buf, w + off, sizeof(double) * (size_t)d);
        if (!orc_alias_initialization(buf, g->adj + off, out + off, d)) {
            ++zero;
            for (int32_t j = 0; j < d; ++j) {
                out[off + j].prob = 0.0;
                out[off + j].first = -1;
                out[off + j].second = -1;
            }
        }
    }
    free(buf);
    return zero;
}

/* inverse_transformation_sampling.h:19-24 per vertex (degree-0 vertices write nothing). */
void orc_its_build(const orc_graph *g, const double *w, double *prefix) {
    for (int32_t v = 0; v < g->nv; ++v) {
        int64_t off = g->offset[v];
        int64_t d = g->offset[v + 1] - off;
        if (d == 0) continue;
        prefix[off] = w[off];
        for (int64_t i = 1; i < d; ++i) prefix[off + i] = prefix[off + i - 1] + w[off + i];
    }
}

/* rejection_sampling.h:12-14 per vertex (degree-0 vertices get 0.0). */
void orc_rs_build(const orc_graph *g, const double *w, double *vmax) {
    for (int32_t v = 0; v < g->nv; ++v) {
        int64_t off = g->offset[v];
        int64_t d = g->offset[v + 1] - off;
        if (d == 0) {
            vmax[v] = 0.0;
            continue;
        }
        double m = w[off];
        for (int64_t i = 1; i < d; ++i)
            if (m < w[off + i]) m = w[off + i];
        vmax[v] = m;
    }
}

/* ------------------------------------------------------------------ */
/* Walks                                                               */
/* ------------------------------------------------------------------ */

/* util/utility.h:18-27 */
int32_t orc_upper_bound(const double *src, int32_t n, double target) {
    const double *base = src;
    while (n > 1) {
        int32_t half = n >> 1;
        base = (base[half] <= target) ? base + half : base;
        n -= half;
    }
    return (int32_t)((*base <= target) + (base - src));
}

/* util/graph/graph.h:180-199: binary search of the larger-degree endpoint in
 * the smaller-degree endpoint's sorted adjacency. */
int32_t orc_is_neighbor(const orc_graph *g, int32_t u, int32_t v) {
    if (g->degree[u] < g->degree[v]) {
        int32_t t = u;
        u = v;
        v = t;
    }
    const int32_t *nbrs = g->adj + g->offset[v];
    int32_t begin = 0, end = g->degree[v] - 1;
    while (begin <= end) {
        int32_t mid = begin + ((end - begin) >> 1);
        if (nbrs[mid] == u) return 1;
        if (nbrs[mid] > u) end = mid - 1;
        else begin = mid + 1;
    }
    return 0;
}

typedef struct {
    const orc_graph *g;
    const orc_params *p;
    double n2v_p, n2v_q, n2v_max;   /* Node2Vec p_, q_, max_ (node2vec.cpp:18-20, 69, 76) */
    double n2v_mw;                  /* weighted Node2Vec: the largest edge weight (0 if none is positive) */
} walk_ctx;

/* F::weight of apps/{deepwalk,ppr,node2vec,metapath,genericwalk}.cpp */
static double app_weight(const walk_ctx *c, int32_t len, int32_t prev, int32_t nbr, int64_t eid) {
    switch (c->p->app) {
    case ORC_APP_PPR:
        return 1.0;                                   /* ppr.cpp:23-25 */
    case ORC_APP_NODE2VEC:                            /* node2vec.cpp:24-38 */
        if (c->p->n2v_weighted) {
            /* The same factor times the edge weight (customwalk.cpp kind 7); w * 1.0 == w. */
            const double ew = c->g->edge_weight[eid];
            if (len == 1) return ew * c->n2v_max;
            if (nbr == prev) return ew * c->n2v_p;
            if (orc_is_neighbor(c->g, prev, nbr)) return ew;
            return ew * c->n2v_q;
        }
        if (len == 1) return c->n2v_max;
        if (nbr == prev) return c->n2v_p;
        if (orc_is_neighbor(c->g, prev, nbr)) return 1.0;
        return c->n2v_q;
    case ORC_APP_METAPATH: {                          /* metapath.cpp:22-30 */
        int32_t pos = (len - 1) % c->p->schema_length;
        return c->g->edge_label[eid] != c->p->schema[pos] ? 0.0 : 1.0;
    }
    default:
        return c->g->edge_weight[eid];                /* deepwalk.cpp:19-21 */
    }
}

/* F::max_weight */
static double app_max_weight(const walk_ctx *c) {
    switch (c->p->app) {
    case ORC_APP_PPR: return 1.0;
    case ORC_APP_NODE2VEC: return c->p->n2v_weighted ? c->n2v_mw * c->n2v_max : c->n2v_max;
    case ORC_APP_METAPATH: return 1.0;
    default: return 5.0;
    }
}

/*
 * One walker at a time, consuming its move draws draws[i*K .. i*K+K) and
 * (PPR) its termination draws term_draws[i*KT ..) in order.  Per iteration
 * this is dynamic_gather (creeper.h:178-296, Dynamic mode only), the move
 * of the selected sampler and update (creeper.h:298-334).
 *   uniform      uniform_sampling.h:47-63
 *   static alias alias_sampling.h:147-172
 *   static ITS   inverse_transformation_sampling.h:259-286
 *   static RS    rejection_sampling.h:334-366
 *   dyn. ITS     inverse_transformation_sampling.h:288-313
 *   dyn. alias   alias_sampling.h:174-199
 *   dyn. RS      rejection_sampling.h:404-436
 *   max-wt RS    rejection_sampling.h:368-402 (Node2Vec)
 * Return: 0, or -1 for an unsupported combination.
 */
int32_t orc_walk(const orc_graph *g, const orc_params *p, const int32_t *sources, int32_t n_walkers,
                 const uint32_t *draws, int64_t K, const uint32_t *term_draws, int64_t KT,
                 int32_t *paths, int32_t *lengths, int32_t *ends, int32_t *status, int32_t *draws_used,
                 uint64_t *steps_out) {
    walk_ctx c;
    c.g = g;
    c.p = p;
    c.n2v_p = 1.0 / p->p;
    c.n2v_q = 1.0 / p->q;
    c.n2v_max = 1.0;
    if (c.n2v_q > c.n2v_max) c.n2v_max = c.n2v_q;
    if (c.n2v_p > c.n2v_max) c.n2v_max = c.n2v_p;
    c.n2v_mw = 0.0;
    if (p->app == ORC_APP_NODE2VEC && p->n2v_weighted) {
        if (!g->edge_weight) return -1;
        for (int64_t e = 0; e < g->ne; ++e)
            if (g->edge_weight[e] > c.n2v_mw) c.n2v_mw = g->edge_weight[e];
    }

    const int32_t exec = p->execution, sm = p->sampling, L = p->length;
    if (exec == ORC_EXEC_STATIC && !(sm == ORC_SM_ALIAS || sm == ORC_SM_ITS || sm == ORC_SM_REJECTION)) return -1;
    if (exec == ORC_EXEC_DYNAMIC && sm == ORC_SM_UNIFORM) return -1;
    if (p->app != ORC_APP_PPR && L < 2) return -1;
    if (p->app == ORC_APP_METAPATH && p->schema_length <= 0) return -1;

    int32_t maxd = 1;
    for (int32_t v = 0; v < g->nv; ++v)
        if (g->degree[v] > maxd) maxd = g->degree[v];
    double *wbuf = (double *)malloc(sizeof(double) * (size_t)maxd);
    orc_alias_slot *abuf = (orc_alias_slot *)malloc(sizeof(orc_alias_slot) * (size_t)maxd);
    uint64_t steps = 0;

    for (int32_t i = 0; i < n_walkers; ++i) {
        const uint32_t *dr = draws ? draws + (int64_t)i * K : NULL;
        const uint32_t *tdr = term_draws ? term_draws + (int64_t)i * KT : NULL;
        int64_t k = 0, kt = 0;
        int32_t cur = sources[i], prev = -1, len = 1, st = ORC_ST_DONE;
        int32_t *seq = paths ? paths + (int64_t)i * (L > 0 ? L : 1) : NULL;
        if (seq && L > 0) seq[0] = cur;
        if (cur < 0 || cur >= g->nv) {
            st = ORC_ST_BAD_SOURCE;
            goto walker_end;
        }

#define NEXT_DRAW(var)                                  \
    do {                                                \
        if (k >= K) { st = ORC_ST_REPLAY_EXHAUSTED; goto walker_end; } \
        (var) = dr[k++];                                \
    } while (0)

        for (;;) {
            int32_t d = g->degree[cur];
            if (d == 0) {
                st = ORC_ST_DEAD_END;
                break;
            }
            int64_t off = g->offset[cur];
            const int32_t *nb = g->adj + off;
            double dyn_max = 0.0;

            /* Static tables of a vertex whose weights sum to zero: the reference
             * reads uninitialised or out-of-segment memory here; the engine defines
             * it as the dynamic gather's termination (creeper.h:267-293). */
            if (exec == ORC_EXEC_STATIC &&
                ((sm == ORC_SM_ALIAS && g->alias[off].first < 0) || (sm == ORC_SM_ITS && g->prefix[off + d - 1] == 0.0) ||
                 (sm == ORC_SM_REJECTION && g->rs_max[cur] == 0.0))) {
                st = ORC_ST_GATHER_TERMINATED;
                break;
            }

            if (exec == ORC_EXEC_DYNAMIC && sm != ORC_SM_MAXREJECTION) {
                for (int32_t j = 0; j < d; ++j) wbuf[j] = app_weight(&c, len, prev, nb[j], off + j);
                if (sm == ORC_SM_ITS) {
                    for (int32_t j = 1; j < d; ++j) wbuf[j] += wbuf[j - 1];
                    if (wbuf[d - 1] == 0.0) { st = ORC_ST_GATHER_TERMINATED; break; }
                } else if (sm == ORC_SM_ALIAS) {
                    if (!orc_alias_initialization(wbuf, nb, abuf, d)) { st = ORC_ST_GATHER_TERMINATED; break; }
                } else {
                    double m = wbuf[0];
                    for (int32_t j = 1; j < d; ++j)
                        if (m < wbuf[j]) m = wbuf[j];
                    if (m == 0) { st = ORC_ST_GATHER_TERMINATED; break; }
                    dyn_max = m;
                }
            }

            int32_t next;
            if (exec == ORC_EXEC_UNIFORM) {
                uint32_t u;
                NEXT_DRAW(u);
                next = nb[u % (uint32_t)d];
            } else if (sm == ORC_SM_ALIAS) {
                uint32_t u0, u1;
                NEXT_DRAW(u0);
                uint32_t col = u0 % (uint32_t)d;
                NEXT_DRAW(u1);
                double x = orc_to_real2(u1);
                const orc_alias_slot *t = (exec == ORC_EXEC_STATIC) ? g->alias + off : abuf;
                next = x <= t[col].prob ? t[col].first : t[col].second;
            } else if (sm == ORC_SM_ITS) {
                uint32_t u;
                NEXT_DRAW(u);
                const double *P = (exec == ORC_EXEC_STATIC) ? g->prefix + off : wbuf;
                double target = orc_to_real2(u) * P[d - 1];
                int32_t pos = orc_upper_bound(P, d, target);
                next = nb[pos < d ? pos : d - 1];   /* pos == d only for negative weights */
            } else if (sm == ORC_SM_REJECTION) {
                const double *W = (exec == ORC_EXEC_STATIC) ? g->rs_weight + off : wbuf;
                double m = (exec == ORC_EXEC_STATIC) ? g->rs_max[cur] : dyn_max;
                uint32_t u, pos;
                double r;
                int32_t trials = 0;
                do {
                    if (p->max_trials > 0 && trials++ >= p->max_trials) { st = ORC_ST_TRIAL_CAP; goto walker_end; }
                    NEXT_DRAW(u);
                    pos = u % (uint32_t)d;
                    NEXT_DRAW(u);
                    r = orc_to_real2(u) * m;
                } while (r > W[pos]);
                next = nb[pos];
            } else { /* max-weight rejection */
                double m = app_max_weight(&c);
                uint32_t u, pos;
                double r, sel;
                int32_t trials = 0;
                do {
                    if (p->max_trials > 0 && trials++ >= p->max_trials) { st = ORC_ST_TRIAL_CAP; goto walker_end; }
                    NEXT_DRAW(u);
                    pos = u % (uint32_t)d;
                    sel = app_weight(&c, len, prev, nb[pos], off + pos);
                    NEXT_DRAW(u);
                    r = orc_to_real2(u) * m;
                } while (r > sel);
                next = nb[pos];
            }

            prev = cur;
            cur = next;
            if (seq && len < L) seq[len] = cur;
            len += 1;
            steps += 1;

            if (p->app == ORC_APP_PPR) {           /* ppr.cpp:27-29 */
                if (kt >= KT) { st = ORC_ST_REPLAY_EXHAUSTED; break; }
                if (orc_to_real1(tdr[kt++]) <= p->stop_prob) break;
            } else if (len == L) {                 /* deepwalk.cpp:23-25 and siblings */
                break;
            }
        }
    walker_end:
#undef NEXT_DRAW
        if (lengths) lengths[i] = len;
        if (ends) ends[i] = cur;
        if (status) status[i] = st;
        if (draws_used) draws_used[i] = (int32_t)k;
    }
    free(wbuf);
    free(abuf);
    if (steps_out) *steps_out = steps;
    return 0;
}

/* ------------------------------------------------------------------ */
/* PPR end-vertex distribution by power iteration (FP64).              */
/* The reference moves before it tests termination (creeper.h:306,     */
/* ppr.cpp:27-29), so the law is sum_{k>=1} a(1-a)^(k-1) x0 P^k.       */
/* weighted = 0: P(u->v) = 1/deg(u);  weighted = 1: w_uv / sum_u w.    */
/* Degree-0 vertices keep their mass (a walker there stops).           */
/* ------------------------------------------------------------------ */
int32_t orc_ppr_power(const orc_graph *g, int32_t weighted, double alpha, const int32_t *sources,
                      int32_t n_sources, double tol, double *out) {
    int32_t n = g->nv;
    double *x = (double *)calloc((size_t)n, sizeof(double));
    double *y = (double *)calloc((size_t)n, sizeof(double));
    if (!x || !y) return -1;
    for (int32_t i = 0; i < n; ++i) out[i] = 0.0;
    for (int32_t i = 0; i < n_sources; ++i) x[sources[i]] += 1.0 / n_sources;
    /* Weighted: each vertex's total weight, summed left to right once (the same value every
     * iteration would recompute). */
    double *strength = NULL;
    if (weighted) {
        strength = (double *)calloc((size_t)n, sizeof(double));
        if (!strength) {
            free(x);
            free(y);
            return -1;
        }
        for (int32_t u = 0; u < n; ++u)
            for (int64_t j = g->offset[u]; j < g->offset[u + 1]; ++j) strength[u] += g->edge_weight[j];
    }
    double coef = alpha, remaining = 1.0;
    int32_t iters = 0;
    while (remaining > tol && iters < 100000) {
        memset(y, 0, sizeof(double) * (size_t)n);
        for (int32_t u = 0; u < n; ++u) {
            double xu = x[u];
            if (xu == 0.0) continue;
            int64_t off = g->offset[u];
            int64_t d = g->offset[u + 1] - off;
            if (d == 0) {
                y[u] += xu;
                continue;
            }
            if (weighted) {
                const double s = strength[u];
                for (int64_t j = 0; j < d; ++j) y[g->adj[off + j]] += xu * (g->edge_weight[off + j] / s);
            } else {
                double share = xu / (double)d;
                for (int64_t j = 0; j < d; ++j) y[g->adj[off + j]] += share;
            }
        }
        for (int32_t v = 0; v < n; ++v) out[v] += coef * y[v];
        double *t = x;
        x = y;
        y = t;
        coef *= (1.0 - alpha);
        remaining *= (1.0 - alpha);
        ++iters;
    }
    free(x);
    free(y);
    free(strength);
    return iters;
}

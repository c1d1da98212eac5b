/*
 * CPU oracle for the ThunderRW Gather-Move-Update step engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * engine is compared against; it is loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by
 * nothing else.  The product (thunderrw_amd/) never links, loads or calls it.
 *
 * Every function restates the reference arithmetic walker-sequentially
 * (the order in which the reference's non-interleaved moves consume the
 * random stream) and cites the reference file:line it follows.
 */
#ifndef TRW_ORACLE_H
#define TRW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AliasSlot layout of config/type_config.h:22-26 (16 bytes). */
typedef struct {
    double prob;
    int32_t first;
    int32_t second;
} orc_alias_slot;

/* SFMT-19937 generator state: 624 32-bit words and a read cursor. */
typedef struct {
    uint32_t st[624];
    int32_t idx;
} orc_sfmt;

/* Read-only CSR view (util/graph/graph.h:51-83). */
typedef struct {
    int32_t nv;
    int64_t ne;
    const int64_t *offset;          /* nv + 1 */
    const int32_t *adj;             /* ne, sorted per vertex */
    const int32_t *degree;          /* nv */
    const double *edge_weight;      /* ne or NULL */
    const int32_t *edge_label;      /* ne or NULL */
    const orc_alias_slot *alias;    /* ne (Static + Alias) */
    const double *prefix;           /* ne (Static + ITS) */
    const double *rs_weight;        /* ne (Static + Rejection) */
    const double *rs_max;           /* nv (Static + Rejection) */
} orc_graph;

/* App ids, enum values of random_walk/types.h and walk parameters. */
enum { ORC_APP_DEEPWALK = 0, ORC_APP_PPR = 1, ORC_APP_NODE2VEC = 2, ORC_APP_METAPATH = 3, ORC_APP_GENERIC = 4 };
enum { ORC_EXEC_UNIFORM = 0, ORC_EXEC_STATIC = 1, ORC_EXEC_DYNAMIC = 2 };
enum { ORC_SM_UNIFORM = 0, ORC_SM_ITS = 1, ORC_SM_ALIAS = 2, ORC_SM_REJECTION = 3, ORC_SM_MAXREJECTION = 4 };
enum { ORC_ST_DONE = 0, ORC_ST_GATHER_TERMINATED = 1, ORC_ST_DEAD_END = 2, ORC_ST_REPLAY_EXHAUSTED = 3,
       ORC_ST_TRIAL_CAP = 4, ORC_ST_BAD_SOURCE = 5 };

typedef struct {
    int32_t app;
    int32_t execution;
    int32_t sampling;
    int32_t length;          /* g_para.length_ */
    double stop_prob;        /* PPR tp */
    double p;                /* Node2Vec -p as given on the command line */
    double q;                /* Node2Vec -q as given on the command line */
    const int32_t *schema;   /* MetaPath -s */
    int32_t schema_length;
    int32_t max_trials;      /* rejection trial cap per step (0 = unlimited) */
    int32_t n2v_weighted;    /* Node2Vec over edge weights: weight w_e * factor, bound W * max_ with W
                                the largest edge weight (no reference app; the functor of kind 7 in
                                tests/dropin_apps/customwalk.cpp) */
} orc_params;

/* ---- SFMT-19937 (dependency/SFMT-src-1.5.1) ---- */
void orc_sfmt_init_gen_rand(orc_sfmt *s, uint32_t seed);
void orc_sfmt_init_by_array(orc_sfmt *s, const uint32_t *key, int32_t key_length);
uint32_t orc_sfmt_genrand_uint32(orc_sfmt *s);
void orc_sfmt_fill(orc_sfmt *s, uint32_t *out, int64_t n);
void orc_sfmt_stream(uint32_t seed, uint64_t skip, int64_t n, uint32_t *out);
double orc_to_real1(uint32_t v);
double orc_to_real2(uint32_t v);

/* ---- Philox-4x32-10 (free-running stream definition) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- table builds ---- */
int32_t orc_alias_initialization(double *src, const int32_t *nbrs, orc_alias_slot *table, int32_t length);
int32_t orc_alias_build(const orc_graph *g, const double *w, orc_alias_slot *out);
void orc_its_build(const orc_graph *g, const double *w, double *prefix);
void orc_rs_build(const orc_graph *g, const double *w, double *vmax);

/* ---- walks ---- */
int32_t orc_upper_bound(const double *src, int32_t n, double target);
int32_t orc_is_neighbor(const orc_graph *g, int32_t u, int32_t v);
int32_t orc_walk(const orc_graph *g, const orc_params *p, const int32_t *sources, int32_t n_walkers,
                 const uint32_t *draws, int64_t draws_per_walker,
                 const uint32_t *term_draws, int64_t term_draws_per_walker,
                 int32_t *paths, int32_t *lengths, int32_t *ends, int32_t *status, int32_t *draws_used,
                 uint64_t *steps);

/* ---- PPR reference distribution ---- */
int32_t orc_ppr_power(const orc_graph *g, int32_t weighted, double alpha, const int32_t *sources,
                      int32_t n_sources, double tol, double *out);

#ifdef __cplusplus
}
#endif

#endif

/*
 * thunderrw.h -- C ABI of the MI355X-native Gather-Move-Update step engine.
 *
 * This is the boundary a caller of the reference's step engine binds to.
 * The reference has no FFI of its own: its apps call the C++ template
 *   template<class F> void compute(Graph&, std::vector<WalkerMeta>&, F)
 * (random_walk/creeper.h:689), which reads the CSR arrays of class Graph
 * (util/graph/graph.h:51-83) and the globals g_para / g_num_threads
 * (random_walk/types.h:72-77,109-110).  Each entry point below names the part
 * of that interface it replaces.  Plain pointers and sizes only; every call
 * returns an int status (TRW_OK = 0, negative on error) and never aborts;
 * trw_last_error() gives a thread-local message for the last failure.
 */
#ifndef THUNDERRW_H
#define THUNDERRW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define TRW_OK 0
#define TRW_E_INVALID (-1)     /* bad argument */
#define TRW_E_HIP (-2)         /* HIP runtime error (message has the HIP error string) */
#define TRW_E_NOMEM (-3)       /* device or host allocation failed */
#define TRW_E_UNSUPPORTED (-4) /* valid in the reference, not provided on the device */
#define TRW_E_STATE (-5)       /* call order: e.g. walking before the tables are built */
#define TRW_E_IO (-6)          /* file missing, short or inconsistent */

/* ExecutionMode, random_walk/types.h:41-45 (values preserved). */
#define TRW_EXEC_UNIFORM 0
#define TRW_EXEC_STATIC 1
#define TRW_EXEC_DYNAMIC 2

/* SamplingMethod, random_walk/types.h:53-59 (values preserved). */
#define TRW_SM_UNIFORM 0
#define TRW_SM_ITS 1
#define TRW_SM_ALIAS 2
#define TRW_SM_REJECTION 3
#define TRW_SM_MAX_WEIGHT_REJECTION 4

/* Built-in apps: the functors of random_walk/apps/{deepwalk,ppr,node2vec,metapath,genericwalk}.cpp. */
#define TRW_APP_DEEPWALK 0
#define TRW_APP_PPR 1
#define TRW_APP_NODE2VEC 2
#define TRW_APP_METAPATH 3
#define TRW_APP_GENERIC 4

/* Random stream.  REPLAY: walker i's move draws are words [i*K, (i+1)*K) of the
 * SFMT-19937 stream init_gen_rand(seed) (PPR termination draws: a second stream),
 * handed to the device by the caller (trw_replay_fill).  PHILOX: counter-based
 * Philox-4x32-10 keyed by (seed, global walker id, step, trial). */
#define TRW_RNG_REPLAY 0
#define TRW_RNG_PHILOX 1

/* Kernel form.  AUTO picks FUSED for fixed-length apps and QUEUE otherwise. */
#define TRW_FORM_AUTO 0
#define TRW_FORM_FUSED 1 /* thread per walker, whole walk in one launch, LDS-staged path tiles */
#define TRW_FORM_QUEUE 2 /* persistent kernel, lanes refill from a device work queue */
#define TRW_FORM_STEP 3  /* plain form: one launch per step over walker state in HBM */
#define TRW_FORM_GRAPH 4 /* launches over a compacted list of live walkers, up to 16 steps each
                            (TRW_TUNE_GRAPH_BURST), replayed from a hipGraph of 6 launches kept
                            with the graph */

/* Per-walker outcome (status[] output). */
#define TRW_WALKER_DONE 0              /* update() ended it (length reached / PPR stop draw) */
#define TRW_WALKER_GATHER_TERMINATED 1 /* all transition weights were zero (creeper.h:267-293) */
#define TRW_WALKER_DEAD_END 2          /* current vertex has degree 0 (undefined in the reference) */
#define TRW_WALKER_REPLAY_EXHAUSTED 3  /* replay draws ran out: an error, never a wrap */
#define TRW_WALKER_TRIAL_CAP 4         /* rejection trial cap hit */
#define TRW_WALKER_BAD_SOURCE 5        /* source vertex outside [0, num_vertices) */
#define TRW_WALKER_STEP_CAP 6          /* drop-in functor engine: the app's update() had not ended the
                                          walk within its move cap (the reference would run on) */

/* Weights F::weight gives static_initialization (creeper.h:24-99): DeepWalk and
 * GenericWalk return edge_weight_[eid] (deepwalk.cpp:19-21), PPR returns 1.0 (ppr.cpp:23-25). */
#define TRW_WEIGHTS_EDGE 0
#define TRW_WEIGHTS_UNIT 1

/* AliasSlot, config/type_config.h:22-26 (16 bytes, identical layout). */
typedef struct {
    double alias_value;
    int32_t first;
    int32_t second;
} trw_alias_slot;

/* g_para (types.h:72-77) plus the app functor's constructor arguments. */
typedef struct {
    int32_t app;            /* TRW_APP_* */
    int32_t execution;      /* TRW_EXEC_*, -em */
    int32_t sampling;       /* TRW_SM_*, -sm */
    int32_t length;         /* g_para.length_ (-l; PPR uses 1, ppr.cpp:51) */
    double stop_probability;/* PPR -sp (ppr.cpp:46-49) */
    double p, q;            /* Node2Vec p_ and q_: return and in-out weights, i.e. 1/p and 1/q of
                               the command line (node2vec.cpp:64-76) */
    int32_t schema[16];     /* MetaPath -s (metapath.cpp:69-83) */
    int32_t schema_length;
    int32_t rng;            /* TRW_RNG_* */
    uint64_t seed;          /* Philox key (PHILOX) */
    int32_t form;           /* TRW_FORM_* */
    int32_t max_trials;     /* rejection trials per step before TRW_WALKER_TRIAL_CAP; 0 = no cap */
    int32_t walkers_per_thread; /* FUSED uniform/alias/static ITS: walkers interleaved per lane (1, 2, 4); 0 = tuned default */
    int32_t weights;        /* TRW_WALK_WEIGHTS_*: the weights Static/Dynamic sampling uses */
} trw_walk_params;

/* trw_walk_params.weights.  APP is the reference: F::weight of the app (edge weights for
 * DeepWalk/GenericWalk, 1.0 for PPR).  EDGE and UNIT override it, e.g. PPR over edge weights. */
#define TRW_WALK_WEIGHTS_APP 0
#define TRW_WALK_WEIGHTS_EDGE 1
#define TRW_WALK_WEIGHTS_UNIT 2

/* Walker batch: all pointers are DEVICE pointers (nullable where noted). */
typedef struct {
    const int32_t *sources;        /* n_walkers start vertices (WalkerMeta.source_, walker.h:9-15) */
    int32_t n_walkers;
    int32_t path_stride;           /* ints per path row, >= max(length, 1) */
    int64_t walker_id_base;        /* global id of sources[0] (keys the Philox stream) */
    const uint32_t *draws;         /* REPLAY move draws, layout [k][n_walkers], k < draws_per_walker */
    int64_t draws_per_walker;
    const uint32_t *term_draws;    /* REPLAY PPR termination draws, [k][n_walkers] */
    int64_t term_draws_per_walker;
    int32_t *paths;                /* nullable: n_walkers x path_stride, seq[0] = source, -1 padded */
    int32_t *lengths;              /* nullable: final WalkerMeta.length_ */
    int32_t *ends;                 /* nullable: final current vertex */
    int32_t *status;               /* nullable: TRW_WALKER_* */
    int32_t *draws_used;           /* nullable: REPLAY move draws consumed */
    uint64_t *counters;            /* nullable: [0] steps [1] gather-terminated [2] dead ends [3] errors; accumulated */
    const uint64_t *walker_ids;    /* nullable: per-walker global ids keying Philox (else walker_id_base + index) */
    int32_t path_flags;            /* TRW_PATHS_*: 0 = rows padded with -1 past the walk's length */
    int32_t path_stride_b;         /* mixed batches: ints per row of the walkers [split, n), whose rows
                                      follow the first split rows; 0 = path_stride (one row matrix) */
} trw_walk_io;

/* trw_walk_io.path_flags.  UNPADDED: row i holds its walk's lengths[i] vertices (at most the
 * kind's length) and nothing past them is written -- the reference's LOG_SEQUENCE output, where
 * a walker's sequence is its length_ entries (creeper.h:310-317), and no -1 padding to write;
 * the caller reads lengths.  The fixed-length FUSED form writes whole rows either way. */
#define TRW_PATHS_PADDED 0
#define TRW_PATHS_UNPADDED 1

typedef struct trw_context trw_context;
typedef struct trw_graph trw_graph;

/* ---- errors and devices ---- */
const char *trw_last_error(void);
int trw_version(int32_t *major, int32_t *minor);
int trw_device_count(int32_t *count);

/* One context per device; replaces the per-process globals (creeper_main.h:12). */
int trw_context_create(int32_t device, trw_context **out);
int trw_context_destroy(trw_context *ctx);
/* The HIP stream the blocking entry points use (hipStream_t as void*). */
int trw_context_stream(trw_context *ctx, void **stream);

/* ---- device memory and timing ----
 * Plumbing with no counterpart in the reference (its arrays live in host memory):
 * buffers in HBM for walker batches, copies (synchronous on the context stream),
 * a zero fill that is a kernel (safe under stream capture), and a device-time
 * measurement with HIP events recorded on a given stream (NULL = context stream). */
int trw_device_alloc(trw_context *ctx, uint64_t bytes, void **out);
int trw_device_free(trw_context *ctx, void *ptr);
int trw_copy_to_device(trw_context *ctx, void *dst, const void *src, uint64_t bytes);
int trw_copy_to_host(trw_context *ctx, void *dst, const void *src, uint64_t bytes);
int trw_device_zero(trw_context *ctx, void *ptr, uint64_t bytes, void *stream);
/* Streaming copy kernel (16-byte vectors; bytes % 16 == 0): the known-size reference
 * kernel the HBM byte counters are calibrated against. */
int trw_device_copy(trw_context *ctx, void *dst, const void *src, uint64_t bytes, void *stream);
int trw_synchronize(trw_context *ctx, void *stream);
int trw_timer_begin(trw_context *ctx, void *stream);
int trw_timer_end(trw_context *ctx, void *stream, double *elapsed_ms);

/* Diagnostic (no reference counterpart): dependent random gathers per second of
 * record_bytes (16: one 16-byte load; 32: two adjacent 16-byte loads, the shape of an
 * alias walk record) over a table of table_bytes, `chains` (1, 2, 4, 8) independent
 * chains per lane at full occupancy, `steps` gathers per chain -- the ceiling for a
 * Move that is one such gather. */
int trw_probe_random_gather(trw_context *ctx, uint64_t table_bytes, int32_t record_bytes, int32_t chains,
                            int32_t steps, double *gathers_per_s);

/* ---- graph ----
 * Graph::load_csr (util/graph/graph.cpp:45-101) from host arrays: degree[nv]
 * (b_degree.bin), adj[ne] sorted per vertex (b_adj.bin), optional FP64 weights
 * (b_edge_weight.bin) and int32 labels (b_edge_label.bin).  offset_ is the
 * exclusive prefix of degree (graph.cpp:258-263).  The library copies the arrays
 * to HBM; the caller keeps ownership of its host memory. */
int trw_graph_upload(trw_context *ctx, int32_t num_vertices, int64_t num_edges, const int32_t *degree,
                     const int32_t *adj, const double *edge_weight, const int32_t *edge_label, trw_graph **out);
int trw_graph_free(trw_graph *g);
int trw_graph_info(const trw_graph *g, int32_t *num_vertices, int64_t *num_edges, int32_t *max_degree);

/* ---- static tables ----
 * static_initialization (creeper.h:102-175) on the device: alias tables
 * (alias_initialization, alias_sampling.h:12-54), ITS prefix sums
 * (its_initialization, inverse_transformation_sampling.h:19-24) or rejection
 * maxima (rs_initialization, rejection_sampling.h:12-14), byte-identical to the
 * reference's.  weights: TRW_WEIGHTS_EDGE or TRW_WEIGHTS_UNIT. */
int trw_build_tables(trw_graph *g, int32_t sampling, int32_t weights);
/* Copies of the built tables into host memory (b_edge_weight_alias_table.bin,
 * b_edge_weight_prefix_sum.bin, b_edge_weight_max.bin layouts, graph.cpp:184-206). */
int trw_copy_alias_table(const trw_graph *g, trw_alias_slot *host_out);
int trw_copy_prefix_sum(const trw_graph *g, double *host_out);
int trw_copy_rejection_max(const trw_graph *g, double *host_out);
/* The reference's table files (graph.cpp:18-30, 184-212; written by alias_preprocess,
 * its_preprocess and rejection_preprocess): b_edge_weight_alias_table.bin (AliasSlot[E]),
 * b_edge_weight_prefix_sum.bin (double[E]), b_edge_weight_max.bin (double[V]), raw arrays.
 * store writes every table that has been built; load reads one into the graph in place of a
 * build, `weights` naming what the file was built from (TRW_WEIGHTS_*). */
int trw_graph_store_tables(const trw_graph *g, const char *dir);
int trw_graph_load_tables(trw_graph *g, const char *dir, int32_t sampling, int32_t weights);
/* Device time of the last trw_build_tables call in milliseconds. */
int trw_last_build_ms(const trw_graph *g, double *ms);

/* static_initialization (creeper.h:102-175) for an arbitrary functor: static_gather
 * (creeper.h:24-99) evaluates F::weight for every edge, here done by the caller's kernel into
 * device_weights (DEVICE array of num_edges doubles, weight of edge e at [e]); the alias, ITS or
 * rejection table is then built from those weights exactly as trw_build_tables builds it from
 * edge weights.  The table is tagged TRW_WEIGHTS_CUSTOM (a later walk of a built-in app that
 * needs edge or unit weights rebuilds it).  Rejection keeps a copy of the weights (its Move
 * reads them, rejection_sampling.h:334-366). */
#define TRW_WEIGHTS_CUSTOM 2
int trw_build_tables_from(trw_graph *g, int32_t sampling, const double *device_weights);

/* The graph's arrays in HBM (DEVICE pointers, null where absent), for kernels compiled outside
 * the library that walk it: the fields of class Graph (util/graph/graph.h:51-83) the reference's
 * Moves and functors read -- offset_ (exclusive prefix of degree_, num_vertices + 1 entries),
 * degree_, adj_, edge_weight_, edge_label_ -- and the static tables (creeper.h:108-170):
 * edge_weight_alias_table_, edge_weight_prefix_sum_, edge_weight_rejection_max_ and the
 * weights the rejection maxima were built from (edge_weight_rejection_; null for unit
 * weights).  A table pointer is null until that table is built; *_weights give the
 * TRW_WEIGHTS_* it was built from (-1: not built).  membership_lines / membership_root: the
 * tables Graph::is_neighbor (graph.h:180-199) can be answered from on the device
 * (trw_dev::member_contains in thunderrw_device.h; null until trw_build_membership).  Valid
 * until the graph is freed or the table rebuilt. */
typedef struct {
    int32_t num_vertices;
    int32_t max_degree;
    int64_t num_edges;
    const int64_t *offset;
    const int32_t *degree;
    const int32_t *adj;
    const double *edge_weight;
    const int32_t *edge_label;
    const trw_alias_slot *alias_table;
    const double *prefix_sum;
    const double *rejection_max;
    const double *rejection_weight;
    int32_t alias_weights, its_weights, rejection_weights;
    int32_t reserved;
    const void *membership_lines;
    const uint32_t *membership_root;
} trw_device_view;
int trw_graph_device_view(const trw_graph *g, trw_device_view *out);

/* Membership tables for is_neighbor on the device (no reference counterpart: graph.h:180-199
 * binary-searches the smaller endpoint's sorted list; these answer the same question in one
 * 128-byte read): a hash set of the neighbours of every vertex with neighbours, in 128-byte
 * buckets, layout in thunderrw_device.h.  Built once per graph (the first Node2Vec walk builds
 * them too); a graph with 2^32 edges or more gets none and its is_neighbor keeps the search. */
int trw_build_membership(trw_graph *g);

/* ---- per-graph tuning ----
 * Launch choices that change speed, never results (no reference counterpart; the
 * reference fixes its ring sizes at compile time, types.h:17-31).  trw_autotune picks
 * them by measurement; 0 means the built-in default. */
#define TRW_TUNE_WALKERS_PER_THREAD 0  /* FUSED uniform/alias/static ITS walkers per lane when params say 0: 1, 2, 4 */
#define TRW_TUNE_ITS_TREE_MIN_DEGREE 1 /* ITS: vertices of at least this degree search their 128-byte-node
                                          tree, the rest binary-search their prefix sums */
#define TRW_TUNE_QUEUE_CHUNK 2         /* QUEUE: walker indices a wavefront takes per atomic (64..4096; at most
                                          the batch's walkers per wavefront of the grid) */
#define TRW_TUNE_ITS_GUIDE_OFF 3       /* ITS: 0 = steps read the guide records where the table has them
                                          (one gather), 1 = they search the prefix-sum structures (the
                                          picks are identical; not autotuned) */
/* Launch shapes, one value per Move (TRW_MOVE_*): */
#define TRW_TUNE_FUSED_BLOCK 4         /* FUSED: threads per block of the Move's kernel: 64, 128 or 256 */
#define TRW_TUNE_FUSED_TILE 5          /* FUSED: path positions a lane stages in LDS before writing them
                                          out: 16, 32 or 64 (0: the whole row when the LDS budget holds it) */
#define TRW_TUNE_QUEUE_BLOCKS_PER_CU 6 /* QUEUE: blocks of the persistent grid per CU, 1..16 (0: as many as
                                          are resident at once) */
#define TRW_TUNE_N2V_BLOCK 7           /* FUSED Node2Vec over membership tables: threads per block, 128 or 256 */
#define TRW_TUNE_GRAPH_BURST 8         /* GRAPH: steps a launch takes each live walker, 1..256 (0: 16) */
#define TRW_TUNE_KNOBS 9
/* The Moves the shape knobs are kept for (the walk kernels' Move of a walk). */
#define TRW_MOVE_UNIFORM 0
#define TRW_MOVE_ALIAS 1
#define TRW_MOVE_ITS 2
#define TRW_MOVE_REJECTION 3
#define TRW_MOVE_NODE2VEC 4
#define TRW_MOVE_METAPATH 5
#define TRW_MOVES 6
/* knobs 0-3 are per graph; a shape knob (4-8) set here applies to every Move, and reads
 * back the uniform Move's value. */
int trw_graph_set_tuning(trw_graph *g, int32_t knob, int64_t value);
int trw_graph_get_tuning(const trw_graph *g, int32_t knob, int64_t *value);
/* A shape knob (4-8) of one Move. */
int trw_graph_set_move_tuning(trw_graph *g, int32_t move, int32_t knob, int64_t value);
int trw_graph_get_move_tuning(const trw_graph *g, int32_t move, int32_t knob, int64_t *value);

/* Autotuner: for the walk `params` describes (rng must be PHILOX), times a sample of
 * sample_walkers walkers from random non-isolated sources (seeded) for each candidate
 * of every knob that applies -- walkers per lane for FUSED uniform/alias/ITS records,
 * the ITS tree degree threshold, the QUEUE chunk, and the launch shapes of the walk's
 * Move: FUSED threads per block and path tile, QUEUE blocks per CU, Node2Vec threads per
 * block, GRAPH steps per launch -- one knob at a time (the QUEUE chunk and blocks per CU as
 * every pair of the two), and keeps the fastest on the graph (shape knobs for the walk's
 * Move).  Each candidate is timed at least three times, and a short sample until 20 ms have
 * been timed (at most 63 runs); the median counts.  Walks are identical under every setting,
 * so tuning never changes results. */
typedef struct {
    int32_t knobs;                 /* entries used below */
    int32_t move;                  /* TRW_MOVE_* of the walk (the shape knobs were kept for it) */
    int32_t knob[TRW_TUNE_KNOBS];  /* TRW_TUNE_* */
    int32_t candidates[TRW_TUNE_KNOBS]; /* values tried per knob */
    int64_t chosen[TRW_TUNE_KNOBS];
    double default_steps_per_s;    /* all knobs at their defaults */
    double tuned_steps_per_s;
} trw_tune_report;
int trw_autotune(trw_graph *g, const trw_walk_params *params, int32_t sample_walkers, uint64_t seed,
                 int32_t with_paths, trw_tune_report *report);
/* The same for a mixed batch as trw_walk_mix_async runs it (QUEUE form, each kind's path rows at
 * its own length when with_paths, io.path_stride_b): the sample's walkers
 * [0, split) walk as params_a, the rest as params_b; the queue knobs are tuned, and the launch
 * shapes kept for params_a's Move (the queue kernel of a mix takes the first kind's). */
int trw_autotune_mix(trw_graph *g, const trw_walk_params *params_a, const trw_walk_params *params_b, int32_t split,
                     int32_t sample_walkers, uint64_t seed, int32_t with_paths, trw_tune_report *report);

/* ---- walks ----
 * compute<F>(graph, walkers, f) (creeper.h:689-831) for a built-in app, on a
 * caller stream (hipStream_t as void*; NULL = the context stream).  Asynchronous:
 * returns after the launches are queued (the STEP and GRAPH forms read a live-walker
 * count back between step chunks, so they return when the walk is done).
 * A graph's walks use its work-queue head and walker-state scratch: walks on one graph
 * must not overlap in time, from different host threads or on different streams (the
 * query server gives each of its slots its own queue head for that reason).  Tables
 * are built lazily by the first walk that needs them unless trw_build_tables ran. */
int trw_walk_async(trw_graph *g, const trw_walk_params *params, const trw_walk_io *io, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the walk
 * trw_walk_async runs, with the same kernels and results, that calls no hipMalloc, hipFree
 * or other allocating routine.  The caller passes inputs and outputs as for trw_walk_async;
 * the tables the walk reads must already be built (else TRW_E_STATE: nothing is built here).
 * Forms AUTO, FUSED, QUEUE and STEP (GRAPH: TRW_E_UNSUPPORTED).  The QUEUE form's work-queue
 * head and the STEP form's live flag and walker state live in a fixed device array of
 * 16 MiB + 256 bytes that each device holds until the process ends; the STEP form therefore
 * takes at most 262,144 walkers (more: TRW_E_INVALID).  Every call on one device shares that
 * array, so calls that overlap in time on one device race. */
int trw_walk_noalloc(trw_graph *g, const trw_walk_params *params, const trw_walk_io *io, void *stream);

/* Whether the library has a Move kernel for params' (app, execution, sampling): TRW_OK, or
 * TRW_E_UNSUPPORTED (TRW_E_INVALID for values outside the reference's enums).  The pairs
 * the reference's apps accept (-em/-sm of apps/<app>.cpp) that the library has no kernel for --
 * walker-dependent weights outside Node2Vec's max-weight rejection and MetaPath's dynamic ITS,
 * or max-weight rejection for DeepWalk/GenericWalk -- run in the drop-in functor engine
 * (thunderrw_amd/dropin/functor_walk.h).  No graph or device is touched. */
int trw_walk_supported(const trw_walk_params *params);

/* Two walker kinds in one batch: walkers [0, split) run params_a, [split, n) run
 * params_b (the uneven PPR + MetaPath mix).  form: QUEUE, STEP or GRAPH.  Path rows: with
 * io.path_stride_b = 0 every walker's row has io.path_stride ints (>= both lengths); otherwise
 * walker w < split has row paths + w * path_stride (>= params_a's length) and walker w >= split
 * row paths + split * path_stride + (w - split) * path_stride_b (>= params_b's length), so each
 * kind's rows are as long as its own walks, as two separate runs would write them (e.g. PPR's
 * length-1 rows as one dense array of sources). */
int trw_walk_mix_async(trw_graph *g, const trw_walk_params *params_a, const trw_walk_params *params_b,
                       int32_t split, const trw_walk_io *io, void *stream);

/* Blocking host-memory convenience: copies sources/draws in and results out.
 * draws/term_draws are host arrays in the [k][n_walkers] layout (may be NULL for PHILOX);
 * paths is n_walkers x max(length,1); outputs are nullable. */
int trw_walk_host(trw_graph *g, const trw_walk_params *params, const int32_t *sources, int32_t n_walkers,
                  const uint32_t *draws, int64_t draws_per_walker, const uint32_t *term_draws,
                  int64_t term_draws_per_walker, int32_t *paths, int32_t *lengths, int32_t *ends, int32_t *status,
                  uint64_t *steps);

/* ---- replay stream (host) ----
 * Walker i's draws are words [i*K, (i+1)*K) of the SFMT-19937 stream
 * init_gen_rand(seed) (SFMT.c:349-362); written for walkers
 * [first_walker, first_walker + n_walkers) into out[k * n_walkers + (i - first_walker)]. */
int trw_replay_fill(uint32_t seed, int64_t first_walker, int32_t n_walkers, int64_t draws_per_walker,
                    uint32_t *host_out);

/* ---- walker init (walker_generator.cpp) ---- */
/* generate_all_source (walker_generator.cpp:9-25): non-isolated vertices,
 * shuffled by std::shuffle with std::mt19937(0) (utility.h:96-100).  out has room for nv. */
int trw_sources_all(const int32_t *degree, int32_t num_vertices, int32_t *out, int32_t *count);
/* generate_single_source (walker_generator.cpp:42-63): n copies of the first max-degree vertex. */
int trw_sources_single(const int32_t *degree, int32_t num_vertices, int32_t n, int32_t *out);
/* generate_random_source (walker_generator.cpp:27-40): SFMT(seed) u % nv, skipping degree 0. */
int trw_sources_random(const int32_t *degree, int32_t num_vertices, int32_t n, uint32_t seed, int32_t *out);

/* ---- host graphs: synthetic generator and the reference's CSR files ---- */
typedef struct trw_host_graph trw_host_graph;
/* Seeded Chung-Lu power-law graph: expected degrees dmin*(1-U)^(-1/(gamma-1)) capped at dmax,
 * endpoint pairs drawn proportional to them until the undirected edge count reaches the target,
 * self-loops and duplicates dropped, symmetrised and sorted as EdgeList2CSR does
 * (graph.cpp:463-519).  Weights U[1,5) and labels U{0..n_labels-1} per undirected edge, mirrored. */
int trw_host_graph_chung_lu(int32_t num_vertices, int64_t undirected_edges, double gamma, double dmin,
                            double dmax, uint64_t seed, int32_t with_weights, int32_t n_labels,
                            trw_host_graph **out);
/* Graph::load_csr file set (graph.cpp:233-426): b_degree.bin, b_adj.bin, optional
 * b_edge_weight.bin / b_edge_label.bin. */
int trw_host_graph_load(const char *dir, int32_t with_weights, int32_t with_labels, trw_host_graph **out);
int trw_host_graph_store(const trw_host_graph *h, const char *dir);
/* Graph::load_edge_list(dir, skip) (graph.cpp:428-519; toolset/edgelist_to_csr.cpp): text
 * lines "src dst" of <dir>/b_edge_list.bin, lines starting with `skip` ignored, pairs ordered
 * (min, max), sorted, duplicates and self-loops dropped, symmetrised; vertex count = largest
 * id + 1.  A line that does not hold two integers, or holds a negative one, is TRW_E_IO (the
 * reference exits). */
int trw_host_graph_load_edge_list(const char *dir, int32_t skip, trw_host_graph **out);
int trw_host_graph_arrays(const trw_host_graph *h, int32_t *num_vertices, int64_t *num_edges,
                          const int32_t **degree, const int32_t **adj, const double **edge_weight,
                          const int32_t **edge_label);
int trw_host_graph_free(trw_host_graph *h);

/* ---- serving: continuous batching of walk queries ----
 * No counterpart in the reference, whose compute<F> (creeper.h:689-831) is one
 * blocking batch over a fixed walker vector.  A server keeps one uploaded graph
 * (tables built beforehand with trw_build_tables) resident and runs a dispatcher
 * thread: queries that arrive while others are on the device are packed, in arrival
 * order, into batches of walkers; each batch runs on one of `streams` slots (own HIP
 * stream, device buffers and pinned staging), so batches overlap on the device.  A
 * batch holds at most max_batch_walkers walkers and path_bytes_per_batch bytes of
 * path rows (walkers x length x 4 B); a query larger than that is split over several
 * batches.  Admission defers a query until a slot frees up, and submit rejects one
 * with TRW_E_BUSY when max_pending_queries are already waiting.
 *
 * Walker i of a query starts at sources[i % n_sources] and draws from the Philox
 * stream of (params.seed, walker_id_base + i), so a query's result does not depend
 * on what else is in flight: alone or in a mix, it is the same. */
#define TRW_E_BUSY (-7)    /* admission: the pending queue is full */
#define TRW_E_TIMEOUT (-8) /* trw_server_wait: not finished within the timeout */

#define TRW_RESULT_PPR 0   /* end-vertex counts (PPR estimate = count / n_walkers) */
#define TRW_RESULT_PATHS 1 /* full paths, n_walkers x max(length, 1) */

#define TRW_QUERY_PENDING 0
#define TRW_QUERY_RUNNING 1
#define TRW_QUERY_DONE 2
#define TRW_QUERY_FAILED 3

typedef struct trw_server trw_server;

typedef struct {
    int32_t streams;              /* slots (batches in flight), 1..16; 0 = 4 */
    int32_t max_batch_walkers;    /* walker capacity of one batch; 0 = 1 << 20 */
    int64_t path_bytes_per_batch; /* path-row capacity of one batch; 0 = 64 MiB */
    int32_t max_pending_queries;  /* queries waiting for admission before submit says BUSY; 0 = 1 << 16 */
    int32_t completion_threads;   /* host threads that copy results out and build PPR counts; 0 = 4 */
    int32_t max_queries_per_batch;/* 0 = 1024 */
    int32_t reserved;
} trw_server_config;

typedef struct {
    const int32_t *sources;   /* host array, copied at submit */
    int32_t n_sources;
    int32_t n_walkers;        /* walker i starts at sources[i % n_sources] */
    uint64_t walker_id_base;  /* Philox id of walker 0; give queries disjoint id ranges */
    trw_walk_params params;   /* rng must be TRW_RNG_PHILOX; form AUTO, FUSED or QUEUE */
    int32_t result;           /* TRW_RESULT_* */
    int32_t reserved;
} trw_query;

typedef struct {
    int32_t state;            /* TRW_QUERY_* */
    int32_t n_walkers;
    uint64_t steps;           /* sum over walkers of (length - 1) */
    uint64_t errors;          /* walkers that ended with status >= TRW_WALKER_REPLAY_EXHAUSTED */
    int64_t ppr_entries;      /* distinct end vertices (TRW_RESULT_PPR) */
    double queue_ms;          /* arrival -> first walkers launched */
    double latency_ms;        /* arrival -> result ready */
    double device_ms;         /* first walkers launched -> the last batch holding its walkers seen done
                                 on the device; latency - queue - device is the host-side completion */
} trw_query_info;

typedef struct {
    uint64_t queries_done;
    uint64_t walkers;
    uint64_t steps;
    uint64_t batches;
    uint64_t rejected;        /* submits refused with TRW_E_BUSY */
    double mean_batch_walkers;
    double mean_batch_queries;
} trw_server_stats;

typedef struct {
    double wall_ms;           /* first scheduled arrival -> last result */
    double offered_qps;
    double achieved_qps;
    uint64_t steps;
    double steps_per_s;
    double p50_ms, p90_ms, p99_ms, p999_ms, mean_ms, max_ms;
    uint64_t batches;
    double mean_batch_queries;
    uint64_t rejected;
} trw_open_loop_report;

/* Allocates the slots, starts the dispatcher, then warms up: two rounds of full batches of
 * internal PPR queries (uniform moves from the highest-degree vertex) on every slot, after which
 * query ids continue and the statistics start from zero.  The first time many full batches run at
 * once the device stalls for several ms, once; this keeps that out of the first queries' latency. */
int trw_server_create(trw_graph *g, const trw_server_config *config, trw_server **out);
/* Stops the dispatcher after the queries in flight finish and frees every result.  No other
 * call on the server may be running; destroy servers before their graph.  The other
 * trw_server_* calls may be made from any number of host threads at once. */
int trw_server_destroy(trw_server *s);
/* Non-blocking; the query's arrival time is the time of the call. */
int trw_server_submit(trw_server *s, const trw_query *q, uint64_t *query_id);
/* Blocks until the query is done or failed (timeout_ms < 0: no limit). */
int trw_server_wait(trw_server *s, uint64_t query_id, double timeout_ms, trw_query_info *info);
int trw_server_query_info(trw_server *s, uint64_t query_id, trw_query_info *info);
/* TRW_RESULT_PPR: the distinct end vertices ascending and their walker counts
 * (capacity >= info.ppr_entries). */
int trw_server_ppr_result(trw_server *s, uint64_t query_id, int32_t *vertices, uint32_t *counts, int64_t capacity);
/* Per-walker outputs of a done TRW_RESULT_PATHS query (nullable): paths (n_walkers x
 * max(length,1), -1 padded), lengths, ends, status.  A PPR query's batch is reduced on
 * the device to end-vertex counts, so it keeps only those (trw_server_ppr_result). */
int trw_server_walk_result(trw_server *s, uint64_t query_id, int32_t *paths, int32_t *lengths, int32_t *ends,
                           int32_t *status);
int trw_server_release(trw_server *s, uint64_t query_id);
int trw_server_stats_get(trw_server *s, trw_server_stats *out);
/* Open-loop load: submits queries[0..n) at Poisson arrival times of rate qps (seeded),
 * waits for all, and reports latency percentiles measured from each scheduled arrival
 * (a submit that runs late still counts from its schedule).  ids and latency_ms
 * (nullable, n entries) receive each query's id and latency; results stay held until
 * trw_server_release. */
int trw_server_run_open_loop(trw_server *s, const trw_query *queries, int32_t n, double qps, uint64_t seed,
                             uint64_t *ids, double *latency_ms, trw_open_loop_report *report);

#ifdef __cplusplus
}
#endif

#endif /* THUNDERRW_H */

/*
 * bingo.h - C ABI of the MI355X (gfx950) bias-factorized random-walk engine.
 *
 * Every entry point takes plain pointers and sizes; no C++ or torch types
 * cross this boundary.  Node ids are 1..n as in the reference, whose loader
 * shifts 0-based input ids by one (include/graph.cuh:22-27).  All calls
 * return 0 or a negative BINGO_E* code; bingo_last_error() describes the
 * last failure of the calling thread.  A handle is single-writer: calls on
 * one handle must not overlap.  Work is ordered on the handle's own HIP
 * stream; host-pointer entry points return after the work has finished.
 *
 * Reference surface replaced (Joeyw02/Bingo):
 *   types ........ include/type.h:4-36 (app, utype, Edges, Deleted)
 *   constants .... include/api.h:5-30 (P, Q, TP, LEN, ORIGINAL, RATE)
 *   build ........ include/operation.cuh:33 build(Edges*, amount, ...)
 *   insert ....... include/operation.cuh:144 insert(Edges*, amount, ...)
 *   deleteE ...... include/operation.cuh:215 deleteE(Deleted*, amount, ...)
 *   randomWalk ... include/graph.cuh:93 randomWalk(app)
 *   instance ..... include/apps.h:4 int instance(app, utype, std::string)
 */
#ifndef BINGO_H
#define BINGO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BINGO_OK 0
#define BINGO_EINVAL (-1)      /* bad argument or batch entry (whole batch rejected) */
#define BINGO_ENOMEM (-2)      /* HBM budget or device allocation exhausted */
#define BINGO_EDEVICE (-3)     /* HIP runtime error */
#define BINGO_ESTATE (-4)      /* call not valid in the handle's state */
#define BINGO_EUNSUPPORTED (-5) /* feature disabled for this handle */

/* ref enum class app (type.h:4-10) */
typedef enum { BINGO_DEEPWALK = 0, BINGO_NODE2VEC = 1, BINGO_PPR = 2, BINGO_SAMPLING = 3 } bingo_app;
/* ref enum class utype (type.h:16-21): insert-only, delete-only, mixed */
typedef enum { BINGO_UPD_INSERT = 0, BINGO_UPD_DELETE = 1, BINGO_UPD_MIXED = 2 } bingo_utype;
/* synthetic weight laws of the workload generators */
typedef enum { BINGO_W_CONST1 = 0, BINGO_W_UNIFORM16 = 1, BINGO_W_LOGUNI32 = 2, BINGO_W_ZIPF24 = 3 } bingo_wkind;

/* ref struct Edges{int u,v; float weight} (type.h:24-29): w is a u32 weight,
 * or, on a handle created with float_weights, the IEEE-754 bits of a float
 * weight in (0, 2^32), taken to 2^-32 resolution (ref wtype::Float,
 * include/Bingo.cuh:65-69) */
typedef struct { int32_t u, v; uint32_t w; } bingo_edge;
/* ref struct Deleted{int u,id} (type.h:33-37): id is the per-node slot ordinal */
typedef struct { int32_t u, slot; } bingo_del;

typedef struct bingo_graph bingo_graph;

typedef struct {
    int32_t device;         /* HIP device ordinal */
    int32_t neighbor_index; /* keep exact neighbour multisets (needed by node2vec; ref NODE2VEC) */
    uint64_t hbm_budget;    /* bytes of HBM the store may use; 0 = 90% of free memory */
    double pool_slack;      /* extra pool room reserved at load, as a fraction (0.25) */
    int32_t min_slots;      /* minimum slot capacity per node (ref ORIGINAL = 16) */
    int32_t fit_load;       /* 1: load sizes every region to its contents plus fit_headroom
                               instead of the next power of two (at least min_slots) */
    double fit_headroom;    /* fit_load's spare room per region, as a fraction (1/16) */
    int32_t float_weights;  /* 1: edge weights are float bits (ref wtype::Float, Bingo.cuh:65-69,
                               254-264); integer part through the bit groups, fraction per slot */
    int32_t grow_in_place;  /* 1: each pool is a reserved address range that grows by mapping
                               more HBM in 1 GiB chunks (no copy, no second copy in HBM while
                               growing, pool addresses stable; at least 1 GiB per pool, so for
                               large stores); 0: a growing pool is copied to a larger
                               allocation.  Falls back to copying where mapping is refused. */
    double dense_rate;      /* ref RATE (api.h:25, operation.cuh:72): 0 (default) or >= 1: every
                               non-empty bit group keeps a member list; in (0, 1): the batch that
                               gives a node its first slots classes its groups, and group j is
                               dense when count_j >= 1 and count_j >= rate * (that batch's arcs of
                               the node), rate taken to 2^-16.  A dense group keeps only its count;
                               walks draw from it by rejection over the node's slots (ref
                               Bingo.cuh:244-252).  Less HBM, costlier draws from dense groups;
                               with it, max_attempts <= 65536. */
} bingo_config;

typedef struct {
    uint64_t seed;          /* Philox key; walker i uses counter (i, step, attempt) */
    double p, q;            /* node2vec return / in-out parameters (ref P=0.5, Q=2) */
    double tp;              /* PPR stop probability per step (ref TP=0.0125) */
    int32_t max_len;        /* path columns, vertices per walk incl. start (ref LEN=80) */
    int32_t max_attempts;   /* node2vec rejection cap per step; last draw accepted (<= 65536
                               with float weights or dense groups) */
    int32_t form;           /* BINGO_FORM_*: how walkers are scheduled (paths are the same) */
    int32_t block;          /* threads per block, 64..512; 0 = tuned or default */
    int32_t walkers_per_lane; /* persistent form: 1, 2 or 4 walker slots per lane; 0 = tuned */
    int32_t lds_depth;      /* persistent form: path staging in LDS, 0, 4 or 16; -1 = tuned */
    int32_t blocks_per_cu;  /* persistent form: resident blocks per CU; 0 = occupancy limit */
    int32_t reserved;
} bingo_walk_params;

/* walk scheduling forms */
#define BINGO_FORM_AUTO 0       /* persistent with the handle's tuned knobs (defaults if untuned) */
#define BINGO_FORM_PLAIN 1      /* one thread per walker, one launch (ref deepwalkKernel style) */
#define BINGO_FORM_PERSISTENT 2 /* persistent lanes with walker refill, knobs as given */

typedef struct {
    double load_ms, insert_ms, delete_ms, walk_ms; /* HIP-event time of the last call */
    double walk_kernel_ms;  /* device time of the last walk kernel */
    uint64_t walk_steps;    /* appended vertices of the last walk */
    uint64_t walk_attempts; /* sample draws of the last walk (node2vec rejections included) */
    uint64_t walk_probes;   /* neighbour-table membership lookups of the last walk (node2vec) */
    uint64_t walk_folds;    /* node2vec steps that looked up the return arc's weight sum
                               (folded return region; only steps with a fold coin below
                               the bound every threshold obeys) */
    uint64_t arcs;          /* live arcs */
    uint64_t bytes_reserved;/* HBM held by the store */
    uint64_t pool_used[4];  /* slot, back-index, member, neighbour pools (elements) */
    double compact_ms;      /* HIP-event time of the last bingo_compact */
    uint64_t walk_dense_trials; /* slot weights read by the last walk's draws from dense groups
                               (rejection trials; dense_rate handles) */
} bingo_stats;

/* Flat export for bit-exact diffs (node-major; see bingo_export). */
typedef struct {
    uint64_t total_slots;
    uint64_t total_members;
    uint32_t *nslots;   /* [n+1] */
    uint64_t *mass;     /* [n+1] sum of live weights */
    uint32_t *bsize;    /* [(n+1)*32] bucket sizes */
    uint64_t *T;        /* [(n+1)*32] alias stay thresholds */
    uint8_t *alias;     /* [(n+1)*32] alias partners */
    uint32_t *slot_v;   /* [total_slots] */
    uint32_t *slot_w;   /* [total_slots] 0 = deleted */
    uint64_t *members;  /* [total_members] (v << 32) | slot, per node, per listed group */
    uint32_t *backpos;  /* [total_members] per live slot, its listed set bits ascending */
    /* float_weights handles only (NULL to skip; slot_f with the slot arrays): */
    uint32_t *slot_f;   /* [total_slots] fraction of each slot's weight, x 2^32 (slot_w: integer part) */
    uint64_t *fmass;    /* [n+1] sum of the live fractions, x 2^32 */
    uint64_t *fthr;     /* [n+1] fractional-part threshold floor(2^64 F / (S 2^32 + F)) */
    uint32_t *dmask;    /* [n+1] dense groups of each node, bit j (NULL to skip); a dense group
                           has no members or back positions in the arrays above */
} bingo_dump;

/* ---- lifetime ---- */
void bingo_config_default(bingo_config *cfg);
void bingo_walk_params_default(bingo_walk_params *p);
int bingo_create(const bingo_config *cfg, bingo_graph **out);
void bingo_destroy(bingo_graph *g);
const char *bingo_last_error(void);
const char *bingo_version(void);

/* ---- graph load: ref build() (operation.cuh:33-121) ----
 * Arcs are taken in array order; the k-th arc of a node gets slot k (as
 * graph.cuh:38-42 numbers nodeIdu).  u, v in 1..n; w >= 1 (float weights:
 * finite, 2^-32 <= w < 2^32).  A load of more than 2^30 arcs is applied as
 * consecutive batches of 2^30 arcs (with dense_rate, a node's classes come
 * from the first of them that holds its arcs); insert and delete batches
 * take at most 2^31 - 1 entries. */
int bingo_load(bingo_graph *g, int32_t n, const bingo_edge *edges, int64_t m);
int bingo_load_device(bingo_graph *g, int32_t n, const bingo_edge *d_edges, int64_t m);

/* ---- batched insert: ref insert() (operation.cuh:144-197) ----
 * Arcs of one node take its next free slots in batch order. */
int bingo_insert(bingo_graph *g, const bingo_edge *edges, int64_t k);
int bingo_insert_device(bingo_graph *g, const bingo_edge *d_edges, int64_t k);

/* ---- batched delete: ref deleteE() (operation.cuh:215-251) ----
 * Each (u, slot) must be live and appear once; the slot is tombstoned. */
int bingo_delete(bingo_graph *g, const bingo_del *dels, int64_t k);
int bingo_delete_device(bingo_graph *g, const bingo_del *d_dels, int64_t k);

/* ---- compaction of tombstoned slots (no reference counterpart: the
 * reference tombstones a deleted slot and never reuses it,
 * include/operation.cuh:209-210, so a node's slot array only grows) ----
 * Every node u with dead = nslots(u) - live(u) >= 1 and
 *     dead * 2^16 >= floor(min_dead_frac * 2^16) * nslots(u)
 * is compacted in place: its k-th live slot (in slot order) becomes slot k
 * and nslots(u) becomes live(u).  When u's dead slots are exactly its first
 * ones (the oldest arcs expired, as in a sliding window), its member lists
 * keep their order and walks draw the same targets as before; otherwise
 * each of its listed groups' member lists is rebuilt in slot order (the
 * k-th member of group j is the k-th live slot with bit j) and the law of a
 * draw is unchanged up to the multiply-shift's 2^-32 steps.  Masses and
 * alias tables are unchanged; deletes after the call name the new slot
 * numbers, and new arcs take slots from live(u) on.  A caller that tracks slot numbers applies the
 * same rule (bingo_updgen_compact does).  nodes_out / slots_out (may be
 * NULL) receive the number of nodes compacted and of slots reclaimed.  The
 * call waits for the device once (the selection sizes its scratch). */
int bingo_compact(bingo_graph *g, double min_dead_frac, uint64_t *nodes_out, uint64_t *slots_out);

/* ---- allocation-free delete (no reference counterpart) ----
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * batch and result as bingo_delete, with every buffer of the batch in a
 * statically sized device array of BINGO_NOALLOC_SCRATCH_BYTES per GPU (set
 * up when the first handle on that GPU is created).  Batches of more than
 * BINGO_NOALLOC_MAX_DELETES deletes are rejected with BINGO_EINVAL.  All
 * calls on one GPU share its array: calls that overlap on one GPU race. */
#define BINGO_NOALLOC_MAX_DELETES 65536
#define BINGO_NOALLOC_SCRATCH_BYTES (16u << 20)
int bingo_delete_noalloc(bingo_graph *g, const bingo_del *dels, int64_t k);

/* ---- walks: ref randomWalk() (graph.cuh:93-150) and the kernels of
 * operation.cuh:252-368 (DeepWalk, node2vec, PPR) and 407-430 (sampling).
 * paths is [nwalkers][params->max_len] int32, -1 padded.  starts NULL:
 * walker i starts at (i % n) + 1.  lens NULL: max_len for every walker.
 * steps_out (may be NULL) receives the number of appended vertices. */
int bingo_walk(bingo_graph *g, int app, const bingo_walk_params *params, const int32_t *starts,
               const int32_t *lens, int64_t nwalkers, int32_t *paths, uint64_t *steps_out);
/* device-pointer form, enqueued on the handle's stream; *d_steps (device
 * uint64) is incremented by the steps.  Returns without waiting.  Starts and
 * lengths are checked on the device: a walker whose start is outside 1..n or
 * whose length is outside 1..max_len gets a row of -1, and the next
 * bingo_sync returns BINGO_EINVAL (the flag clears when the next walk starts;
 * walk graph launches report the same way). */
int bingo_walk_device(bingo_graph *g, int app, const bingo_walk_params *params,
                      const int32_t *d_starts, const int32_t *d_lens, int64_t nwalkers,
                      int32_t *d_paths, uint64_t *d_steps);

/* A walk launch captured once as a HIP graph and replayed (batched form):
 * the reset of the walk counters and the walk kernel, with the device
 * buffers given here.  Replays see the store as it is at replay time; if
 * update batches grew the HBM pools since capture, the launch re-captures
 * transparently.  *d_steps accumulates over replays. */
typedef struct bingo_walk_graph bingo_walk_graph;
int bingo_walk_graph_create(bingo_graph *g, int app, const bingo_walk_params *params, const int32_t *d_starts,
                            const int32_t *d_lens, int64_t nwalkers, int32_t *d_paths, uint64_t *d_steps,
                            bingo_walk_graph **out);
int bingo_walk_graph_launch(bingo_walk_graph *wg);
void bingo_walk_graph_destroy(bingo_walk_graph *wg);

/* ---- captured update batches (batched form; no reference counterpart: the
 * reference runs insert() and deleteE() as host functions that wait for
 * the device mid-batch, include/operation.cuh:144-251) ----
 * An insert batch and a delete batch captured as HIP graphs, one per
 * power-of-two batch size (captured on first use), replayed with the sizes
 * of each launch: d_ins holds nins <= max_ins arcs and d_dels ndels <=
 * max_dels deletes when the launch runs (device buffers given here; a
 * launch reads them in stream order, so writing the next batch with stream-
 * ordered copies after the launch is safe).  A launch applies the inserts,
 * then the deletes, with the same results as bingo_insert_device and
 * bingo_delete_device, and returns without waiting: nothing in it waits for
 * the host.  Pools grow on the device inside room committed at creation
 * (room_bytes, 0 = the store's pool bytes, at least 1 GiB, at most half of
 * the budget left, spread over the pools by size).  A batch that does not
 * fit that room, or is invalid, applies nothing, and neither do the
 * captured batches after it until the next call that waits for the device
 * (bingo_sync, an insert or delete) reports it: BINGO_ENOMEM (room) or
 * BINGO_EINVAL (invalid).  The store then holds every batch before the
 * failed one. */
typedef struct bingo_update_graph bingo_update_graph;
int bingo_update_graph_create(bingo_graph *g, const bingo_edge *d_ins, int64_t max_ins, const bingo_del *d_dels,
                              int64_t max_dels, uint64_t room_bytes, bingo_update_graph **out);
int bingo_update_graph_launch(bingo_update_graph *ug, int64_t nins, int64_t ndels);
void bingo_update_graph_destroy(bingo_update_graph *ug);

/* Autotuner: times the persistent form over block size, walkers per lane
 * and LDS staging depth (and the plain form) on up to `sample` walkers with
 * strided starts, keeps the fastest for `app` on this handle (used by
 * BINGO_FORM_AUTO) and returns it in *chosen (may be NULL).  The graph is
 * not modified. */
int bingo_walk_tune(bingo_graph *g, int app, const bingo_walk_params *params, int64_t sample,
                    bingo_walk_params *chosen);

/* ---- inspection ---- */
int32_t bingo_num_nodes(const bingo_graph *g);
int bingo_stats_get(bingo_graph *g, bingo_stats *out);
/* waits for the handle's stream; BINGO_EINVAL if the last walk had an
 * invalid device-side start or length, BINGO_EDEVICE if a batch overran a pool */
int bingo_sync(bingo_graph *g);
void *bingo_stream(bingo_graph *g);      /* the handle's hipStream_t */
/* multiplicities of live arcs us[i] -> vs[i] (host arrays) */
int bingo_edge_count(bingo_graph *g, const int32_t *us, const int32_t *vs, int64_t k, uint32_t *counts);
/* two calls: first with all pointers NULL to learn the totals.  Per-node
 * arrays (nslots, mass, bsize, T, alias) are filled when non-NULL; the slot
 * and member arrays only when slot_v, slot_w, members and backpos all are. */
int bingo_export(bingo_graph *g, bingo_dump *d);

/* ---- device memory helpers for callers without a HIP runtime binding ---- */
int bingo_dev_alloc(bingo_graph *g, size_t bytes, void **d_ptr);
int bingo_dev_free(bingo_graph *g, void *d_ptr);
int bingo_memcpy_h2d(bingo_graph *g, void *d_dst, const void *src, size_t bytes);
int bingo_memcpy_d2h(bingo_graph *g, void *dst, const void *d_src, size_t bytes);
/* device-to-device copy by a plain streaming kernel (k_copy16) on the
 * handle's stream; also the known-size reference for HBM traffic counters */
int bingo_memcpy_d2d(bingo_graph *g, void *d_dst, const void *d_src, size_t bytes);
/* Measurement probe (no reference counterpart): random reads of `width`
 * bytes (8 = a member, 32 = an alias slot) at hashed positions over the
 * graph's member pool, independent of each other; *reads receives their
 * number and *ms the kernel time.  The rate is the ceiling for a walk's
 * dependent gathers on this store.  width 0: random read-modify-writes of
 * 8-byte words instead, each stored back unchanged (the pool keeps its
 * contents), the ceiling for the update kernels' scattered writes; *reads
 * receives their number.  Overwrites event slots 14 and 15. */
int bingo_gather_probe(bingo_graph *g, int width, uint64_t *reads, float *ms);
/* HIP events on the handle's stream: record into slot (0..15), elapsed ms
 * (bingo_walk_tune and bingo_gather_probe overwrite slots 14 and 15) */
int bingo_event_record(bingo_graph *g, int slot);
int bingo_event_elapsed(bingo_graph *g, int slot_a, int slot_b, float *ms);

/* ---- serving loop (continuous batching of walk queries and update batches) ----
 * A server owns a loaded graph handle (no other calls on that handle while it
 * runs) and a scheduler thread.  Requests take effect in submission order: a
 * walk sees exactly the update batches submitted before it, and an update runs
 * once every walk submitted before it has finished.  Between two updates,
 * walk requests are admitted as device arena (the HBM budget for request
 * inputs and host-bound outputs) frees up and packed in submission order into
 * persistent launches of at most max_launch_walkers walkers (a larger request
 * is cut into chunks).  Up to `streams` launches are in flight on their own HIP
 * streams: the next launch is packed from whatever arrived and enqueued while
 * the current one runs, so work that arrives while others run starts in the
 * next launch and the device never idles between launches.  A walk's paths are those it
 * would get alone (walker i of a request draws from Philox(seed; i, step,
 * attempt), whatever chunk runs it).  An update batch of up to 2^17 inserts
 * and 2^17 deletes is applied from captured graphs (bingo_update_graph_*),
 * for which creating the server commits 1 GiB of pool room (when the budget
 * allows; otherwise, and for larger batches or once that room is used up,
 * batches take bingo_insert / bingo_delete); the results are the same.
 * There is no reference counterpart: the reference runs one walk kernel at a
 * time (include/graph.cuh:93-150). */
typedef struct bingo_server bingo_server;
typedef struct {
    uint64_t arena_bytes;      /* device arena for request staging (HBM admission budget) */
    int64_t max_launch_walkers;/* walkers per launch; a larger request is split into chunks */
    int32_t blocks_per_cu;     /* persistent grid per CU; 0 = occupancy limit */
    int32_t streams;           /* launches in flight at once, 1..4; 0 = 2 */
} bingo_server_config;
typedef struct {
    uint64_t requests, walk_requests, update_requests, rejected;
    uint64_t walk_steps, launches;
    double p50_ms, p90_ms, p99_ms, max_ms, mean_ms; /* submit -> completion, walk requests */
    double busy_s;             /* time the worker spent executing launches and updates */
} bingo_server_stats;
void bingo_server_config_default(bingo_server_config *c);
int bingo_server_create(bingo_graph *g, const bingo_server_config *c, bingo_server **out);
/* starts/lens (host, may be NULL) are copied at submit.  paths is
 * [nwalkers][max_len]; paths_on_device != 0: a device pointer the kernel
 * writes directly, else a host pointer filled before completion. */
int bingo_server_submit_walk(bingo_server *s, int app, const bingo_walk_params *params, const int32_t *starts,
                             const int32_t *lens, int64_t nwalkers, int32_t *paths, int paths_on_device,
                             uint64_t *ticket);
int bingo_server_submit_update(bingo_server *s, const bingo_edge *ins, int64_t ni, const bingo_del *dels, int64_t nd,
                               uint64_t *ticket);
/* blocks until the request completed; returns its status (0 or BINGO_E*) */
int bingo_server_wait(bingo_server *s, uint64_t ticket, uint64_t *steps_out, double *latency_ms);
int bingo_server_stats_get(bingo_server *s, bingo_server_stats *out);
void bingo_server_destroy(bingo_server *s); /* completes every submitted request first */

/* ---- the reference driver: instance() (include/apps.h:4, src/apps.cu:10-42) ----
 * Reads a "u v" edge list (0-based ids, one arc per pair, as graph.cuh:22-27),
 * weights arcs bias(outdeg(v)) + 1, loads, then runs `rounds` x {insert batch,
 * delete batch, one walk per node} for utype I, D or M (M halves the batch),
 * printing the reference's lines to stderr when opts->print is set. */
typedef struct {
    int32_t rounds;     /* ref: 10 */
    int64_t batch;      /* ref BATCHSIZE: 100000 */
    uint64_t seed;
    int32_t walk_len;   /* ref LEN: 80 */
    int32_t device;
    int32_t print;
    int32_t keep_graph; /* 1: the loaded and updated graph is returned in res->graph (caller destroys) */
} bingo_instance_opts;
typedef struct {
    double walk_s;      /* "Random walk in ... s." */
    double eval_s;      /* "Evaluation time: ... s." */
    double memory_gb;   /* "Memory Consumption: ... GB." */
    uint64_t walk_steps;
    int64_t arcs;       /* live arcs at the end */
    int32_t n;
    bingo_graph *graph; /* opts->keep_graph: the graph after the rounds, else NULL */
} bingo_instance_result;
void bingo_instance_opts_default(bingo_instance_opts *o);
/* parse + weight an edge list; edges NULL: only *n_out and *m_out are set */
int bingo_read_edgelist(const char *path, int32_t *n_out, bingo_edge *edges, int64_t cap, int64_t *m_out);
int bingo_instance(int app, int utype, const char *dataset, const bingo_instance_opts *opts,
                   bingo_instance_result *res);
/* The driver's update stream on its own (host only): over a loaded arc list,
 * round after round an insert batch that duplicates random existing arcs
 * (utype I or M; ref insertGraph, graph.cuh:58-77) and a delete batch of
 * random live arcs (D or M; ref deleteGraph, graph.cuh:78-92); mixed mode
 * halves `batch`.  bingo_instance draws its rounds from this stream. */
typedef struct bingo_instgen bingo_instgen;
int bingo_instgen_create(const bingo_edge *base, int64_t m, int32_t n, int utype, int64_t batch, uint64_t seed,
                         bingo_instgen **out);
/* ins / dels: room for one round's batch each; *ni / *nd receive the counts */
int bingo_instgen_round(bingo_instgen *s, bingo_edge *ins, int64_t *ni, bingo_del *dels, int64_t *nd);
int64_t bingo_instgen_live(const bingo_instgen *s);
void bingo_instgen_destroy(bingo_instgen *s);

/* ---- workload generators (host, deterministic; ref graph.cuh:16-92) ----
 * R-MAT arcs of edges [edge_begin, edge_begin + nedges) at the given scale,
 * probabilities a, b, c (d = 1 - a - b - c); undirected emits u->v, v->u. */
int bingo_gen_rmat(int scale, int64_t edge_begin, int64_t nedges, double a, double b, double c,
                   int wkind, uint64_t seed, int undirected, bingo_edge *out);
typedef struct bingo_updgen bingo_updgen;
int bingo_updgen_create(int32_t n, int scale, const bingo_edge *base, int64_t m, int64_t edge_begin,
                        double a, double b, double c, int wkind, uint64_t seed, bingo_updgen **out);
void bingo_updgen_destroy(bingo_updgen *s);
int bingo_updgen_inserts(bingo_updgen *s, int64_t k, bingo_edge *out);
int bingo_updgen_deletes(bingo_updgen *s, int64_t k, bingo_del *out);
/* sliding window: the k oldest live arcs (load order, then inserts); a
 * generator uses either this or bingo_updgen_deletes */
int bingo_updgen_expire(bingo_updgen *s, int64_t k, bingo_del *out);
int64_t bingo_updgen_live(const bingo_updgen *s);
/* renumber the generator's slots as bingo_compact(g, min_dead_frac) renumbers
 * the store's (call both at the same point of the update stream) */
int bingo_updgen_compact(bingo_updgen *s, double min_dead_frac);

#ifdef __cplusplus
}
#endif
#endif

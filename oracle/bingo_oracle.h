/*
 * TEST INFRASTRUCTURE (oracle): a sequential CPU restatement of Bingo's
 * radix-based bias-factorized store, its batched updates and its walk
 * samplers.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library (bingo_amd/libbingo.so) never
 * links, loads or calls it.
 *
 * What each part restates (reference file:line):
 *   store layout ........ include/Bingo.cuh:6-58, 59-94 (edge slots, per-bit
 *                         member lists group[]/idx[], masses num[])
 *   build/load .......... include/operation.cuh:12-30, 33-121; Bingo.cuh:159-181
 *   insert .............. operation.cuh:122-197; Bingo.cuh:182-198
 *   delete .............. operation.cuh:198-251; Bingo.cuh:199-218
 *   alias table ......... Bingo.cuh:111-158 (replaced by an integer-exact table)
 *   sample .............. Bingo.cuh:220-267
 *   DeepWalk/node2vec/PPR operation.cuh:252-368; sampling 407-430
 *   generators .......... graph.cuh:16-92 (replaced by seeded R-MAT + update streams)
 *
 * Deliberate deviations from the reference (each decided, each tested):
 *   - all 32 bit groups of a uint32 weight are kept (ref: LOGT/MAXLOG=16 cap);
 *   - sparse/dense groups (ref RATE, api.h:25, operation.cuh:72) as a policy
 *     flag: off (every non-empty group keeps a member list, the reference's
 *     RATE > 1) or a rate r in (0, 1) in units of 2^-16.  A node's classes
 *     are decided by the batch that gives it its first slots (ref: at
 *     build): group j is dense iff count_j >= 1 and count_j 2^16 >= r len
 *     over that batch's arcs of the node (ref: dense also when empty; here
 *     an empty group keeps a list).  A dense group keeps only its count; a
 *     draw from it is rejection over the node's slots (ref Bingo.cuh:244-252)
 *     from stream blocks (wid; step, DENSE_TAG | it << 16 | attempt), four
 *     trials per block, up to ORC_DENSE_BLOCKS blocks, then the first slot
 *     with the bit after the last trial (cyclic); see node_dense_draw;
 *   - members store (slot, target) so a walk step needs no edge-record read;
 *   - alias table: 32 fixed slots, integer-exact sweep construction with a
 *     64-bit stay threshold (ref: float/EPS table, Bingo.cuh:111-158);
 *   - limits per node: fewer than 2^28 slots and integer weight below 2^59
 *     (an insert batch past either is rejected whole with ORC_ENOMEM);
 *   - randomness: Philox4x32-10 keyed (seed, walker, step, attempt);
 *   - node2vec neighbour test is exact "candidate in N(prev)" (ref hash is
 *     lossy and tests prev in N(cur), operation.cuh:313, Bingo.cuh:39-58);
 *   - node2vec folds the return arc's excess out of the rejection envelope
 *     (integer graphs; see orc_walk_thresholds_ex), same law, fewer draws;
 *   - PPR has a per-walker length cap (ref overwrites slot LEN-1);
 *   - batched delete semantics are order-free (see orc_delete);
 *   - float weights (ref wtype::Float, Bingo.cuh:65-69,176-177,193-194,
 *     216-217,254-264): the integer part goes through the 32 bit groups, the
 *     fraction is kept exactly as a 32-bit fixed-point value per slot; a draw
 *     first picks the fractional part with probability F / (S + F) from an
 *     exact 64-bit threshold, then a uniform slot accepted with probability
 *     equal to its fraction (the reference's rejection on w - floor(w)).
 */
#ifndef BINGO_ORACLE_H
#define BINGO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_GROUPS 32
#define ORC_AUX_BITS 27
#define ORC_AUX_ONE (1u << ORC_AUX_BITS)
#define ORC_MARK 0x80000000u
#define ORC_FRAC_TAG 0x40000000u   /* Philox counter word 3 of the fractional stream */
#define ORC_FRAC_ITERS 4096u       /* fractional rejection rounds per draw */
#define ORC_MAX_ATTEMPTS_FLOAT 65536 /* node2vec cap with float weights (attempt fits 16 bits) */
#define ORC_MAX_SLOTS (1u << 28)      /* slots per node (the store's 28-bit group sizes) */
#define ORC_MAX_MASS (1ull << 59)     /* integer weight per node (alias prefix sums stay in 64 bits) */
#define ORC_MIX_BIT 0x80000000u       /* Philox counter word 2 (step) of the node2vec fold stream */
#define ORC_DENSE_TAG 0x20000000u     /* Philox counter word 3 of the dense-group rejection stream */
#define ORC_DENSE_BLOCKS 1024u        /* its Philox blocks per draw (four trials each) */
#define ORC_MAX_ATTEMPTS_DENSE 65536  /* node2vec cap with dense groups (attempt fits 16 bits) */

enum { ORC_DEEPWALK = 0, ORC_NODE2VEC = 1, ORC_PPR = 2, ORC_SAMPLING = 3 };
enum { ORC_W_CONST1 = 0, ORC_W_UNIFORM16 = 1, ORC_W_LOGUNI32 = 2, ORC_W_ZIPF24 = 3 };

enum {
    ORC_OK = 0,
    ORC_EINVAL = -1,
    ORC_ENOMEM = -2,
};

/* ref Edges (type.h:24-29): w is a u32 weight, or with float weights the
 * IEEE-754 bits of a float in [2^-32, 2^32) taken to 2^-32 resolution */
typedef struct { int32_t u, v; uint32_t w; } orc_edge;
typedef struct { int32_t u, slot; } orc_del;             /* ref Deleted{u,id} (type.h:33-37) */

typedef struct orc_graph orc_graph;

/* Walk parameters after host-side quantisation (shared definition with the
 * product library's bingo_walk_params -> kernel thresholds). */
typedef struct {
    uint64_t seed;
    uint32_t thr_return;   /* node2vec: accept if aux < thr when cand == prev */
    uint32_t thr_neighbor; /* node2vec: accept if aux < thr when cand in N(prev) */
    uint32_t thr_other;    /* node2vec: otherwise */
    uint32_t thr_stop;     /* PPR: stop if aux < thr */
    int32_t max_len;       /* columns of the path matrix */
    int32_t max_attempts;  /* node2vec rejection cap per step */
    uint64_t fold_r;       /* node2vec: folded return region factor x 2^27 (0: no folding) */
} orc_walk_params;

/* graph lifetime; float_weights != 0: edge weights are float bits */
orc_graph *orc_create(int32_t n, int neighbor_index);
orc_graph *orc_create_ex(int32_t n, int neighbor_index, int float_weights);
/* dense_rate16: 0 = every non-empty group keeps a member list; otherwise
 * the sparse/dense rate in units of 2^-16 (ref RATE 0.4 ~ 26214) */
orc_graph *orc_create_rate(int32_t n, int neighbor_index, int float_weights, uint32_t dense_rate16);
void orc_destroy(orc_graph *g);
int32_t orc_num_nodes(const orc_graph *g);

/* updates: 0 on success or a negative code; on error nothing is changed
 * (ORC_ENOMEM: a node would reach ORC_MAX_SLOTS slots or ORC_MAX_MASS weight) */
int orc_insert(orc_graph *g, const orc_edge *e, int64_t k);
int orc_delete(orc_graph *g, const orc_del *d, int64_t k);

/* compaction of tombstoned slots (mirror of the product's bingo_compact):
 * nodes with dead >= 1 and dead 2^16 >= floor(min_dead_frac 2^16) nslots
 * renumber their live slots by rank; a node whose dead slots are its first
 * ones keeps its member lists' order, any other rebuilds them in slot order;
 * returns the number of nodes compacted */
int64_t orc_compact(orc_graph *g, double min_dead_frac, uint64_t *slots_freed);

/* walks: paths is [nwalkers][params->max_len] int32, -1 padded.
 * starts==NULL: walker i starts at node (i % n) + 1.  lens==NULL: max_len.
 * Returns the number of appended vertices (steps) through steps_out. */
int orc_walk(const orc_graph *g, int app, const orc_walk_params *p, const int32_t *starts,
             const int32_t *lens, int64_t nwalkers, int32_t *paths, uint64_t *steps_out,
             int nthreads);

/* the same for a subset of walkers: row i is walker wids[i] (its Philox
 * stream), starting at starts[i] with length lens[i] (starts NULL: node
 * (wids[i] % n) + 1) */
int orc_walk_ids(const orc_graph *g, int app, const orc_walk_params *p, const int32_t *starts,
                 const int32_t *lens, const int64_t *wids, int64_t nwalkers, int32_t *paths,
                 uint64_t *steps_out, int nthreads);

/* quantise float walk parameters exactly as the product does; fold = 0
 * keeps the plain envelope max(1/p, 1, 1/q) (float-weight graphs) */
void orc_walk_thresholds(double p, double q, double tp, orc_walk_params *out);
void orc_walk_thresholds_ex(double p, double q, double tp, int fold, orc_walk_params *out);
/* folded return region threshold (tests) */
uint64_t orc_fold_threshold(uint64_t S, uint64_t wret, uint64_t fold_r);

/* state export for bit-exact diffs (flat, node-major) */
typedef struct {
    uint64_t total_slots;    /* out: sum of nslots over nodes 0..n */
    uint64_t total_members;  /* out: sum of the sizes of groups with member lists */
    uint32_t *nslots;        /* [n+1] */
    uint64_t *mass;          /* [n+1] sum of live weights */
    uint32_t *bsize;         /* [(n+1)*32] */
    uint64_t *T;             /* [(n+1)*32] alias stay thresholds */
    uint8_t *alias;          /* [(n+1)*32] alias partners */
    uint32_t *slot_v;        /* [total_slots] */
    uint32_t *slot_w;        /* [total_slots] 0 = tombstone */
    uint64_t *members;       /* [total_members] (v<<32)|slot, node-major, group-major (lists only) */
    uint32_t *backpos;       /* [total_members] live slots in order, set bits of listed groups ascending */
    /* float weights only (may be NULL): */
    uint32_t *slot_f;        /* [total_slots] fraction of each slot's weight, x 2^32 */
    uint64_t *fmass;         /* [n+1] sum of the fractions, x 2^32 */
    uint64_t *fthr;          /* [n+1] fractional-part threshold against a 64-bit coin */
    uint32_t *dmask;         /* [n+1] dense groups of each node (bit j); may be NULL */
} orc_dump;
/* pass all pointers NULL to learn the totals, then again with buffers */
int orc_export(const orc_graph *g, orc_dump *d);
/* multiplicity of live arcs u->v (the node2vec neighbour test) */
uint32_t orc_edge_count(const orc_graph *g, int32_t u, int32_t v);
/* alias table of an arbitrary mass vector (tests) */
void orc_alias_build(const uint64_t mass[32], uint64_t T[32], uint8_t alias[32]);
/* fractional-part threshold of a node with integer mass S and fraction mass
 * F (units of 2^-32): floor(2^64 F / (S 2^32 + F)) (tests) */
uint64_t orc_frac_threshold(uint64_t S, uint64_t F);
/* integer part and 32-bit fraction of float weight bits; 0 if invalid */
int orc_split_float(uint32_t bits, uint32_t *ipart, uint32_t *frac);
/* raw Philox4x32-10 block (tests) */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* one sampling draw on node u with explicit random words (tests) */
int32_t orc_sample_words(const orc_graph *g, int32_t u, const uint32_t x[4]);

/* generators (workload; mirror of the product's bingo_gen_*) */
void orc_rmat_thresholds(double a, double b, double c, uint32_t thr[3]);
void orc_gen_rmat(int scale, int64_t edge_begin, int64_t nedges, const uint32_t thr[3], int wkind,
                  uint64_t seed, int undirected, orc_edge *out, int nthreads);
uint32_t orc_gen_weight(int wkind, uint64_t seed, uint64_t e);

typedef struct orc_updgen orc_updgen;
orc_updgen *orc_updgen_create(int32_t n, int scale, const orc_edge *base, int64_t m,
                              int64_t edge_begin, const uint32_t thr[3], int wkind, uint64_t seed);
void orc_updgen_destroy(orc_updgen *s);
void orc_updgen_inserts(orc_updgen *s, int64_t k, orc_edge *out);
int orc_updgen_deletes(orc_updgen *s, int64_t k, orc_del *out);
int64_t orc_updgen_live(const orc_updgen *s);

#ifdef __cplusplus
}
#endif
#endif

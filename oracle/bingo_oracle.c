/*
 * TEST INFRASTRUCTURE (oracle): CPU restatement of the bias-factorized store.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this file's library.  The product never links or calls it.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   Philox ........... Random123 KATs + rocRAND cross-check (golden fixture)
 *   buckets .......... brute-force recomputation from the live slot set,
 *                      idx∘group = identity, masses = count_j * 2^j
 *   delete order ..... convention (holes filled from the tail), pinned by
 *                      invariants and a brute-force restatement
 *   alias ............ exact rational enumeration of implied probabilities
 *                      against mass_j / sum (|rel err| <= 1e-6)
 *   sampler .......... exact law w/sum(w) by enumeration + chi-square
 *   node2vec ......... closed-form second-order law by brute force
 *   PPR .............. geometric length law with cap
 *   generators ....... R-MAT quadrant frequencies, weight ranges
 *   float weights .... exact law (1 - F/(S+F)) I/S + (F/(S+F)) f/F = w/sum(w)
 *                      from the exported state, chi-square of draws,
 *                      threshold against 128-bit division, float splitting
 *   dense groups ..... class rule restated from each node's first batch,
 *                      brute-force store, exact law of a rejection draw by
 *                      enumeration, chi-square, the trial fallback
 *   compaction ....... brute-force drop of the dead slots; a dead prefix
 *                      keeps the lists' order (slot fields lowered), other
 *                      nodes' lists equal the live slots with bit j in slot
 *                      order; the one-draw law moves by less than 2^-26
 */
#include "bingo_oracle.h"
#include "philox.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

typedef struct { uint32_t key, count; uint64_t wsum; } hent;  /* target, arcs, their weight */

typedef struct {
    uint32_t nslots, slot_cap;
    uint32_t *v, *w, *boff;          /* per slot: target, weight (0 = tombstone), back offset */
    uint32_t *back;                  /* bucket positions of each slot's set bits */
    uint32_t back_used, back_cap;
    uint64_t *mem[ORC_GROUPS];       /* member lists: (v << 32) | slot */
    uint32_t size[ORC_GROUPS], cap[ORC_GROUPS];
    hent *hash;                      /* multiset of live targets (node2vec) */
    uint32_t hcap, hused;
    uint64_t mass;
    uint64_t T[ORC_GROUPS];
    uint8_t alias[ORC_GROUPS];
    uint32_t *fr;                    /* float weights: per slot fraction x 2^32 */
    uint64_t fmass, fthr;            /* sum of fractions; fractional-part threshold */
    uint32_t dmask;                  /* dense groups: count only, no member list */
} node_t;

struct orc_graph {
    int32_t n;
    int nbr;
    int fw;                          /* weights are float bits */
    uint32_t rate16;                 /* sparse/dense rate x 2^16 (0: every group listed) */
    node_t *nodes;
};

static void *xrealloc(void *p, size_t bytes)
{
    void *q = realloc(p, bytes ? bytes : 1);
    if (!q) abort();
    return q;
}

static uint32_t grow_cap(uint32_t need, uint32_t minc)
{
    uint64_t c = minc;
    while (c < need) c <<= 1;
    return (uint32_t)(c > 0xFFFFFFFFu ? 0xFFFFFFFFu : c);
}

static inline int popc32(uint32_t x) { return __builtin_popcount(x); }
static inline int hibit32(uint32_t x) { return 31 - __builtin_clz(x); }

/* position of slot s inside bucket j (s must have bit j set, j listed):
 * a slot's back entries are those of its listed set bits, ascending */
static inline uint32_t *backref(node_t *x, uint32_t s, int j)
{
    return &x->back[x->boff[s] + popc32(x->w[s] & ~x->dmask & ((1u << j) - 1u))];
}

/* ---------------- neighbour multiset (node2vec membership) ---------------- */

static inline uint32_t hslot(uint32_t key, uint32_t cap)
{
    return (uint32_t)(((uint64_t)(key * 0x9E3779B1u) * cap) >> 32);
}

static void hash_put(node_t *x, uint32_t key, int32_t delta, int64_t wdelta);

static void hash_rehash(node_t *x, uint32_t newcap)
{
    hent *old = x->hash;
    uint32_t oldcap = x->hcap;
    x->hash = (hent *)calloc(newcap, sizeof(hent));
    if (!x->hash) abort();
    x->hcap = newcap;
    x->hused = 0;
    for (uint32_t i = 0; i < oldcap; ++i)
        if (old[i].key && old[i].count) hash_put(x, old[i].key, (int32_t)old[i].count, (int64_t)old[i].wsum);
    free(old);
}

/* key's arc count += delta and weight sum += wdelta (integer parts) */
static void hash_put(node_t *x, uint32_t key, int32_t delta, int64_t wdelta)
{
    if (delta > 0 && (uint64_t)(x->hused + 1) * 2 > x->hcap)
        hash_rehash(x, x->hcap ? x->hcap * 2 : 16);
    uint32_t mask = x->hcap - 1, i = hslot(key, x->hcap);
    for (;;) {
        if (x->hash[i].key == key) {
            x->hash[i].count += (uint32_t)delta;
            x->hash[i].wsum += (uint64_t)wdelta;
            return;
        }
        if (x->hash[i].key == 0) {
            if (delta < 0) abort(); /* removing an absent key: store corrupted */
            x->hash[i].key = key;
            x->hash[i].count = (uint32_t)delta;
            x->hash[i].wsum = (uint64_t)wdelta;
            x->hused++;
            return;
        }
        i = (i + 1) & mask;
    }
}

static const hent *hash_find(const node_t *x, uint32_t key)
{
    if (!x->hcap) return NULL;
    uint32_t mask = x->hcap - 1, i = hslot(key, x->hcap);
    for (;;) {
        if (x->hash[i].key == key) return &x->hash[i];
        if (x->hash[i].key == 0) return NULL;
        i = (i + 1) & mask;
    }
}
static uint32_t hash_get(const node_t *x, uint32_t key)
{
    const hent *h = hash_find(x, key);
    return h ? h->count : 0;
}
static uint64_t hash_wsum(const node_t *x, uint32_t key)
{
    const hent *h = hash_find(x, key);
    return h ? h->wsum : 0;
}

/* ---------------- alias table ---------------- */

/* Integer-exact alias table over 32 fixed slots (slot j <-> bit group j),
 * built by a sweep.  s_j = 32 * M_j, slot capacity S = sum M_j.  Light =
 * {s_j < S}, heavy = the rest, each in ascending index order.  One heavy is
 * current, with residual r (initially its s).  Each light i in turn takes
 * tau_i = s_i and alias_i = the current heavy h, whose residual drops by the
 * light's deficit S - s_i; whenever r < S the current heavy turns light on
 * the spot: tau_h = r, alias_h = the next heavy, whose residual drops by
 * S - r and which becomes current.  Heavies never turned light keep
 * tau = S, alias = self.  The stay threshold is the 64-bit fraction
 * T_j = floor(tau_j * 2^64 / S) (all-ones when tau_j == S), used as
 * "stay iff coin64 < T_j".  Replaces Bingo.cuh:111-158 (float, EPS=1e-3). */
void orc_alias_build(const uint64_t M[32], uint64_t T[32], uint8_t alias[32])
{
    u128 S = 0, s[32], tau[32];
    int light[32], heavy[32], nl = 0, nh = 0;
    for (int j = 0; j < 32; ++j) S += M[j];
    for (int j = 0; j < 32; ++j) {
        s[j] = (u128)M[j] * 32u;
        tau[j] = S;
        alias[j] = (uint8_t)j;
        if (s[j] < S) light[nl++] = j; else heavy[nh++] = j;
    }
    int k = 0;
    u128 r = nh ? s[heavy[0]] : 0;
    for (int t = 0; t < nl && k < nh; ++t) {
        const int i = light[t];
        tau[i] = s[i];
        alias[i] = (uint8_t)heavy[k];
        r -= S - s[i];
        while (r < S && k + 1 < nh) {
            const int h = heavy[k];
            tau[h] = r;
            alias[h] = (uint8_t)heavy[k + 1];
            r = s[heavy[++k]] - (S - tau[h]);
        }
    }
    for (int j = 0; j < 32; ++j)
        T[j] = (tau[j] == S) ? UINT64_MAX : (uint64_t)((tau[j] << 64) / S);
}

/* Fractional part of a draw: the 64-bit fraction floor(F 2^64 / D) of
 * F / D, D = S 2^32 + F, F in units of 2^-32 (S < 2^60, F < 2^60: D < 2^93,
 * F 2^64 < 2^124).  S = 0: always the fractional part. */
uint64_t orc_frac_threshold(uint64_t S, uint64_t F)
{
    if (!F) return 0;
    if (!S) return UINT64_MAX;
    u128 D = ((u128)S << 32) + F;
    return (uint64_t)(((u128)F << 64) / D);
}

/* Float weight bits -> integer part and fraction x 2^32 (floor).  Valid:
 * finite, 0 < w < 2^32, and at least 2^-32 after the floor. */
int orc_split_float(uint32_t bits, uint32_t *ipart, uint32_t *frac)
{
    float x;
    memcpy(&x, &bits, 4);
    double d = (double)x;
    if (!(d > 0.0) || !(d < 4294967296.0)) return 0;
    double fl = floor(d);
    uint32_t ip = (uint32_t)fl, fr = (uint32_t)floor((d - fl) * 4294967296.0);
    if (!ip && !fr) return 0;
    *ipart = ip;
    *frac = fr;
    return 1;
}

static void node_alias(node_t *x)
{
    uint64_t M[32], S = 0;
    for (int j = 0; j < 32; ++j) { M[j] = (uint64_t)x->size[j] << j; S += M[j]; }
    x->mass = S;
    orc_alias_build(M, x->T, x->alias);
    x->fthr = orc_frac_threshold(S, x->fmass);
}

/* ---------------- graph ---------------- */

orc_graph *orc_create(int32_t n, int neighbor_index) { return orc_create_rate(n, neighbor_index, 0, 0); }

orc_graph *orc_create_ex(int32_t n, int neighbor_index, int float_weights)
{
    return orc_create_rate(n, neighbor_index, float_weights, 0);
}

orc_graph *orc_create_rate(int32_t n, int neighbor_index, int float_weights, uint32_t dense_rate16)
{
    if (n < 1) return NULL;
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    g->n = n;
    g->nbr = neighbor_index;
    g->fw = float_weights != 0;
    g->rate16 = dense_rate16;
    g->nodes = (node_t *)calloc((size_t)n + 1, sizeof(node_t));
    if (!g->nodes) abort();
    for (int32_t u = 0; u <= n; ++u) node_alias(&g->nodes[u]);
    return g;
}

void orc_destroy(orc_graph *g)
{
    if (!g) return;
    for (int32_t u = 0; u <= g->n; ++u) {
        node_t *x = &g->nodes[u];
        free(x->v); free(x->w); free(x->boff); free(x->back); free(x->hash); free(x->fr);
        for (int j = 0; j < 32; ++j) free(x->mem[j]);
    }
    free(g->nodes);
    free(g);
}

int32_t orc_num_nodes(const orc_graph *g) { return g->n; }

/* Append one arc as a new slot: slot = nslots++, and for each set bit j in
 * ascending order the pair (v, slot) is appended to bucket j
 * (Bingo.cuh:182-198 insert; Bingo.cuh:159-181 build uses the same order);
 * a dense group only counts the arc (no member, no back entry). */
static void node_append(node_t *x, uint32_t v, uint32_t w, uint32_t f, int fw, int nbr)
{
    uint32_t s = x->nslots;
    if (s == x->slot_cap) {
        x->slot_cap = grow_cap(s + 1, 16);
        x->v = (uint32_t *)xrealloc(x->v, (size_t)x->slot_cap * 4);
        x->w = (uint32_t *)xrealloc(x->w, (size_t)x->slot_cap * 4);
        x->boff = (uint32_t *)xrealloc(x->boff, (size_t)x->slot_cap * 4);
        if (fw) x->fr = (uint32_t *)xrealloc(x->fr, (size_t)x->slot_cap * 4);
    }
    x->v[s] = v;
    x->w[s] = w;
    if (fw) {
        x->fr[s] = f;
        x->fmass += f;
    }
    x->boff[s] = x->back_used;
    uint32_t pc = (uint32_t)popc32(w & ~x->dmask);
    if (x->back_used + pc > x->back_cap) {
        x->back_cap = grow_cap(x->back_used + pc, 16);
        x->back = (uint32_t *)xrealloc(x->back, (size_t)x->back_cap * 4);
    }
    for (uint32_t bits = w; bits; bits &= bits - 1) {
        int j = __builtin_ctz(bits);
        if (x->dmask >> j & 1u) {
            x->size[j]++;
            continue;
        }
        if (x->size[j] == x->cap[j]) {
            x->cap[j] = grow_cap(x->size[j] + 1, 4);
            x->mem[j] = (uint64_t *)xrealloc(x->mem[j], (size_t)x->cap[j] * 8);
        }
        uint32_t pos = x->size[j]++;
        x->mem[j][pos] = ((uint64_t)v << 32) | s;
        x->back[x->back_used++] = pos;
    }
    x->nslots = s + 1;
    if (nbr) hash_put(x, v, 1, (int64_t)w);
}

/* Stable grouping of k items by node id: order[] lists item indices grouped
 * by node, batch order kept inside a node; runs[] holds run starts. */
typedef struct { int64_t *order; int64_t *runs; int32_t *run_node; int64_t nruns; } grouping;

static grouping group_by_node(int32_t n, int64_t k, const int32_t *u_of, size_t stride)
{
    grouping G;
    int64_t *cnt = (int64_t *)calloc((size_t)n + 2, sizeof(int64_t));
    for (int64_t i = 0; i < k; ++i) cnt[*(const int32_t *)((const char *)u_of + i * stride) + 1]++;
    int64_t nr = 0;
    for (int32_t u = 0; u <= n; ++u) if (cnt[u + 1]) nr++;
    for (int32_t u = 0; u <= n; ++u) cnt[u + 1] += cnt[u];
    G.order = (int64_t *)malloc((size_t)(k ? k : 1) * sizeof(int64_t));
    G.runs = (int64_t *)malloc((size_t)(nr + 1) * sizeof(int64_t));
    G.run_node = (int32_t *)malloc((size_t)(nr ? nr : 1) * sizeof(int32_t));
    int64_t *pos = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    memcpy(pos, cnt, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < k; ++i) {
        int32_t u = *(const int32_t *)((const char *)u_of + i * stride);
        G.order[pos[u]++] = i;
    }
    nr = 0;
    for (int32_t u = 0; u <= n; ++u)
        if (cnt[u + 1] > cnt[u]) { G.runs[nr] = cnt[u]; G.run_node[nr] = u; nr++; }
    G.runs[nr] = k;
    G.nruns = nr;
    free(pos);
    free(cnt);
    return G;
}

static void grouping_free(grouping *G) { free(G->order); free(G->runs); free(G->run_node); }

/* Batched insert (operation.cuh:144-197 host, 122-143 kernel).  Arcs are
 * grouped by source node keeping batch order (the reference's stable_sort,
 * operation.cuh:147); each node then appends its arcs in that order and
 * rebuilds its alias table once. */
int orc_insert(orc_graph *g, const orc_edge *e, int64_t k)
{
    if (k < 0 || (k > 0 && !e)) return ORC_EINVAL;
    for (int64_t i = 0; i < k; ++i) {
        uint32_t ip, fr;
        if (e[i].u < 1 || e[i].u > g->n || e[i].v < 1 || e[i].v > g->n) return ORC_EINVAL;
        if (g->fw ? !orc_split_float(e[i].w, &ip, &fr) : e[i].w == 0) return ORC_EINVAL;
    }
    if (!k) return ORC_OK;
    grouping G = group_by_node(g->n, k, &e[0].u, sizeof(orc_edge));
    for (int64_t r = 0; r < G.nruns; ++r) {
        node_t *x = &g->nodes[G.run_node[r]];
        uint64_t add = 0;  /* integer weight the run adds */
        for (int64_t t = G.runs[r]; t < G.runs[r + 1]; ++t) {
            uint32_t w = e[G.order[t]].w, f;
            if (g->fw) orc_split_float(e[G.order[t]].w, &w, &f);
            add += w;
        }
        if ((uint64_t)x->nslots + (uint64_t)(G.runs[r + 1] - G.runs[r]) >= ORC_MAX_SLOTS ||
            x->mass + add >= ORC_MAX_MASS) {
            grouping_free(&G);
            return ORC_ENOMEM;
        }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < G.nruns; ++r) {
        node_t *x = &g->nodes[G.run_node[r]];
        if (!x->nslots) {
            /* the node's first slots: its group classes, from this run
             * (ref build, operation.cuh:57-72): dense iff count_j >= 1 and
             * count_j 2^16 >= rate16 len */
            uint32_t cnt[32] = {0};
            const uint64_t len = (uint64_t)(G.runs[r + 1] - G.runs[r]);
            for (int64_t t = G.runs[r]; t < G.runs[r + 1]; ++t) {
                uint32_t w = e[G.order[t]].w, f;
                if (g->fw) orc_split_float(e[G.order[t]].w, &w, &f);
                for (uint32_t b = w; b; b &= b - 1) cnt[__builtin_ctz(b)]++;
            }
            x->dmask = 0;
            for (int j = 0; g->rate16 && j < 32; ++j)
                if (cnt[j] && ((uint64_t)cnt[j] << 16) >= (uint64_t)g->rate16 * len) x->dmask |= 1u << j;
        }
        for (int64_t t = G.runs[r]; t < G.runs[r + 1]; ++t) {
            const orc_edge *a = &e[G.order[t]];
            uint32_t w = a->w, f = 0;
            if (g->fw) orc_split_float(a->w, &w, &f);
            node_append(x, (uint32_t)a->v, w, f, g->fw, g->nbr);
        }
        node_alias(x);
    }
    grouping_free(&G);
    return ORC_OK;
}

/* Deletes of one node, slots ascending and distinct.
 *
 * Semantics per bit group j (order-free within a batch): let D_j be the
 * deleted slots with bit j, S the bucket size and S' = S - |D_j|.  Deleted
 * members at positions < S' leave holes; the survivors at positions >= S'
 * fill them: the k-th hole, holes taken in ascending slot order, receives
 * the k-th survivor in ascending position.  The moved member's back index
 * is rewritten.  This is the batched form of the reference's per-delete
 * swap-with-last (Bingo.cuh:199-218): each hole is filled by one member from
 * the end of the list, and the result does not depend on the order of the
 * deletes inside the batch (the reference's unstable std::sort leaves that
 * order unspecified, operation.cuh:218).  The slot itself is tombstoned
 * (w = 0, v kept), as operation.cuh:209-210 does. */
static void node_delete_run(node_t *x, const uint32_t *slots, int64_t cnt, int nbr)
{
    for (int j = 0; j < 32; ++j) {
        uint32_t bit = 1u << j, kj = 0;
        for (int64_t t = 0; t < cnt; ++t) kj += (x->w[slots[t]] & bit) != 0;
        if (!kj) continue;
        if (x->dmask & bit) {  /* a dense group only counts */
            x->size[j] -= kj;
            continue;
        }
        uint32_t S = x->size[j], S2 = S - kj;
        uint64_t *m = x->mem[j];
        for (int64_t t = 0; t < cnt; ++t)
            if (x->w[slots[t]] & bit) m[*backref(x, slots[t], j)] |= ORC_MARK;
        uint32_t tail = S2;
        for (int64_t t = 0; t < cnt; ++t) {
            if (!(x->w[slots[t]] & bit)) continue;
            uint32_t pos = *backref(x, slots[t], j);
            if (pos >= S2) continue;
            while (m[tail] & ORC_MARK) tail++;
            uint64_t mv = m[tail++];
            m[pos] = mv;
            *backref(x, (uint32_t)mv, j) = pos;
        }
        x->size[j] = S2;
    }
    for (int64_t t = 0; t < cnt; ++t) {
        if (nbr) hash_put(x, x->v[slots[t]], -1, -(int64_t)x->w[slots[t]]);
        x->w[slots[t]] = 0;
        if (x->fr) {
            x->fmass -= x->fr[slots[t]];
            x->fr[slots[t]] = 0;
        }
    }
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* Batched delete (operation.cuh:215-251 host, 198-214 kernel).  Every entry
 * must name a live slot; duplicates are rejected.  Either the whole batch is
 * applied or nothing is. */
int orc_delete(orc_graph *g, const orc_del *d, int64_t k)
{
    if (k < 0 || (k > 0 && !d)) return ORC_EINVAL;
    for (int64_t i = 0; i < k; ++i) {
        if (d[i].u < 1 || d[i].u > g->n || d[i].slot < 0) return ORC_EINVAL;
        node_t *x = &g->nodes[d[i].u];
        if ((uint32_t)d[i].slot >= x->nslots) return ORC_EINVAL;
        if (x->w[d[i].slot] == 0 && !(g->fw && x->fr[d[i].slot])) return ORC_EINVAL;
    }
    if (!k) return ORC_OK;
    uint64_t *key = (uint64_t *)malloc((size_t)k * 8);
    for (int64_t i = 0; i < k; ++i) key[i] = ((uint64_t)(uint32_t)d[i].u << 32) | (uint32_t)d[i].slot;
    qsort(key, (size_t)k, 8, cmp_u64);
    for (int64_t i = 1; i < k; ++i)
        if (key[i] == key[i - 1]) { free(key); return ORC_EINVAL; }
    int64_t nr = 0;
    int64_t *runs = (int64_t *)malloc((size_t)(k + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < k; ++i)
        if (!i || (key[i] >> 32) != (key[i - 1] >> 32)) runs[nr++] = i;
    runs[nr] = k;
    uint32_t *slots = (uint32_t *)malloc((size_t)k * 4);
    for (int64_t i = 0; i < k; ++i) slots[i] = (uint32_t)key[i];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < nr; ++r) {
        node_t *x = &g->nodes[key[runs[r]] >> 32];
        node_delete_run(x, slots + runs[r], runs[r + 1] - runs[r], g->nbr);
        node_alias(x);
    }
    free(slots);
    free(runs);
    free(key);
    return ORC_OK;
}

/* Compaction of tombstoned slots (the reference never reclaims them,
 * operation.cuh:209-210).  Node u is compacted when dead = nslots - live >= 1
 * and dead * 2^16 >= floor(min_dead_frac * 2^16) * nslots: its live slots
 * keep their order and take the numbers 0..live-1.  When its dead slots are
 * exactly its first `dead` (a sliding window's expiries), the live slots
 * move down by dead with their back entries, and the member lists keep their
 * order, each member's slot field lowered by dead.  Otherwise each of its
 * listed groups' member lists is rebuilt in slot order (the k-th member of
 * group j is the k-th live slot with bit j), with the back entries to match.
 * Masses and the alias table do not change.  Returns the number of nodes
 * compacted (ORC_EINVAL for a fraction outside [0, 1]). */
int64_t orc_compact(orc_graph *g, double min_dead_frac, uint64_t *slots_freed)
{
    if (!(min_dead_frac >= 0 && min_dead_frac <= 1)) return ORC_EINVAL;
    const uint64_t thr16 = (uint64_t)floor(min_dead_frac * 65536.0);
    int64_t nodes = 0;
    uint64_t freed = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nodes, freed)
    for (int32_t u = 1; u <= g->n; ++u) {
        node_t *x = &g->nodes[u];
        uint32_t live = 0, lead = 0;
        for (uint32_t s = 0; s < x->nslots; ++s) live += (x->w[s] || (x->fr && x->fr[s])) ? 1u : 0u;
        while (lead < x->nslots && !x->w[lead] && !(x->fr && x->fr[lead])) ++lead;
        const uint32_t dead = x->nslots - live;
        if (!dead || (uint64_t)dead * 65536u < thr16 * x->nslots) continue;
        if (lead == dead) {  /* a dead prefix */
            const uint32_t D = lead < x->nslots ? x->boff[lead] : x->back_used;
            for (uint32_t s = lead; s < x->nslots; ++s) {
                x->v[s - lead] = x->v[s];
                x->w[s - lead] = x->w[s];
                if (x->fr) x->fr[s - lead] = x->fr[s];
                x->boff[s - lead] = x->boff[s] - D;
            }
            if (x->back_used > D) memmove(x->back, x->back + D, (size_t)(x->back_used - D) * sizeof(uint32_t));
            x->back_used -= D;
            for (int j = 0; j < ORC_GROUPS; ++j)
                if (!((x->dmask >> j) & 1u))
                    for (uint32_t k = 0; k < x->size[j]; ++k) x->mem[j][k] -= lead;
            x->nslots -= lead;
        } else {
            uint32_t ns = 0, nb = 0, pos[32] = {0};
            for (uint32_t s = 0; s < x->nslots; ++s) {
                const uint32_t w = x->w[s];
                if (!w && !(x->fr && x->fr[s])) continue;
                x->v[ns] = x->v[s];
                x->w[ns] = w;
                if (x->fr) x->fr[ns] = x->fr[s];
                x->boff[ns] = nb;
                for (uint32_t bits = w & ~x->dmask; bits; bits &= bits - 1) {  /* ns <= s, nb <= old offset */
                    const int j = __builtin_ctz(bits);
                    x->mem[j][pos[j]] = ((uint64_t)x->v[ns] << 32) | ns;
                    x->back[nb++] = pos[j]++;
                }
                ++ns;
            }
            x->nslots = ns;
            x->back_used = nb;
        }
        nodes += 1;
        freed += dead;
    }
    if (slots_freed) *slots_freed = freed;
    return nodes;
}

uint32_t orc_edge_count(const orc_graph *g, int32_t u, int32_t v)
{
    if (u < 1 || u > g->n || v < 1 || v > g->n) return 0;
    const node_t *x = &g->nodes[u];
    if (g->nbr) return hash_get(x, (uint32_t)v);
    uint32_t c = 0;
    for (uint32_t s = 0; s < x->nslots; ++s) c += ((x->w[s] || (g->fw && x->fr[s])) && x->v[s] == (uint32_t)v);
    return c;
}

/* ---------------- sampling ---------------- */

/* A draw from dense group g of node x (no member list): rejection over the
 * node's slots, as the reference samples a group classed dense
 * (Bingo.cuh:244-252: a uniform slot until its weight has bit g).  Trials
 * come four per Philox block (wid; step, DENSE_TAG | it << 16 | attempt),
 * words 0..3 in order, slot floor(word * nslots / 2^32); tombstones have no
 * bits.  After ORC_DENSE_BLOCKS blocks without a hit (probability below
 * (1 - count/nslots)^4096), the first slot with bit g after the last
 * trial's slot, cyclically. */
static int32_t node_dense_draw(const node_t *x, int g, const uint32_t key[2], uint64_t wid, uint32_t step,
                               uint32_t att)
{
    uint32_t b[4], ctr[4] = {(uint32_t)wid, (uint32_t)(wid >> 32), step, 0};
    uint32_t sl = 0;
    for (uint32_t it = 0; it < ORC_DENSE_BLOCKS; ++it) {
        ctr[3] = ORC_DENSE_TAG | (it << 16) | att;
        orc_philox4x32_10(ctr, key, b);
        for (int q = 0; q < 4; ++q) {
            sl = (uint32_t)(((uint64_t)b[q] * x->nslots) >> 32);
            if (x->w[sl] >> g & 1u) return (int32_t)x->v[sl];
        }
    }
    for (uint32_t t = 1; t <= x->nslots; ++t) {
        uint32_t s2 = sl + t;
        if (s2 >= x->nslots) s2 -= x->nslots;
        if (x->w[s2] >> g & 1u) return (int32_t)x->v[s2];
    }
    return -1;
}

/* One draw (Bingo.cuh:220-267): slot i = top 5 bits of x0, 64-bit coin
 * (x1:x2) against T_i picks group i or its alias, then member
 * floor(x3 * size / 2^32) of that group, or for a dense group the rejection
 * draw of node_dense_draw.  Returns -1 on a node without live arcs (the
 * reference's aliasTableSZ == 0 case). */
static inline int32_t node_sample(const node_t *x, const uint32_t r[4], const uint32_t key[2], uint64_t wid,
                                  uint32_t step, uint32_t att)
{
    uint32_t i = r[0] >> 27;
    uint64_t c = ((uint64_t)r[1] << 32) | r[2];
    uint32_t gsel = (c < x->T[i]) ? i : x->alias[i];
    uint32_t sz = x->size[gsel];
    if (!sz) return -1;
    if (x->dmask >> gsel & 1u) return node_dense_draw(x, (int)gsel, key, wid, step, att);
    uint32_t pos = (uint32_t)(((uint64_t)r[3] * sz) >> 32);
    return (int32_t)(x->mem[gsel][pos] >> 32);
}

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    orc_philox4x32_10(ctr, key, out);
}

/* One draw with float weights: with probability F / (S + F) (coin64 of the
 * fractional stream below the threshold, or no integer mass) the fractional
 * part: a uniform slot, accepted with probability fraction / 2^32, up to
 * ORC_FRAC_ITERS rounds from stream blocks (wid; step, FRAC_TAG | it << 16 |
 * attempt); otherwise (or when every round rejects) the integer draw. */
static int32_t node_draw(const orc_graph *g, const node_t *x, const uint32_t r[4], const uint32_t key[2],
                         uint64_t wid, uint32_t step, uint32_t att)
{
    if (g->fw && x->fthr) {
        uint32_t b[4], ctr[4] = {(uint32_t)wid, (uint32_t)(wid >> 32), step, ORC_FRAC_TAG | att};
        orc_philox4x32_10(ctr, key, b);
        uint64_t coin = ((uint64_t)b[0] << 32) | b[1];
        if (!x->mass || coin < x->fthr) {
            for (uint32_t it = 0; it < ORC_FRAC_ITERS; ++it) {
                if (it) {
                    ctr[3] = ORC_FRAC_TAG | (it << 16) | att;
                    orc_philox4x32_10(ctr, key, b);
                }
                uint32_t sl = (uint32_t)(((uint64_t)b[2] * x->nslots) >> 32);
                if (b[3] < x->fr[sl]) return (int32_t)x->v[sl];
            }
        }
    }
    return node_sample(x, r, key, wid, step, att);
}

/* dense groups draw from the stream of walker 0, step 0, attempt 0, key 0 */
int32_t orc_sample_words(const orc_graph *g, int32_t u, const uint32_t x[4])
{
    static const uint32_t key0[2] = {0, 0};
    if (u < 1 || u > g->n) return -1;
    return node_sample(&g->nodes[u], x, key0, 0, 0, 0);
}

static uint32_t thr27(double prob)
{
    if (!(prob > 0)) return 0;
    if (prob >= 1.0) return ORC_AUX_ONE;
    return (uint32_t)floor(prob * (double)ORC_AUX_ONE);
}

/* node2vec bias alpha = 1/p (return), 1 (distance 1), 1/q (distance 2)
 * (operation.cuh:291,302-325; lim there is max(1/p, 1/q)).  A proposal
 * x ~ w is accepted with alpha(x) / env, env = max(1, 1/q).  When 1/p > env
 * the return arc's excess is folded out of the envelope (outlier folding,
 * Yang et al., KnightKing, SOSP'19): the proposal accepts a return
 * candidate always, and each attempt first takes the return arc outright
 * with the probability of an extra region of mass w_ret (1/p - env) / env
 * next to the proposal mass S (fold_r = that factor in units of 2^-27).
 * Without folding (fold = 0, 1/p <= env, or a factor (1/p - env) / env of
 * 2^32 or more) the envelope is max(1/p, 1, 1/q).
 * PPR stops with probability tp after each step (operation.cuh:359). */
void orc_walk_thresholds_ex(double p, double q, double tp, int fold, orc_walk_params *out)
{
    double ip = 1.0 / p, iq = 1.0 / q, env = iq > 1.0 ? iq : 1.0;
    if (fold && ip > env && (ip - env) / env < 4294967296.0) {
        out->thr_return = ORC_AUX_ONE;
        out->thr_neighbor = thr27(1.0 / env);
        out->thr_other = thr27(iq / env);
        out->fold_r = (uint64_t)floor((ip - env) / env * (double)ORC_AUX_ONE);
    } else {
        double lim = ip > env ? ip : env;
        out->thr_return = thr27(ip / lim);
        out->thr_neighbor = thr27(1.0 / lim);
        out->thr_other = thr27(iq / lim);
        out->fold_r = 0;
    }
    out->thr_stop = thr27(tp);
}

void orc_walk_thresholds(double p, double q, double tp, orc_walk_params *out)
{
    orc_walk_thresholds_ex(p, q, tp, 1, out);
}

/* Probability of the folded return region as a 64-bit threshold:
 * floor(2^64 num / den), num = w_ret fold_r, den = S 2^27 + num, both cut to
 * their top 60 bits by the same shift (the ratio moves by at most 2^-59).  0 when
 * num = 0; all-ones when num >= den after the cut. */
uint64_t orc_fold_threshold(uint64_t S, uint64_t wret, uint64_t fold_r)
{
    u128 num = (u128)wret * fold_r, den = ((u128)S << 27) + num;
    if (!num) return 0;
    int bits = 0;
    for (u128 t = den; t; t >>= 1) ++bits;
    if (bits > 60) {
        num >>= bits - 60;
        den >>= bits - 60;
    }
    if (num >= den) return UINT64_MAX;
    return (uint64_t)((num << 64) / den);
}

static uint64_t walk_one(const orc_graph *g, int app, const orc_walk_params *p, uint64_t wid,
                         int32_t start, int32_t len, int32_t *row)
{
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    uint64_t steps = 0;
    int32_t cur = start, prev = -1, j = 1;
    row[0] = start;
    for (; j < len; ++j) {
        const node_t *x = &g->nodes[cur];
        uint32_t r[4], ctr[4] = {(uint32_t)wid, (uint32_t)(wid >> 32), (uint32_t)j, 0};
        int32_t nx;
        if (app == ORC_NODE2VEC && j >= 2) {
            const node_t *px = &g->nodes[prev];
            /* folded return region: per attempt a 64-bit coin from stream
             * (wid; step | 2^31, attempt) below tout takes the return arc */
            const uint64_t tout =
                p->fold_r ? orc_fold_threshold(x->mass, g->nbr ? hash_wsum(x, (uint32_t)prev) : 0, p->fold_r) : 0;
            for (uint32_t a = 0;; ++a) {
                ctr[3] = a;
                orc_philox4x32_10(ctr, key, r);
                nx = node_draw(g, x, r, key, wid, (uint32_t)j, a);
                if (tout) {
                    uint32_t m[4], mc[4] = {(uint32_t)wid, (uint32_t)(wid >> 32), (uint32_t)j | ORC_MIX_BIT, a};
                    orc_philox4x32_10(mc, key, m);
                    if ((((uint64_t)m[0] << 32) | m[1]) < tout) {
                        nx = prev;
                        break;
                    }
                }
                if (nx < 0) break;
                if ((int64_t)a + 1 >= p->max_attempts) break;
                uint32_t aux = r[0] & (ORC_AUX_ONE - 1u), thr;
                if (nx == prev) thr = p->thr_return;
                else thr = (g->nbr ? hash_get(px, (uint32_t)nx) : orc_edge_count(g, prev, nx))
                               ? p->thr_neighbor : p->thr_other;
                if (aux < thr) break;
            }
        } else {
            orc_philox4x32_10(ctr, key, r);
            nx = node_draw(g, x, r, key, wid, (uint32_t)j, 0);
        }
        if (nx < 0) break;
        row[j] = nx;
        steps++;
        prev = cur;
        cur = nx;
        if (app == ORC_PPR && (r[0] & (ORC_AUX_ONE - 1u)) < p->thr_stop) { ++j; break; }
    }
    for (; j < p->max_len; ++j) row[j] = -1;
    return steps;
}

int orc_walk(const orc_graph *g, int app, const orc_walk_params *p, const int32_t *starts,
             const int32_t *lens, int64_t nwalkers, int32_t *paths, uint64_t *steps_out,
             int nthreads)
{
    return orc_walk_ids(g, app, p, starts, lens, NULL, nwalkers, paths, steps_out, nthreads);
}

int orc_walk_ids(const orc_graph *g, int app, const orc_walk_params *p, const int32_t *starts,
                 const int32_t *lens, const int64_t *wids, int64_t nwalkers, int32_t *paths,
                 uint64_t *steps_out, int nthreads)
{
    if (!p || p->max_len < 1 || nwalkers < 0 || (nwalkers && !paths) || app < 0 || app > 3)
        return ORC_EINVAL;
    if (app == ORC_NODE2VEC && p->max_attempts < 1) return ORC_EINVAL;
    if (g->fw && p->max_attempts > ORC_MAX_ATTEMPTS_FLOAT) return ORC_EINVAL;
    if (g->rate16 && p->max_attempts > ORC_MAX_ATTEMPTS_DENSE) return ORC_EINVAL;
    if (app == ORC_NODE2VEC && p->fold_r && (g->fw || !g->nbr)) return ORC_EINVAL;  /* folding needs weight sums */
    for (int64_t i = 0; starts && i < nwalkers; ++i)
        if (starts[i] < 1 || starts[i] > g->n) return ORC_EINVAL;
    for (int64_t i = 0; lens && i < nwalkers; ++i)
        if (lens[i] < 1 || lens[i] > p->max_len) return ORC_EINVAL;
    uint64_t total = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total) num_threads(nthreads)
    for (int64_t i = 0; i < nwalkers; ++i) {
        const int64_t w = wids ? wids[i] : i;
        int32_t s = starts ? starts[i] : (int32_t)(w % g->n) + 1;
        int32_t L = lens ? lens[i] : p->max_len;
        if (app == ORC_SAMPLING && L > 2) L = 2;
        total += walk_one(g, app, p, (uint64_t)w, s, L, paths + i * (int64_t)p->max_len);
    }
    if (steps_out) *steps_out = total;
    return ORC_OK;
}

/* ---------------- export ---------------- */

int orc_export(const orc_graph *g, orc_dump *d)
{
    uint64_t ts = 0, tm = 0;
    for (int32_t u = 0; u <= g->n; ++u) {
        const node_t *x = &g->nodes[u];
        ts += x->nslots;
        for (int j = 0; j < 32; ++j) tm += (x->dmask >> j & 1u) ? 0 : x->size[j];
    }
    d->total_slots = ts;
    d->total_members = tm;
    if (!d->nslots) return ORC_OK;
    uint64_t so = 0, mo = 0, bo = 0;
    for (int32_t u = 0; u <= g->n; ++u) {
        const node_t *x = &g->nodes[u];
        d->nslots[u] = x->nslots;
        d->mass[u] = x->mass;
        for (int j = 0; j < 32; ++j) {
            d->bsize[(size_t)u * 32 + j] = x->size[j];
            d->T[(size_t)u * 32 + j] = x->T[j];
            d->alias[(size_t)u * 32 + j] = x->alias[j];
            if (x->dmask >> j & 1u) continue;
            for (uint32_t t = 0; t < x->size[j]; ++t) d->members[mo++] = x->mem[j][t];
        }
        if (d->fmass) d->fmass[u] = x->fmass;
        if (d->fthr) d->fthr[u] = x->fthr;
        if (d->dmask) d->dmask[u] = x->dmask;
        for (uint32_t s = 0; s < x->nslots; ++s) {
            d->slot_v[so] = x->v[s];
            d->slot_w[so] = x->w[s];
            if (d->slot_f) d->slot_f[so] = x->fr ? x->fr[s] : 0u;
            so++;
            uint32_t w = x->w[s] & ~x->dmask;
            for (uint32_t k = 0; k < (uint32_t)popc32(w); ++k) d->backpos[bo++] = x->back[x->boff[s] + k];
        }
    }
    return ORC_OK;
}

/* ---------------- generators ---------------- */

static uint32_t q32(double x)
{
    if (!(x > 0)) return 0;
    if (x >= 1.0) return 0xFFFFFFFFu;
    return (uint32_t)floor(x * 4294967296.0);
}

void orc_rmat_thresholds(double a, double b, double c, uint32_t thr[3])
{
    thr[0] = q32(a);
    thr[1] = q32(a + b);
    thr[2] = q32(a + b + c);
}

#define ORC_TAG_RMAT 0x524D4154u
#define ORC_TAG_WGHT 0x57474854u
#define ORC_TAG_DELS 0x44454C53u

/* R-MAT arc of edge index e: one quadrant per level, most significant bit
 * first, from word (level & 3) of the Philox block (seed; e, level/4, tag). */
static void rmat_edge(int scale, uint64_t e, const uint32_t thr[3], uint64_t seed, int32_t *u,
                      int32_t *v)
{
    uint32_t x[4] = {0, 0, 0, 0}, uu = 0, vv = 0;
    for (int l = 0; l < scale; ++l) {
        if ((l & 3) == 0) orc_philox_block(seed, (uint32_t)e, (uint32_t)(e >> 32), (uint32_t)(l >> 2), ORC_TAG_RMAT, x);
        uint32_t r = x[l & 3];
        uint32_t bu = r >= thr[1];
        uint32_t bv = (r >= thr[0] && r < thr[1]) || r >= thr[2];
        uu = (uu << 1) | bu;
        vv = (vv << 1) | bv;
    }
    *u = (int32_t)uu + 1;
    *v = (int32_t)vv + 1;
}

uint32_t orc_gen_weight(int wkind, uint64_t seed, uint64_t e)
{
    uint32_t x[4];
    orc_philox_block(seed, (uint32_t)e, (uint32_t)(e >> 32), 0xFFFFFFFFu, ORC_TAG_WGHT, x);
    switch (wkind) {
    case ORC_W_UNIFORM16:
        return 1u + (uint32_t)(((uint64_t)x[0] * 65535u) >> 32);
    case ORC_W_LOGUNI32: {
        uint32_t L = 1u + (uint32_t)(((uint64_t)x[0] * 32u) >> 32);
        uint32_t top = 1u << (L - 1);
        return top | (x[1] & (top - 1u));
    }
    case ORC_W_ZIPF24: {
        /* Zipf-like law, s = 1.2 over [1, 2^24]: inversion of the
         * continuous density y^-1.2 on [0.5, 2^24 + 0.5], rounded to the
         * nearest integer, so P(k) ~ k^-1.2.  Workload shape only. */
        double u01 = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6)) / 9007199254740992.0;
        double lo = pow(0.5, -0.2), hi = pow(16777216.5, -0.2);
        double y = pow(lo - u01 * (lo - hi), -5.0);
        uint64_t w = (uint64_t)floor(y + 0.5);
        if (w < 1) w = 1;
        if (w > 16777216u) w = 16777216u;
        return (uint32_t)w;
    }
    default:
        return 1u;
    }
}

void orc_gen_rmat(int scale, int64_t edge_begin, int64_t nedges, const uint32_t thr[3], int wkind,
                  uint64_t seed, int undirected, orc_edge *out, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < nedges; ++i) {
        uint64_t e = (uint64_t)(edge_begin + i);
        int32_t u, v;
        rmat_edge(scale, e, thr, seed, &u, &v);
        uint32_t w = orc_gen_weight(wkind, seed, e);
        if (undirected) {
            out[2 * i].u = u; out[2 * i].v = v; out[2 * i].w = w;
            out[2 * i + 1].u = v; out[2 * i + 1].v = u; out[2 * i + 1].w = w;
        } else {
            out[i].u = u; out[i].v = v; out[i].w = w;
        }
    }
}

struct orc_updgen {
    int32_t n;
    int scale, wkind;
    uint32_t thr[3];
    uint64_t seed;
    uint64_t next_edge, draw;
    uint32_t *deg;
    uint64_t *live;
    int64_t nlive, cap;
};

static void live_push(orc_updgen *s, int32_t u, uint32_t slot)
{
    if (s->nlive == s->cap) {
        s->cap = s->cap ? s->cap * 2 : 1024;
        s->live = (uint64_t *)xrealloc(s->live, (size_t)s->cap * 8);
    }
    s->live[s->nlive++] = ((uint64_t)(uint32_t)u << 32) | slot;
}

/* Update stream (graph.cuh:58-92 roles): inserts are fresh R-MAT arcs with
 * fresh weights that take the next slot of their source node; deletes are
 * uniformly random live arcs (no repeats), drawn with swap-remove. */
orc_updgen *orc_updgen_create(int32_t n, int scale, const orc_edge *base, int64_t m,
                              int64_t edge_begin, const uint32_t thr[3], int wkind, uint64_t seed)
{
    orc_updgen *s = (orc_updgen *)calloc(1, sizeof(orc_updgen));
    s->n = n;
    s->scale = scale;
    s->wkind = wkind;
    memcpy(s->thr, thr, sizeof(s->thr));
    s->seed = seed;
    s->next_edge = (uint64_t)edge_begin;
    s->deg = (uint32_t *)calloc((size_t)n + 1, 4);
    for (int64_t i = 0; i < m; ++i) live_push(s, base[i].u, s->deg[base[i].u]++);
    return s;
}

void orc_updgen_destroy(orc_updgen *s)
{
    if (!s) return;
    free(s->deg);
    free(s->live);
    free(s);
}

void orc_updgen_inserts(orc_updgen *s, int64_t k, orc_edge *out)
{
    for (int64_t i = 0; i < k; ++i) {
        uint64_t e = s->next_edge++;
        int32_t u, v;
        rmat_edge(s->scale, e, s->thr, s->seed, &u, &v);
        out[i].u = u;
        out[i].v = v;
        out[i].w = orc_gen_weight(s->wkind, s->seed, e);
        live_push(s, u, s->deg[u]++);
    }
}

int orc_updgen_deletes(orc_updgen *s, int64_t k, orc_del *out)
{
    if (k > s->nlive) return ORC_EINVAL;
    for (int64_t i = 0; i < k; ++i) {
        uint32_t x[4];
        orc_philox_block(s->seed, (uint32_t)s->draw, (uint32_t)(s->draw >> 32), 0, ORC_TAG_DELS, x);
        s->draw++;
        uint64_t r64 = ((uint64_t)x[1] << 32) | x[0];
        int64_t idx = (int64_t)(((u128)r64 * (uint64_t)s->nlive) >> 64);
        uint64_t a = s->live[idx];
        s->live[idx] = s->live[--s->nlive];
        out[i].u = (int32_t)(a >> 32);
        out[i].slot = (int32_t)(uint32_t)a;
    }
    return ORC_OK;
}

int64_t orc_updgen_live(const orc_updgen *s) { return s->nlive; }

// Deterministic workload generators (host): R-MAT graphs and update streams.
//
// They take the role of the reference's loadGraph/insertGraph/deleteGraph
// generators (include/graph.cuh:16-92), which read a text edge list and
// draw updates with rand() seeded from a heap address.  Here every arc and
// every update is a pure function of (seed, index) through Philox4x32-10,
// so a graph can be regenerated identically anywhere.
//   arc of edge e:  level l of scale S takes word (l & 3) of
//                   Philox(seed; e_lo, e_hi, l >> 2, 'RMAT'); the quadrant
//                   thresholds are floor(a 2^32), floor((a+b) 2^32), ...
//   weight of e:    Philox(seed; e_lo, e_hi, 0xFFFFFFFF, 'WGHT')
//   k-th delete:    Philox(seed; k_lo, k_hi, 0, 'DELS') picks a live arc
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/bingo.h"

namespace {

struct Philox {
    static void block(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4])
    {
        uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
            c0 = n0;
            c1 = (uint32_t)p1;
            c2 = n2;
            c3 = (uint32_t)p0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        out[0] = c0;
        out[1] = c1;
        out[2] = c2;
        out[3] = c3;
    }
};

constexpr uint32_t kTagRmat = 0x524D4154u, kTagWeight = 0x57474854u, kTagDel = 0x44454C53u;

uint32_t frac32(double x)
{
    if (!(x > 0)) return 0;
    if (x >= 1.0) return 0xFFFFFFFFu;
    return (uint32_t)std::floor(x * 4294967296.0);
}

struct RmatSpec {
    int scale;
    uint32_t t0, t1, t2;
    int wkind;
    uint64_t seed;
};

void rmat_arc(const RmatSpec &sp, uint64_t e, int32_t *u, int32_t *v)
{
    uint32_t x[4] = {0, 0, 0, 0}, uu = 0, vv = 0;
    for (int l = 0; l < sp.scale; ++l) {
        if ((l & 3) == 0) Philox::block(sp.seed, (uint32_t)e, (uint32_t)(e >> 32), (uint32_t)(l >> 2), kTagRmat, x);
        const uint32_t r = x[l & 3];
        const uint32_t right = (r >= sp.t0 && r < sp.t1) || r >= sp.t2;  // quadrants b and d
        const uint32_t lower = r >= sp.t1;                              // quadrants c and d
        uu = (uu << 1) | lower;
        vv = (vv << 1) | right;
    }
    *u = (int32_t)uu + 1;
    *v = (int32_t)vv + 1;
}

uint32_t weight_of(int wkind, uint64_t seed, uint64_t e)
{
    uint32_t x[4];
    Philox::block(seed, (uint32_t)e, (uint32_t)(e >> 32), 0xFFFFFFFFu, kTagWeight, x);
    if (wkind == BINGO_W_UNIFORM16) return 1u + (uint32_t)(((uint64_t)x[0] * 65535u) >> 32);
    if (wkind == BINGO_W_LOGUNI32) {
        const uint32_t bitlen = 1u + (uint32_t)(((uint64_t)x[0] * 32u) >> 32);
        const uint32_t top = 1u << (bitlen - 1);
        return top | (x[1] & (top - 1u));
    }
    if (wkind == BINGO_W_ZIPF24) {
        // continuous y^-1.2 on [0.5, 2^24 + 0.5] inverted, rounded to nearest
        const double u01 = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6)) / 9007199254740992.0;
        const double lo = std::pow(0.5, -0.2), hi = std::pow(16777216.5, -0.2);
        const double y = std::pow(lo - u01 * (lo - hi), -5.0);
        uint64_t w = (uint64_t)std::floor(y + 0.5);
        if (w < 1) w = 1;
        if (w > 16777216u) w = 16777216u;
        return (uint32_t)w;
    }
    return 1u;
}

bool spec_ok(int scale, double a, double b, double c, int wkind)
{
    return scale >= 1 && scale <= 30 && a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0 && wkind >= 0 && wkind <= 3;
}

}  // namespace

// Slots are kept as raw numbers (assigned per node in arrival order, never
// reused); a node's current slot number is raw - base[u].  Compaction (the
// rule of bingo_compact) renumbers a node's live slots by rank: under expiry
// the dead slots of a node are always its lowest raw numbers, so base[u]
// grows by the dead count; after random deletes the live arcs of compacted
// nodes are renumbered one by one.
struct bingo_updgen {
    RmatSpec sp;
    int32_t n;
    uint64_t next_edge, draws;
    std::vector<uint32_t> nslots;  // current slot count per node (tombstones included)
    std::vector<uint32_t> dead;    // tombstones per node since its last compaction
    std::vector<uint32_t> base;    // raw - current slot number
    std::vector<uint64_t> live;    // (u << 32) | raw slot; live[head..] in arrival order until a random delete
    std::vector<uint64_t> dels;    // random deletes since the last compaction: (u << 32) | current slot
    size_t head = 0;               // arcs expired so far (sliding window)
    int mode = 0;                  // 0 no deletes yet, 1 random deletes, 2 expiry

    uint32_t cur(uint64_t e) const { return (uint32_t)e - base[e >> 32]; }
    bingo_del del_of(uint64_t e) const { return bingo_del{(int32_t)(e >> 32), (int32_t)cur(e)}; }
    void push(int32_t u)
    {
        live.push_back(((uint64_t)(uint32_t)u << 32) | (nslots[u] + base[u]));
        ++nslots[u];
    }
};

extern "C" {

int bingo_gen_rmat(int scale, int64_t edge_begin, int64_t nedges, double a, double b, double c, int wkind,
                   uint64_t seed, int undirected, bingo_edge *out)
{
    if (!spec_ok(scale, a, b, c, wkind) || edge_begin < 0 || nedges < 0 || (nedges && !out)) return BINGO_EINVAL;
    const RmatSpec sp{scale, frac32(a), frac32(a + b), frac32(a + b + c), wkind, seed};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nedges; ++i) {
        const uint64_t e = (uint64_t)(edge_begin + i);
        int32_t u, v;
        rmat_arc(sp, e, &u, &v);
        const uint32_t w = weight_of(wkind, seed, e);
        if (undirected) {
            out[2 * i] = bingo_edge{u, v, w};
            out[2 * i + 1] = bingo_edge{v, u, w};
        } else {
            out[i] = bingo_edge{u, v, w};
        }
    }
    return BINGO_OK;
}

int bingo_updgen_create(int32_t n, int scale, const bingo_edge *base, int64_t m, int64_t edge_begin, double a,
                        double b, double c, int wkind, uint64_t seed, bingo_updgen **out)
{
    if (!out || n < 1 || m < 0 || (m && !base) || edge_begin < 0 || !spec_ok(scale, a, b, c, wkind))
        return BINGO_EINVAL;
    bingo_updgen *s = new bingo_updgen();
    s->sp = RmatSpec{scale, frac32(a), frac32(a + b), frac32(a + b + c), wkind, seed};
    s->n = n;
    s->next_edge = (uint64_t)edge_begin;
    s->draws = 0;
    s->nslots.assign((size_t)n + 1, 0);
    s->dead.assign((size_t)n + 1, 0);
    s->base.assign((size_t)n + 1, 0);
    s->live.reserve((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
        const int32_t u = base[i].u;
        if (u < 1 || u > n) {
            delete s;
            return BINGO_EINVAL;
        }
        s->push(u);
    }
    *out = s;
    return BINGO_OK;
}

void bingo_updgen_destroy(bingo_updgen *s) { delete s; }

// Inserts: fresh R-MAT arcs (next edge indices) with fresh weights; each
// takes the next slot of its source node, as graph.cuh:66 assigns d[u]++.
int bingo_updgen_inserts(bingo_updgen *s, int64_t k, bingo_edge *out)
{
    if (!s || k < 0 || (k && !out)) return BINGO_EINVAL;
    const uint64_t e0 = s->next_edge;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < k; ++i) {  // arcs are pure functions of their edge index
        const uint64_t e = e0 + (uint64_t)i;
        int32_t u, v;
        rmat_arc(s->sp, e, &u, &v);
        bad |= u > s->n || v > s->n;
        out[i] = bingo_edge{u, v, weight_of(s->sp.wkind, s->sp.seed, e)};
    }
    if (bad) return BINGO_EINVAL;
    s->next_edge = e0 + (uint64_t)k;
    for (int64_t i = 0; i < k; ++i) s->push(out[i].u);  // slots in arrival order
    return BINGO_OK;
}

// Deletes: uniformly random live arcs without repetition (graph.cuh:78-92
// retries dead positions; swap-remove gives the same law).
int bingo_updgen_deletes(bingo_updgen *s, int64_t k, bingo_del *out)
{
    if (!s || k < 0 || (k && !out) || k > (int64_t)s->live.size() || s->mode == 2) return BINGO_EINVAL;
    if (k) s->mode = 1;
    for (int64_t i = 0; i < k; ++i) {
        uint32_t x[4];
        Philox::block(s->sp.seed, (uint32_t)s->draws, (uint32_t)(s->draws >> 32), 0, kTagDel, x);
        ++s->draws;
        const uint64_t r64 = ((uint64_t)x[1] << 32) | x[0];
        const uint64_t nl = s->live.size();
        const uint64_t idx = (uint64_t)(((unsigned __int128)r64 * nl) >> 64);
        const uint64_t pick = s->live[idx];
        s->live[idx] = s->live.back();
        s->live.pop_back();
        out[i] = s->del_of(pick);
        ++s->dead[pick >> 32];
        s->dels.push_back(((pick >> 32) << 32) | s->cur(pick));
    }
    return BINGO_OK;
}

// Expiry: the k oldest live arcs in arrival order (load order, then inserts),
// the deletes of a sliding window.  Not mixed with random deletes.
int bingo_updgen_expire(bingo_updgen *s, int64_t k, bingo_del *out)
{
    if (!s || k < 0 || (k && !out) || s->mode == 1 || k > (int64_t)(s->live.size() - s->head)) return BINGO_EINVAL;
    if (k) s->mode = 2;
    for (int64_t i = 0; i < k; ++i) {
        const uint64_t pick = s->live[s->head++];
        out[i] = s->del_of(pick);
        ++s->dead[pick >> 32];
    }
    return BINGO_OK;
}

int64_t bingo_updgen_live(const bingo_updgen *s) { return s ? (int64_t)(s->live.size() - s->head) : 0; }

// The selection rule of bingo_compact: dead >= 1 and dead 2^16 >= thr16 nslots.
int bingo_updgen_compact(bingo_updgen *s, double min_dead_frac)
{
    if (!s || !(min_dead_frac >= 0 && min_dead_frac <= 1)) return BINGO_EINVAL;
    const uint64_t thr16 = (uint64_t)std::floor(min_dead_frac * 65536.0);
    const int32_t n = s->n;
    std::vector<uint8_t> pick((size_t)n + 1, 0);
    bool any = false;
    for (int32_t u = 1; u <= n; ++u)
        if (s->dead[u] && (uint64_t)s->dead[u] * 65536u >= thr16 * s->nslots[u]) {
            pick[u] = 1;
            any = true;
        }
    if (!any) return BINGO_OK;
    if (s->mode == 1) {
        // random deletes: a live arc's new number is its current one minus the
        // node's tombstones below it
        std::sort(s->dels.begin(), s->dels.end());
        std::vector<uint64_t> first((size_t)n + 2, 0);  // dels of node u: [first[u], first[u + 1])
        for (uint64_t d : s->dels) ++first[(d >> 32) + 1];
        for (int32_t u = 0; u <= n; ++u) first[u + 1] += first[u];
        const std::vector<uint64_t> &D = s->dels;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)s->live.size(); ++i) {
            const uint64_t e = s->live[i];
            const uint32_t u = (uint32_t)(e >> 32);
            if (!pick[u]) continue;
            const uint32_t c = s->cur(e);
            const auto lo = D.begin() + (int64_t)first[u], hi = D.begin() + (int64_t)first[u + 1];
            const uint32_t below = (uint32_t)(std::lower_bound(lo, hi, ((uint64_t)u << 32) | c) - lo);
            s->live[i] = e - below;
        }
        std::vector<uint64_t> keep;
        for (uint64_t d : s->dels)
            if (!pick[d >> 32]) keep.push_back(d);
        s->dels.swap(keep);
    }
    for (int32_t u = 1; u <= n; ++u)
        if (pick[u]) {
            if (s->mode == 2) s->base[u] += s->dead[u];  // expiry: the dead are the lowest raw numbers
            s->nslots[u] -= s->dead[u];
            s->dead[u] = 0;
        }
    return BINGO_OK;
}

}  // extern "C"

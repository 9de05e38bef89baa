// The reference's application driver on top of the C ABI.
//
// bingo_read_edgelist  = loadGraph's parsing and weighting (include/graph.cuh:16-57):
//                        whitespace-separated "u v" pairs, 0-based, shifted to
//                        1-based; the k-th arc of u gets slot k; weight of
//                        u->v = bias(outdeg(v)) + 1 with bias(d) = d, or
//                        floor(d^0.3) when n > 1e7 (include/api.h:29-30).
// bingo_instance       = instance() (include/apps.h:4, src/apps.cu:10-42):
//                        load, then `rounds` x {insert batch, delete batch,
//                        one walk per node of length LEN}; mixed mode halves
//                        the batch (apps.cu:22-23).  Insert batches duplicate
//                        a random existing arc (graph.cuh:58-77), deletes pick
//                        random live arcs (graph.cuh:78-92).  Randomness is a
//                        seeded Philox stream instead of rand() seeded from a
//                        heap address, and an inserted weight is clamped to
//                        >= 1 (the reference may insert weight 0).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bingo.h"

namespace {

void philox_block(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4])
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

struct Stream {  // sequential uniform integers from Philox (seed, counter, tag)
    uint64_t seed, ctr = 0;
    uint32_t tag;
    uint64_t below(uint64_t n)
    {
        uint32_t x[4];
        philox_block(seed, (uint32_t)ctr, (uint32_t)(ctr >> 32), 0, tag, x);
        ++ctr;
        const uint64_t r = ((uint64_t)x[1] << 32) | x[0];
        return (uint64_t)(((unsigned __int128)r * n) >> 64);
    }
};

uint32_t bias_weight(uint32_t d, bool large)
{
    // ref: Int weights are bias(d[v]) + 1 (graph.cuh:47-50)
    const uint32_t b = large ? (uint32_t)std::floor(std::pow((double)d, 0.3)) : d;
    return b + 1u;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// The driver's update stream (ref insertGraph/deleteGraph, graph.cuh:58-92):
// a host mirror of every arc's slot; an insert duplicates a random existing
// arc as the source node's next slot with weight d[v] (clamped to >= 1), a
// delete picks a random live arc.  Mixed mode halves the batch of each kind
// (apps.cu:22-23).
struct bingo_instgen {
    struct Arc {
        int32_t u, v, slot;
    };
    std::vector<Arc> arcs;
    std::vector<uint32_t> d;
    int utype = 0;
    int64_t batch = 0;
    Stream rs{0, 0, 0x494E5354u};
};

extern "C" {

int bingo_instgen_create(const bingo_edge *base, int64_t m, int32_t n, int utype, int64_t batch, uint64_t seed,
                         bingo_instgen **out)
{
    if (!out || m < 0 || (m && !base) || n < 1 || utype < 0 || utype > 2 || batch < 0) return BINGO_EINVAL;
    auto *s = new bingo_instgen();
    s->utype = utype;
    s->batch = utype == BINGO_UPD_MIXED ? batch >> 1 : batch;
    s->rs = Stream{seed, 0, 0x494E5354u};
    s->d.assign((size_t)n + 1, 0);
    s->arcs.resize((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
        if (base[i].u < 1 || base[i].u > n) {
            delete s;
            return BINGO_EINVAL;
        }
        s->arcs[i] = bingo_instgen::Arc{base[i].u, base[i].v, (int32_t)s->d[base[i].u]++};
    }
    *out = s;
    return BINGO_OK;
}

void bingo_instgen_destroy(bingo_instgen *s) { delete s; }

int bingo_instgen_round(bingo_instgen *s, bingo_edge *ins, int64_t *ni, bingo_del *dels, int64_t *nd)
{
    if (!s || !ni || !nd) return BINGO_EINVAL;
    const int64_t batch = s->batch;
    *ni = *nd = 0;
    if (s->utype == BINGO_UPD_MIXED || s->utype == BINGO_UPD_INSERT) {
        if (batch && !ins) return BINGO_EINVAL;
        const int64_t last = (int64_t)s->arcs.size();
        if (batch && !last) return BINGO_EINVAL;
        for (int64_t i = 0; i < batch; ++i) {
            const bingo_instgen::Arc a = s->arcs[s->rs.below((uint64_t)last)];
            s->arcs.push_back(bingo_instgen::Arc{a.u, a.v, (int32_t)s->d[a.u]++});
            const uint32_t w = s->d[a.v];  // graph.cuh:70 passes bias(d[v]) = d[v], without the +1
            ins[i] = bingo_edge{a.u, a.v, w ? w : 1u};
        }
        *ni = batch;
    }
    if (s->utype == BINGO_UPD_MIXED || s->utype == BINGO_UPD_DELETE) {
        if (batch && !dels) return BINGO_EINVAL;
        int64_t live = 0;
        for (const auto &a : s->arcs) live += a.slot >= 0;
        if (live < batch) return BINGO_EINVAL;
        for (int64_t i = 0; i < batch; ++i) {
            uint64_t pos;
            do pos = s->rs.below(s->arcs.size());
            while (s->arcs[pos].slot < 0);
            dels[i] = bingo_del{s->arcs[pos].u, s->arcs[pos].slot};
            s->arcs[pos].slot = -1;
        }
        *nd = batch;
    }
    return BINGO_OK;
}

int64_t bingo_instgen_live(const bingo_instgen *s)
{
    if (!s) return 0;
    int64_t live = 0;
    for (const auto &a : s->arcs) live += a.slot >= 0;
    return live;
}

void bingo_instance_opts_default(bingo_instance_opts *o)
{
    if (!o) return;
    o->rounds = 10;        // apps.cu:24
    o->batch = 100000;     // BATCHSIZE (api.h:27)
    o->seed = 1;
    o->walk_len = 80;      // LEN (api.h:12)
    o->device = 0;
    o->print = 1;
    o->keep_graph = 0;
}

int bingo_read_edgelist(const char *path, int32_t *n_out, bingo_edge *edges, int64_t cap, int64_t *m_out)
{
    if (!path || !n_out || !m_out) return BINGO_EINVAL;
    FILE *f = std::fopen(path, "r");
    if (!f) return BINGO_EINVAL;
    std::vector<int32_t> us, vs;
    long long u, v;
    int32_t n = 0;
    while (std::fscanf(f, "%lld%lld", &u, &v) == 2) {
        if (u < 0 || v < 0 || u >= (1ll << 30) - 1 || v >= (1ll << 30) - 1) {
            std::fclose(f);
            return BINGO_EINVAL;
        }
        us.push_back((int32_t)u + 1);
        vs.push_back((int32_t)v + 1);
        n = std::max(n, (int32_t)std::max(u, v) + 1);
    }
    std::fclose(f);
    const int64_t m = (int64_t)us.size();
    *n_out = n;
    *m_out = m;
    if (!edges) return BINGO_OK;  // size query
    if (cap < m) return BINGO_EINVAL;
    std::vector<uint32_t> d((size_t)n + 1, 0);
    for (int64_t i = 0; i < m; ++i) d[us[i]]++;
    const bool large = n > 10000000;
    for (int64_t i = 0; i < m; ++i) edges[i] = bingo_edge{us[i], vs[i], bias_weight(d[vs[i]], large)};
    return BINGO_OK;
}

int bingo_instance(int app, int utype, const char *dataset, const bingo_instance_opts *opts,
                   bingo_instance_result *res)
{
    bingo_instance_opts o;
    if (opts) o = *opts; else bingo_instance_opts_default(&o);
    if (app < 0 || app > 3 || utype < 0 || utype > 2 || o.rounds < 0 || o.batch < 0 || o.walk_len < 1)
        return BINGO_EINVAL;
    const double t_total = now_s();
    int32_t n = 0;
    int64_t m = 0;
    int rc = bingo_read_edgelist(dataset, &n, nullptr, 0, &m);
    if (rc) return rc;
    std::vector<bingo_edge> base((size_t)m);
    if ((rc = bingo_read_edgelist(dataset, &n, base.data(), m, &m))) return rc;
    if (o.print) {
        std::fprintf(stderr, "Graph dataset: %s.\n", dataset);
        std::fprintf(stderr, "Edge number: %lld.\n", (long long)m);
    }
    bingo_config cfg;
    bingo_config_default(&cfg);
    cfg.device = o.device;
    cfg.neighbor_index = app == BINGO_NODE2VEC ? 1 : 0;  // ref NODE2VEC (apps.cu:14-15)
    bingo_graph *g = nullptr;
    if ((rc = bingo_create(&cfg, &g))) return rc;
    if ((rc = bingo_load(g, n, base.data(), m))) {
        bingo_destroy(g);
        return rc;
    }
    // the update stream, over a host mirror of every arc's slot (EdgeData, graph.cuh:12)
    bingo_instgen *gen = nullptr;
    if ((rc = bingo_instgen_create(base.data(), m, n, utype, o.batch, o.seed, &gen))) {
        bingo_destroy(g);
        return rc;
    }
    base.clear();
    base.shrink_to_fit();
    bingo_walk_params wp;
    bingo_walk_params_default(&wp);
    wp.seed = o.seed;
    wp.max_len = o.walk_len;
    void *d_paths = nullptr, *d_steps = nullptr;
    if ((rc = bingo_dev_alloc(g, (size_t)n * o.walk_len * 4, &d_paths)) || (rc = bingo_dev_alloc(g, 8, &d_steps))) {
        bingo_instgen_destroy(gen);
        bingo_destroy(g);
        return rc;
    }
    uint64_t zero = 0, steps_total = 0;
    double walk_s = 0;
    const double t_eval = now_s();
    std::vector<bingo_edge> ins((size_t)gen->batch);
    std::vector<bingo_del> del((size_t)gen->batch);
    for (int r = 0; r < o.rounds && !rc; ++r) {
        int64_t ni = 0, nd = 0;
        rc = bingo_instgen_round(gen, ins.data(), &ni, del.data(), &nd);
        if (!rc && ni) rc = bingo_insert(g, ins.data(), ni);
        if (!rc && nd) rc = bingo_delete(g, del.data(), nd);
        if (rc) break;
        bingo_memcpy_h2d(g, d_steps, &zero, 8);
        const double tw = now_s();
        rc = bingo_walk_device(g, app, &wp, nullptr, nullptr, n, (int32_t *)d_paths, (uint64_t *)d_steps);
        if (!rc) rc = bingo_sync(g);
        walk_s += now_s() - tw;
        uint64_t st = 0;
        if (!rc) rc = bingo_memcpy_d2h(g, &st, d_steps, 8);
        steps_total += st;
    }
    const double eval_s = now_s() - t_eval;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    const double mem_gb = (double)(tot - fr) / 1073741824.0;
    if (o.print && !rc) {
        std::fprintf(stderr, "Random walk in %g s.\n", walk_s);       // apps.cu:33
        std::fprintf(stderr, "Evaluation time: %g s.\n", eval_s);     // apps.cu:34
        std::fprintf(stderr, "Total time: %g s.\n", now_s() - t_total);
        std::fprintf(stderr, "Memory Consumption: %g GB.\n", mem_gb); // apps.cu:38
    }
    if (res) {
        res->walk_s = walk_s;
        res->eval_s = eval_s;
        res->memory_gb = mem_gb;
        res->walk_steps = steps_total;
        res->n = n;
        res->arcs = bingo_instgen_live(gen);
        res->graph = nullptr;
    }
    bingo_instgen_destroy(gen);
    bingo_dev_free(g, d_paths);
    bingo_dev_free(g, d_steps);
    if (!rc && o.keep_graph && res) res->graph = g;  // the caller destroys it
    else bingo_destroy(g);
    return rc;
}

}  // extern "C"

// Serving loop: continuous batching of walk queries and update batches on one
// graph handle.
//
// A worker thread takes requests in submission order.  An update batch runs
// alone through the normal insert/delete pipeline.  A run of consecutive walk
// requests is admitted into one persistent launch (k_walk_serve) while the
// launch's walker budget and the device arena (HBM for request inputs and
// host-bound outputs) allow; the next launch takes whatever queued up
// meanwhile.  Every request records submit -> completion latency.
//
// There is no counterpart in the reference, which launches one walk kernel at
// a time and synchronises after it (include/graph.cuh:93-150).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handle.h"

using namespace bingo;

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Req {
    uint64_t ticket = 0;
    bool walk = true;
    int app = 0;
    WalkArgs args{};
    std::vector<int32_t> starts, lens;
    int64_t nwalkers = 0;
    int32_t *paths = nullptr;
    bool paths_dev = false;
    std::vector<bingo_edge> ins;
    std::vector<bingo_del> dels;
    size_t arena_need = 0;
    double t_submit = 0, t_done = 0;
    bool done = false;
    int status = BINGO_OK;
    uint64_t steps = 0;
};

}  // namespace

struct bingo_server {
    bingo_graph *g = nullptr;
    bingo_server_config cfg{};
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
    uint64_t next_ticket = 1;
    bool stopping = false;
    char *arena = nullptr;
    ServeReq *d_table = nullptr;
    unsigned long long *d_counts = nullptr;
    std::vector<double> lat;
    bingo_server_stats st{};
    int blocks_per_cu = 1;
};

static int run_launch(bingo_server *sv, std::vector<std::shared_ptr<Req>> &batch)
{
    bingo_graph *g = sv->g;
    std::vector<ServeReq> table(batch.size());
    size_t off = 0;
    unsigned long long base = 0;
    for (size_t i = 0; i < batch.size(); ++i) {
        Req &r = *batch[i];
        ServeReq q{};
        q.seed = r.args.seed;
        q.thr_return = r.args.thr_return;
        q.thr_neighbor = r.args.thr_neighbor;
        q.thr_other = r.args.thr_other;
        q.thr_stop = r.args.thr_stop;
        q.max_len = r.args.max_len;
        q.max_attempts = r.args.max_attempts;
        q.app = r.app;
        q.nwalkers = r.nwalkers;
        q.base = base;
        q.steps = sv->d_counts + i;
        if (!r.starts.empty()) {
            q.starts = reinterpret_cast<const int32_t *>(sv->arena + off);
            HIPCHK(hipMemcpyAsync(sv->arena + off, r.starts.data(), r.starts.size() * 4, hipMemcpyHostToDevice, g->st));
            off += align256(r.starts.size() * 4);
        }
        if (!r.lens.empty()) {
            q.lens = reinterpret_cast<const int32_t *>(sv->arena + off);
            HIPCHK(hipMemcpyAsync(sv->arena + off, r.lens.data(), r.lens.size() * 4, hipMemcpyHostToDevice, g->st));
            off += align256(r.lens.size() * 4);
        }
        if (r.paths_dev) {
            q.paths = r.paths;
        } else {
            q.paths = reinterpret_cast<int32_t *>(sv->arena + off);
            off += align256((size_t)r.nwalkers * r.args.max_len * 4);
        }
        base += (unsigned long long)r.nwalkers;
        table[i] = q;
    }
    std::vector<unsigned long long> counts(batch.size(), 0);
    HIPCHK(hipMemcpyAsync(sv->d_table, table.data(), table.size() * sizeof(ServeReq), hipMemcpyHostToDevice, g->st));
    HIPCHK(hipMemcpyAsync(sv->d_counts, counts.data(), counts.size() * 8, hipMemcpyHostToDevice, g->st));
    launch_reset_walk(g->s.ctl, g->st);
    const int64_t per_block = 256 * 2;
    const int64_t blocks =
        std::max<int64_t>(1, std::min<int64_t>((int64_t)g->num_cu * sv->blocks_per_cu, ((int64_t)base + per_block - 1) / per_block));
    launch_walk_serve(g->s, sv->d_table, (int)batch.size(), base, &g->s.ctl->cursor, (int)blocks, g->st);
    HIPCHK(hipGetLastError());
    for (size_t i = 0; i < batch.size(); ++i)
        if (!batch[i]->paths_dev)
            HIPCHK(hipMemcpyAsync(batch[i]->paths, table[i].paths, (size_t)batch[i]->nwalkers * table[i].max_len * 4,
                                  hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipMemcpyAsync(counts.data(), sv->d_counts, counts.size() * 8, hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    for (size_t i = 0; i < batch.size(); ++i) batch[i]->steps = counts[i];
    return BINGO_OK;
}

static void worker_main(bingo_server *sv)
{
    (void)hipSetDevice(sv->g->device);
    for (;;) {
        std::vector<std::shared_ptr<Req>> batch;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] { return sv->stopping || !sv->queue.empty(); });
            if (sv->queue.empty()) return;  // stopping and drained
            if (!sv->queue.front()->walk) {
                batch.push_back(sv->queue.front());
                sv->queue.pop_front();
            } else {  // admit consecutive walks while slots and arena last
                size_t used = 0;
                int64_t walkers = 0;
                while (!sv->queue.empty() && sv->queue.front()->walk && batch.size() < (size_t)kMaxServeReqs) {
                    const auto &r = sv->queue.front();
                    if (!batch.empty() && (used + r->arena_need > sv->cfg.arena_bytes ||
                                           walkers + r->nwalkers > sv->cfg.max_launch_walkers))
                        break;
                    used += r->arena_need;
                    walkers += r->nwalkers;
                    batch.push_back(r);
                    sv->queue.pop_front();
                }
            }
        }
        const double t0 = now_ms();
        int rc;
        if (!batch[0]->walk) {
            Req &u = *batch[0];
            rc = bingo_insert(sv->g, u.ins.data(), (int64_t)u.ins.size());
            if (!rc) rc = bingo_delete(sv->g, u.dels.data(), (int64_t)u.dels.size());
        } else {
            rc = run_launch(sv, batch);
        }
        const double t1 = now_ms();
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->st.busy_s += (t1 - t0) / 1e3;
        if (batch[0]->walk) sv->st.launches++;
        for (auto &r : batch) {
            r->status = rc;
            r->t_done = t1;
            r->done = true;
            sv->st.requests++;
            if (r->walk) {
                sv->st.walk_requests++;
                sv->st.walk_steps += r->steps;
                sv->lat.push_back(r->t_done - r->t_submit);
            } else {
                sv->st.update_requests++;
            }
        }
        sv->cv_done.notify_all();
    }
}

static int submit(bingo_server *sv, std::shared_ptr<Req> r, uint64_t *ticket)
{
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stopping) return fail(BINGO_ESTATE, "server is stopping");
    r->ticket = sv->next_ticket++;
    r->t_submit = now_ms();
    sv->live[r->ticket] = r;
    sv->queue.push_back(r);
    if (ticket) *ticket = r->ticket;
    sv->cv_work.notify_one();
    return BINGO_OK;
}

extern "C" {

void bingo_server_config_default(bingo_server_config *c)
{
    if (!c) return;
    c->arena_bytes = 4ull << 30;
    c->max_launch_walkers = 1ll << 23;
    c->blocks_per_cu = 0;
    c->reserved = 0;
}

int bingo_server_create(bingo_graph *g, const bingo_server_config *c, bingo_server **out)
{
    if (!out) return fail(BINGO_EINVAL, "null output");
    *out = nullptr;
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    bingo_server *sv = new bingo_server();
    sv->g = g;
    if (c) sv->cfg = *c; else bingo_server_config_default(&sv->cfg);
    if (sv->cfg.arena_bytes < 4096 || sv->cfg.max_launch_walkers < 1) {
        delete sv;
        return fail(BINGO_EINVAL, "server arena or walker budget too small");
    }
    const int occ = serve_max_blocks_per_cu();
    sv->blocks_per_cu = sv->cfg.blocks_per_cu > 0 ? std::min(sv->cfg.blocks_per_cu, occ) : occ;
    if (hipMalloc((void **)&sv->arena, sv->cfg.arena_bytes) != hipSuccess ||
        hipMalloc((void **)&sv->d_table, kMaxServeReqs * sizeof(ServeReq)) != hipSuccess ||
        hipMalloc((void **)&sv->d_counts, kMaxServeReqs * 8) != hipSuccess) {
        (void)hipGetLastError();
        if (sv->arena) (void)hipFree(sv->arena);
        if (sv->d_table) (void)hipFree(sv->d_table);
        delete sv;
        return fail(BINGO_ENOMEM, "server arena allocation failed");
    }
    sv->worker = std::thread(worker_main, sv);
    *out = sv;
    return BINGO_OK;
}

int bingo_server_submit_walk(bingo_server *sv, int app, const bingo_walk_params *params, const int32_t *starts,
                             const int32_t *lens, int64_t nwalkers, int32_t *paths, int paths_on_device,
                             uint64_t *ticket)
{
    if (!sv) return fail(BINGO_EINVAL, "null server");
    bingo_graph *g = sv->g;
    auto r = std::make_shared<Req>();
    int rc = fill_walk_args(params, app, &r->args);
    if (rc) return rc;
    if (nwalkers < 1 || !paths) return fail(BINGO_EINVAL, "bad walk request");
    if (app == BINGO_NODE2VEC && !g->s.nbr) return fail(BINGO_EUNSUPPORTED, "node2vec needs neighbor_index=1");
    for (int64_t i = 0; starts && i < nwalkers; ++i)
        if (starts[i] < 1 || starts[i] > g->s.n) return fail(BINGO_EINVAL, "start node outside 1..n");
    for (int64_t i = 0; lens && i < nwalkers; ++i)
        if (lens[i] < 1 || lens[i] > params->max_len) return fail(BINGO_EINVAL, "walk length outside 1..max_len");
    r->walk = true;
    r->app = app;
    r->nwalkers = nwalkers;
    r->paths = paths;
    r->paths_dev = paths_on_device != 0;
    if (starts) r->starts.assign(starts, starts + nwalkers);
    if (lens) r->lens.assign(lens, lens + nwalkers);
    r->arena_need = (starts ? align256((size_t)nwalkers * 4) : 0) + (lens ? align256((size_t)nwalkers * 4) : 0) +
                    (r->paths_dev ? 0 : align256((size_t)nwalkers * params->max_len * 4));
    if (r->arena_need > sv->cfg.arena_bytes) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->st.rejected++;
        return fail(BINGO_ENOMEM, "walk request needs more than the server arena");
    }
    return submit(sv, r, ticket);
}

int bingo_server_submit_update(bingo_server *sv, const bingo_edge *ins, int64_t ni, const bingo_del *dels, int64_t nd,
                               uint64_t *ticket)
{
    if (!sv) return fail(BINGO_EINVAL, "null server");
    if (ni < 0 || nd < 0 || (ni && !ins) || (nd && !dels)) return fail(BINGO_EINVAL, "bad update request");
    auto r = std::make_shared<Req>();
    r->walk = false;
    if (ni) r->ins.assign(ins, ins + ni);
    if (nd) r->dels.assign(dels, dels + nd);
    return submit(sv, r, ticket);
}

int bingo_server_wait(bingo_server *sv, uint64_t ticket, uint64_t *steps_out, double *latency_ms)
{
    if (!sv) return fail(BINGO_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->live.find(ticket);
    if (it == sv->live.end()) return fail(BINGO_EINVAL, "unknown or already collected ticket");
    std::shared_ptr<Req> r = it->second;
    sv->cv_done.wait(lk, [&] { return r->done; });
    sv->live.erase(ticket);
    if (steps_out) *steps_out = r->steps;
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (r->status) return fail(r->status, "request failed");
    return BINGO_OK;
}

int bingo_server_stats_get(bingo_server *sv, bingo_server_stats *out)
{
    if (!sv || !out) return fail(BINGO_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    *out = sv->st;
    std::vector<double> l = sv->lat;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) { return l[std::min(l.size() - 1, (size_t)(p * (double)(l.size() - 1) + 0.5))]; };
        out->p50_ms = pct(0.50);
        out->p90_ms = pct(0.90);
        out->p99_ms = pct(0.99);
        out->max_ms = l.back();
        double s = 0;
        for (double x : l) s += x;
        out->mean_ms = s / (double)l.size();
    }
    return BINGO_OK;
}

void bingo_server_destroy(bingo_server *sv)
{
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
        sv->cv_work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipStreamSynchronize(sv->g->st);
    (void)hipFree(sv->arena);
    (void)hipFree(sv->d_table);
    (void)hipFree(sv->d_counts);
    delete sv;
}

}  // extern "C"

// Serving loop: continuous batching of walk queries and update batches on one
// graph handle.
//
// A scheduler thread owns the device.  Requests take effect in submission
// order; an update batch is a barrier: it runs once every walk submitted
// before it has finished, and walks submitted after it start only once it is
// applied.  Between barriers, walk requests are admitted as the device arena
// frees up (their starts/lens are staged there, and host-bound outputs are
// written there) and packed in submission order into persistent launches
// (k_walk_serve) of at most max_launch_walkers walkers; a larger request is
// cut into chunks.  Up to `streams` launches are in flight, each on its own
// HIP stream with its own request table and counters, and a launch's
// completion is observed through its event: the next launch is packed from
// whatever arrived and enqueued while the current one runs, so the device
// goes from one to the next without a gap and a request that arrives while
// others run starts in the next launch.  While launches run, the scheduler
// stages the inputs of arrived walks, so packing a launch waits for no
// copies, and an update batch replays captured graphs (no host round trip
// inside it) when it fits them.  Every request records submit -> completion
// latency.
//
// There is no counterpart in the reference, which launches one walk kernel at
// a time and synchronises after it (include/graph.cuh:93-150).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
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

constexpr int kMaxStreams = 4;
// update batches of up to kCapturedMax arcs each way replay captured graphs
// (bingo_update_graph_*), whose pools grow inside kCapturedRoom committed at
// creation; larger batches, and a batch past the room, take bingo_insert /
// bingo_delete (which grow the pools)
constexpr int64_t kCapturedMax = 1 << 17;
constexpr uint64_t kCapturedRoom = 1ull << 30;  // (test hook: BINGO_SERVE_ROOM bytes)

struct Req {
    uint64_t ticket = 0;
    bool walk = true;
    int app = 0;
    WalkArgs args{};
    std::vector<int32_t> starts, lens;
    int64_t nwalkers = 0;
    int32_t *paths = nullptr;  // caller's output, host or device
    bool paths_dev = false;
    std::vector<bingo_edge> ins;
    std::vector<bingo_del> dels;
    size_t arena_need = 0;
    // scheduler state (scheduler thread only)
    bool admitted = false;
    bool has_arena = false;  // holds [arena_off, arena_off + arena_need)
    size_t arena_off = 0;
    const int32_t *d_starts = nullptr, *d_lens = nullptr;
    int32_t *d_paths = nullptr;
    int64_t next = 0;      // first walker not yet given to a launch
    int64_t finished = 0;  // walkers whose launch completed
    uint64_t acc_steps = 0;
    // completion (guarded by the server mutex)
    double t_submit = 0, t_done = 0;
    bool done = false;
    int status = BINGO_OK;
    uint64_t steps = 0;
};

struct Part {
    std::shared_ptr<Req> r;
    int64_t first, take;  // walkers [first, first + take) of r
    int64_t base;         // at launch indices [base, base + take)
};

struct Launch {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    ServeReq *d_table = nullptr, *h_table = nullptr;             // h_table: pinned
    unsigned long long *d_counts = nullptr, *h_counts = nullptr;  // [kMaxServeReqs] steps, then the cursor
    std::vector<Part> parts;
    int64_t total = 0;
    bool busy = false;
};

// first-fit placement of request staging in the device arena
struct Arena {
    char *base = nullptr;
    size_t bytes = 0;
    std::map<size_t, size_t> used;  // offset -> size
    bool alloc(size_t need, size_t *off)
    {
        size_t at = 0;
        for (const auto &u : used) {
            if (u.first - at >= need) break;
            at = u.first + u.second;
        }
        if (at + need > bytes) return false;
        used[at] = need;
        *off = at;
        return true;
    }
    void release(size_t off) { used.erase(off); }
};

}  // namespace

struct bingo_server {
    bingo_graph *g = nullptr;
    bingo_server_config cfg{};
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> inbox;  // submitted, not yet seen by the scheduler
    std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
    uint64_t next_ticket = 1;
    bool stopping = false;
    Arena arena;
    hipStream_t copy_st = nullptr;
    Launch ln[kMaxStreams];
    int nstreams = 2;
    std::vector<double> lat;
    bingo_server_stats st{};
    int blocks_per_cu[3] = {1, 1, 1};  // per launch class: other apps, node2vec, mixed
    bingo_update_graph *ug = nullptr;  // captured update batches (null: not available)
    bingo_edge *d_ins = nullptr;       // their staging buffers
    bingo_del *d_dels = nullptr;
    // scheduler-local
    std::deque<std::shared_ptr<Req>> order;  // unfinished requests, submission order
    double busy_since = -1;
};

static void complete(bingo_server *sv, const std::shared_ptr<Req> &r, int rc)
{
    const double t = now_ms();
    std::lock_guard<std::mutex> lk(sv->mu);
    r->status = rc;
    r->t_done = t;
    r->steps = r->acc_steps;
    r->done = true;
    sv->st.requests++;
    if (r->walk) {
        sv->st.walk_requests++;
        sv->st.walk_steps += r->acc_steps;
        sv->lat.push_back(r->t_done - r->t_submit);
    } else {
        sv->st.update_requests++;
    }
    sv->cv_done.notify_all();
}

static void retire(bingo_server *sv, const std::shared_ptr<Req> &r)
{
    auto it = std::find(sv->order.begin(), sv->order.end(), r);
    if (it != sv->order.end()) sv->order.erase(it);
}

static int busy_launches(const bingo_server *sv)
{
    int b = 0;
    for (int i = 0; i < sv->nstreams; ++i) b += sv->ln[i].busy;
    return b;
}

static void busy_clock(bingo_server *sv)
{
    const bool busy = busy_launches(sv) > 0;
    const double t = now_ms();
    if (busy && sv->busy_since < 0) sv->busy_since = t;
    if (!busy && sv->busy_since >= 0) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->st.busy_s += (t - sv->busy_since) / 1e3;
        sv->busy_since = -1;
    }
}

// a walk whose last chunk completed: host-bound paths come back, arena frees
static void finish_walk(bingo_server *sv, const std::shared_ptr<Req> &r)
{
    int rc = BINGO_OK;
    if (!r->paths_dev) {
        const size_t bytes = (size_t)r->nwalkers * r->args.max_len * 4;
        if (hipMemcpyAsync(r->paths, r->d_paths, bytes, hipMemcpyDeviceToHost, sv->copy_st) != hipSuccess ||
            hipStreamSynchronize(sv->copy_st) != hipSuccess)
            rc = fail(BINGO_EDEVICE, "copying walk output to the host failed");
    }
    if (r->has_arena) sv->arena.release(r->arena_off);
    r->has_arena = false;
    retire(sv, r);
    complete(sv, r, rc);
}

// A launch ended: its parts' walkers are done; finished requests complete.
static void harvest(bingo_server *sv, Launch &L, hipError_t e)
{
    for (size_t i = 0; i < L.parts.size(); ++i) {
        Part &pt = L.parts[i];
        auto &r = pt.r;
        r->finished += pt.take;
        r->acc_steps += e == hipSuccess ? L.h_counts[i] : 0;
        if (e != hipSuccess) r->status = BINGO_EDEVICE;
        if (r->finished == r->nwalkers) {
            if (r->status) {
                if (r->has_arena) sv->arena.release(r->arena_off);
                r->has_arena = false;
                retire(sv, r);
                complete(sv, r, fail(BINGO_EDEVICE, "walk launch failed"));
            } else {
                finish_walk(sv, r);
            }
        }
    }
    L.parts.clear();
    L.busy = false;
}

// stage a request's inputs in the arena (on the copy stream, which the caller
// synchronises before any launch uses them); false if the arena is full for now
static bool admit(bingo_server *sv, Req &r)
{
    const hipStream_t st = sv->copy_st;
    size_t off = 0;
    if (r.arena_need) {
        if (!sv->arena.alloc(r.arena_need, &off)) return false;
        r.has_arena = true;
    }
    r.arena_off = off;
    r.admitted = true;
    char *p = sv->arena.base + off;
    if (!r.starts.empty()) {
        (void)hipMemcpyAsync(p, r.starts.data(), r.starts.size() * 4, hipMemcpyHostToDevice, st);
        r.d_starts = reinterpret_cast<const int32_t *>(p);
        p += align256(r.starts.size() * 4);
    }
    if (!r.lens.empty()) {
        (void)hipMemcpyAsync(p, r.lens.data(), r.lens.size() * 4, hipMemcpyHostToDevice, st);
        r.d_lens = reinterpret_cast<const int32_t *>(p);
        p += align256(r.lens.size() * 4);
    }
    r.d_paths = r.paths_dev ? r.paths : reinterpret_cast<int32_t *>(p);
    return true;
}

// Pack the schedulable walks (those before the first pending update) into
// launch L in submission order: whole requests while the launch's walker
// budget (max_launch_walkers) and the arena allow, a request larger than the
// budget in chunks.  A launch whose requests are all node2vec, or hold no
// node2vec request, runs the kernel variant of that class (fewer registers,
// more resident waves); a mixed launch runs the variant for any app.  The launch is enqueued while the previous one still
// runs, so under load launches are large and back to back, and a request
// that arrives while the device is busy goes into the next one.  False if
// there was nothing to run.
static bool schedule(bingo_server *sv, Launch &L)
{
    const int64_t budget = sv->cfg.max_launch_walkers;
    int64_t total = 0;
    bool staged = false;
    auto add = [&](const std::shared_ptr<Req> &r, int64_t first, int64_t take) {
        ServeReq q{};
        q.starts = r->d_starts;
        q.lens = r->d_lens;
        q.paths = r->d_paths;
        q.steps = L.d_counts + L.parts.size();
        q.base = (unsigned long long)total;
        q.first = first;
        q.nwalkers = take;
        q.seed = r->args.seed;
        q.thr_return = r->args.thr_return;
        q.thr_neighbor = r->args.thr_neighbor;
        q.thr_other = r->args.thr_other;
        q.thr_stop = r->args.thr_stop;
        q.max_len = r->args.max_len;
        q.max_attempts = r->args.max_attempts;
        q.fold_r = r->args.fold_r;
        q.fold_cap = r->args.fold_cap;
        q.app = r->app;
        L.h_table[L.parts.size()] = q;
        L.parts.push_back(Part{r, first, take, total});
        total += take;
    };
    for (auto &r : sv->order) {
        if (!r->walk || total >= budget || (int)L.parts.size() >= kMaxServeReqs) break;
        if (r->next >= r->nwalkers) continue;  // already in flight
        if (!r->admitted) {
            if (!admit(sv, *r)) break;  // arena full: later requests wait their turn
            staged = true;
        }
        const int64_t take = std::min(r->nwalkers - r->next, budget - total);
        add(r, r->next, take);
        r->next += take;
    }
    if (L.parts.empty()) return false;
    L.total = total;
    int n2v = 0;
    for (const Part &pt : L.parts) n2v += pt.r->app == BINGO_NODE2VEC;
    const int cls = n2v == 0 ? 0 : n2v == (int)L.parts.size() ? 1 : 2;
    // inputs staged here may be read by chunks on any stream from now on
    hipError_t e = staged ? hipStreamSynchronize(sv->copy_st) : hipSuccess;
    bingo_graph *g = sv->g;
    const int n = (int)L.parts.size();
    const int64_t per_block = 256 * serve_walkers_per_lane(cls);
    const int64_t full = (int64_t)g->num_cu * sv->blocks_per_cu[cls];
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(full, (total + per_block - 1) / per_block));
    if (e == hipSuccess) e = hipMemcpyAsync(L.d_table, L.h_table, n * sizeof(ServeReq), hipMemcpyHostToDevice, L.st);
    if (e == hipSuccess) e = hipMemsetAsync(L.d_counts, 0, (kMaxServeReqs + 1) * sizeof(unsigned long long), L.st);
    if (e == hipSuccess) {
        launch_walk_serve(g->s, L.d_table, n, (unsigned long long)total, L.d_counts + kMaxServeReqs, cls, (int)blocks,
                          L.st);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(L.h_counts, L.d_counts, (kMaxServeReqs + 1) * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, L.st);
    if (e == hipSuccess) e = hipEventRecord(L.ev, L.st);
    L.busy = true;
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(L.st);
        harvest(sv, L, e);
        return true;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->st.launches++;
    return true;
}

// One side of an update batch (inserts, or deletes) replayed from the
// captured graphs: staged on the handle's stream, launched, and waited for;
// BINGO_ENOMEM if it did not fit the committed room (then it applied
// nothing), BINGO_EINVAL if it was invalid (likewise).
static int captured_side(bingo_server *sv, bool del, const void *host, int64_t k)
{
    bingo_graph *g = sv->g;
    void *dst = del ? static_cast<void *>(sv->d_dels) : static_cast<void *>(sv->d_ins);
    const size_t bytes = (size_t)k * (del ? sizeof(bingo_del) : sizeof(bingo_edge));
    if (hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, g->st) != hipSuccess) {
        (void)hipGetLastError();
        return fail(BINGO_EDEVICE, "staging an update batch failed");
    }
    int rc = bingo_update_graph_launch(sv->ug, del ? 0 : k, del ? k : 0);
    if (rc) return rc;
    if ((rc = update_wait(g))) g->arcs += del ? k : -k;  // the launch counted arcs it did not apply
    return rc;
}

// An update batch: its inserts, then its deletes, each side applied whole or
// not at all; a side that fails leaves the deletes unapplied too (as
// bingo_insert then bingo_delete).  Each side replays the captured graphs
// when it fits them (no host round trip inside the batch), else -- or when
// the captured side ran past the committed room -- it is applied by
// bingo_insert / bingo_delete, which grow the pools.
static void run_update(bingo_server *sv, const std::shared_ptr<Req> &u)
{
    const double t0 = now_ms();
    bingo_graph *g = sv->g;
    const int64_t ni = (int64_t)u->ins.size(), nd = (int64_t)u->dels.size();
    auto side = [&](bool del, int64_t k) {
        if (!k) return (int)BINGO_OK;
        int rc = BINGO_ENOMEM;
        if (sv->ug && k <= kCapturedMax)
            rc = captured_side(sv, del, del ? static_cast<const void *>(u->dels.data()) : u->ins.data(), k);
        if (rc == BINGO_ENOMEM) rc = del ? bingo_delete(g, u->dels.data(), k) : bingo_insert(g, u->ins.data(), k);
        return rc;
    };
    int rc = side(false, ni);
    if (!rc) rc = side(true, nd);
    if (hipStreamSynchronize(g->st) != hipSuccess && !rc) rc = fail(BINGO_EDEVICE, "update failed on the device");
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->st.busy_s += (now_ms() - t0) / 1e3;
    }
    retire(sv, u);
    complete(sv, u, rc);
}

static void worker_main(bingo_server *sv)
{
    (void)hipSetDevice(sv->g->device);
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            while (!sv->inbox.empty()) {
                sv->order.push_back(sv->inbox.front());
                sv->inbox.pop_front();
            }
            if (sv->stopping && sv->order.empty() && !busy_launches(sv)) return;
        }
        bool did = false;
        for (int i = 0; i < sv->nstreams; ++i) {  // completed launches
            Launch &L = sv->ln[i];
            if (!L.busy) continue;
            const hipError_t e = hipEventQuery(L.ev);
            if (e == hipErrorNotReady) continue;
            harvest(sv, L, e);
            did = true;
        }
        if (!sv->order.empty() && !sv->order.front()->walk) {
            if (!busy_launches(sv)) {  // every earlier walk has finished
                run_update(sv, sv->order.front());
                did = true;
            }
        } else {
            for (int i = 0; i < sv->nstreams; ++i)
                if (!sv->ln[i].busy && schedule(sv, sv->ln[i])) did = true;
        }
        busy_clock(sv);
        if (did) continue;
        // nothing to start now: stage the inputs of the walks that have
        // arrived (submission order, as the arena allows) while launches
        // run, so that packing their launch waits for no copies
        bool staged = false;
        for (auto &r : sv->order) {
            if (!r->walk || r->admitted) continue;
            if (!admit(sv, *r)) break;
            staged = true;
        }
        if (staged) {
            (void)hipStreamSynchronize(sv->copy_st);
            continue;
        }
        const int busy = busy_launches(sv);
        std::unique_lock<std::mutex> lk(sv->mu);
        if (busy)  // poll completions, waking at once for a new request (to stage it)
            sv->cv_work.wait_for(lk, std::chrono::microseconds(50), [&] { return !sv->inbox.empty(); });
        else
            sv->cv_work.wait(lk, [&] { return sv->stopping || !sv->inbox.empty(); });
    }
}

static int submit(bingo_server *sv, std::shared_ptr<Req> r, uint64_t *ticket)
{
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stopping) return fail(BINGO_ESTATE, "server is stopping");
    r->ticket = sv->next_ticket++;
    r->t_submit = now_ms();
    sv->live[r->ticket] = r;
    sv->inbox.push_back(r);
    if (ticket) *ticket = r->ticket;
    sv->cv_work.notify_one();
    return BINGO_OK;
}

static void release_server(bingo_server *sv)
{
    for (auto &L : sv->ln) {
        if (L.st) (void)hipStreamSynchronize(L.st);
        if (L.ev) (void)hipEventDestroy(L.ev);
        if (L.st) (void)hipStreamDestroy(L.st);
        if (L.d_table) (void)hipFree(L.d_table);
        if (L.d_counts) (void)hipFree(L.d_counts);
        if (L.h_table) (void)hipHostFree(L.h_table);
        if (L.h_counts) (void)hipHostFree(L.h_counts);
    }
    if (sv->ug) bingo_update_graph_destroy(sv->ug);
    if (sv->d_ins) (void)hipFree(sv->d_ins);
    if (sv->d_dels) (void)hipFree(sv->d_dels);
    if (sv->copy_st) (void)hipStreamDestroy(sv->copy_st);
    if (sv->arena.base) (void)hipFree(sv->arena.base);
    delete sv;
}

extern "C" {

void bingo_server_config_default(bingo_server_config *c)
{
    if (!c) return;
    c->arena_bytes = 4ull << 30;
    c->max_launch_walkers = 1ll << 22;
    c->blocks_per_cu = 0;
    c->streams = 2;
}

int bingo_server_create(bingo_graph *g, const bingo_server_config *c, bingo_server **out)
{
    if (!out) return fail(BINGO_EINVAL, "null output");
    *out = nullptr;
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    bingo_server *sv = new bingo_server();
    sv->g = g;
    if (c) sv->cfg = *c; else bingo_server_config_default(&sv->cfg);
    if (sv->cfg.arena_bytes < 4096 || sv->cfg.max_launch_walkers < 1 || sv->cfg.streams < 0 ||
        sv->cfg.streams > kMaxStreams) {
        delete sv;
        return fail(BINGO_EINVAL, "server arena, walker budget or stream count out of range");
    }
    sv->nstreams = sv->cfg.streams ? sv->cfg.streams : 2;
    for (int c = 0; c < 3; ++c) {
        const int occ = serve_max_blocks_per_cu(g->s.fw, c, g->s.rate16 != 0);
        sv->blocks_per_cu[c] = sv->cfg.blocks_per_cu > 0 ? std::min(sv->cfg.blocks_per_cu, occ) : occ;
    }
    (void)hipSetDevice(g->device);
    bool ok = hipMalloc((void **)&sv->arena.base, sv->cfg.arena_bytes) == hipSuccess &&
              hipStreamCreateWithFlags(&sv->copy_st, hipStreamNonBlocking) == hipSuccess;
    sv->arena.bytes = sv->cfg.arena_bytes;
    for (int i = 0; ok && i < sv->nstreams; ++i) {
        Launch &L = sv->ln[i];
        ok = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&L.ev, hipEventDisableTiming) == hipSuccess &&
             hipMalloc((void **)&L.d_table, kMaxServeReqs * sizeof(ServeReq)) == hipSuccess &&
             hipMalloc((void **)&L.d_counts, (kMaxServeReqs + 1) * sizeof(unsigned long long)) == hipSuccess &&
             hipHostMalloc((void **)&L.h_table, kMaxServeReqs * sizeof(ServeReq), hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void **)&L.h_counts, (kMaxServeReqs + 1) * sizeof(unsigned long long),
                           hipHostMallocDefault) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        release_server(sv);
        return fail(BINGO_ENOMEM, "server arena, stream or table allocation failed");
    }
    uint64_t room = kCapturedRoom;
    if (const char *e = std::getenv("BINGO_SERVE_ROOM")) room = std::max<uint64_t>(1, std::strtoull(e, nullptr, 0));
    if (hipMalloc((void **)&sv->d_ins, kCapturedMax * sizeof(bingo_edge)) == hipSuccess &&
        hipMalloc((void **)&sv->d_dels, kCapturedMax * sizeof(bingo_del)) == hipSuccess &&
        bingo_update_graph_create(g, sv->d_ins, kCapturedMax, sv->d_dels, kCapturedMax, room, &sv->ug) != BINGO_OK)
        sv->ug = nullptr;  // update batches then all take bingo_insert / bingo_delete
    (void)hipGetLastError();
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
    int rc = fill_walk_args(params, app, &r->args, g->s.fw, g->s.rate16 != 0);
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
    release_server(sv);
}

}  // extern "C"

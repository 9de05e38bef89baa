// Host runtime behind include/bingo.h: handle, HBM pools under a budget,
// batch staging, sort + run encoding of update batches, kernel launches.
//
// Reference roles: build/insert/deleteE host drivers (include/operation.cuh:
// 33-121, 144-197, 215-251), randomWalk (include/graph.cuh:93-150) and the
// MemPool bump allocator (include/gpuMem.cuh:4-32).  Differences by design:
// no per-node host round trips (capacity growth is planned on the device and
// applied in one pass), errors are returned instead of exit(), and all work
// is ordered on one HIP stream per handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bingo.h"
#include "handle.h"
#include "store.h"

using namespace bingo;

namespace bingo {

static thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

}  // namespace bingo

static int ensure_scratch(bingo_graph *g, int idx, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (g->sc_bytes[idx] >= bytes) return BINGO_OK;
    if (g->sc[idx]) {
        HIPCHK(hipStreamSynchronize(g->st));
        HIPCHK(hipFree(g->sc[idx]));
        g->sc[idx] = nullptr;
        g->sc_bytes[idx] = 0;
    }
    // room to spare for batches of varying size (at least 4 MiB, so batches
    // up to a few 10^5 arcs never reallocate); a large buffer is taken exact
    const size_t want = bytes < (256ull << 20) ? std::max<size_t>(bytes + bytes / 4, 4ull << 20) : bytes;
    if (hipMalloc(&g->sc[idx], want) != hipSuccess) {
        (void)hipGetLastError();
        g->sc[idx] = nullptr;
        return fail(BINGO_ENOMEM, "scratch allocation of " + std::to_string(want) + " bytes failed");
    }
    g->sc_bytes[idx] = want;
    return BINGO_OK;
}

static int release_scratch(bingo_graph *g, std::initializer_list<int> which)
{
    HIPCHK(hipStreamSynchronize(g->st));
    for (int idx : which)
        if (g->sc[idx]) {
            HIPCHK(hipFree(g->sc[idx]));
            g->sc[idx] = nullptr;
            g->sc_bytes[idx] = 0;
        }
    return BINGO_OK;
}

template <class T>
static T *scp(bingo_graph *g, int idx)
{
    return static_cast<T *>(g->sc[idx]);
}

static uint64_t store_bytes(const bingo_graph *g)
{
    uint64_t b = g->node_bytes;
    for (int p = 0; p < NPOOL; ++p) b += g->s.cap[p] * kElem[p];
    return b;
}

static void phase_begin(bingo_graph *g, int ph) { (void)hipEventRecord(g->ph[ph][0], g->st); }
static void phase_end(bingo_graph *g, int ph)
{
    (void)hipEventRecord(g->ph[ph][1], g->st);
    g->ph_pending[ph] = true;
}
static void phase_collect(bingo_graph *g)
{
    for (int p = 0; p < NPHASE; ++p) {
        if (!g->ph_pending[p]) continue;
        float ms = 0.f;
        if (hipEventSynchronize(g->ph[p][1]) == hipSuccess &&
            hipEventElapsedTime(&ms, g->ph[p][0], g->ph[p][1]) == hipSuccess)
            g->ph_ms[p] = ms;
        g->ph_pending[p] = false;
    }
}

static int read_ctl(bingo_graph *g)
{
    HIPCHK(hipMemcpyAsync(g->ctl_host, g->s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

// A captured update batch that applied nothing (read by read_ctl) is reported
// once: the status is cleared and its error returned.
static int take_ustatus(bingo_graph *g)
{
    const unsigned u = g->ctl_host->ustatus;
    if (!u) return BINGO_OK;
    launch_clear_ustatus(g->s.ctl, g->st);
    g->ctl_host->ustatus = 0;
    HIPCHK(hipStreamSynchronize(g->st));
    const std::string rest = "; it and the captured batches after it were not applied (an insert or delete that "
                             "reports this is not applied either)";
    if (u & UST_INVAL) return fail(BINGO_EINVAL, "a captured update batch was invalid" + rest);
    return fail(BINGO_ENOMEM, "a captured update batch did not fit the pools' committed room" + rest);
}

namespace bingo {
int update_wait(bingo_graph *g)
{
    int rc = read_ctl(g);
    if (rc) return rc;
    if (g->ctl_host->error & ERR_OVERRUN) return fail(BINGO_EDEVICE, "pool overrun detected");
    return take_ustatus(g);
}
}  // namespace bingo

// grow_in_place: pool p is a reserved address range (the HBM budget, at
// most 2^36 elements) mapped from its start in 1 GiB chunks; growing maps
// more chunks of physical HBM behind the mapped part.  Returns BINGO_OK, a negative error,
// or 1 when the device refuses the reservation (the caller copies instead).
static int grow_mapped(bingo_graph *g, int p, uint64_t cap_elems)
{
    auto &v = g->vp[p];
    hipMemAllocationProp prop{};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = g->device;
    size_t gran = 0;
    if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended) != hipSuccess || !gran) {
        (void)hipGetLastError();
        // refused before anything was mapped: the caller copies instead; a
        // mapped pool cannot move to a copied allocation
        return v.va ? fail(BINGO_EDEVICE, "pool mapping granularity query failed") : 1;
    }
    // the range and every chunk in whole 1 GiB units at 1 GiB offsets (chunks
    // of other sizes after the first were refused on this part; the reported
    // granularity is a 4 KiB page)
    gran = std::max<size_t>(gran, 1ull << 30);
    auto up = [gran](uint64_t x) { return (x + gran - 1) / gran * gran; };
    if (!v.va) {
        if (g->refuse_vmm & (1u << p)) return 1;  // test hook (BINGO_REFUSE_VMM)
        const uint64_t limit = std::min<uint64_t>(g->budget, (1ull << kOffBits) * kElem[p]);
        if (hipMemAddressReserve(&v.va, up(limit), gran, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            v.va = nullptr;
            return 1;
        }
        v.reserved = up(limit);
    }
    const uint64_t want = up(cap_elems * kElem[p]);
    if (want > v.reserved) return fail(BINGO_ENOMEM, "pool would exceed its reserved address range");
    while (want > v.mapped) {
        const size_t delta = gran;
        char *at = static_cast<char *>(v.va) + v.mapped;
        hipMemGenericAllocationHandle_t h{};
        if (hipMemCreate(&h, delta, &prop, 0) != hipSuccess) {
            (void)hipGetLastError();
            return fail(BINGO_ENOMEM, "pool growth of " + std::to_string(delta) + " bytes failed");
        }
        hipMemAccessDesc acc{};
        acc.location = prop.location;
        acc.flags = hipMemAccessFlagsProtReadWrite;
        if (hipMemMap(at, delta, 0, h, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemRelease(h);
            return fail(BINGO_EDEVICE, "pool mapping failed");
        }
        const hipError_t ea = hipMemSetAccess(at, delta, &acc, 1);
        if (ea != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemUnmap(at, delta);
            (void)hipMemRelease(h);
            return fail(BINGO_EDEVICE, std::string("pool access setup failed (") + hipGetErrorString(ea) + ", pool " +
                                           std::to_string(p) + ", mapped " + std::to_string(v.mapped) + " + " +
                                           std::to_string(delta) + " bytes, " + std::to_string(v.chunks.size()) +
                                           " chunks)");
        }
        v.chunks.emplace_back(h, delta);
        v.mapped += delta;
    }
    return BINGO_OK;
}

static void release_mapped(bingo_graph *g)
{
    for (auto &v : g->vp) {
        if (!v.va) continue;
        size_t at = 0;
        for (auto &c : v.chunks) {
            (void)hipMemUnmap(static_cast<char *>(v.va) + at, c.second);
            (void)hipMemRelease(c.first);
            at += c.second;
        }
        (void)hipMemAddressFree(v.va, v.reserved);
        v = bingo_graph::VPool{};
    }
}

// Whether pool p grows by mapping (a reserved range that is mapped already,
// or a first allocation of a grow_in_place handle whose device has not
// refused a reservation).  A pool that is mapped stays mapped: it can never
// move to a copied allocation.
static bool pool_mapped(const bingo_graph *g, int p, const void *cur)
{
    return g->vp[p].va || (g->cfg.grow_in_place && !g->vmm_refused && !cur);
}

// HBM a pool of cap elements holds: mapped pools take whole 1 GiB chunks
static uint64_t pool_footprint(const bingo_graph *g, int p, const void *cur, uint64_t cap)
{
    const uint64_t b = cap * kElem[p];
    if (!pool_mapped(g, p, cur)) return b;
    const uint64_t gib = 1ull << 30;
    return (b + gib - 1) / gib * gib;
}

static void **pool_slot(bingo_graph *g, int p)
{
    return p == POOL_SLOT   ? (void **)&g->s.slot
           : p == POOL_BACK ? (void **)&g->s.back
           : p == POOL_MEM  ? (void **)&g->s.mem
                            : (void **)&g->s.hash;
}

// Grow pools so that top + need fits; a grown pool is either mapped further
// in place (grow_in_place) or a new allocation with the old contents copied
// (region offsets stay valid either way).  Refuses growth past the HBM
// budget with BINGO_ENOMEM before anything is modified; a batch that grows
// no pool is never refused.
static int ensure_pools(bingo_graph *g)
{
    const Ctl &c = *g->ctl_host;
    for (int p = 0; p < NPOOL; ++p)  // alias slots address members and tables with 36 bits
        if (c.top[p] + c.need[p] > (1ull << kOffBits))
            return fail(BINGO_ENOMEM, "a pool would exceed 2^36 elements");
    uint64_t newcap[NPOOL];
    bool grows = false;
    for (int p = 0; p < NPOOL; ++p) {
        const uint64_t want = c.top[p] + c.need[p];
        newcap[p] = g->s.cap[p];
        if (want > g->s.cap[p]) {
            // a copied pool grows by at least half (amortised copies); a
            // mapped one by at least an eighth (fewer mapping calls)
            const uint64_t slack = (uint64_t)((double)want * g->cfg.pool_slack);
            const uint64_t step = g->vp[p].va ? g->s.cap[p] / 8 : g->s.cap[p] / 2;
            newcap[p] = std::min<uint64_t>(std::max<uint64_t>(want + slack + 1024, g->s.cap[p] + step),
                                           1ull << kOffBits);
            grows = true;
        }
    }
    if (!grows) return BINGO_OK;
    auto total_of = [&](const uint64_t *cap) {
        uint64_t t = g->node_bytes;
        for (int p = 0; p < NPOOL; ++p) t += pool_footprint(g, p, *pool_slot(g, p), cap[p]);
        return t;
    };
    uint64_t total = total_of(newcap);
    if (total > g->budget) {  // retry without slack
        for (int p = 0; p < NPOOL; ++p) {
            const uint64_t want = c.top[p] + c.need[p];
            if (want > g->s.cap[p]) newcap[p] = want;
        }
        total = total_of(newcap);
        if (total > g->budget)
            return fail(BINGO_ENOMEM, "store would need " + std::to_string(total) + " bytes, budget is " +
                                          std::to_string(g->budget));
    }
    for (int p = 0; p < NPOOL; ++p) {
        if (newcap[p] == g->s.cap[p]) continue;
        void **slotp = pool_slot(g, p);
        if (pool_mapped(g, p, *slotp)) {
            const int rc = grow_mapped(g, p, newcap[p]);
            if (rc < 0) return rc;
            if (rc == 0) {
                const bool first = !*slotp;
                *slotp = g->vp[p].va;
                g->s.cap[p] = std::min<uint64_t>(g->vp[p].mapped / kElem[p], 1ull << kOffBits);
                if (first) ++g->store_gen;
                continue;
            }
            g->vmm_refused = true;  // refused before mapping: this and later new pools are copied
        }
        void *fresh = nullptr;
        if (hipMalloc(&fresh, newcap[p] * kElem[p]) != hipSuccess) {
            (void)hipGetLastError();
            return fail(BINGO_ENOMEM, "pool allocation of " + std::to_string(newcap[p] * kElem[p]) + " bytes failed");
        }
        if (*slotp) {
            if (c.top[p])
                HIPCHK(hipMemcpyAsync(fresh, *slotp, c.top[p] * kElem[p], hipMemcpyDeviceToDevice, g->st));
            HIPCHK(hipStreamSynchronize(g->st));
            HIPCHK(hipFree(*slotp));
        }
        *slotp = fresh;
        g->s.cap[p] = newcap[p];
        ++g->store_gen;  // captured walk graphs hold the old pool addresses
    }
    return BINGO_OK;
}

static int bits_for(uint64_t maxval)
{
    int b = 1;
    while (b < 64 && (maxval >> b)) ++b;
    return b;
}

// Keys are node ids in 0..n (0 for a rejected entry), so a batch has at most
// min(k, n + 1) runs.
static int64_t max_runs(const bingo_graph *g, int64_t k) { return std::min<int64_t>(k, (int64_t)g->s.n + 1); }

// rocPRIM temporary storage for run-length encoding k sorted keys into at
// most kr runs and scanning the run counts (sizes only: nothing is touched)
static size_t encode_runs_tmp(int64_t k, int64_t kr, hipStream_t st)
{
    return std::max(run_length_encode_u32(nullptr, 0, nullptr, k, nullptr, nullptr, nullptr, st),
                    exclusive_scan_u32(nullptr, 0, nullptr, nullptr, kr, st));
}

// run-length encode sorted node keys into the Runs arrays (node, start,
// count), with rocPRIM's temporary storage in tmp (encode_runs_tmp bytes)
static int encode_runs_in(bingo_graph *g, const uint32_t *keys_sorted, int64_t k, uint32_t *uniq, uint32_t *counts,
                          uint32_t *starts, uint32_t *nruns, void *tmp, size_t tmp_bytes, Runs *r)
{
    const int64_t kr = max_runs(g, k);
    run_length_encode_u32(tmp, tmp_bytes, keys_sorted, k, uniq, counts, nruns, g->st);
    exclusive_scan_u32(tmp, tmp_bytes, counts, starts, kr, g->st);
    HIPCHK(hipGetLastError());
    r->node = uniq;
    r->start = starts;
    r->count = counts;
    r->nruns = nruns;
    return BINGO_OK;
}

static int encode_runs(bingo_graph *g, const uint32_t *keys_sorted, int64_t k, Runs *r)
{
    int rc;
    if ((rc = ensure_scratch(g, SC_TMP, encode_runs_tmp(k, max_runs(g, k), g->st)))) return rc;
    return encode_runs_in(g, keys_sorted, k, scp<uint32_t>(g, SC_UNIQ), scp<uint32_t>(g, SC_COUNTS),
                          scp<uint32_t>(g, SC_STARTS), scp<uint32_t>(g, SC_NRUNS), g->sc[SC_TMP], g->sc_bytes[SC_TMP],
                          r);
}

// scratch of one insert (del false) or delete batch of k entries
static int ensure_batch_scratch(bingo_graph *g, int64_t k, bool del)
{
    int rc;
    const size_t k4 = (size_t)k * 4, k8 = (size_t)k * 8, r4 = (size_t)max_runs(g, k) * 4;
    if ((rc = ensure_scratch(g, SC_KEYS2, k4)) || (rc = ensure_scratch(g, SC_VALS2, k4)) ||
        (rc = ensure_scratch(g, SC_UNIQ, r4)) || (rc = ensure_scratch(g, SC_COUNTS, r4)) ||
        (rc = ensure_scratch(g, SC_STARTS, r4)) || (rc = ensure_scratch(g, SC_NRUNS, 16)))
        return rc;
    if (del) {
        if ((rc = ensure_scratch(g, SC_U64A, k8)) || (rc = ensure_scratch(g, SC_U64B, k8))) return rc;
    } else {
        if ((rc = ensure_scratch(g, SC_KEYS, k4)) || (rc = ensure_scratch(g, SC_VALS, k4)) ||
            (rc = ensure_scratch(g, SC_EDGES, (size_t)k * sizeof(DevEdge))) ||
            (g->s.fw && (rc = ensure_scratch(g, SC_FRAC, k4))))
            return rc;
    }
    return BINGO_OK;
}

// the next free lists (s.fl2) hold at least `entries`; their old content is
// not kept (the merge rewrites them whole)
static int ensure_fl2(bingo_graph *g, uint64_t entries)
{
    if (entries <= g->flcap[1] && g->flbuf[1]) return BINGO_OK;
    if (g->flbuf[1]) {
        HIPCHK(hipStreamSynchronize(g->st));
        HIPCHK(hipFree(g->flbuf[1]));
        g->flbuf[1] = nullptr;
        g->flcap[1] = 0;
    }
    const uint64_t want = std::max<uint64_t>(entries + entries / 4, 1u << 16);
    if (hipMalloc((void **)&g->flbuf[1], want * sizeof(uint64_t)) != hipSuccess) {
        (void)hipGetLastError();
        g->flbuf[1] = nullptr;
        return fail(BINGO_ENOMEM, "free-list allocation failed");
    }
    g->flcap[1] = want;
    return BINGO_OK;
}

static void set_lists(bingo_graph *g)
{
    g->s.fl = g->flbuf[0];
    g->s.fl2 = g->flbuf[1];
}

// arcs or deletes per batch (a batch indexes its entries in 31 bits); arcs
// per batch of a load
constexpr int64_t kMaxBatch = (1ll << 31) - 1;
constexpr int64_t kLoadChunk = 1ll << 30;

static int do_insert(bingo_graph *g, const DevEdge *d_e, int64_t k)
{
    int rc;
    if (k < 0 || k > kMaxBatch) return fail(BINGO_EINVAL, "batch size out of range");
    if (k == 0) return BINGO_OK;
    if ((rc = ensure_batch_scratch(g, k, false))) return rc;
    if ((rc = ensure_scratch(g, SC_WCLS, insert_wcls_bytes(k)))) return rc;
    g->s.wcls = scp<uint32_t>(g, SC_WCLS);
    launch_alloc_reset(g->s.fm, g->st);
    launch_reset_ctl(g->s.ctl, g->st);
    launch_insert_keys(d_e, k, nullptr, g->s.n, g->s.fw, scp<uint32_t>(g, SC_KEYS), scp<uint32_t>(g, SC_VALS), g->s.ctl,
                       g->st);
    const int bits = bits_for((uint64_t)g->s.n);
    uint32_t *k0 = scp<uint32_t>(g, SC_KEYS), *k1 = scp<uint32_t>(g, SC_KEYS2);
    uint32_t *v0 = scp<uint32_t>(g, SC_VALS), *v1 = scp<uint32_t>(g, SC_VALS2);
    size_t tb = sort_pairs_u32(nullptr, 0, k0, k1, v0, v1, k, bits, g->st, nullptr);
    if ((rc = ensure_scratch(g, SC_TMP, tb))) return rc;
    bool second = false;
    sort_pairs_u32(g->sc[SC_TMP], g->sc_bytes[SC_TMP], k0, k1, v0, v1, k, bits, g->st, &second);
    uint32_t *frac = g->s.fw ? scp<uint32_t>(g, SC_FRAC) : nullptr;
    launch_gather_edges(d_e, second ? v1 : v0, k, nullptr, scp<DevEdge>(g, SC_EDGES), frac, g->st);
    Runs r;
    if ((rc = encode_runs(g, second ? k1 : k0, k, &r))) return rc;
    launch_insert_plan(g->s, scp<DevEdge>(g, SC_EDGES), r, k, g->st);
    launch_alloc_resolve(g->s, g->s.fw, g->st);
    HIPCHK(hipGetLastError());
    if ((rc = read_ctl(g)) || (rc = take_ustatus(g))) return rc;
    if (g->ctl_host->error & ERR_INVAL)
        return fail(BINGO_EINVAL, g->s.fw ? "insert batch has an arc with u or v outside 1..n or a float weight that is "
                                            "not finite, not in (0, 2^32) or below 2^-32"
                                          : "insert batch has an arc with u or v outside 1..n or w == 0");
    if (g->ctl_host->error & ERR_NOMEM)
        return fail(BINGO_ENOMEM, "a node would reach the per-node limits (2^28 slots, integer weight 2^59)");
    if (g->ctl_host->error & ERR_OVERRUN)
        return fail(BINGO_EDEVICE, "an earlier batch overran a pool; the store is no longer consistent");
    if (g->ctl_host->error) return fail(BINGO_EDEVICE, "store reported an internal error");
    if ((rc = ensure_pools(g))) return rc;
    if ((rc = ensure_scratch(g, SC_PF, g->ctl_host->pend_total * sizeof(uint64_t))) ||
        (rc = ensure_fl2(g, g->ctl_host->fl_total)))
        return rc;
    g->s.pf = scp<uint64_t>(g, SC_PF);
    set_lists(g);
    launch_insert_apply(g->s, scp<DevEdge>(g, SC_EDGES), frac, r, k, g->st);
    // moves of large regions (hubs) by the whole grid, before the neighbour
    // tables take the batch's arcs
    if (g->ctl_host->nbigcopy) launch_copy_tasks(g->s, g->st);
    if (g->s.nbr) launch_hash_insert(g->s, scp<DevEdge>(g, SC_EDGES), k, g->st);
    // the batch's frees join the lists (after the copies that read them)
    launch_freelist_merge(g->s, g->st);
    std::swap(g->flbuf[0], g->flbuf[1]);
    std::swap(g->flcap[0], g->flcap[1]);
    set_lists(g);
    HIPCHK(hipGetLastError());
    g->arcs += k;
    return BINGO_OK;
}

// Device buffers of one delete batch of k entries.
struct DelWork {
    unsigned long long *ka, *kb;  // keys (u << 32 | slot), then sorted
    uint32_t *us, *slots;         // checked node and slot of each sorted delete
    uint32_t *uniq, *counts, *starts, *nruns;  // runs per node
    uint32_t *big;                // long runs
    void *tmp;                    // rocPRIM temporary storage
    size_t tmp_bytes;
};

static size_t del_tmp_bytes(bingo_graph *g, int64_t k)
{
    const int bits = 32 + bits_for((uint64_t)g->s.n);
    return std::max(sort_keys_u64(nullptr, 0, nullptr, nullptr, k, bits, g->st), encode_runs_tmp(k, max_runs(g, k), g->st));
}

// the buffers of a delete batch from the handle's scratch (allocated as needed)
static int del_work_scratch(bingo_graph *g, int64_t k, DelWork *w)
{
    int rc;
    if ((rc = ensure_batch_scratch(g, k, true)) || (rc = ensure_scratch(g, SC_TMP, del_tmp_bytes(g, k))) ||
        (rc = ensure_scratch(g, SC_BIG, ((size_t)k / kBigDelete + 1) * sizeof(uint32_t))))
        return rc;
    *w = DelWork{scp<unsigned long long>(g, SC_U64A), scp<unsigned long long>(g, SC_U64B), scp<uint32_t>(g, SC_KEYS2),
                 scp<uint32_t>(g, SC_VALS2),          scp<uint32_t>(g, SC_UNIQ),             scp<uint32_t>(g, SC_COUNTS),
                 scp<uint32_t>(g, SC_STARTS),         scp<uint32_t>(g, SC_NRUNS),            scp<uint32_t>(g, SC_BIG),
                 g->sc[SC_TMP],                       g->sc_bytes[SC_TMP]};
    return BINGO_OK;
}

static int do_delete(bingo_graph *g, const int32_t *d_pairs, int64_t k, const DelWork &w)
{
    int rc;
    launch_reset_ctl(g->s.ctl, g->st);
    launch_delete_keys(d_pairs, k, nullptr, g->s.n, g->s.ctl, w.ka, g->st);
    const int bits = 32 + bits_for((uint64_t)g->s.n);
    sort_keys_u64(w.tmp, w.tmp_bytes, w.ka, w.kb, k, bits, g->st);
    launch_delete_check(g->s, w.kb, k, 0, w.us, w.slots, g->st);
    Runs r;
    if ((rc = encode_runs_in(g, w.us, k, w.uniq, w.counts, w.starts, w.nruns, w.tmp, w.tmp_bytes, &r))) return rc;
    // long runs (more than kBigDelete deletes)
    launch_delete_big_list(r, k, w.big, g->s.ctl, g->st);
    if ((rc = read_ctl(g)) || (rc = take_ustatus(g))) return rc;
    if (g->ctl_host->error & ERR_INVAL)
        return fail(BINGO_EINVAL, "delete batch names a node outside 1..n, a slot that is not live, or a slot twice");
    if (g->ctl_host->error & ERR_OVERRUN)
        return fail(BINGO_EDEVICE, "an earlier batch overran a pool; the store is no longer consistent");
    if (g->ctl_host->error) return fail(BINGO_EDEVICE, "store reported an internal error");
    // neighbour entries first: they subtract each deleted slot's weight,
    // which the apply kernel then tombstones
    if (g->s.nbr) launch_hash_delete(g->s, w.us, w.slots, k, g->st);
    launch_delete_apply(g->s, w.slots, r, w.big, k, g->st);
    HIPCHK(hipGetLastError());
    g->arcs -= k;
    return BINGO_OK;
}

static int do_delete(bingo_graph *g, const int32_t *d_pairs, int64_t k)
{
    if (k < 0 || k > kMaxBatch) return fail(BINGO_EINVAL, "batch size out of range");
    if (k == 0) return BINGO_OK;
    DelWork w;
    int rc = del_work_scratch(g, k, &w);
    return rc ? rc : do_delete(g, d_pairs, k, w);
}

// ------------------------------------------------------------------------
// Allocation-free deletes: the batch's buffers live in `scratch`, one
// statically sized device array per GPU (each device holds its own copy at
// its own address, looked up once per device by hipGetSymbolAddress when the
// first handle on that device is created).  The host-side name `scratch` is
// only a stand-in for the symbol: host code passes the looked-up address.
// Every call on a device shares that device's array: calls that overlap on
// one GPU (two handles on the same device, from two threads) race.  Nothing
// in it needs to start at zero (the batch's counters live in the handle).
// ------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[BINGO_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// the device's scratch address (the device current); looked up once per device
static int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(BINGO_EINVAL, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return BINGO_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(BINGO_EDEVICE, std::string("scratch symbol lookup failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return BINGO_OK;
}

// the buffers of a delete batch carved from the device's scratch; k must fit
static int del_work_static(bingo_graph *g, int64_t k, DelWork *w, void **stage)
{
    unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[g->device]);
    if (!base) return fail(BINGO_ESTATE, "no scratch on this device");
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t k4 = al((size_t)k * 4), k8 = al((size_t)k * 8), r4 = al((size_t)max_runs(g, k) * 4);
    const size_t tmp = del_tmp_bytes(g, k);
    const size_t need = 3 * k8 + 2 * k4 + 3 * r4 + 256 + al(((size_t)k / kBigDelete + 1) * 4) + al(tmp);
    if (need > BINGO_NOALLOC_SCRATCH_BYTES)
        return fail(BINGO_EINVAL, "delete batch needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                      std::to_string(BINGO_NOALLOC_SCRATCH_BYTES) + " of the allocation-free path");
    unsigned char *p = base;
    auto take = [&](size_t b) {
        unsigned char *q = p;
        p += b;
        return q;
    };
    *stage = take(k8);
    w->ka = reinterpret_cast<unsigned long long *>(take(k8));
    w->kb = reinterpret_cast<unsigned long long *>(take(k8));
    w->us = reinterpret_cast<uint32_t *>(take(k4));
    w->slots = reinterpret_cast<uint32_t *>(take(k4));
    w->uniq = reinterpret_cast<uint32_t *>(take(r4));
    w->counts = reinterpret_cast<uint32_t *>(take(r4));
    w->starts = reinterpret_cast<uint32_t *>(take(r4));
    w->nruns = reinterpret_cast<uint32_t *>(take(256));
    w->big = reinterpret_cast<uint32_t *>(take(al(((size_t)k / kBigDelete + 1) * 4)));
    w->tmp = take(al(tmp));
    w->tmp_bytes = al(tmp);
    return BINGO_OK;
}

// How a walk call is scheduled: the plain form (one thread per walker) or
// the persistent form with explicit, tuned or default knobs.
struct WalkChoice {
    int form, block, wpl, depth;
};

static bool choice_ok(const WalkChoice &c, int32_t max_len)
{
    const bool blk = c.block == 64 || c.block == 128 || c.block == 256 || c.block == 512;
    if (c.form == BINGO_FORM_PLAIN) return blk;
    return blk && (c.wpl == 1 || c.wpl == 2 || c.wpl == 4) && (c.depth == 0 || c.depth == 4 || c.depth == 16) &&
           (c.depth == 0 || max_len % c.depth == 0);
}

static int resolve_walk(bingo_graph *g, int app, const bingo_walk_params *p, WalkChoice *c)
{
    if (p->form == BINGO_FORM_PLAIN) {
        *c = WalkChoice{BINGO_FORM_PLAIN, p->block ? p->block : 256, 1, 0};
    } else if (p->form == BINGO_FORM_AUTO && g->tuned[app].valid && g->tuned[app].max_len == p->max_len) {
        const auto &t = g->tuned[app];
        *c = WalkChoice{t.form, t.block, t.wpl, t.depth};
    } else if (p->form == BINGO_FORM_AUTO || p->form == BINGO_FORM_PERSISTENT) {
        *c = WalkChoice{BINGO_FORM_PERSISTENT, p->block ? p->block : 256, p->walkers_per_lane ? p->walkers_per_lane : 2,
                        p->lds_depth >= 0 ? p->lds_depth : 0};
    } else {
        return fail(BINGO_EINVAL, "unknown walk form");
    }
    if (!choice_ok(*c, p->max_len))
        return fail(BINGO_EINVAL, "walk knobs: block 64..512 (power of two), walkers_per_lane 1/2/4, lds_depth "
                                  "0/4/16 dividing max_len");
    return BINGO_OK;
}

static int launch_choice(bingo_graph *g, int app, const WalkChoice &c, const WalkArgs &a, int32_t blocks_per_cu)
{
    if (c.form == BINGO_FORM_PLAIN) {
        launch_walk(g->s, app, a, c.block, g->st);
    } else {
        int nb = persist_max_blocks_per_cu(app, c.wpl, c.depth, c.block, g->s.fw, g->s.rate16 != 0);
        if (nb < 1) return fail(BINGO_EINVAL, "persistent walk configuration does not fit a CU");
        if (blocks_per_cu > 0 && blocks_per_cu < nb) nb = blocks_per_cu;
        const int64_t per_block = (int64_t)c.block * c.wpl;
        int64_t blocks = (int64_t)g->num_cu * nb;
        blocks = std::min<int64_t>(blocks, (a.nwalkers + per_block - 1) / per_block);
        PersistCfg pc{c.block, c.wpl, c.depth, (int)std::max<int64_t>(blocks, 1)};
        launch_walk_persist(g->s, app, a, pc, g->st);
    }
    HIPCHK(hipGetLastError());
    return BINGO_OK;
}

static uint32_t thr27(double prob)
{
    if (!(prob > 0)) return 0;
    if (prob >= 1.0) return 1u << 27;
    return (uint32_t)std::floor(prob * 134217728.0);
}

int bingo::fill_walk_args(const bingo_walk_params *p, int app, WalkArgs *a, int fw, int dense)
{
    if (!p) return fail(BINGO_EINVAL, "null walk params");
    if (app < 0 || app > 3) return fail(BINGO_EINVAL, "unknown app");
    if (p->max_len < 1) return fail(BINGO_EINVAL, "max_len must be >= 1");
    if (app == BINGO_NODE2VEC && (p->max_attempts < 1 || !(p->p > 0) || !(p->q > 0)))
        return fail(BINGO_EINVAL, "node2vec needs p, q > 0 and max_attempts >= 1");
    // the fractional stream's counter holds the attempt in 16 bits
    if (fw && p->max_attempts > kMaxAttemptsFloat)
        return fail(BINGO_EINVAL, "max_attempts above 65536 with float weights");
    // so does the dense-group rejection stream's
    if (dense && p->max_attempts > kMaxAttemptsDense)
        return fail(BINGO_EINVAL, "max_attempts above 65536 with dense groups (dense_rate)");
    if (app == BINGO_PPR && !(p->tp >= 0 && p->tp <= 1)) return fail(BINGO_EINVAL, "tp must be in [0, 1]");
    // node2vec: alpha = 1/p (return), 1 (distance 1), 1/q (distance 2)
    // (ref operation.cuh:291,302-325; lim there is max(1/p, 1/q)).  A proposal
    // is accepted with alpha / env, env = max(1, 1/q); when 1/p > env the
    // return arc's excess is folded into a region of its own (fold_r, x 2^27)
    // that each attempt takes outright with probability w_ret fold_r /
    // (S 2^27 + w_ret fold_r).  Float-weight stores, and factors
    // (1/p - env) / env of 2^32 or more, keep the plain envelope
    // max(1/p, 1, 1/q): folding needs integer weight sums.
    const double ip = 1.0 / p->p, iq = 1.0 / p->q, env = std::max(1.0, iq);
    if (!fw && ip > env && (ip - env) / env < 4294967296.0) {
        a->thr_return = 1u << 27;
        a->thr_neighbor = thr27(1.0 / env);
        a->thr_other = thr27(iq / env);
        a->fold_r = (uint64_t)std::floor((ip - env) / env * 134217728.0);
        // a return weight is at most its node's weight, so every threshold
        // is at most 2^64 R / (2^27 + R) (+ 33 for the 60-bit cut); the
        // kernels skip the weight lookup for coins at or above this bound
        typedef unsigned __int128 u128;
        const u128 q = ((u128)a->fold_r << 64) / (((u128)1 << 27) + a->fold_r);
        a->fold_cap = q >= ((u128)1 << 64) - 128 ? ~0ull : (uint64_t)q + 128u;
    } else {
        const double lim = std::max(ip, env);
        a->thr_return = thr27(ip / lim);
        a->thr_neighbor = thr27(1.0 / lim);
        a->thr_other = thr27(iq / lim);
        a->fold_r = 0;
        a->fold_cap = 0;
    }
    a->thr_stop = thr27(p->tp);
    a->seed = p->seed;
    a->max_len = p->max_len;
    a->max_attempts = p->max_attempts;
    return BINGO_OK;
}

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------
extern "C" {

void bingo_config_default(bingo_config *cfg)
{
    if (!cfg) return;
    cfg->device = 0;
    cfg->neighbor_index = 1;
    cfg->hbm_budget = 0;
    cfg->pool_slack = 0.25;
    cfg->min_slots = 16;  // ref ORIGINAL (include/api.h:13)
    cfg->fit_load = 0;
    cfg->fit_headroom = 1.0 / 16;
    cfg->grow_in_place = 0;
    cfg->float_weights = 0;
    cfg->dense_rate = 0.0;  // every non-empty group keeps a member list (ref RATE > 1)
}

void bingo_walk_params_default(bingo_walk_params *p)
{
    if (!p) return;
    p->seed = 1;
    p->p = 0.5;          // ref P (include/api.h:8)
    p->q = 2.0;          // ref Q (include/api.h:9)
    p->tp = 0.0125;      // ref TP (include/api.h:20)
    p->max_len = 80;     // ref LEN (include/api.h:12)
    p->max_attempts = 4096;
    p->form = BINGO_FORM_AUTO;
    p->block = 0;
    p->walkers_per_lane = 0;
    p->lds_depth = -1;
    p->blocks_per_cu = 0;
    p->reserved = 0;
}

const char *bingo_last_error(void) { return bingo::g_err.c_str(); }
const char *bingo_version(void) { return "bingo-amd 0.1 gfx950"; }

int bingo_create(const bingo_config *cfg, bingo_graph **out)
{
    if (!out) return fail(BINGO_EINVAL, "null output handle");
    *out = nullptr;
    bingo_graph *g = new bingo_graph();
    if (cfg) g->cfg = *cfg; else bingo_config_default(&g->cfg);
    if (g->cfg.min_slots < 1 || g->cfg.min_slots > (1 << 20) || (g->cfg.min_slots & (g->cfg.min_slots - 1))) {
        delete g;
        return fail(BINGO_EINVAL, "min_slots must be a power of two in [1, 2^20]");
    }
    if (!(g->cfg.pool_slack >= 0) || g->cfg.pool_slack > 16) g->cfg.pool_slack = 0.25;
    if (g->cfg.fit_load && !(g->cfg.fit_headroom >= 0 && g->cfg.fit_headroom <= 1)) {
        delete g;
        return fail(BINGO_EINVAL, "fit_headroom must be in [0, 1]");
    }
    if (!(g->cfg.dense_rate >= 0)) {
        delete g;
        return fail(BINGO_EINVAL, "dense_rate must be >= 0");
    }
    g->device = g->cfg.device;
    // test hook: treat the address reservation of the pools in this bit mask
    // as refused (exercises the copy fallback next to mapped pools)
    if (const char *e = std::getenv("BINGO_REFUSE_VMM")) g->refuse_vmm = (unsigned)std::strtoul(e, nullptr, 0);
    // test hook: a smaller sbase limit (exercises the lowering of member slot
    // fields by dead-prefix compactions on small graphs)
    g->s.sbase_max = kSbaseMax;
    if (const char *e = std::getenv("BINGO_SBASE_MAX")) {
        const long long c = std::strtoll(e, nullptr, 0);
        if (c >= 0 && c <= (long long)kSbaseMax) g->s.sbase_max = (uint32_t)c;
    }
    // test hook: a smaller load batch (exercises the batched load on small graphs)
    g->load_chunk = kLoadChunk;
    if (const char *e = std::getenv("BINGO_LOAD_CHUNK")) {
        const long long c = std::strtoll(e, nullptr, 0);
        if (c >= 1 && c <= kLoadChunk) g->load_chunk = c;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete g;
        return fail(BINGO_EDEVICE, "no HIP device");
    }
    if (g->device < 0 || g->device >= ndev) {
        delete g;
        return fail(BINGO_EINVAL, "device ordinal out of range");
    }
    if (hipSetDevice(g->device) != hipSuccess || hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return fail(BINGO_EDEVICE, "cannot create HIP stream");
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    (void)hipDeviceGetAttribute(&g->num_cu, hipDeviceAttributeMultiprocessorCount, g->device);
    if (g->num_cu < 1) g->num_cu = 1;
    g->budget = g->cfg.hbm_budget ? g->cfg.hbm_budget : (uint64_t)(0.9 * (double)fr);
    bool ok = hipHostMalloc((void **)&g->ctl_host, sizeof(Ctl), hipHostMallocDefault) == hipSuccess &&
              hipMalloc((void **)&g->s.ctl, sizeof(Ctl)) == hipSuccess &&
              hipMalloc((void **)&g->s.tasks, kMaxCopyTasks * sizeof(CopyTask)) == hipSuccess &&
              hipMalloc((void **)&g->s.fm, sizeof(FreeMeta)) == hipSuccess &&
              hipMalloc((void **)&g->d_steps, sizeof(unsigned long long)) == hipSuccess;
    for (int i = 0; ok && i < kEvents; ++i) ok = hipEventCreate(&g->ev[i]) == hipSuccess;
    for (int p = 0; ok && p < NPHASE; ++p)
        ok = hipEventCreate(&g->ph[p][0]) == hipSuccess && hipEventCreate(&g->ph[p][1]) == hipSuccess;
    if (!ok) {
        bingo_destroy(g);
        return fail(BINGO_EDEVICE, "handle allocation failed");
    }
    if (int rc = scratch_setup(g->device)) {
        bingo_destroy(g);
        return rc;
    }
    (void)hipMemset(g->s.ctl, 0, sizeof(Ctl));
    (void)hipMemset(g->s.fm, 0, sizeof(FreeMeta));
    g->s.min_slots = g->cfg.min_slots;
    g->s.nbr = g->cfg.neighbor_index ? 1 : 0;
    g->s.fw = g->cfg.float_weights ? 1 : 0;
    // the rate in units of 2^-16; 0 (off) for 0 and for 1 or more, where
    // every non-empty group keeps a list (a count never exceeds its run)
    const double r = g->cfg.dense_rate;
    g->s.rate16 = (r > 0 && r < 1) ? (uint32_t)std::min(std::max((long long)(r * 65536.0 + 0.5), 1ll), 65535ll) : 0u;
    *out = g;
    return BINGO_OK;
}

void bingo_destroy(bingo_graph *g)
{
    if (!g) return;
    if (g->st) (void)hipStreamSynchronize(g->st);
    void *bufs[] = {g->s.hdr,  g->s.grp,    g->s.alias,   g->s.fnode,   g->s.nmass,
                    g->s.ctl,  g->s.tasks,  g->d_steps,   g->s.fm,      g->flbuf[0],
                    g->flbuf[1]};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    void *pools[NPOOL] = {g->s.slot, g->s.back, g->s.mem, g->s.hash};
    for (int p = 0; p < NPOOL; ++p)
        if (pools[p] && pools[p] != g->vp[p].va) (void)hipFree(pools[p]);
    release_mapped(g);
    for (int i = 0; i < NSCRATCH; ++i)
        if (g->sc[i]) (void)hipFree(g->sc[i]);
    for (int i = 0; i < kEvents; ++i)
        if (g->ev[i]) (void)hipEventDestroy(g->ev[i]);
    for (int p = 0; p < NPHASE; ++p)
        for (int e = 0; e < 2; ++e)
            if (g->ph[p][e]) (void)hipEventDestroy(g->ph[p][e]);
    if (g->ctl_host) (void)hipHostFree(g->ctl_host);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
}

int bingo_load_device(bingo_graph *g, int32_t n, const bingo_edge *d_edges, int64_t m)
{
    if (!g) return fail(BINGO_EINVAL, "null handle");
    if (g->loaded) return fail(BINGO_ESTATE, "graph already loaded; create a new handle");
    if (n < 1 || n >= (1 << 30)) return fail(BINGO_EINVAL, "n must be in [1, 2^30)");
    if (m < 0 || (m > 0 && !d_edges)) return fail(BINGO_EINVAL, "bad edge array");
    HIPCHK(hipSetDevice(g->device));
    const size_t n1 = (size_t)n + 1;
    const size_t nb = n1 * sizeof(NodeHdr) + n1 * kGroups * (sizeof(GroupHdr) + sizeof(AliasEnt)) +
                      (g->s.fw ? n1 * sizeof(FNode) : 0) + (g->s.nbr ? n1 * sizeof(uint64_t) : 0);
    if (nb > g->budget) return fail(BINGO_ENOMEM, "node records exceed the HBM budget");
    bool ok = hipMalloc((void **)&g->s.hdr, n1 * sizeof(NodeHdr)) == hipSuccess &&
              hipMalloc((void **)&g->s.grp, n1 * kGroups * sizeof(GroupHdr)) == hipSuccess &&
              hipMalloc((void **)&g->s.alias, n1 * kGroups * sizeof(AliasEnt)) == hipSuccess &&
              (!g->s.fw || hipMalloc((void **)&g->s.fnode, n1 * sizeof(FNode)) == hipSuccess) &&
              (!g->s.nbr || hipMalloc((void **)&g->s.nmass, n1 * sizeof(uint64_t)) == hipSuccess);
    if (!ok) {
        (void)hipGetLastError();
        return fail(BINGO_ENOMEM, "node record allocation failed");
    }
    g->node_bytes = nb;
    g->s.n = n;
    phase_begin(g, PH_LOAD);
    launch_init_nodes(g->s, g->st);
    HIPCHK(hipGetLastError());
    g->loaded = true;
    if (g->cfg.fit_load) g->s.fit = 1u + (uint32_t)std::llround(g->cfg.fit_headroom * 65536.0);
    // larger loads go in consecutive batches of kLoadChunk arcs (same slots:
    // arcs keep array order), which bounds the load's sort scratch (28 bytes
    // per arc of a batch)
    const DevEdge *de = reinterpret_cast<const DevEdge *>(d_edges);
    int rc = m ? BINGO_OK : do_insert(g, de, 0);
    for (int64_t at = 0; !rc && at < m; at += g->load_chunk) rc = do_insert(g, de + at, std::min(g->load_chunk, m - at));
    phase_end(g, PH_LOAD);
    if (rc) {
        g->loaded = false;  // leave a handle that can be destroyed, not reused
        return rc;
    }
    // a load's sort buffers are sized for the whole graph: give them back
    return release_scratch(
        g, {SC_KEYS, SC_KEYS2, SC_VALS, SC_VALS2, SC_EDGES, SC_UNIQ, SC_COUNTS, SC_STARTS, SC_TMP, SC_FRAC});
}

int bingo_load(bingo_graph *g, int32_t n, const bingo_edge *edges, int64_t m)
{
    if (!g) return fail(BINGO_EINVAL, "null handle");
    if (m < 0 || (m > 0 && !edges)) return fail(BINGO_EINVAL, "bad edge array");
    int rc;
    if ((rc = ensure_scratch(g, SC_STAGE, (size_t)m * sizeof(bingo_edge)))) return rc;
    if (m) HIPCHK(hipMemcpyAsync(g->sc[SC_STAGE], edges, (size_t)m * sizeof(bingo_edge), hipMemcpyHostToDevice, g->st));
    rc = bingo_load_device(g, n, static_cast<const bingo_edge *>(g->sc[SC_STAGE]), m);
    if (rc) return rc;
    return release_scratch(g, {SC_STAGE});
}

int bingo_insert_device(bingo_graph *g, const bingo_edge *d_edges, int64_t k)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (k > 0 && !d_edges) return fail(BINGO_EINVAL, "null edges");
    phase_begin(g, PH_INSERT);
    int rc = do_insert(g, reinterpret_cast<const DevEdge *>(d_edges), k);
    phase_end(g, PH_INSERT);
    return rc;
}

int bingo_insert(bingo_graph *g, const bingo_edge *edges, int64_t k)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (k < 0 || (k > 0 && !edges)) return fail(BINGO_EINVAL, "bad edge array");
    int rc;
    if ((rc = ensure_scratch(g, SC_STAGE, (size_t)k * sizeof(bingo_edge)))) return rc;
    if (k) HIPCHK(hipMemcpyAsync(g->sc[SC_STAGE], edges, (size_t)k * sizeof(bingo_edge), hipMemcpyHostToDevice, g->st));
    rc = bingo_insert_device(g, static_cast<const bingo_edge *>(g->sc[SC_STAGE]), k);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

int bingo_delete_device(bingo_graph *g, const bingo_del *d_dels, int64_t k)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (k > 0 && !d_dels) return fail(BINGO_EINVAL, "null deletes");
    phase_begin(g, PH_DELETE);
    int rc = do_delete(g, reinterpret_cast<const int32_t *>(d_dels), k);
    phase_end(g, PH_DELETE);
    return rc;
}

int bingo_delete(bingo_graph *g, const bingo_del *dels, int64_t k)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (k < 0 || (k > 0 && !dels)) return fail(BINGO_EINVAL, "bad delete array");
    int rc;
    if ((rc = ensure_scratch(g, SC_STAGE, (size_t)k * sizeof(bingo_del)))) return rc;
    if (k) HIPCHK(hipMemcpyAsync(g->sc[SC_STAGE], dels, (size_t)k * sizeof(bingo_del), hipMemcpyHostToDevice, g->st));
    rc = bingo_delete_device(g, static_cast<const bingo_del *>(g->sc[SC_STAGE]), k);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

int bingo_delete_noalloc(bingo_graph *g, const bingo_del *dels, int64_t k)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (k < 0 || (k > 0 && !dels)) return fail(BINGO_EINVAL, "bad delete array");
    if (k > BINGO_NOALLOC_MAX_DELETES)
        return fail(BINGO_EINVAL, "batch of " + std::to_string(k) + " deletes above the allocation-free limit of " +
                                      std::to_string(BINGO_NOALLOC_MAX_DELETES));
    if (k == 0) return BINGO_OK;
    DelWork w;
    void *stage = nullptr;
    int rc = del_work_static(g, k, &w, &stage);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(stage, dels, (size_t)k * sizeof(bingo_del), hipMemcpyHostToDevice, g->st));
    phase_begin(g, PH_DELETE);
    rc = do_delete(g, static_cast<const int32_t *>(stage), k, w);
    phase_end(g, PH_DELETE);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

int bingo_walk_device(bingo_graph *g, int app, const bingo_walk_params *params, const int32_t *d_starts,
                      const int32_t *d_lens, int64_t nwalkers, int32_t *d_paths, uint64_t *d_steps)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (nwalkers < 0 || (nwalkers > 0 && (!d_paths || !d_steps))) return fail(BINGO_EINVAL, "bad walk buffers");
    if (app == BINGO_NODE2VEC && !g->s.nbr) return fail(BINGO_EUNSUPPORTED, "node2vec needs neighbor_index=1");
    WalkArgs a{};
    int rc = fill_walk_args(params, app, &a, g->s.fw, g->s.rate16 != 0);
    if (rc) return rc;
    if (nwalkers == 0) return BINGO_OK;
    a.starts = d_starts;
    a.lens = d_lens;
    a.paths = d_paths;
    a.steps = reinterpret_cast<unsigned long long *>(d_steps);
    a.cursor = &g->s.ctl->cursor;
    a.nwalkers = nwalkers;
    WalkChoice c;
    if ((rc = resolve_walk(g, app, params, &c))) return rc;
    launch_reset_walk(g->s.ctl, g->st);
    phase_begin(g, PH_WALK);
    rc = launch_choice(g, app, c, a, params->blocks_per_cu);
    phase_end(g, PH_WALK);
    return rc;
}

}  // extern "C"

struct bingo_walk_graph {
    bingo_graph *g = nullptr;
    int app = 0;
    bingo_walk_params params{};
    WalkArgs args{};
    WalkChoice choice{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
};

static void walk_graph_release(bingo_walk_graph *wg)
{
    if (wg->exec) (void)hipGraphExecDestroy(wg->exec);
    if (wg->graph) (void)hipGraphDestroy(wg->graph);
    wg->exec = nullptr;
    wg->graph = nullptr;
}

// Capture reset + walk on the handle's stream.  Nothing captured allocates
// or issues a memset: the counters are cleared by a kernel.
static int walk_graph_capture(bingo_walk_graph *wg)
{
    bingo_graph *g = wg->g;
    walk_graph_release(wg);
    HIPCHK(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
    launch_reset_walk(g->s.ctl, g->st);
    int rc = launch_choice(g, wg->app, wg->choice, wg->args, wg->params.blocks_per_cu);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(g->st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(BINGO_EDEVICE, std::string("walk graph capture: ") + hipGetErrorString(e));
    wg->graph = graph;
    if (hipGraphInstantiate(&wg->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        walk_graph_release(wg);
        return fail(BINGO_EDEVICE, "walk graph instantiate failed");
    }
    wg->gen = g->store_gen;
    return BINGO_OK;
}

extern "C" {

int bingo_walk_graph_create(bingo_graph *g, int app, const bingo_walk_params *params, const int32_t *d_starts,
                            const int32_t *d_lens, int64_t nwalkers, int32_t *d_paths, uint64_t *d_steps,
                            bingo_walk_graph **out)
{
    if (!out) return fail(BINGO_EINVAL, "null output");
    *out = nullptr;
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (nwalkers < 1 || !d_paths || !d_steps) return fail(BINGO_EINVAL, "bad walk buffers");
    if (app == BINGO_NODE2VEC && !g->s.nbr) return fail(BINGO_EUNSUPPORTED, "node2vec needs neighbor_index=1");
    bingo_walk_graph *wg = new bingo_walk_graph();
    wg->g = g;
    wg->app = app;
    int rc = fill_walk_args(params, app, &wg->args, g->s.fw, g->s.rate16 != 0);
    if (!rc) rc = resolve_walk(g, app, params, &wg->choice);
    if (rc) {
        delete wg;
        return rc;
    }
    wg->params = *params;
    wg->args.starts = d_starts;
    wg->args.lens = d_lens;
    wg->args.paths = d_paths;
    wg->args.steps = reinterpret_cast<unsigned long long *>(d_steps);
    wg->args.cursor = &g->s.ctl->cursor;
    wg->args.nwalkers = nwalkers;
    if ((rc = walk_graph_capture(wg))) {
        delete wg;
        return rc;
    }
    *out = wg;
    return BINGO_OK;
}

int bingo_walk_graph_launch(bingo_walk_graph *wg)
{
    if (!wg || !wg->g || !wg->g->loaded) return fail(BINGO_ESTATE, "no walk graph");
    bingo_graph *g = wg->g;
    if (wg->gen != g->store_gen) {  // same walk, new pool addresses
        int rc = walk_graph_capture(wg);
        if (rc) return rc;
    }
    phase_begin(g, PH_WALK);
    HIPCHK(hipGraphLaunch(wg->exec, g->st));
    phase_end(g, PH_WALK);
    return BINGO_OK;
}

void bingo_walk_graph_destroy(bingo_walk_graph *wg)
{
    if (!wg) return;
    if (wg->g && wg->g->st) (void)hipStreamSynchronize(wg->g->st);
    walk_graph_release(wg);
    delete wg;
}

}  // extern "C"

// ------------------------------------------------------------------------
// Captured update batches.  One HIP graph per power-of-two batch size
// (captured on first use) holds a whole insert batch -- keys, stable merge
// sort, gather, runs, plan, resolve, apply, region moves, neighbour tables,
// free-list merge -- and one per delete size the whole delete batch.  Batch
// sizes are set on the device before the replays (entries past them are
// padding, ignored), so one graph serves every size up to its bucket.
// Nothing waits for the host: pools grow on the device inside the room
// committed when the update graph was created, and a batch that does not
// fit that room, or is invalid, applies nothing and raises Ctl::ustatus,
// which makes every later captured batch apply nothing either until a call
// that waits for the device reports it.  No captured node is a memset
// (checked after every capture).
// ------------------------------------------------------------------------
struct bingo_update_graph {
    bingo_graph *g = nullptr;
    const DevEdge *ins = nullptr;
    const int32_t *dels = nullptr;
    int64_t max_ins = 0, max_dels = 0, max_k = 0;
    void *buf = nullptr;  // the captured batches' own scratch, carved below
    uint32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *vals2 = nullptr, *frac = nullptr;
    uint32_t *uniq = nullptr, *counts = nullptr, *starts = nullptr, *nruns = nullptr, *us = nullptr, *slots = nullptr;
    uint32_t *big = nullptr, *wcls = nullptr;
    unsigned long long *ka = nullptr, *kb = nullptr;
    DevEdge *edges = nullptr;
    uint64_t *pf = nullptr, *fl2 = nullptr;
    uint64_t pfcap = 0, flcap = 0;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    struct Bucket {
        int64_t size = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Bucket> ib, db;  // insert and delete buckets
    DevStore key{};               // the store as the captured graphs saw it
};

// a free-list buffer of the handle's first list slot holding at least
// `entries`, its contents kept (the live lists)
static int ensure_fl0(bingo_graph *g, uint64_t entries)
{
    if (entries <= g->flcap[0] && g->flbuf[0]) return BINGO_OK;
    uint64_t *fresh = nullptr;
    if (hipMalloc((void **)&fresh, entries * sizeof(uint64_t)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(BINGO_ENOMEM, "free-list allocation failed");
    }
    HIPCHK(hipStreamSynchronize(g->st));
    if (g->flbuf[0]) {
        if (g->flcap[0]) HIPCHK(hipMemcpy(fresh, g->flbuf[0], g->flcap[0] * sizeof(uint64_t), hipMemcpyDeviceToDevice));
        HIPCHK(hipFree(g->flbuf[0]));
    }
    g->flbuf[0] = fresh;
    g->flcap[0] = entries;
    set_lists(g);
    return BINGO_OK;
}

static int64_t ug_bucket(const bingo_update_graph *ug, int64_t k, int64_t max)
{
    int64_t b = 1024;
    while (b < k) b <<= 1;
    return std::min(b, max);
}

// the store view of the captured batches: the handle's store with the update
// graph's own scratch and second free list
static DevStore ug_store(const bingo_update_graph *ug)
{
    DevStore cs = ug->g->s;
    cs.wcls = ug->wcls;
    cs.pf = ug->pf;
    cs.fl = ug->g->flbuf[0];
    cs.fl2 = ug->fl2;
    return cs;
}

static void ug_release(std::vector<bingo_update_graph::Bucket> &v)
{
    for (auto &b : v) {
        if (b.exec) (void)hipGraphExecDestroy(b.exec);
        if (b.graph) (void)hipGraphDestroy(b.graph);
    }
    v.clear();
}

// no memset node may be replayed (such nodes have given wrong results here
// on replays after the first)
static int check_no_memset(hipGraph_t graph)
{
    size_t n = 0;
    if (hipGraphGetNodes(graph, nullptr, &n) != hipSuccess) return fail(BINGO_EDEVICE, "graph node query failed");
    std::vector<hipGraphNode_t> nodes(n);
    if (n && hipGraphGetNodes(graph, nodes.data(), &n) != hipSuccess)
        return fail(BINGO_EDEVICE, "graph node query failed");
    for (auto nd : nodes) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nd, &t) != hipSuccess) return fail(BINGO_EDEVICE, "graph node query failed");
        if (t == hipGraphNodeTypeMemset)
            return fail(BINGO_EUNSUPPORTED, "a captured update batch recorded a memset node");
    }
    return BINGO_OK;
}

static int ug_capture(bingo_update_graph *ug, bool del, int64_t B, bingo_update_graph::Bucket *out)
{
    bingo_graph *g = ug->g;
    const DevStore cs = ug_store(ug);
    Ctl *ctl = g->s.ctl;
    const int64_t kr = std::min<int64_t>(B, (int64_t)cs.n + 1);
    Runs r{ug->uniq, ug->starts, ug->counts, ug->nruns};
    HIPCHK(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
    if (!del) {
        launch_alloc_reset(cs.fm, g->st);
        launch_reset_ctl(ctl, g->st);
        launch_insert_keys(ug->ins, B, &ctl->bk[0], cs.n, cs.fw, ug->keys, ug->vals, ctl, g->st);
        merge_sort_pairs_u32(ug->tmp, ug->tmp_bytes, ug->keys, ug->keys2, ug->vals, ug->vals2, B, g->st);
        launch_gather_edges(ug->ins, ug->vals2, B, &ctl->bk[0], ug->edges, cs.fw ? ug->frac : nullptr, g->st);
        run_length_encode_u32(ug->tmp, ug->tmp_bytes, ug->keys2, B, ug->uniq, ug->counts, ug->nruns, g->st);
        exclusive_scan_u32(ug->tmp, ug->tmp_bytes, ug->counts, ug->starts, kr, g->st);
        launch_insert_plan(cs, ug->edges, r, B, g->st);
        launch_alloc_resolve(cs, cs.fw, g->st, 1, ug->flcap, ug->pfcap);
        launch_insert_apply(cs, ug->edges, cs.fw ? ug->frac : nullptr, r, B, g->st);
        launch_copy_tasks(cs, g->st);
        if (cs.nbr) launch_hash_insert(cs, ug->edges, B, g->st);
        launch_freelist_merge(cs, g->st);
        launch_freelist_copyback(cs, g->st);
    } else {
        launch_reset_ctl(ctl, g->st);
        launch_delete_keys(ug->dels, B, &ctl->bk[1], cs.n, ctl, ug->ka, g->st);
        merge_sort_keys_u64(ug->tmp, ug->tmp_bytes, ug->ka, ug->kb, B, g->st);
        launch_delete_check(cs, ug->kb, B, 1, ug->us, ug->slots, g->st);
        run_length_encode_u32(ug->tmp, ug->tmp_bytes, ug->us, B, ug->uniq, ug->counts, ug->nruns, g->st);
        exclusive_scan_u32(ug->tmp, ug->tmp_bytes, ug->counts, ug->starts, kr, g->st);
        launch_delete_big_list(r, B, ug->big, ctl, g->st);
        launch_delete_guard(ctl, g->st);
        if (cs.nbr) launch_hash_delete(cs, ug->us, ug->slots, B, g->st);
        launch_delete_apply(cs, ug->slots, r, ug->big, B, g->st);
    }
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(g->st, &graph);
    if (e != hipSuccess || !graph) {
        (void)hipGetLastError();
        if (graph) (void)hipGraphDestroy(graph);
        return fail(BINGO_EDEVICE, std::string("update batch capture: ") + hipGetErrorString(e));
    }
    int rc = check_no_memset(graph);
    hipGraphExec_t exec = nullptr;
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(BINGO_EDEVICE, "update graph instantiate failed");
    }
    if (rc) {
        (void)hipGraphDestroy(graph);
        return rc;
    }
    *out = bingo_update_graph::Bucket{B, graph, exec};
    return BINGO_OK;
}

// the bucket's replayable graph (captured on first use, or again after the
// store moved, grew or swapped its free lists)
static int ug_exec(bingo_update_graph *ug, bool del, int64_t k, hipGraphExec_t *out)
{
    DevStore cs = ug_store(ug);
    if (std::memcmp(&cs, &ug->key, sizeof(DevStore)) != 0) {
        // an uncaptured insert may have swapped in a smaller list buffer:
        // the captured merge copies its lists back into it
        int rc = ensure_fl0(ug->g, ug->flcap);
        if (rc) return rc;
        ug_release(ug->ib);
        ug_release(ug->db);
        cs = ug_store(ug);
        ug->key = cs;
    }
    auto &v = del ? ug->db : ug->ib;
    const int64_t B = ug_bucket(ug, k, del ? ug->max_dels : ug->max_ins);
    for (auto &b : v)
        if (b.size == B) {
            *out = b.exec;
            return BINGO_OK;
        }
    bingo_update_graph::Bucket nb;
    int rc = ug_capture(ug, del, B, &nb);
    if (rc) return rc;
    v.push_back(nb);
    *out = nb.exec;
    return BINGO_OK;
}

static void ug_free(bingo_update_graph *ug)
{
    ug_release(ug->ib);
    ug_release(ug->db);
    if (ug->buf) (void)hipFree(ug->buf);
    ug->buf = nullptr;
}

extern "C" {

int bingo_update_graph_create(bingo_graph *g, const bingo_edge *d_ins, int64_t max_ins, const bingo_del *d_dels,
                              int64_t max_dels, uint64_t room_bytes, bingo_update_graph **out)
{
    if (!out) return fail(BINGO_EINVAL, "null output");
    *out = nullptr;
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (max_ins < 0 || max_dels < 0 || max_ins > kMaxBatch || max_dels > kMaxBatch || (max_ins && !d_ins) ||
        (max_dels && !d_dels) || (!max_ins && !max_dels))
        return fail(BINGO_EINVAL, "bad update graph buffers or sizes");
    int rc;
    if ((rc = read_ctl(g)) || (rc = take_ustatus(g))) return rc;
    // the committed room: `room_bytes` (0: the store's pool bytes, at least
    // 1 GiB, at most half the budget left) spread over the pools by size
    uint64_t pool_bytes = 0, w[NPOOL];
    for (int p = 0; p < NPOOL; ++p) {
        w[p] = std::max<uint64_t>(g->ctl_host->top[p] * kElem[p], 1ull << 20);
        pool_bytes += g->ctl_host->top[p] * kElem[p];
    }
    const uint64_t used = store_bytes(g);
    uint64_t room = room_bytes;
    if (!room) {
        const uint64_t left = g->budget > used ? g->budget - used : 0;
        room = std::min<uint64_t>(std::max<uint64_t>(pool_bytes, 1ull << 30), left / 2);
    }
    uint64_t wsum = 0;
    for (int p = 0; p < NPOOL; ++p) wsum += w[p];
    for (int p = 0; p < NPOOL; ++p) {
        const uint64_t share = (uint64_t)((double)room * (double)w[p] / (double)wsum);
        const uint64_t free_elems = g->s.cap[p] > g->ctl_host->top[p] ? g->s.cap[p] - g->ctl_host->top[p] : 0;
        const uint64_t want = share / kElem[p];
        g->ctl_host->need[p] = want > free_elems ? want : 0;
    }
    if ((rc = ensure_pools(g))) return rc;
    bingo_update_graph *ug = new bingo_update_graph();
    ug->g = g;
    ug->ins = reinterpret_cast<const DevEdge *>(d_ins);
    ug->dels = reinterpret_cast<const int32_t *>(d_dels);
    ug->max_ins = max_ins;
    ug->max_dels = max_dels;
    const int64_t K = std::max(max_ins, max_dels);
    ug->max_k = K;
    const int64_t kr = std::min<int64_t>(K, (int64_t)g->s.n + 1);
    // regions a batch frees: at most the slot, back and table regions and 32
    // member lists of each run (node); free lists: room for the current ones
    // and a good many batches' frees
    ug->pfcap = std::max<uint64_t>(35ull * (uint64_t)std::min<int64_t>(max_ins, (int64_t)g->s.n + 1), 1024);
    ug->flcap = std::max<uint64_t>({4ull * g->ctl_host->fl_total, 1ull << 22, 2 * ug->pfcap});
    if ((rc = ensure_fl0(g, ug->flcap))) {
        delete ug;
        return rc;
    }
    const size_t tb = std::max({merge_sort_pairs_u32(nullptr, 0, nullptr, nullptr, nullptr, nullptr, K, g->st),
                                merge_sort_keys_u64(nullptr, 0, nullptr, nullptr, K, g->st),
                                run_length_encode_u32(nullptr, 0, nullptr, K, nullptr, nullptr, nullptr, g->st),
                                exclusive_scan_u32(nullptr, 0, nullptr, nullptr, kr, g->st)});
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t k4 = al((size_t)K * 4), k8 = al((size_t)K * 8), r4 = al((size_t)kr * 4);
    const size_t sizes[] = {k4,  k4, k4, k4, g->s.fw ? k4 : 0, r4, r4, r4, 256, k4, k4, al(((size_t)K / kBigDelete + 1) * 4),
                            al(insert_wcls_bytes(K)), k8, k8, al((size_t)K * sizeof(DevEdge)),
                            al(ug->pfcap * 8), al(ug->flcap * 8), al(tb)};
    size_t total = 0;
    for (size_t b : sizes) total += b;
    if (hipMalloc(&ug->buf, total) != hipSuccess) {
        (void)hipGetLastError();
        delete ug;
        return fail(BINGO_ENOMEM, "update graph scratch allocation failed");
    }
    unsigned char *p = static_cast<unsigned char *>(ug->buf);
    void *parts[19];
    for (int i = 0; i < 19; ++i) {
        parts[i] = sizes[i] ? p : nullptr;
        p += sizes[i];
    }
    ug->keys = (uint32_t *)parts[0];
    ug->keys2 = (uint32_t *)parts[1];
    ug->vals = (uint32_t *)parts[2];
    ug->vals2 = (uint32_t *)parts[3];
    ug->frac = (uint32_t *)parts[4];
    ug->uniq = (uint32_t *)parts[5];
    ug->counts = (uint32_t *)parts[6];
    ug->starts = (uint32_t *)parts[7];
    ug->nruns = (uint32_t *)parts[8];
    ug->us = (uint32_t *)parts[9];
    ug->slots = (uint32_t *)parts[10];
    ug->big = (uint32_t *)parts[11];
    ug->wcls = (uint32_t *)parts[12];
    ug->ka = (unsigned long long *)parts[13];
    ug->kb = (unsigned long long *)parts[14];
    ug->edges = (DevEdge *)parts[15];
    ug->pf = (uint64_t *)parts[16];
    ug->fl2 = (uint64_t *)parts[17];
    ug->tmp = parts[18];
    ug->tmp_bytes = sizes[18];
    ug->key = ug_store(ug);
    *out = ug;
    return BINGO_OK;
}

int bingo_update_graph_launch(bingo_update_graph *ug, int64_t nins, int64_t ndels)
{
    if (!ug || !ug->g || !ug->g->loaded) return fail(BINGO_ESTATE, "no update graph");
    if (nins < 0 || nins > ug->max_ins || ndels < 0 || ndels > ug->max_dels)
        return fail(BINGO_EINVAL, "batch sizes outside the update graph's maxima");
    bingo_graph *g = ug->g;
    hipGraphExec_t ei = nullptr, ed = nullptr;
    int rc;
    if ((nins && (rc = ug_exec(ug, false, nins, &ei))) || (ndels && (rc = ug_exec(ug, true, ndels, &ed)))) return rc;
    launch_set_batch(g->s.ctl, nins, ndels, g->st);
    if (ei) {
        phase_begin(g, PH_INSERT);
        HIPCHK(hipGraphLaunch(ei, g->st));
        phase_end(g, PH_INSERT);
    }
    if (ed) {
        phase_begin(g, PH_DELETE);
        HIPCHK(hipGraphLaunch(ed, g->st));
        phase_end(g, PH_DELETE);
    }
    g->arcs += nins - ndels;
    return BINGO_OK;
}

void bingo_update_graph_destroy(bingo_update_graph *ug)
{
    if (!ug) return;
    if (ug->g && ug->g->st) (void)hipStreamSynchronize(ug->g->st);
    ug_free(ug);
    delete ug;
}


int bingo_walk_tune(bingo_graph *g, int app, const bingo_walk_params *params, int64_t sample,
                    bingo_walk_params *chosen)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (app == BINGO_NODE2VEC && !g->s.nbr) return fail(BINGO_EUNSUPPORTED, "node2vec needs neighbor_index=1");
    WalkArgs a{};
    int rc = fill_walk_args(params, app, &a, g->s.fw, g->s.rate16 != 0);
    if (rc) return rc;
    if (sample < 1) return fail(BINGO_EINVAL, "sample must be >= 1");
    const int32_t n = g->s.n;
    std::vector<int32_t> starts((size_t)sample);
    const int64_t stride = std::max<int64_t>(1, n / sample);
    for (int64_t i = 0; i < sample; ++i) starts[(size_t)i] = (int32_t)((i * stride) % n) + 1;
    void *dstarts = nullptr, *dpaths = nullptr, *dsteps = nullptr;
    auto cleanup = [&]() {
        for (void *p : {dstarts, dpaths, dsteps})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc(&dstarts, (size_t)sample * 4) || hipMalloc(&dpaths, (size_t)sample * params->max_len * 4) ||
        hipMalloc(&dsteps, 8)) {
        (void)hipGetLastError();
        cleanup();
        return fail(BINGO_ENOMEM, "tuner buffers");
    }
    (void)hipMemcpyAsync(dstarts, starts.data(), (size_t)sample * 4, hipMemcpyHostToDevice, g->st);
    a.starts = static_cast<int32_t *>(dstarts);
    a.lens = nullptr;
    a.paths = static_cast<int32_t *>(dpaths);
    a.steps = static_cast<unsigned long long *>(dsteps);
    a.cursor = &g->s.ctl->cursor;
    a.nwalkers = sample;
    std::vector<WalkChoice> cand;
    for (int b : {256, 512, 128}) cand.push_back(WalkChoice{BINGO_FORM_PLAIN, b, 1, 0});
    for (int b : {128, 256, 512})
        for (int w : {1, 2, 4})
            for (int d : {0, 4, 16}) {
                WalkChoice c{BINGO_FORM_PERSISTENT, b, w, d};
                if (choice_ok(c, params->max_len) && persist_max_blocks_per_cu(app, w, d, b, g->s.fw, g->s.rate16 != 0) > 0)
                    cand.push_back(c);
            }
    float best = 1e30f;
    WalkChoice pick = cand[0];
    for (const WalkChoice &c : cand) {
        // the first launch warms up; the faster of the next two counts
        float ms = 0, fastest = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            launch_reset_walk(g->s.ctl, g->st);
            (void)hipEventRecord(g->ev[kEvents - 2], g->st);
            if ((rc = launch_choice(g, app, c, a, 0))) {
                cleanup();
                return rc;
            }
            (void)hipEventRecord(g->ev[kEvents - 1], g->st);
            if (hipEventSynchronize(g->ev[kEvents - 1]) != hipSuccess ||
                hipEventElapsedTime(&ms, g->ev[kEvents - 2], g->ev[kEvents - 1]) != hipSuccess) {
                cleanup();
                return fail(BINGO_EDEVICE, "tuner timing failed");
            }
            if (rep) fastest = std::min(fastest, ms);
        }
        if (fastest < best) {
            best = fastest;
            pick = c;
        }
    }
    cleanup();
    auto &t = g->tuned[app];
    t.valid = true;
    t.max_len = params->max_len;
    t.form = pick.form;
    t.block = pick.block;
    t.wpl = pick.wpl;
    t.depth = pick.depth;
    if (chosen) {
        *chosen = *params;
        chosen->form = pick.form;
        chosen->block = pick.block;
        chosen->walkers_per_lane = pick.wpl;
        chosen->lds_depth = pick.depth;
        chosen->blocks_per_cu = 0;
    }
    return BINGO_OK;
}

int bingo_walk(bingo_graph *g, int app, const bingo_walk_params *params, const int32_t *starts, const int32_t *lens,
               int64_t nwalkers, int32_t *paths, uint64_t *steps_out)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (!params) return fail(BINGO_EINVAL, "null walk params");
    if (nwalkers < 0 || (nwalkers > 0 && !paths)) return fail(BINGO_EINVAL, "bad path buffer");
    for (int64_t i = 0; starts && i < nwalkers; ++i)
        if (starts[i] < 1 || starts[i] > g->s.n) return fail(BINGO_EINVAL, "start node outside 1..n");
    for (int64_t i = 0; lens && i < nwalkers; ++i)
        if (lens[i] < 1 || lens[i] > params->max_len) return fail(BINGO_EINVAL, "walk length outside 1..max_len");
    int rc;
    const size_t pbytes = (size_t)nwalkers * (size_t)std::max(params->max_len, 1) * 4;
    if ((rc = ensure_scratch(g, SC_AUX1, pbytes)) || (rc = ensure_scratch(g, SC_AUX2, (size_t)nwalkers * 4)) ||
        (rc = ensure_scratch(g, SC_AUX3, (size_t)nwalkers * 4)))
        return rc;
    const int32_t *ds = nullptr, *dl = nullptr;
    if (starts && nwalkers) {
        HIPCHK(hipMemcpyAsync(g->sc[SC_AUX2], starts, (size_t)nwalkers * 4, hipMemcpyHostToDevice, g->st));
        ds = scp<int32_t>(g, SC_AUX2);
    }
    if (lens && nwalkers) {
        HIPCHK(hipMemcpyAsync(g->sc[SC_AUX3], lens, (size_t)nwalkers * 4, hipMemcpyHostToDevice, g->st));
        dl = scp<int32_t>(g, SC_AUX3);
    }
    HIPCHK(hipMemsetAsync(g->d_steps, 0, 8, g->st));
    rc = bingo_walk_device(g, app, params, ds, dl, nwalkers, scp<int32_t>(g, SC_AUX1),
                           reinterpret_cast<uint64_t *>(g->d_steps));
    if (rc) return rc;
    if (nwalkers) HIPCHK(hipMemcpyAsync(paths, g->sc[SC_AUX1], pbytes, hipMemcpyDeviceToHost, g->st));
    unsigned long long steps = 0;
    HIPCHK(hipMemcpyAsync(&steps, g->d_steps, 8, hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    g->last_walk_steps = steps;
    if (steps_out) *steps_out = steps;
    return BINGO_OK;
}

int bingo_compact(bingo_graph *g, double min_dead_frac, uint64_t *nodes_out, uint64_t *slots_out)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (!(min_dead_frac >= 0 && min_dead_frac <= 1)) return fail(BINGO_EINVAL, "min_dead_frac must be in [0, 1]");
    const uint32_t thr16 = (uint32_t)std::floor(min_dead_frac * 65536.0);
    const size_t n1 = (size_t)g->s.n + 1;
    int rc;
    if ((rc = ensure_scratch(g, SC_CLIST, 3 * n1 * sizeof(uint32_t)))) return rc;
    uint32_t *lists = scp<uint32_t>(g, SC_CLIST);
    phase_begin(g, PH_COMPACT);
    launch_compact_select(g->s, thr16, lists, lists + n1, lists + 2 * n1, g->st);
    HIPCHK(hipGetLastError());
    // the listed counts size the big nodes' staging scratch
    if ((rc = read_ctl(g))) return rc;
    const Ctl c = *g->ctl_host;
    const uint32_t nsmall = c.ncomp[0], nbig = c.ncomp[1], nzs = c.ncomp[2], nzb = c.ncomp[3];
    CompactBig cb{};
    if (nbig) {
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t sz[] = {al((nbig + 1) * 4ull),          al(c.comp_nseg * 4), al(c.comp_nseg * 4),
                             al(c.comp_nseg * 4 * kGroups),    al(nbig * 8ull),     al(nbig * 4ull),
                             al(nbig * 4ull),                  al(c.comp_bigslots * 16)};
        size_t tot = 0;
        for (size_t x : sz) tot += x;
        if ((rc = ensure_scratch(g, SC_CBIG, tot))) return rc;
        char *p = static_cast<char *>(g->sc[SC_CBIG]);
        void **dst[] = {(void **)&cb.segbase, (void **)&cb.seg_live, (void **)&cb.seg_pc, (void **)&cb.seg_gc,
                        (void **)&cb.trec,    (void **)&cb.nlive,    (void **)&cb.nback,  (void **)&cb.rec};
        for (int i = 0; i < 8; ++i) {
            *dst[i] = p;
            p += sz[i];
        }
    }
    launch_compact_apply(g->s, lists, lists + n1, lists + 2 * n1, nsmall, nbig, nzs, nzb, c.comp_nseg, cb, g->st);
    phase_end(g, PH_COMPACT);
    HIPCHK(hipGetLastError());
    if (nodes_out) *nodes_out = (uint64_t)nsmall + nbig + nzs + nzb;
    if (slots_out) *slots_out = c.comp_freed;
    return BINGO_OK;
}

int32_t bingo_num_nodes(const bingo_graph *g) { return g && g->loaded ? g->s.n : 0; }

int bingo_sync(bingo_graph *g)
{
    if (!g) return fail(BINGO_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(g->st));
    if (g->loaded) {
        int rc = read_ctl(g);
        if (rc) return rc;
        if (g->ctl_host->error & ERR_OVERRUN) return fail(BINGO_EDEVICE, "pool overrun detected");
        if ((rc = take_ustatus(g))) return rc;
        // device-side walk inputs are checked by the walk kernels: a walker
        // with a start outside 1..n or a length outside 1..max_len gets a row
        // of -1 and is reported here (until the next walk starts)
        if (g->ctl_host->werr & ERR_INVAL)
            return fail(BINGO_EINVAL, "a walk had a start node outside 1..n or a length outside 1..max_len");
    }
    return BINGO_OK;
}

void *bingo_stream(bingo_graph *g) { return g ? (void *)g->st : nullptr; }

int bingo_stats_get(bingo_graph *g, bingo_stats *out)
{
    if (!g || !out) return fail(BINGO_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(g->st));
    phase_collect(g);
    std::memset(out, 0, sizeof(*out));
    out->load_ms = g->ph_ms[PH_LOAD];
    out->insert_ms = g->ph_ms[PH_INSERT];
    out->delete_ms = g->ph_ms[PH_DELETE];
    out->walk_ms = g->ph_ms[PH_WALK];
    out->compact_ms = g->ph_ms[PH_COMPACT];
    out->walk_kernel_ms = g->ph_ms[PH_WALK];
    out->walk_steps = g->last_walk_steps;
    out->arcs = (uint64_t)g->arcs;
    out->bytes_reserved = store_bytes(g);
    if (g->loaded) {
        int rc = read_ctl(g);
        if (rc) return rc;
        for (int p = 0; p < NPOOL; ++p) out->pool_used[p] = g->ctl_host->top[p];
        out->walk_steps = g->ctl_host->steps;
        out->walk_attempts = g->ctl_host->attempts;
        out->walk_probes = g->ctl_host->probes;
        out->walk_folds = g->ctl_host->folds;
        out->walk_dense_trials = g->ctl_host->dtrials;
    }
    return BINGO_OK;
}

int bingo_edge_count(bingo_graph *g, const int32_t *us, const int32_t *vs, int64_t k, uint32_t *counts)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (k < 0 || (k > 0 && (!us || !vs || !counts))) return fail(BINGO_EINVAL, "bad query arrays");
    if (!k) return BINGO_OK;
    int rc;
    if ((rc = ensure_scratch(g, SC_AUX1, (size_t)k * 4)) || (rc = ensure_scratch(g, SC_AUX2, (size_t)k * 4)) ||
        (rc = ensure_scratch(g, SC_AUX3, (size_t)k * 4)))
        return rc;
    HIPCHK(hipMemcpyAsync(g->sc[SC_AUX2], us, (size_t)k * 4, hipMemcpyHostToDevice, g->st));
    HIPCHK(hipMemcpyAsync(g->sc[SC_AUX3], vs, (size_t)k * 4, hipMemcpyHostToDevice, g->st));
    launch_edge_count(g->s, scp<int32_t>(g, SC_AUX2), scp<int32_t>(g, SC_AUX3), k, scp<uint32_t>(g, SC_AUX1), g->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(counts, g->sc[SC_AUX1], (size_t)k * 4, hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

int bingo_export(bingo_graph *g, bingo_dump *d)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if (!d) return fail(BINGO_EINVAL, "null dump");
    const size_t n1 = (size_t)g->s.n + 1, ng = n1 * kGroups;
    std::vector<uint32_t> nsl(n1), bsz(ng), dmk(n1);
    void *dn = nullptr, *dm = nullptr, *db = nullptr, *dt = nullptr, *da = nullptr;
    void *dfm = nullptr, *dft = nullptr, *ddm = nullptr;
    auto cleanup = [&]() {
        for (void *p : {dn, dm, db, dt, da, dfm, dft, ddm})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc(&dn, n1 * 4) || hipMalloc(&dm, n1 * 8) || hipMalloc(&db, ng * 4) || hipMalloc(&dt, ng * 8) ||
        hipMalloc(&da, ng) || hipMalloc(&ddm, n1 * 4)) {
        cleanup();
        return fail(BINGO_ENOMEM, "export buffers");
    }
    if (g->s.fw && (hipMalloc(&dfm, n1 * 8) || hipMalloc(&dft, n1 * 8))) {  // fraction masses, thresholds
        cleanup();
        return fail(BINGO_ENOMEM, "export buffers");
    }
    launch_export_a(g->s, (uint32_t *)dn, (uint64_t *)dm, (uint32_t *)db, (uint64_t *)dt, (uint8_t *)da,
                    (uint64_t *)dfm, (uint64_t *)dft, (uint32_t *)ddm, g->st);
    (void)hipMemcpyAsync(nsl.data(), dn, n1 * 4, hipMemcpyDeviceToHost, g->st);
    (void)hipMemcpyAsync(bsz.data(), db, ng * 4, hipMemcpyDeviceToHost, g->st);
    (void)hipMemcpyAsync(dmk.data(), ddm, n1 * 4, hipMemcpyDeviceToHost, g->st);
    if (hipStreamSynchronize(g->st) != hipSuccess) {
        cleanup();
        return fail(BINGO_EDEVICE, "export");
    }
    std::vector<uint64_t> soff(n1), moff(n1);
    uint64_t ts = 0, tm = 0;
    for (size_t u = 0; u < n1; ++u) {
        soff[u] = ts;
        moff[u] = tm;
        ts += nsl[u];
        for (int j = 0; j < kGroups; ++j) tm += ((dmk[u] >> j) & 1u) ? 0u : bsz[u * kGroups + j];  // listed groups
    }
    d->total_slots = ts;
    d->total_members = tm;
    // per-node arrays are copied when their pointer is given; the slot and
    // member arrays only when all four of them are given
    if (d->nslots) std::memcpy(d->nslots, nsl.data(), n1 * 4);
    if (d->bsize) std::memcpy(d->bsize, bsz.data(), ng * 4);
    if (d->dmask) std::memcpy(d->dmask, dmk.data(), n1 * 4);
    if (d->mass) (void)hipMemcpyAsync(d->mass, dm, n1 * 8, hipMemcpyDeviceToHost, g->st);
    if (d->T) (void)hipMemcpyAsync(d->T, dt, ng * 8, hipMemcpyDeviceToHost, g->st);
    if (d->alias) (void)hipMemcpyAsync(d->alias, da, ng, hipMemcpyDeviceToHost, g->st);
    if (g->s.fw && d->fmass) (void)hipMemcpyAsync(d->fmass, dfm, n1 * 8, hipMemcpyDeviceToHost, g->st);
    if (g->s.fw && d->fthr) (void)hipMemcpyAsync(d->fthr, dft, n1 * 8, hipMemcpyDeviceToHost, g->st);
    if (!d->slot_v || !d->slot_w || !d->members || !d->backpos) {
        const hipError_t e0 = hipStreamSynchronize(g->st);
        cleanup();
        return e0 == hipSuccess ? BINGO_OK : fail(BINGO_EDEVICE, "export");
    }
    void *dso = nullptr, *dmo = nullptr, *dsv = nullptr, *dsw = nullptr, *dmem = nullptr, *dbp = nullptr,
         *dsf = nullptr;
    auto cleanup2 = [&]() {
        for (void *p : {dso, dmo, dsv, dsw, dmem, dbp, dsf})
            if (p) (void)hipFree(p);
    };
    const bool want_f = g->s.fw && d->slot_f;
    if (hipMalloc(&dso, n1 * 8) || hipMalloc(&dmo, n1 * 8) || hipMalloc(&dsv, ts * 4 + 4) ||
        hipMalloc(&dsw, ts * 4 + 4) || hipMalloc(&dmem, tm * 8 + 8) || hipMalloc(&dbp, tm * 4 + 4) ||
        (want_f && hipMalloc(&dsf, ts * 4 + 4))) {
        cleanup();
        cleanup2();
        return fail(BINGO_ENOMEM, "export buffers");
    }
    (void)hipMemcpyAsync(dso, soff.data(), n1 * 8, hipMemcpyHostToDevice, g->st);
    (void)hipMemcpyAsync(dmo, moff.data(), n1 * 8, hipMemcpyHostToDevice, g->st);
    launch_export_b(g->s, (uint64_t *)dso, (uint64_t *)dmo, (uint32_t *)dsv, (uint32_t *)dsw, (uint64_t *)dmem,
                    (uint32_t *)dbp, (uint32_t *)dsf, g->st);
    if (ts) {
        (void)hipMemcpyAsync(d->slot_v, dsv, ts * 4, hipMemcpyDeviceToHost, g->st);
        (void)hipMemcpyAsync(d->slot_w, dsw, ts * 4, hipMemcpyDeviceToHost, g->st);
        if (want_f) (void)hipMemcpyAsync(d->slot_f, dsf, ts * 4, hipMemcpyDeviceToHost, g->st);
    }
    if (tm) {
        (void)hipMemcpyAsync(d->members, dmem, tm * 8, hipMemcpyDeviceToHost, g->st);
        (void)hipMemcpyAsync(d->backpos, dbp, tm * 4, hipMemcpyDeviceToHost, g->st);
    }
    hipError_t e = hipStreamSynchronize(g->st);
    cleanup();
    cleanup2();
    if (e != hipSuccess) return fail(BINGO_EDEVICE, std::string("export: ") + hipGetErrorString(e));
    return BINGO_OK;
}

int bingo_dev_alloc(bingo_graph *g, size_t bytes, void **d_ptr)
{
    if (!g || !d_ptr) return fail(BINGO_EINVAL, "null argument");
    HIPCHK(hipSetDevice(g->device));
    if (hipMalloc(d_ptr, bytes ? bytes : 8) != hipSuccess) {
        (void)hipGetLastError();
        *d_ptr = nullptr;
        return fail(BINGO_ENOMEM, "device allocation failed");
    }
    return BINGO_OK;
}

int bingo_dev_free(bingo_graph *g, void *d_ptr)
{
    if (!g) return fail(BINGO_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(g->st));
    if (d_ptr) HIPCHK(hipFree(d_ptr));
    return BINGO_OK;
}

int bingo_memcpy_h2d(bingo_graph *g, void *d_dst, const void *src, size_t bytes)
{
    if (!g) return fail(BINGO_EINVAL, "null handle");
    HIPCHK(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

int bingo_memcpy_d2h(bingo_graph *g, void *dst, const void *d_src, size_t bytes)
{
    if (!g) return fail(BINGO_EINVAL, "null handle");
    HIPCHK(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
    return BINGO_OK;
}

int bingo_memcpy_d2d(bingo_graph *g, void *d_dst, const void *d_src, size_t bytes)
{
    if (!g || (bytes && (!d_dst || !d_src))) return fail(BINGO_EINVAL, "bad copy arguments");
    if (((uintptr_t)d_dst | (uintptr_t)d_src) & 15) return fail(BINGO_EINVAL, "copy needs 16-byte aligned pointers");
    if (bytes) launch_copy(d_dst, d_src, bytes, g->st);
    HIPCHK(hipGetLastError());
    return BINGO_OK;
}

int bingo_gather_probe(bingo_graph *g, int width, uint64_t *reads, float *ms)
{
    if (!g || !g->loaded) return fail(BINGO_ESTATE, "graph not loaded");
    if ((width != 0 && width != 8 && width != 32) || !reads || !ms)
        return fail(BINGO_EINVAL, "width must be 0, 8 or 32");
    int rc;
    if ((rc = read_ctl(g))) return rc;
    const uint64_t bytes = g->ctl_host->top[POOL_MEM] * kElem[POOL_MEM];
    const uint64_t units = bytes / (uint64_t)(width ? width : 8);
    if (units < 1) return fail(BINGO_EINVAL, "member pool is empty");
    if ((rc = ensure_scratch(g, SC_AUX1, 64))) return rc;
    hipEvent_t e0 = g->ev[kEvents - 2], e1 = g->ev[kEvents - 1];
    HIPCHK(hipEventRecord(e0, g->st));
    *reads = launch_gather_probe(g->s.mem, units, width, scp<unsigned long long>(g, SC_AUX1), g->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, g->st));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    return BINGO_OK;
}

int bingo_event_record(bingo_graph *g, int slot)
{
    if (!g || slot < 0 || slot >= kEvents) return fail(BINGO_EINVAL, "bad event slot");
    HIPCHK(hipEventRecord(g->ev[slot], g->st));
    return BINGO_OK;
}

int bingo_event_elapsed(bingo_graph *g, int a, int b, float *ms)
{
    if (!g || !ms || a < 0 || b < 0 || a >= kEvents || b >= kEvents) return fail(BINGO_EINVAL, "bad event slot");
    HIPCHK(hipEventSynchronize(g->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, g->ev[a], g->ev[b]));
    return BINGO_OK;
}

}  // extern "C"

// Internal view of a graph handle, shared by the C ABI (api.cpp) and the
// serving loop (server.cpp).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/bingo.h"
#include "store.h"

namespace bingo {

// records msg as the calling thread's last error and returns code
int fail(int code, const std::string &msg);
// host-side quantisation of walk parameters into kernel thresholds (fw:
// float weights; dense: the store has dense groups)
int fill_walk_args(const bingo_walk_params *p, int app, WalkArgs *a, int fw, int dense);
// waits for the handle's stream and reports what the update batches since
// the last wait left: a pool overrun, or a captured batch that applied
// nothing (invalid, or past the committed room); unlike bingo_sync it does
// not report walk inputs
int update_wait(bingo_graph *g);

}  // namespace bingo

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bingo::fail(BINGO_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace bingo {

enum Scratch {
    SC_KEYS,
    SC_KEYS2,
    SC_VALS,
    SC_VALS2,
    SC_EDGES,
    SC_UNIQ,
    SC_COUNTS,
    SC_STARTS,
    SC_NRUNS,
    SC_TMP,
    SC_U64A,
    SC_U64B,
    SC_STAGE,
    SC_AUX1,
    SC_AUX2,
    SC_AUX3,
    SC_BIG,   // delete batch: long runs
    SC_FRAC,  // insert batch with float weights: fractions in node order
    SC_CLIST, // compaction: the listed nodes (three lists of up to n + 1)
    SC_CBIG,  // compaction: big nodes' segment tables and staging
    SC_WCLS,  // insert batch: each plan workgroup's class index ranges
    SC_PF,    // insert batch: the regions it frees
    NSCRATCH
};

constexpr size_t kElem[NPOOL] = {4, 4, 8, 8};
constexpr int kEvents = 16;

enum Phase { PH_LOAD, PH_INSERT, PH_DELETE, PH_WALK, PH_COMPACT, NPHASE };

}  // namespace bingo

struct bingo_graph {
    bingo_config cfg{};
    int device = 0;
    hipStream_t st = nullptr;
    bingo::DevStore s{};
    bool loaded = false;
    int64_t arcs = 0;
    size_t node_bytes = 0;
    uint64_t budget = 0;
    bingo::Ctl *ctl_host = nullptr;  // pinned
    void *sc[bingo::NSCRATCH] = {};
    size_t sc_bytes[bingo::NSCRATCH] = {};
    hipEvent_t ev[bingo::kEvents] = {};
    hipEvent_t ph[bingo::NPHASE][2] = {};
    bool ph_pending[bingo::NPHASE] = {};
    double ph_ms[bingo::NPHASE] = {};
    uint64_t last_walk_steps = 0;
    unsigned long long *d_steps = nullptr;
    int num_cu = 0;
    uint64_t store_gen = 0;  // bumped whenever a pool moves
    // grow_in_place pools: a reserved address range per pool, mapped chunk by chunk
    struct VPool {
        void *va = nullptr;
        size_t reserved = 0, mapped = 0;
        std::vector<std::pair<hipMemGenericAllocationHandle_t, size_t>> chunks;
    } vp[bingo::NPOOL];
    uint64_t *flbuf[2] = {nullptr, nullptr};  // free lists: current (s.fl) and next (s.fl2)
    size_t flcap[2] = {0, 0};                  // their capacities (entries)
    bool vmm_refused = false;  // the device refused a reservation: pools are copied instead
    unsigned refuse_vmm = 0;   // test hook: pools (bit p) whose reservation is treated as refused
    int64_t load_chunk = 0;    // arcs per batch of a load (kLoadChunk; smaller as a test hook)
    struct Tuned {
        bool valid = false;
        int32_t max_len = 0;
        int form = BINGO_FORM_PERSISTENT, block = 256, wpl = 2, depth = 0;
    } tuned[4];
};

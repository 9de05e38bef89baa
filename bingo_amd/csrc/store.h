// Internal layout of the HBM-resident store and the kernel launchers.
//
// Per node u (ids 1..n; record 0 is an always-empty node):
//   NodeHdr  hdr[u]        slot region, back-index region, neighbour table
//   GroupHdr grp[u*32 + j] member list of bit group j: pool offset, size, capacity
//   AliasEnt alias[u*32+i] 32-slot alias table; slot i carries the (offset,
//                          size) of group i and of its alias partner, so one
//                          32-byte read decides the group of a walk step
// Pools (one device allocation each, regions handed out by atomic bump):
//   slot  u32: per node [v x cap | w x cap | back-offset x cap], with float
//              weights also [fraction x cap] (w holds the integer part)
//   back  u32: per slot, the member positions of its set bits, ascending
//   mem   u64: member entries (v << 32) | slot
//   hash  u64: neighbour multiset, 16-byte entries {key | count << 32, weight
//              sum of those arcs} (linear probing); a table of cap entries
//              takes 2 cap words
//
// Reference roles: NodeData (include/Bingo.cuh:6-58) held edge[SZ],
// group[BITS*SZ], idx[BITS*SZ], num[], aliasTable[] per node from a bump pool
// (include/gpuMem.cuh:4-32); here every bucket has its own capacity and the
// alias table stores group locators instead of bit indices.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bingo {

constexpr int kGroups = 32;
constexpr uint32_t kMark = 0x80000000u;
constexpr uint32_t kAuxMask = (1u << 27) - 1u;

enum Pool { POOL_SLOT = 0, POOL_BACK = 1, POOL_MEM = 2, POOL_HASH = 3, NPOOL = 4 };
enum ErrFlag { ERR_INVAL = 1, ERR_NOMEM = 2, ERR_OVERRUN = 4 };

struct NodeHdr {
    uint64_t slot_off;
    uint32_t nslots, slot_cap;  // slots ever assigned (tombstones included), capacity
    uint64_t back_off;
    uint32_t back_used, back_cap;
    uint64_t hash_off;
    uint32_t hash_cap, hash_used;
    uint32_t nlive;              // live slots (nslots - nlive tombstones, reclaimed by compaction)
    uint32_t dmask;              // dense bit groups (count only, no member list; DevStore::rate16)
    uint32_t sbase;              // a member word's slot field is its slot + sbase: a compaction that
                                 // drops a dead prefix of d slots adds d here instead of rewriting them
    uint32_t dend;               // one past the highest slot deleted since the last compaction: the
                                 // dead slots are exactly the first ones iff dend = nslots - nlive
};

struct GroupHdr {
    uint64_t off;
    uint32_t size, cap;
};

// Float weights only: a node's fraction mass F (sum of the slots' fractions,
// units of 2^-32) and the threshold thr = floor(2^64 F / (S 2^32 + F)) a draw
// compares its fractional-stream coin with (all-ones when the integer mass S
// is 0: every draw takes the fractional part).
struct FNode {
    uint64_t F;
    uint64_t thr;
};
constexpr uint32_t kFracTag = 0x40000000u;     // Philox counter word 3 of the fractional stream
constexpr uint32_t kMixBit = 0x80000000u;      // Philox counter word 2 (step) of node2vec's fold stream
constexpr uint32_t kFracIters = 4096u;         // fractional rejection rounds per draw
constexpr int32_t kMaxAttemptsFloat = 65536;   // node2vec cap with float weights (attempt in 16 bits)
constexpr uint32_t kThrHiMax = (1u << 17) - 1;  // the alias slot's 17-bit copy of thr saturates here
constexpr uint64_t kHmetaMask = (1ull << 42) - 1;
constexpr uint32_t kDenseTag = 0x20000000u;      // Philox counter word 3 of the dense-group rejection stream
constexpr uint32_t kDenseBlocks = 1024u;         // its Philox blocks per draw (four trials each)
constexpr int32_t kMaxAttemptsDense = 65536;     // node2vec cap with dense groups (attempt in 16 bits)

// One 32-byte alias slot i of a node:
//   T      stay (group i) iff coin64 < T, else the partner
//   first  locator of the likelier outcome: group i if T >= 2^63, else the
//          partner (member offset (36 bits) | size << 36; for a dense group
//          kDenseLoc | its index, size field 0; 0 for an empty group)
//   second locator of the other outcome
// (the first 16 bytes decide most draws alone: a draw reads the second
// locator only when the less likely outcome comes up)
//   meta  partner index (5 bits) | the node's neighbour-table locator << 5
//         (42 bits: (log2 cap + 1) | table offset << 6, 0: no table)
//         | << 47, 17 bits: with float weights, the node's fractional-part
//         threshold rounded up (ceil(thr / 2^47), at most 2^17 - 1; 0 iff
//         thr = 0); with integer weights, the node's integer weight S coded
//         as exponent e (6 bits) << 11 | the top 11 bits m of S (S exact
//         below 2^11; else m 2^e <= S < (m + 1) 2^e)
// so a walk step learns the group, its members, (for node2vec) where the
// node's neighbour table lives and almost always whether to take node2vec's
// folded return region or (float weights) the fractional part, from a
// single read.
struct AliasEnt {
    uint64_t T;
    uint64_t first;
    uint64_t second;
    uint64_t meta;
};
constexpr int kOffBits = 36;
constexpr uint64_t kOffMask36 = (1ull << kOffBits) - 1ull;
constexpr uint64_t kDenseLoc = 1ull << 35;  // a locator with size field 0 and this bit: a dense group
constexpr uint32_t kMaxGroupSize = 1u << 28;  // sizes live in 28 bits
static_assert(sizeof(AliasEnt) == 32, "alias entry is one 32-byte read");
static_assert(sizeof(NodeHdr) == 64, "node header layout");

// slot arrays per node region: v, w, back offset (and fraction)
__host__ __device__ inline uint32_t slot_arrays(int32_t fw) { return fw ? 4u : 3u; }

struct Ctl {
    unsigned long long top[NPOOL];   // pool cursors (elements)
    unsigned long long need[NPOOL];  // planned growth of the pending batch
    unsigned int error;              // ErrFlag bits
    unsigned int nbig;               // delete batch: runs split into per-group tasks
    unsigned long long steps;        // last walk: appended vertices
    unsigned long long attempts;     // last walk: sample draws (node2vec rejections included)
    unsigned long long probes;       // last walk: neighbour-table membership lookups
    unsigned long long folds;        // last walk: node2vec return-weight lookups (neighbour entries)
    unsigned long long dtrials;      // last walk: slot weights read by dense-group rejection draws
    unsigned long long cursor;       // persistent walk: next unclaimed walker
    unsigned int nbigcopy;           // insert plan: runs that move a region of kBigCopyBytes or more
    unsigned int werr;               // walks since the last walk reset: ERR_INVAL (a start or length out of range)
    unsigned long long taskctr;      // insert apply: region moves left to k_copy_tasks << 32 | their chunks
    unsigned int ncomp[4];           // compaction: nodes listed to be rebuilt (wave path, segment path) and to
                                     // drop a dead prefix (wave path, workgroup path)
    unsigned long long comp_freed;   // compaction: tombstoned slots reclaimed
    unsigned long long comp_bigslots;  // compaction: slots of the listed big nodes
    unsigned long long comp_nseg;      // compaction: segments of the listed big nodes
    unsigned long long fl_total;     // insert batch: free-list entries after the batch
    unsigned long long pend_total;   // insert batch: regions the batch frees
    // Captured update batches (update graphs).  bk: the replay's insert and
    // delete counts (set by a kernel before the graph); entries past them are
    // padding.  ustatus (sticky until the host reads it): a captured batch
    // could not apply (UST_*); every apply-side kernel returns at once while it
    // is set, so later captured batches apply nothing either.
    unsigned long long bk[2];
    unsigned int ustatus;
    unsigned int pad_ustatus;
};
enum UStatus { UST_INVAL = 1, UST_NOROOM = 2 };

// Region sizes and recycling.  Every region of pool p has a capacity of a
// size class: class k = 8 q + r (r < 8) holds S(k) = floor((8 + r) 2^q / 8)
// units -- 8 classes per power of two, neighbours at most 1/8 apart -- and a
// unit is kMult[p] elements (slot regions: one per slot array; neighbour
// tables: 2 pool words per entry).  A region that moves out is freed to the
// list of its class.  An insert batch's plan counts, per workgroup, its
// allocations and frees per (pool, class); the batch's allocations of a
// class take the class's free regions first and bump the pool past them;
// its frees become available to the next batch (a moved region is still
// read by the batch's deferred copies).  Index t = pool * kClasses + class.
constexpr int kClasses = 256;
constexpr int kPC = NPOOL * kClasses;
__host__ __device__ constexpr uint32_t class_size(uint32_t k)
{
    return (uint32_t)((((uint64_t)8u + (k & 7u)) << (k >> 3)) >> 3);
}
// the largest class whose size is at most c (c >= 1)
__host__ __device__ inline uint32_t class_down(uint32_t c)
{
    const uint32_t q = 31u - (uint32_t)__builtin_clz(c);
    const uint32_t r = q >= 3 ? (c >> (q - 3)) - 8u : (c << (3 - q)) - 8u;
    return 8u * q + r;
}
// the smallest class whose size is at least c (1 <= c <= S(255) = 15 2^28)
__host__ __device__ inline uint32_t class_up(uint32_t c)
{
    const uint32_t k = class_down(c);
    return class_size(k) == c ? k : k + 1u;
}
struct FreeMeta {
    uint32_t avail[kPC];   // free regions of the class: list entries [fbase, fbase + avail)
    uint32_t take[kPC];    // this batch: allocations of the class (plan)
    uint32_t freed[kPC];   // this batch: regions of the class freed
    uint32_t nav[kPC];     // this batch: free regions after it
    uint64_t fbase[kPC];   // list segment of the class
    uint64_t nfbase[kPC];  // list segment after the batch
    uint64_t bump[kPC];    // this batch: pool offset of the class's allocations past avail
    uint64_t pseg[kPC];    // this batch: pending-list segment of the class's frees
};

// A region move of an insert batch that is too large for the run's own wave:
// the apply kernel records it and k_copy_tasks carries it out with the whole
// grid before the batch's neighbour-table kernel.  Nothing reads a moved
// region's old part within the apply kernel (a run only appends past it), so
// the move may finish later.  A task is cut into chunks of kCopyChunk bytes
// of its first phase (copy, or zeroing the new table), one workgroup each;
// the tasks' chunk ranges [cbase, cbase + nch) follow each other in task
// order (one 64-bit atomic hands out both).
enum CopyKind : uint32_t { TASK_COPY = 0, TASK_REHASH = 1 };
struct CopyTask {
    uint64_t dst, src;  // device addresses
    uint64_t n;         // TASK_COPY: bytes; TASK_REHASH: the new table's capacity (entries)
    uint32_t kind;
    int32_t u;          // TASK_REHASH: the node (its header counts the re-inserted keys)
    uint32_t old_cap;   // TASK_REHASH: the old table's capacity
    uint32_t cbase, nch;
    uint32_t pad;
};
constexpr uint64_t kBigCopyBytes = 64ull << 10;
constexpr uint64_t kCopyChunk = 16ull << 10;
constexpr uint32_t kMaxCopyTasks = 4096;

struct DevEdge {
    int32_t u, v;
    uint32_t w;
};

struct DevStore {
    int32_t n;
    int32_t nbr;
    int32_t min_slots;
    // Growth rule.  0: a region that must grow gets the next power of two
    // (slots: at least min_slots).  Otherwise (a fitted store) it gets the
    // smallest size class holding its contents plus ceil(size * (fit - 1) /
    // 2^16); neighbour tables stay powers of two.
    uint32_t fit;
    int32_t fw;      // weights are float bits: slot regions carry a fraction array
    // Sparse/dense groups (ref RATE, include/api.h:25, operation.cuh:72): 0,
    // or the rate x 2^16.  The batch that gives a node its first slots sets
    // its dense groups: j with count_j >= 1 and count_j 2^16 >= rate16 len
    // over that batch's arcs of the node.  A dense group keeps only its count
    // (GroupHdr::size, cap 0) and no back entries; draws from it reject over
    // the node's slots (dense_draw).
    uint32_t rate16;
    // a dead-prefix compaction that would take a node's sbase past this
    // lowers the node's member slot fields instead and returns sbase to 0
    // (kSbaseMax; smaller only as a test hook)
    uint32_t sbase_max;
    FNode *fnode;    // [n+1] when fw
    uint64_t *nmass; // [n+1] when nbr: each node's integer weight S (node2vec's folded return region)
    NodeHdr *hdr;
    GroupHdr *grp;
    AliasEnt *alias;
    uint32_t *slot;
    uint32_t *back;
    uint64_t *mem;
    uint64_t *hash;
    unsigned long long cap[NPOOL];
    Ctl *ctl;
    CopyTask *tasks;  // [kMaxCopyTasks]
    FreeMeta *fm;
    uint64_t *fl;     // free regions, per class segments (FreeMeta::fbase)
    uint64_t *fl2;    // the lists after the current batch
    uint64_t *pf;     // regions freed by the current batch, per class segments (FreeMeta::pseg)
    uint32_t *wcls;   // insert batch: per workgroup, its first allocation and free index per class (2 kPC)
};

struct WalkArgs {
    const int32_t *starts;
    const int32_t *lens;
    int32_t *paths;
    unsigned long long *steps;
    unsigned long long *cursor;  // persistent form: next unclaimed walker
    int64_t nwalkers;
    uint64_t seed;
    uint32_t thr_return, thr_neighbor, thr_other, thr_stop;
    int32_t max_len;
    int32_t max_attempts;
    uint64_t fold_r;    // node2vec: folded return region factor x 2^27 (0: plain envelope)
    uint64_t fold_cap;  // an upper bound of every node's fold threshold (return weight <= node weight)
};

// One admitted walk request of the serving loop.  Its walkers occupy the
// global index range [base, base + nwalkers) of a serving launch; walker i
// of the request draws from Philox(seed; i, step, attempt), exactly as the
// request would alone.
// One entry of a serving launch: walkers [first, first + nwalkers) of a
// request, at launch indices [base, base + nwalkers).  starts/lens/paths are
// the request's whole arrays, indexed by the request's own walker id.
struct ServeReq {
    const int32_t *starts;
    const int32_t *lens;
    int32_t *paths;
    union {
        unsigned long long *steps;  // the request's step counter (device memory), as passed
        unsigned long long count;   // a workgroup's LDS copy of the entry: its steps for the request
    };
    unsigned long long base;
    int64_t first;
    int64_t nwalkers;
    uint64_t seed;
    uint32_t thr_return, thr_neighbor, thr_other, thr_stop;
    int32_t max_len, max_attempts;
    int32_t app, pad;
    uint64_t fold_r, fold_cap;
};
constexpr int kMaxServeReqs = 128;

struct Runs {          // per-node runs of a sorted batch
    const uint32_t *node;
    const uint32_t *start;
    const uint32_t *count;
    const uint32_t *nruns;  // device scalar
};

// ---- launchers (kernels.hip); all enqueue on `st` ----
void launch_reset_ctl(Ctl *ctl, hipStream_t st);
void launch_reset_walk(Ctl *ctl, hipStream_t st);
void launch_init_nodes(DevStore s, hipStream_t st);
// fw: weights are float bits (validated as such; the gather then writes the
// integer part as the arc's weight and the fraction to frac[i])
// dk (captured batches): the batch's true size on the device; entries at
// or past it are padding (key 0, no error; gathered as an empty arc)
void launch_insert_keys(const DevEdge *e, int64_t k, const unsigned long long *dk, int32_t n, int fw, uint32_t *keys,
                        uint32_t *vals, Ctl *ctl, hipStream_t st);
void launch_gather_edges(const DevEdge *e, const uint32_t *order, int64_t k, const unsigned long long *dk, DevEdge *out,
                         uint32_t *frac, hipStream_t st);
// The plan counts each workgroup's region allocations and frees per class
// (s.wcls: insert_wcls_bytes(k)); launch_alloc_resolve then sets Ctl::need
// (the batch's bump space per pool), fl_total and pend_total; the apply,
// launched for the same k, takes its regions and moves the pool tops past
// the bump space; after it (and its copy tasks) launch_freelist_merge writes
// the free lists to s.fl2 (the caller swaps fl and fl2).
size_t insert_wcls_bytes(int64_t k);
void launch_alloc_reset(FreeMeta *fm, hipStream_t st);  // before a batch's plan
// captured != 0: a batch whose growth does not fit the pools' capacity, whose
// free lists would pass flcap entries or whose frees pfcap, or that is
// invalid, sets Ctl::ustatus instead of asking the host
void launch_alloc_resolve(DevStore s, int fw, hipStream_t st, int captured = 0, uint64_t flcap = 0,
                          uint64_t pfcap = 0);
// captured batches: the merged lists (s.fl2) copied back over s.fl
void launch_freelist_copyback(DevStore s, hipStream_t st);
void launch_set_batch(Ctl *ctl, int64_t kins, int64_t kdel, hipStream_t st);
// after a captured delete batch's check: an invalid batch sets Ctl::ustatus
void launch_delete_guard(Ctl *ctl, hipStream_t st);
// host clears a reported status (stream-ordered)
void launch_clear_ustatus(Ctl *ctl, hipStream_t st);
void launch_freelist_merge(DevStore s, hipStream_t st);
void launch_insert_plan(DevStore s, const DevEdge *sorted, Runs r, int64_t k, hipStream_t st);
void launch_insert_apply(DevStore s, const DevEdge *sorted, const uint32_t *frac, Runs r, int64_t k, hipStream_t st);
// neighbour multisets (nbr stores): the batch's arcs, one thread each, after
// the insert apply kernel; the batch's deletes before the delete apply kernel
// (their slot weights are read before the tombstones)
void launch_hash_insert(DevStore s, const DevEdge *sorted, int64_t k, hipStream_t st);
void launch_hash_delete(DevStore s, const uint32_t *us, const uint32_t *slots, int64_t k, hipStream_t st);
void launch_delete_keys(const int32_t *dels, int64_t k, const unsigned long long *dk, int32_t n, Ctl *ctl,
                        unsigned long long *keys, hipStream_t st);
// pads: key 0 is padding (captured batches), not an invalid delete
void launch_delete_check(DevStore s, const unsigned long long *keys, int64_t k, int pads, uint32_t *us, uint32_t *slots,
                         hipStream_t st);
// Delete runs longer than kBigDelete are listed (big: run indices, room for
// k / kBigDelete + 1 entries), handled by one wave per bit group and then
// finished (tombstones, alias table) by a second kernel.
constexpr uint32_t kBigDelete = 128;
void launch_delete_big_list(Runs r, int64_t k, uint32_t *big, Ctl *ctl, hipStream_t st);
void launch_delete_apply(DevStore s, const uint32_t *slots, Runs r, const uint32_t *big, int64_t k, hipStream_t st);
void launch_walk(DevStore s, int app, const WalkArgs &a, int block, hipStream_t st);
struct PersistCfg {
    int block, wpl, depth, blocks;  // blocks = grid size
};
size_t persist_lds_bytes(int block, int wpl, int depth);
int persist_max_blocks_per_cu(int app, int wpl, int depth, int block, int fw, int dense);
void launch_walk_persist(DevStore s, int app, const WalkArgs &a, const PersistCfg &c, hipStream_t st);
// cls: 0 = the launch holds no node2vec request, 1 = only node2vec
// requests, 2 = any mix (a kernel variant per class)
void launch_walk_serve(DevStore s, const ServeReq *reqs, int nreq, unsigned long long total,
                       unsigned long long *cursor, int cls, int blocks, hipStream_t st);
int serve_max_blocks_per_cu(int fw, int cls, int dense);
int serve_walkers_per_lane(int cls);
// fmass/fthr (export_a) and slot_f (export_b) are null unless fw
void launch_export_a(DevStore s, uint32_t *nslots, uint64_t *mass, uint32_t *bsize, uint64_t *T, uint8_t *alias,
                     uint64_t *fmass, uint64_t *fthr, uint32_t *dmask, hipStream_t st);
void launch_export_b(DevStore s, const uint64_t *slot_out_off, const uint64_t *mem_out_off, uint32_t *slot_v,
                     uint32_t *slot_w, uint64_t *members, uint32_t *backpos, uint32_t *slot_f, hipStream_t st);
// Compaction of tombstoned slots: nodes with dead = nslots - nlive >= 1 and
// dead 2^16 >= thr16 nslots are listed (room for n + 1 entries in each list)
// by launch_compact_select; Ctl::ncomp, comp_bigslots and comp_nseg then
// size the big-node scratch (CompactBig) for launch_compact_apply.  Nodes of
// at most kBigCompact slots are rewritten by one wave each; larger ones in
// segments of kCompactSeg slots spread over the grid (a wave each), their
// records staged through the scratch.
constexpr uint32_t kBigCompact = 8192;
// a dead prefix is dropped without rewriting member words while the node's
// sbase stays at most this (member slot fields, slot + sbase < 2^30 + 2^28,
// must stay clear of the delete kernels' kMark bit); past it the member slot
// fields are lowered in place and sbase returns to 0
constexpr uint32_t kSbaseMax = 1u << 30;
constexpr uint32_t kCompactSeg = 2048;
struct CompactBig {
    uint32_t *segbase;  // [nbig + 1] first segment of each big node
    uint32_t *seg_live; // [nseg] live slots per segment, then (in place) the new id of its first
    uint32_t *seg_pc;   // [nseg] back entries per segment, then the new offset of its first
    uint32_t *seg_gc;   // [nseg][32] members of each listed group per segment, then its first position
    uint64_t *trec;     // [nbig] staging offset of the node's records
    uint32_t *nlive;    // [nbig] live slots after compaction
    uint32_t *nback;    // [nbig] back entries after compaction
    uint4 *rec;         // staged records {v, w, fraction, back offset} (comp_bigslots)
};
// shift_list (n + 1 entries) holds the nodes whose dead slots are their
// first ones: those of at most kBigCompact slots from its start, larger ones
// from its end (Ctl::ncomp[2], ncomp[3])
void launch_compact_select(DevStore s, uint32_t thr16, uint32_t *small_list, uint32_t *big_list, uint32_t *shift_list,
                           hipStream_t st);
void launch_compact_apply(DevStore s, const uint32_t *small_list, const uint32_t *big_list, const uint32_t *shift_list,
                          uint32_t nsmall, uint32_t nbig, uint32_t nzs, uint32_t nzb, uint64_t nseg,
                          const CompactBig &cb, hipStream_t st);
void launch_edge_count(DevStore s, const int32_t *us, const int32_t *vs, int64_t k, uint32_t *counts,
                       hipStream_t st);

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
void launch_copy_tasks(const DevStore &s, hipStream_t st);
// random reads of `width` (8 or 32) bytes over units * width bytes at base;
// returns the number of reads
uint64_t launch_gather_probe(void *base, uint64_t units, int width, unsigned long long *sink, hipStream_t st);

// rocPRIM wrappers (temp storage sized by a null call)
// Pairs sort over double buffers: input in (k0, v0), (k1, v1) is the other
// half; *second says whether the sorted output ended in (k1, v1).
size_t sort_pairs_u32(void *tmp, size_t tmp_bytes, uint32_t *k0, uint32_t *k1, uint32_t *v0, uint32_t *v1, int64_t k,
                      int bits, hipStream_t st, bool *second);
size_t sort_keys_u64(void *tmp, size_t tmp_bytes, const unsigned long long *kin, unsigned long long *kout, int64_t k,
                     int bits, hipStream_t st);
size_t run_length_encode_u32(void *tmp, size_t tmp_bytes, const uint32_t *keys, int64_t k, uint32_t *uniq,
                             uint32_t *counts, uint32_t *nruns, hipStream_t st);
size_t exclusive_scan_u32(void *tmp, size_t tmp_bytes, const uint32_t *in, uint32_t *out, int64_t k,
                          hipStream_t st);
// stable merge sorts for captured batches (no memset is recorded: the radix
// sort's hipMemsetAsync has given wrong results on graph replays)
size_t merge_sort_pairs_u32(void *tmp, size_t tmp_bytes, const uint32_t *kin, uint32_t *kout, const uint32_t *vin,
                            uint32_t *vout, int64_t k, hipStream_t st);
size_t merge_sort_keys_u64(void *tmp, size_t tmp_bytes, const unsigned long long *kin, unsigned long long *kout,
                           int64_t k, hipStream_t st);

}  // namespace bingo

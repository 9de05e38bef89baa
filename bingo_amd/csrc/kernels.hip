// gfx950 kernels of the bias-factorized walk engine.
//
// Execution model: one wave64 owns one node of a sorted update batch (lanes
// 0..31 own bit groups 0..31, all 64 lanes cooperate on arc chunks); one
// lane owns one walker.  Nothing here is a dense contraction, so there is no
// MFMA: every kernel is bound by HBM latency/bandwidth of dependent gathers.
//
// Reference kernels replaced: buildKernel / insertKernel / deleteKernel
// (include/operation.cuh:12-30, 122-143, 198-214), NodeData::build / insert /
// deleteE / buildAliasTable / sample (include/Bingo.cuh:111-267), and the
// walk kernels deepwalkKernel / node2vecKernel / pprKernel / samplingKernel
// (operation.cuh:252-368, 407-430).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "store.h"

namespace bingo {

// ------------------------------------------------------------------------
// wave helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint32_t rl32(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, int l)
{
    return ((uint64_t)rl32((uint32_t)(v >> 32), l) << 32) | rl32((uint32_t)v, l);
}

// LDS-permute forms, correct whatever lanes are active (walk kernel tails,
// export)
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
// inclusive prefix sum over the wave
__device__ __forceinline__ uint32_t wave_incl32(uint32_t v)
{
    const int ln = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (ln >= d) v += t;
    }
    return v;
}

// dst[0, n) = src[0, n) by one wave, kCU loads in flight per lane
template <typename T>
__device__ __forceinline__ void wave_copy(T *__restrict__ dst, const T *__restrict__ src, uint32_t n, int ln)
{
    constexpr uint32_t kCU = 8;
    uint32_t i = (uint32_t)ln;
    for (; i + (kCU - 1) * 64 < n; i += kCU * 64) {
        T x[kCU];
#pragma unroll
        for (uint32_t q = 0; q < kCU; ++q) x[q] = src[i + q * 64];
#pragma unroll
        for (uint32_t q = 0; q < kCU; ++q) dst[i + q * 64] = x[q];
    }
    for (; i < n; i += 64) dst[i] = src[i];
}

// DPP forms of the wave scans for the update kernels, where all 64 lanes are
// active: row shifts 1, 2, 4, 8 scan each 16-lane row, then row broadcasts of
// lanes 15 and 31 carry the row totals; a few VALU steps instead of a chain of
// LDS permutes.  Lane 63 ends with the total.
#define BINGO_DPP(v, ctrl, rows) ((uint32_t)__builtin_amdgcn_update_dpp(0, (int)(v), (ctrl), (rows), 0xf, false))
__device__ __forceinline__ uint32_t dpp_incl_add(uint32_t v)
{
    v += BINGO_DPP(v, 0x111, 0xf);
    v += BINGO_DPP(v, 0x112, 0xf);
    v += BINGO_DPP(v, 0x114, 0xf);
    v += BINGO_DPP(v, 0x118, 0xf);
    v += BINGO_DPP(v, 0x142, 0xa);
    v += BINGO_DPP(v, 0x143, 0xc);
    return v;
}
__device__ __forceinline__ uint32_t dpp_incl_or(uint32_t v)
{
    v |= BINGO_DPP(v, 0x111, 0xf);
    v |= BINGO_DPP(v, 0x112, 0xf);
    v |= BINGO_DPP(v, 0x114, 0xf);
    v |= BINGO_DPP(v, 0x118, 0xf);
    v |= BINGO_DPP(v, 0x142, 0xa);
    v |= BINGO_DPP(v, 0x143, 0xc);
    return v;
}
#undef BINGO_DPP
__device__ __forceinline__ uint32_t dpp_sum32(uint32_t v) { return rl32(dpp_incl_add(v), 63); }
__device__ __forceinline__ uint32_t dpp_or32(uint32_t v) { return rl32(dpp_incl_or(v), 63); }

// 64-bit wave sum / inclusive prefix sum, exact when the result is below
// 2^64: three DPP scans of the 24-, 24- and 16-bit pieces of each value, none
// of which can overflow 32 bits over 64 lanes
__device__ __forceinline__ uint64_t dpp_sum64(uint64_t v)
{
    const uint64_t a = dpp_sum32((uint32_t)v & 0xFFFFFFu), b = dpp_sum32((uint32_t)(v >> 24) & 0xFFFFFFu),
                   c = dpp_sum32((uint32_t)(v >> 48));
    return a + (b << 24) + (c << 48);
}
__device__ __forceinline__ uint64_t dpp_incl64(uint64_t v)
{
    const uint64_t a = dpp_incl_add((uint32_t)v & 0xFFFFFFu), b = dpp_incl_add((uint32_t)(v >> 24) & 0xFFFFFFu),
                   c = dpp_incl_add((uint32_t)(v >> 48));
    return a + (b << 24) + (c << 48);
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t x)
{
    return x <= 1u ? 1u : (1u << (32 - __clz(x - 1u)));
}

// ------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter (walker lo, walker hi,
// step, attempt), key = seed.  Replaces the reference's xoshiro256++ stream
// (include/utils.cuh:60-72).
// ------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                        uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// ------------------------------------------------------------------------
// neighbour multiset (node2vec's exact "x in N(prev)" test)
// ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hslot(uint32_t key, uint32_t cap)
{
    return (uint32_t)(((uint64_t)(key * 0x9E3779B1u) * cap) >> 32);
}

// A neighbour entry is 16 bytes: key | arc count << 32, then the arcs'
// weight sum (integer parts); a table of cap entries takes 2 cap pool words.
// hash_add adds `cnt` arcs of total weight `w` to key; returns 1 if the key
// was new.  A new key is one 64-bit CAS (key and count together); a present
// key adds to the count half; the weight sum is one 64-bit atomic add.
__device__ uint32_t hash_add(uint64_t *tab, uint32_t cap, uint32_t key, uint32_t cnt, uint64_t w)
{
    uint32_t i = hslot(key, cap);
    const uint32_t mask = cap - 1u;
    const unsigned long long add = (unsigned long long)cnt << 32;
    for (;;) {
        unsigned long long *p = reinterpret_cast<unsigned long long *>(tab + 2ull * i);
        const unsigned long long old = atomicCAS(p, 0ull, add | key);
        if (old == 0ull || (uint32_t)old == key) {
            if (old) atomicAdd(p, add);
            if (w) atomicAdd(p + 1, (unsigned long long)w);
            return old == 0ull;
        }
        i = (i + 1u) & mask;
    }
}

__device__ void hash_dec(uint64_t *tab, uint32_t cap, uint32_t key, uint64_t w)
{
    uint32_t i = hslot(key, cap);
    const uint32_t mask = cap - 1u;
    for (uint32_t probe = 0; probe < cap; ++probe) {
        uint32_t *kp = reinterpret_cast<uint32_t *>(tab + 2ull * i);
        if (*reinterpret_cast<volatile uint32_t *>(kp) == key) {
            atomicSub(kp + 1, 1u);
            if (w) atomicSub(reinterpret_cast<unsigned long long *>(tab + 2ull * i + 1), (unsigned long long)w);
            return;
        }
        i = (i + 1u) & mask;
    }
}

// neighbour-table locator kept in every alias slot of its node:
// (log2 cap + 1) | offset << 6, 0 when the node has no table
__device__ __forceinline__ uint64_t make_hmeta(uint64_t off, uint32_t cap)
{
    return cap ? ((uint64_t)(32 - __clz(cap)) | (off << 6)) : 0ull;
}

// hm: a table locator, possibly with the alias slot's bits above it (masked)
__device__ __forceinline__ bool is_neighbor(const uint64_t *__restrict__ hash, uint64_t hm, uint32_t key)
{
    hm &= kHmetaMask;
    const uint32_t lg = (uint32_t)(hm & 63);
    if (!lg) return false;
    const uint32_t cap = 1u << (lg - 1u), mask = cap - 1u;
    const uint64_t *tab = hash + (hm >> 6);
    uint32_t i = hslot(key, cap);
    for (;;) {
        const uint64_t e = tab[2ull * i];
        const uint32_t k = (uint32_t)e;
        if (k == key) return (e >> 32) != 0;
        if (k == 0u) return false;
        i = (i + 1u) & mask;
    }
}

// weight sum of the live arcs to `key` in the table located by hm (0 if none)
__device__ __forceinline__ uint64_t neighbor_weight(const uint64_t *__restrict__ hash, uint64_t hm, uint32_t key)
{
    hm &= kHmetaMask;
    const uint32_t lg = (uint32_t)(hm & 63);
    if (!lg) return 0ull;
    const uint32_t cap = 1u << (lg - 1u), mask = cap - 1u;
    const uint4 *tab = reinterpret_cast<const uint4 *>(hash + (hm >> 6));
    uint32_t i = hslot(key, cap);
    for (;;) {
        const uint4 e = tab[i];  // one 16-byte entry
        if (e.x == key) return ((uint64_t)e.w << 32) | e.z;
        if (e.x == 0u) return 0ull;
        i = (i + 1u) & mask;
    }
}

// ------------------------------------------------------------------------
// float weights (ref wtype::Float, include/Bingo.cuh:65-69,176-177,193-194,
// 216-217,254-264): an arc's weight w (IEEE-754 single, 0 < w < 2^32) is
// split into its integer part, which goes through the bit groups like a u32
// weight, and its fraction to 2^-32 resolution, kept per slot.  Valid iff
// finite, in range and not 0 after the floor.  Every step is exact in double
// precision, so the split is the same on any IEEE-754 machine.
// ------------------------------------------------------------------------
__device__ __forceinline__ bool split_float(uint32_t bits, uint32_t &ip, uint32_t &fr)
{
    const double d = (double)__uint_as_float(bits);
    ip = fr = 0u;
    if (!(d > 0.0) || !(d < 4294967296.0)) return false;
    const double fl = floor(d);
    ip = (uint32_t)fl;
    fr = (uint32_t)floor((d - fl) * 4294967296.0);
    return ip || fr;
}

// Fractional-part threshold floor(2^64 F / (S 2^32 + F)) for integer mass
// S < 2^60 and fraction mass F < 2^60 (units of 2^-32): restoring division,
// one quotient bit per step, remainder below the divisor (< 2^93).  All-ones
// for S = 0 (every draw takes the fractional part), 0 for F = 0.
__device__ uint64_t frac_threshold(uint64_t S, uint64_t F)
{
    typedef unsigned __int128 u128;
    if (!F) return 0ull;
    if (!S) return ~0ull;
    const u128 D = ((u128)S << 32) + F;
    u128 r = F;
    uint64_t q = 0;
    for (int i = 0; i < 64; ++i) {
        r <<= 1;
        q <<= 1;
        if (r >= D) {
            r -= D;
            q |= 1ull;
        }
    }
    return q;
}

// ------------------------------------------------------------------------
// alias table: integer-exact sweep over 32 fixed slots, built by one wave.
// Lane j holds group j's size and member offset.  Masses M_j = size_j << j,
// s_j = 32 M_j against slot capacity S = sum M_j.  Lights (s_j < S) and
// heavies (the rest), each in index order.  The sequential sweep -- each
// light takes the current heavy as alias and charges it its deficit S - s_j;
// a heavy whose residual falls below S turns light on the spot, with the next
// heavy as alias -- is computed lane-parallel from two prefix sums: A_j, the
// lights' deficits before light j, and E_k, the heavies' excesses s - S up to
// and including heavy k.  Light j's alias is the first heavy with E_k >= A_j;
// heavy k splits iff E_k falls inside a light's deficit range [A_j, A_j + d_j),
// with tau_k = S - (A_j + d_j - E_k) and the next heavy as alias; other
// heavies keep tau = S.  Stay threshold T_j = floor(tau_j 2^64 / S).
// Replaces the float/EPS table of NodeData::buildAliasTable
// (include/Bingo.cuh:111-158).
//
// Bounds: group sizes < 2^28 and weights < 2^32 give M_j < 2^59; the node
// limit S < 2^59 (inserts past it are rejected) keeps the prefix sums, at most
// 31 S, below 2^64.
// ------------------------------------------------------------------------
constexpr uint64_t kMaxNodeMass = 1ull << 59;

// floor(tau 2^64 / S) for tau < S < 2^60: a double-precision estimate (off by
// at most a few thousand), one refinement from the exact 128-bit remainder,
// then at most a step onto the exact quotient (the loops only finish the job;
// 2*10^7 random and edge-case pairs checked against 128-bit division needed
// one step at most).
__device__ __forceinline__ uint64_t frac64(uint64_t tau, uint64_t S)
{
    typedef __int128 i128;
    const double two64 = 18446744073709551616.0;
    const double dS = (double)S;
    const double est = (double)tau / dS * two64;
    uint64_t q = est >= two64 ? ~0ull : (uint64_t)est;
    i128 r = ((i128)tau << 64) - (i128)((unsigned __int128)q * S);
    // |r| < 2^73: high part exact, low 32 bits added once (no cancellation)
    const double dr = (double)(int64_t)(r >> 32) * 4294967296.0 + (double)(uint32_t)r;
    const int64_t d = (int64_t)floor(dr / dS);
    q += (uint64_t)d;
    r -= (i128)d * (i128)S;
    while (r < 0) {
        --q;
        r += S;
    }
    while (r >= (i128)S) {
        ++q;
        r -= S;
    }
    return q;
}

// node2vec's folded return region as a 64-bit threshold: floor(2^64 num /
// den), num = w_ret fold_r, den = S 2^27 + num, both cut to their top 60 bits
// by the same shift; 0 when num = 0, all-ones when num >= den after the cut.
// S < 2^59 and fold_r < 2^59 keep num and den below 2^119.
__device__ __forceinline__ uint64_t fold_threshold(uint64_t S, uint64_t wret, uint64_t fold_r)
{
    typedef unsigned __int128 u128;
    u128 num = (u128)wret * fold_r;
    if (!num) return 0ull;
    u128 den = ((u128)S << 27) + num;
    const uint64_t hi = (uint64_t)(den >> 64), lo = (uint64_t)den;
    const int bits = hi ? 128 - __clzll(hi) : 64 - __clzll(lo);
    if (bits > 60) {
        num >>= bits - 60;
        den >>= bits - 60;
    }
    if (num >= den) return ~0ull;
    return frac64((uint64_t)num, (uint64_t)den);
}

// A node's integer weight S in 17 bits (kept in every alias slot of an
// integer-weight store): exact below 2^11, else exponent e << 11 | the top 11
// bits m, with m 2^e <= S < (m + 1) 2^e.
__device__ __forceinline__ uint32_t mass_code(uint64_t S)
{
    const int bl = S ? 64 - __clzll((long long)S) : 0;
    if (bl <= 11) return (uint32_t)S;
    const int e = bl - 11;
    return ((uint32_t)e << 11) | (uint32_t)(S >> e);
}

// Bracket [lo, hi) around the fold threshold tout for a node whose mass
// code is `code` and whose return arc weighs w: a coin below lo is below
// tout, a coin at or above hi is not.  tout is within 33 of 2^64 num / den
// (the 60-bit cut moves the ratio by at most 2^-59), which falls as S grows,
// so the thresholds at the range's ends widened by 256 bracket every S in
// it.  lo = hi = tout when the code holds S exactly.
__device__ __forceinline__ void fold_bracket(uint32_t code, uint64_t w, uint64_t fold_r, uint64_t &lo, uint64_t &hi)
{
    const uint32_t e = code >> 11;
    const uint64_t slo = (uint64_t)(code & 2047u) << e, shi = slo + ((1ull << e) - 1ull);
    if (slo == shi) {
        lo = hi = fold_threshold(slo, w, fold_r);
        return;
    }
    const uint64_t a = fold_threshold(shi, w, fold_r), b = fold_threshold(slo, w, fold_r);
    lo = a > 256u ? a - 256u : 0ull;
    hi = b < ~0ull - 256u ? b + 256u : ~0ull;
}

// A walker's fold state for one step.  Until the return arc's weight is
// looked up (known false), a coin at or above fold_cap -- a bound on every
// node's threshold, since a return weight never exceeds its node's weight
// -- cannot take the return arc, and no lookup is made for it.
struct FoldState {
    uint64_t lo, hi, w;
    bool known;
};

// Whether the fold coin of (walker, step, attempt) takes the return arc.
// The first coin below fold_cap looks up the return arc's weight in cur's
// neighbour table and brackets the threshold from the slot's mass code; the
// node's exact mass is read only when a coin falls inside the bracket,
// which then collapses to the exact threshold.
__device__ __forceinline__ bool fold_step(const DevStore &s, int32_t cur, int32_t prev, uint64_t hcur,
                                          uint64_t fold_r, uint64_t fold_cap, FoldState &f, uint32_t &lookups,
                                          uint32_t w0, uint32_t w1, uint32_t step, uint32_t att, uint32_t k0,
                                          uint32_t k1)
{
    const uint4 m = philox(w0, w1, step | kMixBit, att, k0, k1);
    const uint64_t c = ((uint64_t)m.x << 32) | m.y;
    if (!f.known) {
        if (c >= fold_cap) return false;
        f.known = true;
        ++lookups;
        f.w = neighbor_weight(s.hash, hcur, (uint32_t)prev);
        fold_bracket((uint32_t)(hcur >> 42), f.w, fold_r, f.lo, f.hi);
    }
    if (!f.hi) return false;  // tout = 0
    if (c < f.lo) return true;
    if (c >= f.hi) return false;
    f.lo = f.hi = fold_threshold(s.nmass[cur], f.w, fold_r);
    return c < f.lo;
}

// Returns the node's integer mass S (all lanes).  All 64 lanes active.  With
// float weights, F is the node's new fraction mass: the node's FNode is
// written and every slot carries the threshold's 17-bit ceiling.  dm: the
// node's dense groups, whose locators are kDenseLoc | j.
__device__ uint64_t alias_rebuild(const DevStore &s, int32_t u, uint32_t size, uint64_t off, uint64_t hmeta,
                                  uint64_t F, uint32_t dm, int ln)
{
    const bool own = ln < kGroups;
    const uint64_t M = own ? ((uint64_t)size << ln) : 0ull;
    const uint64_t S = dpp_sum64(M);
    const uint64_t sj = M << 5;
    const bool light = own && sj < S, heavy = own && !light;
    const uint64_t d = light ? S - sj : 0ull, e = heavy ? sj - S : 0ull;
    const uint64_t A = dpp_incl64(d) - d;  // lights' deficits before this lane
    const uint64_t E = dpp_incl64(e);      // heavies' excesses up to this lane
    const uint64_t Aend = A + d;
    uint64_t tau = light ? sj : S;
    int ali = ln;
    bool found = false;
    for (uint32_t hb = S ? (uint32_t)__ballot(heavy) : 0u; hb;) {  // S = 0: no lights
        const int k = __ffs(hb) - 1;
        hb &= hb - 1u;
        const uint64_t Ek = rl64(E, k);
        if (light && !found && A <= Ek) {
            ali = k;
            found = true;
        }
        const uint64_t cont = __ballot(light && A <= Ek && Ek < Aend);
        if (cont && hb) {  // heavy k turns light inside that light's range
            const uint64_t end = rl64(Aend, __ffsll((unsigned long long)cont) - 1);
            if (ln == k) {
                tau = S - (end - Ek);
                ali = __ffs(hb) - 1;
            }
        }
    }
    uint64_t T = ~0ull;
    if (own && tau != S) T = frac64(tau, S);
    uint64_t thr_hi = 0;
    if (s.fw) {
        uint64_t thr = 0;
        if (ln == 0) {
            thr = frac_threshold(S, F);
            FNode f;
            f.F = F;
            f.thr = thr;
            s.fnode[u] = f;
        }
        thr = rl64(thr, 0);
        thr_hi = (thr >> 47) + ((thr & ((1ull << 47) - 1)) != 0);
        thr_hi = thr_hi > kThrHiMax ? kThrHiMax : thr_hi;
    }
    if (!s.fw) thr_hi = mass_code(S);
    if (s.nmass && ln == 0) s.nmass[u] = S;
    // this lane's group as an outcome: its members, or a dense group's index
    // (0 when empty: such an outcome is never drawn)
    const uint64_t stay = !size ? 0ull
                          : ((dm >> (ln & 31)) & 1u) ? kDenseLoc | (uint64_t)ln
                                                     : off | ((uint64_t)size << kOffBits);
    const uint64_t alias = __shfl(stay, ali, 64);
    if (own) {
        AliasEnt ae;
        ae.T = T;
        const bool stay_first = (T >> 63) != 0;  // the likelier outcome first
        ae.first = stay_first ? stay : alias;
        ae.second = stay_first ? alias : stay;
        ae.meta = (uint64_t)ali | (hmeta << 5) | (thr_hi << 47);
        s.alias[(size_t)u * kGroups + ln] = ae;
    }
    return S;
}

// ------------------------------------------------------------------------
// small kernels
// ------------------------------------------------------------------------
__global__ void k_reset_ctl(Ctl *ctl)
{
    const int t = threadIdx.x;
    if (t < NPOOL) ctl->need[t] = 0;
    if (t == 0) {
        ctl->error &= (unsigned)ERR_OVERRUN;  // sticky: a batch overran a pool
        ctl->nbig = 0;
        ctl->nbigcopy = 0;
        ctl->taskctr = 0;
    }
}

__global__ void k_reset_walk(Ctl *ctl)
{
    if (threadIdx.x == 0) {
        ctl->steps = 0;
        ctl->attempts = 0;
        ctl->probes = 0;
        ctl->folds = 0;
        ctl->dtrials = 0;
        ctl->cursor = 0;
        ctl->werr = 0;
    }
}

__global__ void k_init_nodes(DevStore s)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t total = ((int64_t)s.n + 1) * kGroups;
    if (i >= total) return;
    const int j = (int)(i & (kGroups - 1));
    AliasEnt e;
    e.T = ~0ull;
    e.first = 0;
    e.second = 0;
    e.meta = (uint64_t)j;
    s.alias[i] = e;
    GroupHdr g = {0, 0, 0};
    s.grp[i] = g;
    if (j == 0) {
        NodeHdr h = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        s.hdr[i >> 5] = h;
        if (s.fw) {
            FNode f = {0, 0};
            s.fnode[i >> 5] = f;
        }
        if (s.nmass) s.nmass[i >> 5] = 0ull;
    }
}

__global__ void k_insert_keys(const DevEdge *__restrict__ e, int64_t k, const unsigned long long *__restrict__ dk,
                              int32_t n, int fw, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, Ctl *ctl)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    if (dk && (uint64_t)i >= *dk) {  // padding of a captured batch
        keys[i] = 0u;
        vals[i] = (uint32_t)i;
        return;
    }
    const DevEdge x = e[i];
    uint32_t ip, fr;
    const bool wbad = fw ? !split_float(x.w, ip, fr) : x.w == 0u;
    const bool bad = x.u < 1 || x.u > n || x.v < 1 || x.v > n || wbad;
    if (bad) atomicOr(&ctl->error, (unsigned)ERR_INVAL);
    keys[i] = bad ? 0u : (uint32_t)x.u;
    vals[i] = (uint32_t)i;
}

// Arcs in node order.  With float weights (frac != null) the gathered arc
// carries the integer part as its weight and frac[i] its fraction.
__global__ void k_gather_edges(const DevEdge *__restrict__ e, const uint32_t *__restrict__ order, int64_t k,
                               const unsigned long long *__restrict__ dk, DevEdge *__restrict__ out,
                               uint32_t *__restrict__ frac)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    const uint32_t o = order[i];
    DevEdge x = (dk && o >= *dk) ? DevEdge{0, 0, 0u} : e[o];  // padding: an empty arc (node 0)
    if (frac) {
        uint32_t ip, fr;
        split_float(x.w, ip, fr);
        x.w = ip;
        frac[i] = fr;
    }
    out[i] = x;
}

// ------------------------------------------------------------------------
// insert: plan (capacity growth per node) and apply
// ------------------------------------------------------------------------
struct Growth {
    uint32_t cnt, new_n, new_back, pcsum;
    uint32_t dm;      // the node's dense groups after the run
    bool heavy_node;  // the node's integer weight would reach kMaxNodeMass
    uint32_t cap_slots, cap_back, cap_hash;
    bool g_slots, g_back, g_hash;
    // per lane (group j = lane)
    uint32_t addj, cap_g;
    bool g_g;
};

// capacity of a region that has to grow to hold `need` elements (DevStore::fit)
__device__ __forceinline__ uint32_t grow_cap(const DevStore &s, uint32_t need, uint32_t least)
{
    if (s.fit) return class_size(class_up(need + (uint32_t)(((uint64_t)need * (s.fit - 1u) + 0xFFFFu) >> 16)));
    return pow2_at_least(max(need, least));
}

// Runs of at most kSmallRun arcs (most runs of a batch spread over many
// nodes) take a lane-per-group path: lane j walks the arcs and tests bit j,
// instead of one ballot per set bit.  Both paths give the same counts, and the
// same member and back-index order (arcs in run order within each group).
constexpr uint32_t kSmallRun = 4;

// counts of the run and the growth decisions; identical in plan and apply.
// w0: this lane's arc of the first chunk (already loaded; 0 past the run).
// A node without slots takes its dense groups from the run (DevStore::rate16);
// back entries and member regions are for its listed groups only.
template <bool DN>
__device__ Growth plan_run(const DevStore &s, const DevEdge *__restrict__ e, uint32_t cnt, const NodeHdr &h,
                           const GroupHdr &gh, uint32_t w0, int ln)
{
    Growth G;
    uint32_t addj = 0;
    uint64_t wtot = 0;
    if (cnt <= kSmallRun) {  // lane j tests bit j of each arc in turn
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t wi = rl32(w0, (int)i);
            if (ln < kGroups) addj += (wi >> ln) & 1u;
            wtot += wi;
        }
    } else {  // 64 arcs at a time, one ballot per set bit
        uint64_t wsum = 0;
        for (uint32_t c = 0; c < cnt; c += 64) {
            const uint32_t i = c + (uint32_t)ln;
            const uint32_t w = c == 0 ? w0 : (i < cnt ? e[i].w : 0u);
            wsum += w;
            const uint32_t bits = dpp_or32(w);
            for (uint32_t b = bits; b;) {
                const int j = __ffs(b) - 1;
                b &= b - 1u;
                const uint64_t m = __ballot((w >> j) & 1u);
                if (ln == j) addj += (uint32_t)__popcll(m);
            }
        }
        wtot = dpp_sum64(wsum);
    }
    uint32_t dm = 0;
    if (DN) {
        dm = (uint32_t)__builtin_amdgcn_readfirstlane((int)h.dmask);
        if (h.nslots == 0)  // the node's first slots: its classes (ref build, operation.cuh:57-72)
            dm = (uint32_t)__ballot(ln < kGroups && addj && ((uint64_t)addj << 16) >= (uint64_t)s.rate16 * cnt);
    }
    const bool dense = ln < kGroups && ((dm >> ln) & 1u);
    const uint32_t pcsum = dpp_sum32(ln < kGroups && !dense ? addj : 0u);  // back entries of the run
    G.dm = dm;
    G.cnt = cnt;
    G.addj = addj;
    G.pcsum = pcsum;
    // integer mass before (group sizes << j) plus the run's weights; the
    // run's sum is below 2^28 * 2^32 and the mass below 2^59, so no overflow
    G.heavy_node = dpp_sum64(ln < kGroups ? (uint64_t)gh.size << ln : 0ull) + wtot >= kMaxNodeMass;
    G.new_n = h.nslots + cnt;
    G.g_slots = G.new_n > h.slot_cap;
    G.cap_slots = G.g_slots ? grow_cap(s, G.new_n, (uint32_t)s.min_slots) : h.slot_cap;
    G.new_back = h.back_used + pcsum;
    G.g_back = G.new_back > h.back_cap;
    G.cap_back = G.g_back ? grow_cap(s, G.new_back, 16u) : h.back_cap;
    const uint32_t need_g = gh.size + addj;
    G.g_g = ln < kGroups && addj && !dense && need_g > gh.cap;
    G.cap_g = G.g_g ? grow_cap(s, need_g, 4u) : gh.cap;
    const uint32_t hneed = 2u * (h.hash_used + cnt);
    G.g_hash = s.nbr && hneed > h.hash_cap;
    G.cap_hash = G.g_hash ? pow2_at_least(max(hneed, 16u)) : h.hash_cap;
    return G;
}

// whether the run moves a region of kBigCopyBytes or more (wave-uniform)
__device__ __forceinline__ bool run_moves_big(const NodeHdr &h, const GroupHdr &gh, const Growth &G)
{
    const bool grp = __ballot(G.g_g && (uint64_t)gh.size * 8u >= kBigCopyBytes) != 0ull;
    return grp || (G.g_slots && (uint64_t)h.nslots * 4u >= kBigCopyBytes) ||
           (G.g_back && (uint64_t)h.back_used * 4u >= kBigCopyBytes) ||
           (G.g_hash && (uint64_t)h.hash_cap * 16u >= kBigCopyBytes);
}

__device__ __forceinline__ bool run_too_big(const NodeHdr &h, const Growth &G)
{
    // group sizes must fit the 28-bit size field of an alias slot; the node's
    // integer weight stays below kMaxNodeMass
    return (uint64_t)h.nslots + G.cnt >= kMaxGroupSize || (uint64_t)h.back_used + G.pcsum >= (1ull << 31) ||
           2ull * ((uint64_t)h.hash_used + G.cnt) >= (1ull << 31) || G.heavy_node;
}

// ------------------------------------------------------------------------
// Run loops are software-pipelined: while a wave works on one run it already
// holds the next runs' node headers, group headers and first 64 arcs (two
// runs ahead in the insert kernels, one in the delete kernel), and the run
// table entry of the one after, so the dependent round trips at the start
// of a run (table -> node -> records) overlap with the previous runs' work.  Runs name distinct nodes, so reading the next
// node's records early sees the same values as reading them later.
// ------------------------------------------------------------------------
struct RunMeta {
    int32_t u;  // 0: no run (past the end, or an invalid arc's key)
    uint32_t start, cnt;
};

__device__ __forceinline__ RunMeta run_meta(const DevStore &s, const Runs &r, uint32_t ri, uint32_t nruns)
{
    RunMeta m = {0, 0, 0};
    if (ri < nruns) {
        const int32_t u = (int32_t)r.node[ri];
        if (u >= 1 && u <= s.n) {
            m.u = u;
            m.start = r.start[ri];
            m.cnt = r.count[ri];
        }
    }
    return m;
}

struct RunPre {
    NodeHdr h;
    GroupHdr gh;  // this lane's group (lanes < 32)
    DevEdge x;    // insert: this lane's arc of the first chunk
    uint32_t f;   // insert with float weights: its fraction; delete: this lane's slot
};

template <bool DEL>
__device__ __forceinline__ RunPre run_pre(const DevStore &s, const DevEdge *__restrict__ E,
                                          const uint32_t *__restrict__ aux, const RunMeta &m, int ln)
{
    RunPre p;
    p.h = NodeHdr{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    p.gh = GroupHdr{0, 0, 0};
    p.x = DevEdge{0, 0, 0u};
    p.f = 0;
    if (m.u) {
        p.h = s.hdr[m.u];
        if (ln < kGroups) p.gh = s.grp[(size_t)m.u * kGroups + ln];
        if ((uint32_t)ln < m.cnt) {
            if (DEL) {
                p.f = aux[m.start + ln];
            } else {
                p.x = E[m.start + ln];
                if (aux) p.f = aux[m.start + ln];
            }
        }
    }
    return p;
}

// The plan also reserves each workgroup's regions: the apply kernel runs with
// the same grid, so every wave sees the same runs in the same order and makes
// the same growth decisions.  Per workgroup, one global atomic per class used
// (counted in LDS first) instead of one per growing region: a single
// contended 64-bit atomic costs ~12 ns on this part.
// class index of a region of capacity cap (its size class, or for a free
// region of any capacity the largest class it holds)
__device__ __forceinline__ uint32_t cls_of(int pool, uint32_t cap) { return (uint32_t)pool * kClasses + class_down(cap); }

// The plan counts each workgroup's region allocations and frees per class in
// LDS and reserves the workgroup's index ranges with one atomic per used
// class (s.wcls); the apply kernel, with the same grid, hands out indices
// from those ranges (any order within a workgroup: the regions of a class
// are interchangeable).
template <bool DN>
__global__ __launch_bounds__(256) void k_insert_plan(DevStore s, const DevEdge *__restrict__ E, Runs r)
{
    __shared__ uint32_t lc[2 * kPC];  // allocations, then frees, per class
    for (int i = threadIdx.x; i < 2 * kPC; i += blockDim.x) lc[i] = 0u;
    __syncthreads();
    const int ln = lane_id();
    const uint32_t nruns = *r.nruns;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    uint32_t nbigcopy = 0;
    RunMeta m0 = run_meta(s, r, wave, nruns), m1 = run_meta(s, r, wave + nwaves, nruns),
            m2 = run_meta(s, r, wave + 2 * nwaves, nruns);
    RunPre p0 = run_pre<false>(s, E, nullptr, m0, ln), p1 = run_pre<false>(s, E, nullptr, m1, ln);
    for (uint32_t ri = wave; ri < nruns; ri += nwaves) {
        const RunMeta m3 = run_meta(s, r, ri + 3 * nwaves, nruns);
        const RunPre p2 = run_pre<false>(s, E, nullptr, m2, ln);
        const RunMeta m = m0;
        const NodeHdr h = p0.h;
        const GroupHdr gh = p0.gh;
        const uint32_t w0 = p0.x.w;
        m0 = m1;
        m1 = m2;
        m2 = m3;
        p0 = p1;
        p1 = p2;
        if (!m.u) continue;  // invalid arcs were flagged already
        const uint32_t cnt = m.cnt;
        const Growth G = plan_run<DN>(s, E + m.start, cnt, h, gh, w0, ln);
        if (run_too_big(h, G)) {
            if (ln == 0) atomicOr(&s.ctl->error, (unsigned)ERR_NOMEM);
            continue;
        }
        nbigcopy += run_moves_big(h, gh, G) ? 1u : 0u;
        if (ln == 0) {
            if (G.g_slots) {
                atomicAdd(&lc[cls_of(POOL_SLOT, G.cap_slots)], 1u);
                if (h.slot_cap) atomicAdd(&lc[kPC + cls_of(POOL_SLOT, h.slot_cap)], 1u);
            }
            if (G.g_back) {
                atomicAdd(&lc[cls_of(POOL_BACK, G.cap_back)], 1u);
                if (h.back_cap) atomicAdd(&lc[kPC + cls_of(POOL_BACK, h.back_cap)], 1u);
            }
            if (G.g_hash) {
                atomicAdd(&lc[cls_of(POOL_HASH, G.cap_hash)], 1u);
                if (h.hash_cap) atomicAdd(&lc[kPC + cls_of(POOL_HASH, h.hash_cap)], 1u);
            }
        }
        if (G.g_g) {
            atomicAdd(&lc[cls_of(POOL_MEM, G.cap_g)], 1u);
            if (gh.cap) atomicAdd(&lc[kPC + cls_of(POOL_MEM, gh.cap)], 1u);
        }
    }
    if (ln == 0 && nbigcopy) atomicAdd(&s.ctl->nbigcopy, nbigcopy);
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * kPC; i += blockDim.x) {
        const uint32_t c = lc[i];
        const int t = i % kPC;
        s.wcls[(size_t)blockIdx.x * 2 * kPC + i] = c ? atomicAdd(i < kPC ? &s.fm->take[t] : &s.fm->freed[t], c) : 0u;
    }
}

// After the plan of a batch: per class, the allocations served by free
// regions and the bump space past them (classes of a pool in order after the
// pool's top), the pending segments of the batch's frees, and the lists'
// layout after the batch.  One workgroup of kPC threads; pool p is waves
// 4p .. 4p + 3.
__global__ __launch_bounds__(kPC) void k_alloc_resolve(DevStore s, int fw, int captured, uint64_t flcap, uint64_t pfcap)
{
    constexpr int kW = kPC / 64, kWP = kClasses / 64;  // waves, waves per pool
    __shared__ uint64_t part[3][kW];
    const int t = threadIdx.x, p = t / kClasses, k = t % kClasses, ln = lane_id(), w = t >> 6;
    FreeMeta *fm = s.fm;
    const uint32_t av = fm->avail[t], tk = fm->take[t], fr = fm->freed[t];
    const uint32_t from = min(tk, av);
    const uint64_t mult = p == POOL_SLOT ? slot_arrays(fw) : p == POOL_HASH ? 2u : 1u;
    const uint64_t ex = (uint64_t)(tk - from) * mult * class_size((uint32_t)k);
    const uint64_t nv = (uint64_t)(av - from) + fr;
    const uint64_t einc = dpp_incl64(ex), pinc = dpp_incl64((uint64_t)fr), ninc = dpp_incl64(nv);
    if (ln == 63) {
        part[0][w] = einc;
        part[1][w] = pinc;
        part[2][w] = ninc;
    }
    __syncthreads();
    uint64_t eb = 0, etot = 0, pb = 0, ptot = 0, nb = 0, ntot = 0;
    for (int i = 0; i < kW; ++i) {
        if (i / kWP == p) {  // the excess prefix stays inside the pool
            eb += i < w ? part[0][i] : 0ull;
            etot += part[0][i];
        }
        pb += i < w ? part[1][i] : 0ull;
        ptot += part[1][i];
        nb += i < w ? part[2][i] : 0ull;
        ntot += part[2][i];
    }
    fm->bump[t] = s.ctl->top[p] + eb + einc - ex;
    if (k == 0) s.ctl->need[p] = etot;
    // a captured batch grows nothing from the host: it applies only if the
    // pools, lists and pending frees have room (and the batch is valid)
    if (captured && k == 0 && s.ctl->top[p] + etot > s.cap[p]) atomicOr(&s.ctl->ustatus, (unsigned)UST_NOROOM);
    fm->pseg[t] = pb + pinc - fr;
    fm->nfbase[t] = nb + ninc - nv;
    fm->nav[t] = (uint32_t)nv;
    if (t == 0) {
        s.ctl->pend_total = ptot;
        s.ctl->fl_total = ntot;
        if (captured) {
            const unsigned e = s.ctl->error;
            if (e & (ERR_INVAL | ERR_OVERRUN)) atomicOr(&s.ctl->ustatus, (unsigned)UST_INVAL);
            if ((e & ERR_NOMEM) || ntot > flcap || ptot > pfcap) atomicOr(&s.ctl->ustatus, (unsigned)UST_NOROOM);
        }
    }
}

// After a batch (and its copy tasks): class t's list becomes its unused free
// regions followed by the batch's frees, at [nfbase[t], nfbase[t] + nav[t])
// of s.fl2.  The grid strides over all the new lists' entries (one class may
// hold millions of free regions), each finding its class by a search of
// nfbase (ascending in t; an empty class shares its base with the next).
__global__ __launch_bounds__(256) void k_freelist_merge(DevStore s)
{
    __shared__ uint64_t nfb[kPC];
    if (s.ctl->ustatus) return;  // a captured batch that applied nothing
    const FreeMeta *fm = s.fm;
    for (int t = threadIdx.x; t < kPC; t += blockDim.x) nfb[t] = fm->nfbase[t];
    __syncthreads();
    const uint64_t total = s.ctl->fl_total;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = kPC - 1;  // the largest t with nfb[t] <= i
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (nfb[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const uint64_t q = i - nfb[lo];
        const uint32_t av = fm->avail[lo], from = min(fm->take[lo], av), rem = av - from;
        s.fl2[i] = q < rem ? s.fl[fm->fbase[lo] + from + q] : s.pf[fm->pseg[lo] + (q - rem)];
    }
}

// then the lists' bookkeeping (after every read of the old one)
__global__ void k_freelist_commit(DevStore s)
{
    if (s.ctl->ustatus) return;
    FreeMeta *fm = s.fm;
    const int t = threadIdx.x;
    if (t < kPC) {
        fm->avail[t] = fm->nav[t];
        fm->fbase[t] = fm->nfbase[t];
        fm->take[t] = 0u;
        fm->freed[t] = 0u;
    }
}

// A workgroup's class allocator in the apply kernel:
// the plan reserved index ranges [wcls[t], wcls[t] + count) per class for
// allocations and frees; li hands out indices inside them (LDS counters).
struct ClassAlloc {
    const uint32_t *base;  // this workgroup's 2 kPC range starts
    uint32_t *li;          // 2 kPC LDS counters
    // a region of class `cap` of pool p (any lane that calls); ~0 on overrun
    __device__ __forceinline__ uint64_t take(const DevStore &s, int p, uint32_t cap) const
    {
        const uint32_t t = cls_of(p, cap);
        const uint32_t g = base[t] + atomicAdd(&li[t], 1u);
        const FreeMeta *fm = s.fm;
        const uint64_t elems = (uint64_t)(p == POOL_SLOT ? slot_arrays(s.fw) : p == POOL_HASH ? 2u : 1u) * cap;
        const uint32_t av = fm->avail[t];
        const uint64_t off = g < av ? s.fl[fm->fbase[t] + g] : fm->bump[t] + (uint64_t)(g - av) * elems;
        if (off == ~0ull || off + elems > s.cap[p]) {
            atomicOr(&s.ctl->error, (unsigned)ERR_OVERRUN);
            return ~0ull;
        }
        return off;
    }
    // the region of capacity `cap` at `off` of pool p, free from the next batch on
    __device__ __forceinline__ void give(const DevStore &s, int p, uint32_t cap, uint64_t off) const
    {
        const uint32_t t = cls_of(p, cap);
        s.pf[s.fm->pseg[t] + base[kPC + t] + atomicAdd(&li[kPC + t], 1u)] = off;
    }
};

// the batch's reservations are used: the pool tops move past them
__global__ void k_commit_tops(Ctl *ctl)
{
    if (ctl->ustatus) return;
    const int t = threadIdx.x;
    if (t < NPOOL) ctl->top[t] += ctl->need[t];
}

// Hands a region move of kBigCopyBytes or more to k_copy_tasks (true), unless
// the task table is full (false: the caller's wave moves it).  Wave-uniform.
__device__ __forceinline__ bool defer_task(const DevStore &s, CopyTask t, int ln)
{
    const uint64_t bytes = t.kind == TASK_COPY ? t.n : 16u * t.n;
    t.nch = (uint32_t)((bytes + kCopyChunk - 1) / kCopyChunk);
    unsigned long long at = 0;
    if (ln == 0) at = atomicAdd(&s.ctl->taskctr, (1ull << 32) | t.nch);
    at = __shfl(at, 0, 64);
    const uint32_t ti = (uint32_t)(at >> 32);
    if (ti >= kMaxCopyTasks) return false;
    t.cbase = (uint32_t)at;
    if (ln == 0) s.tasks[ti] = t;
    return true;
}

template <typename T>
__device__ __forceinline__ void region_copy(const DevStore &s, T *dst, const T *src, uint32_t n, int ln)
{
    const uint64_t bytes = (uint64_t)n * sizeof(T);
    if (bytes >= kBigCopyBytes) {
        const CopyTask t = {(uint64_t)dst, (uint64_t)src, bytes, TASK_COPY, 0, 0u, 0u, 0u, 0u};
        if (defer_task(s, t, ln)) return;
    }
    wave_copy(dst, src, n, ln);
}

// `fr` (float weights only) holds the run's fractions, arc for arc; `p` the
// node's records and this lane's first arc, loaded by the run loop.
template <bool DN>
__device__ void insert_run(const DevStore &s, const DevEdge *__restrict__ e, const uint32_t *__restrict__ fr,
                           uint32_t cnt, int32_t u, const RunPre &p, const ClassAlloc &ca, int ln)
{
    NodeHdr h = p.h;
    const GroupHdr gh = p.gh;
    const Growth G = plan_run<DN>(s, e, cnt, h, gh, p.x.w, ln);
    if (run_too_big(h, G)) return;  // rejected by the plan

    // ---- regions: each from its class; the region it replaces is freed ----
    uint64_t a_slot = h.slot_off, a_back = h.back_off, a_hash = h.hash_off;
    uint64_t goff = gh.off;
    if (ln == 0) {
        if (G.g_slots) {
            a_slot = ca.take(s, POOL_SLOT, G.cap_slots);
            if (h.slot_cap) ca.give(s, POOL_SLOT, h.slot_cap, h.slot_off);
        }
        if (G.g_back) {
            a_back = ca.take(s, POOL_BACK, G.cap_back);
            if (h.back_cap) ca.give(s, POOL_BACK, h.back_cap, h.back_off);
        }
        if (G.g_hash) {
            a_hash = ca.take(s, POOL_HASH, G.cap_hash);
            if (h.hash_cap) ca.give(s, POOL_HASH, h.hash_cap, h.hash_off);
        }
    }
    a_slot = rl64(a_slot, 0);
    a_back = rl64(a_back, 0);
    a_hash = rl64(a_hash, 0);
    if (G.g_g) {
        goff = ca.take(s, POOL_MEM, G.cap_g);
        if (gh.cap) ca.give(s, POOL_MEM, gh.cap, gh.off);
    }
    // an overrun (never, with the pools sized from the plan) leaves the store
    // inconsistent and is reported sticky
    const bool fail = a_slot == ~0ull || a_back == ~0ull || a_hash == ~0ull || __any(goff == ~0ull);
    if (fail) return;  // node left as it was

    const uint32_t n0 = h.nslots;
    uint32_t *nv = s.slot + a_slot;
    uint32_t *nw = nv + G.cap_slots;
    uint32_t *nb = nv + 2ull * G.cap_slots;
    uint32_t *nf = nv + 3ull * G.cap_slots;  // float weights only
    if (G.g_slots && n0) {
        const uint32_t *ov = s.slot + h.slot_off;
        region_copy(s, nv, ov, n0, ln);
        region_copy(s, nw, ov + h.slot_cap, n0, ln);
        region_copy(s, nb, ov + 2ull * h.slot_cap, n0, ln);
        if (s.fw) region_copy(s, nf, ov + 3ull * h.slot_cap, n0, ln);
    }
    uint32_t *back = s.back + a_back;
    if (G.g_back && h.back_used) region_copy(s, back, s.back + h.back_off, h.back_used, ln);
    for (uint32_t gm = (uint32_t)__ballot(G.g_g && gh.size); gm;) {
        const int j = __ffs(gm) - 1;
        gm &= gm - 1u;
        region_copy(s, s.mem + rl64(goff, j), s.mem + rl64(gh.off, j), rl32(gh.size, j), ln);
    }
    uint32_t hused = h.hash_used;
    uint64_t *htab = s.hash + a_hash;
    bool rehash = G.g_hash;
    if (rehash && (uint64_t)h.hash_cap * 16u >= kBigCopyBytes) {
        // zeroing and re-inserting by k_copy_tasks, which counts the keys
        // into the header's hash_used
        const CopyTask t = {(uint64_t)htab, (uint64_t)(s.hash + h.hash_off), G.cap_hash, TASK_REHASH, u,
                            h.hash_cap, 0u, 0u, 0u};
        if (defer_task(s, t, ln)) {
            rehash = false;
            hused = 0;
        }
    }
    if (rehash) {
        // Only this wave touches the new table until the kernel ends, so the
        // zeros need only be complete before its own CAS: a workgroup-scope
        // fence (a wait on the stores).  A device-scope fence here would write
        // back and invalidate this XCD's L2 once per growing table.
        for (uint32_t i = ln; i < 2u * G.cap_hash; i += 64) htab[i] = 0ull;
        __threadfence_block();
        uint32_t fresh = 0;
        if (h.hash_cap) {
            const uint64_t *old = s.hash + h.hash_off;
            for (uint32_t i = ln; i < h.hash_cap; i += 64) {
                const uint64_t x = old[2ull * i], w = old[2ull * i + 1];
                if ((uint32_t)x && (x >> 32)) fresh += hash_add(htab, G.cap_hash, (uint32_t)x, (uint32_t)(x >> 32), w);
            }
        }
        hused = dpp_sum32(fresh);
    }

    // ---- append arcs in batch order: ballot + prefix rank per bit group ----
    uint32_t szj = gh.size;
    uint32_t boff_run = h.back_used;
    uint64_t fsum = 0;
    if (cnt <= kSmallRun) {
        // slot records by lanes < cnt; then lane j appends group j's members
        // arc by arc (arc i's back entries start at boff_i, bits ascending)
        const DevEdge x = p.x;  // zero past the run
        const bool act = (uint32_t)ln < cnt;
        const uint32_t sm = ~G.dm;  // listed groups
        const uint32_t pc = (uint32_t)__popc(x.w & sm);
        const uint32_t my_boff = boff_run + dpp_incl_add(pc) - pc;
        if (act) {
            nv[n0 + ln] = (uint32_t)x.v;
            nw[n0 + ln] = x.w;
            nb[n0 + ln] = my_boff;
            if (s.fw) {
                nf[n0 + ln] = p.f;
                fsum = p.f;
            }
        }
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t wi = rl32(x.w, (int)i), vi = rl32((uint32_t)x.v, (int)i), bi = rl32(my_boff, (int)i);
            if (ln < kGroups && ((wi >> ln) & 1u)) {
                if ((sm >> ln) & 1u) {  // a dense group only counts
                    s.mem[goff + szj] = ((uint64_t)vi << 32) | (n0 + i + h.sbase);
                    back[bi + __popc(wi & sm & ((1u << ln) - 1u))] = szj;
                }
                ++szj;
            }
        }
    } else {
        for (uint32_t c = 0; c < cnt; c += 64) {
            const uint32_t i = c + (uint32_t)ln;
            const bool act = i < cnt;
            DevEdge x = {0, 0, 0u};
            if (c == 0) x = p.x;
            else if (act) x = e[i];
            const uint32_t slot = n0 + i;
            const uint32_t pc = (uint32_t)__popc(x.w & ~G.dm);
            const uint32_t pincl = dpp_incl_add(pc);
            const uint32_t my_boff = boff_run + pincl - pc;
            if (act) {
                nv[slot] = (uint32_t)x.v;
                nw[slot] = x.w;
                nb[slot] = my_boff;
                if (s.fw) {
                    const uint32_t f = c == 0 ? p.f : fr[i];
                    nf[slot] = f;
                    fsum += f;
                }
            }
            const uint64_t lt = (1ull << ln) - 1ull;
            uint32_t rank = 0;
            const uint64_t entry = ((uint64_t)(uint32_t)x.v << 32) | (slot + h.sbase);
            for (uint32_t b = dpp_or32(x.w); b;) {
                const int j = __ffs(b) - 1;
                b &= b - 1u;
                const bool has = (x.w >> j) & 1u;
                const uint64_t m = __ballot(has);
                if (!((G.dm >> j) & 1u)) {  // a dense group only counts
                    const uint32_t base = rl32(szj, j);
                    const uint64_t go = rl64(goff, j);
                    if (has) {
                        const uint32_t pos = base + (uint32_t)__popcll(m & lt);
                        s.mem[go + pos] = entry;
                        back[my_boff + rank] = pos;
                        ++rank;
                    }
                }
                if (ln == j) szj += (uint32_t)__popcll(m);
            }
            boff_run += rl32(pincl, 63);
        }
    }
    // the run's arcs enter the neighbour table in k_hash_insert (one thread
    // per arc), which also counts the new keys into hash_used

    // ---- headers ----
    if (ln < kGroups) {
        GroupHdr ng;
        ng.off = goff;
        ng.size = szj;
        ng.cap = G.cap_g;
        s.grp[(size_t)u * kGroups + ln] = ng;
    }
    if (ln == 0) {
        h.slot_off = a_slot;
        h.slot_cap = G.cap_slots;
        h.nslots = G.new_n;
        h.nlive += cnt;
        h.back_off = a_back;
        h.back_cap = G.cap_back;
        h.back_used = G.new_back;
        h.dmask = G.dm;
        if (s.nbr) {
            h.hash_off = a_hash;
            h.hash_cap = G.cap_hash;
            h.hash_used = hused;
        }
        s.hdr[u] = h;
    }
    const uint64_t F = s.fw ? s.fnode[u].F + dpp_sum64(fsum) : 0ull;
    alias_rebuild(s, u, ln < kGroups ? szj : 0u, goff, s.nbr ? make_hmeta(a_hash, G.cap_hash) : 0ull, F, G.dm, ln);
}

// Runs by grid stride in node order (neighbouring runs touch neighbouring
// node records; a shared work counter measured slower).
template <bool DN>
__global__ __launch_bounds__(256) void k_insert_apply(DevStore s, const DevEdge *__restrict__ E,
                                                      const uint32_t *__restrict__ FR, Runs r)
{
    if (s.ctl->ustatus) return;  // a captured batch that does not apply
    __shared__ uint32_t li[2 * kPC];
    for (int i = threadIdx.x; i < 2 * kPC; i += blockDim.x) li[i] = 0u;
    __syncthreads();
    const ClassAlloc ca{s.wcls + (size_t)blockIdx.x * 2 * kPC, li};
    const int ln = lane_id();
    const uint32_t nruns = *r.nruns;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    RunMeta m0 = run_meta(s, r, wave, nruns), m1 = run_meta(s, r, wave + nwaves, nruns),
            m2 = run_meta(s, r, wave + 2 * nwaves, nruns);
    RunPre p0 = run_pre<false>(s, E, FR, m0, ln), p1 = run_pre<false>(s, E, FR, m1, ln);
    for (uint32_t ri = wave; ri < nruns; ri += nwaves) {
        const RunMeta m3 = run_meta(s, r, ri + 3 * nwaves, nruns);
        const RunPre p2 = run_pre<false>(s, E, FR, m2, ln);
        if (m0.u) insert_run<DN>(s, E + m0.start, FR ? FR + m0.start : nullptr, m0.cnt, m0.u, p0, ca, ln);
        m0 = m1;
        m1 = m2;
        m2 = m3;
        p0 = p1;
        p1 = p2;
    }
}

// ------------------------------------------------------------------------
// delete: check (all-or-nothing) and apply
// ------------------------------------------------------------------------
__global__ void k_delete_keys(const int32_t *__restrict__ d, int64_t k, const unsigned long long *__restrict__ dk,
                              int32_t n, Ctl *ctl, unsigned long long *__restrict__ keys)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    if (dk) {  // captured: key 0 marks padding, so a real node-0 delete is flagged here
        if ((uint64_t)i >= *dk) {
            keys[i] = 0ull;
            return;
        }
        if (d[2 * i] < 1 || d[2 * i] > n) atomicOr(&ctl->error, (unsigned)ERR_INVAL);
    }
    keys[i] = ((unsigned long long)(uint32_t)d[2 * i] << 32) | (uint32_t)d[2 * i + 1];
}

__global__ void k_delete_check(DevStore s, const unsigned long long *__restrict__ keys, int64_t k, int pads,
                               uint32_t *__restrict__ us, uint32_t *__restrict__ slots)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    const unsigned long long key = keys[i];
    if (pads && key == 0ull) {  // padding of a captured batch
        us[i] = 0u;
        slots[i] = 0u;
        return;
    }
    const int32_t u = (int32_t)(key >> 32);
    const uint32_t sl = (uint32_t)key;
    bool bad = u < 1 || u > s.n || (i > 0 && keys[i - 1] == key);
    if (!bad) {
        const NodeHdr h = s.hdr[u];
        // live: a weight, or with float weights a weight or a fraction
        bad = sl >= h.nslots ||
              (s.slot[h.slot_off + h.slot_cap + sl] == 0u && !(s.fw && s.slot[h.slot_off + 3ull * h.slot_cap + sl]));
    }
    if (bad) atomicOr(&s.ctl->error, (unsigned)ERR_INVAL);
    us[i] = bad ? 0u : (uint32_t)u;
    slots[i] = sl;
}

// Batched delete of one node (slots ascending, distinct, live).  Per bit
// group j, with k_j deleted members and new size S2 = size - k_j: the deleted
// members below S2 are holes, and the k-th hole (holes in slot order) takes
// the k-th surviving member of the tail [S2, size) in position order; the
// moved member's back index is rewritten.  This is the batch form of
// NodeData::deleteE's swap-with-last (include/Bingo.cuh:199-218).
//
// Work split: the deletes are staged 64 at a time, lane t holding delete t's
// slot, weight and back offset; lane j < 32 owns group j and reads a staged
// delete with readlane.  A group lane issues its back-index gathers, tail
// reads and moved-member gathers kDB at a time, so a hub's thousands of
// deletes cost a few dependent round trips per kDB deletes rather than a
// chain per delete.  The tail is compacted in place first (survivors to
// [S2, S2 + holes)), which lets the k-th hole read its survivor at S2 + k.
constexpr int kDB = 8;

template <bool DN>
__device__ void delete_run(const DevStore &s, const uint32_t *__restrict__ slots, uint32_t cnt, int32_t u,
                           const RunPre &p, int ln)
{
    const NodeHdr h = p.h;
    uint32_t *V = s.slot + h.slot_off;
    uint32_t *W = V + h.slot_cap;
    const uint32_t *B = V + 2ull * h.slot_cap;
    uint32_t *back = s.back + h.back_off;
    const bool own = ln < kGroups;
    // bit counts this lane's group; lbit marks list work (0 for a dense
    // group, which only counts); a slot's back entries are its listed bits'
    const uint32_t dm = DN ? (uint32_t)__builtin_amdgcn_readfirstlane((int)h.dmask) : 0u, sm = ~dm;  // (uniform)
    const uint32_t bit = own ? 1u << ln : 0u, lbit = bit & sm, below = (bit - 1u) & sm;
    const GroupHdr gh = p.gh;
    uint64_t *m = s.mem + gh.off;
    const bool single = cnt <= 64u;  // one staged chunk serves both passes

    uint32_t csl = 0, cw = 0, cbo = 0;
    auto stage = [&](uint32_t c0) {
        const uint32_t t = c0 + (uint32_t)ln;
        const bool in = t < cnt;
        csl = c0 == 0 ? p.f : (in ? slots[t] : 0u);
        cw = in ? W[csl] : 0u;  // 0 past the end: such a delete holds no group
        cbo = in ? B[csl] : 0u;
    };
    // back-index positions in this lane's group of staged deletes q0..q0+kDB-1
    auto positions = [&](uint32_t q0, uint32_t(&pos)[kDB]) {
#pragma unroll
        for (int q = 0; q < kDB; ++q) {
            const uint32_t w = rl32(cw, (int)q0 + q), bo = rl32(cbo, (int)q0 + q);
            pos[q] = (w & lbit) ? back[bo + __popc(w & below)] : 0u;
        }
    };

    // pass 1: count this group's deletes and mark their members
    uint32_t kj = 0;
    uint32_t keep[kDB];  // positions of a run of at most kDB deletes, reused by pass 2
    for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {
        stage(c0);
        const uint32_t nq = min(64u, cnt - c0);
        for (uint32_t q0 = 0; q0 < nq; q0 += kDB) {
            uint32_t pos[kDB];
            positions(q0, pos);
#pragma unroll
            for (int q = 0; q < kDB; ++q) {
                keep[q] = pos[q];
                const uint32_t w = rl32(cw, (int)q0 + q);
                kj += (w & bit) ? 1u : 0u;
                if (w & lbit) reinterpret_cast<uint32_t *>(m + pos[q])[0] = rl32(csl, (int)q0 + q) | kMark;
            }
        }
    }

    // compact the tail's survivors to [S2, S2 + nsurv), in position order
    // (a dense group has no list: nothing to compact, no holes)
    const uint32_t S = gh.size, S2 = S - kj, Slist = lbit ? S : S2;
    uint32_t nsurv = 0;
    for (uint32_t p0 = S2; p0 < Slist; p0 += kDB) {
        uint64_t e[kDB];
#pragma unroll
        for (int q = 0; q < kDB; ++q) e[q] = p0 + q < Slist ? m[p0 + q] : (uint64_t)kMark;
#pragma unroll
        for (int q = 0; q < kDB; ++q)
            if (!((uint32_t)e[q] & kMark)) m[S2 + nsurv++] = e[q];
    }

    // pass 2: fill holes; then the slots' neighbour counts and tombstones
    uint32_t hr = 0;
    uint64_t fsub = 0;
    bool bad = false;
    for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {
        if (!single) stage(c0);
        const uint32_t nq = min(64u, cnt - c0);
        for (uint32_t q0 = 0; q0 < nq; q0 += kDB) {
            uint32_t pos[kDB];
            if (cnt <= (uint32_t)kDB) {
#pragma unroll
                for (int q = 0; q < kDB; ++q) pos[q] = keep[q];
            } else {
                positions(q0, pos);
            }
            uint32_t rk[kDB];
            uint64_t mv[kDB];
#pragma unroll
            for (int q = 0; q < kDB; ++q) {
                const bool hole = (rl32(cw, (int)q0 + q) & lbit) && pos[q] < S2;
                rk[q] = hole ? hr : ~0u;
                hr += hole;
                bad |= hole && rk[q] >= nsurv;  // no survivor left: the store is inconsistent
                mv[q] = rk[q] < nsurv ? m[S2 + rk[q]] : 0ull;
            }
            uint32_t mb[kDB], mw[kDB];
#pragma unroll
            for (int q = 0; q < kDB; ++q) {
                const uint32_t ms = (uint32_t)mv[q] - h.sbase;
                mb[q] = rk[q] < nsurv ? B[ms] : 0u;
                mw[q] = rk[q] < nsurv ? W[ms] : 0u;
                if (rk[q] < nsurv) m[pos[q]] = mv[q];
            }
#pragma unroll
            for (int q = 0; q < kDB; ++q)
                if (rk[q] < nsurv) back[mb[q] + __popc(mw[q] & below)] = pos[q];
        }
        if (c0 + (uint32_t)ln < cnt) {
            W[csl] = 0u;
            if (s.fw) {
                uint32_t *FR = V + 3ull * h.slot_cap;
                fsub += FR[csl];
                FR[csl] = 0u;
            }
        }
    }
    if (bad) atomicOr(&s.ctl->error, (unsigned)ERR_OVERRUN);
    if (kj) s.grp[(size_t)u * kGroups + ln].size = S2;
    if (ln == 0) {
        s.hdr[u].nlive = h.nlive - cnt;
        s.hdr[u].dend = max(h.dend, slots[cnt - 1] + 1u);  // a run's slots ascend
    }
    const uint64_t F = s.fw ? s.fnode[u].F - dpp_sum64(fsub) : 0ull;
    alias_rebuild(s, u, own ? S2 : 0u, gh.off, s.nbr ? make_hmeta(h.hash_off, h.hash_cap) : 0ull, F, dm, ln);
}

// A run of at most K (<= kDB) deletes: the same result as delete_run in fewer
// dependent round trips.  Each group lane prefetches its list's last
// min(cnt, size) members together with the deletes' slot records, tells the
// deleted tail members from the survivors by comparing slots in registers,
// pairs the k-th hole with the k-th survivor there, and only gathers the
// moved members' back offsets afterwards.  Instantiated for K = 1, 2, 4, 8
// (the loops are unrolled over K; most runs of a batch delete one arc).
template <int K, bool DN>
__device__ void delete_small(const DevStore &s, uint32_t cnt, int32_t u, const RunPre &p, int ln)
{
    const NodeHdr h = p.h;
    const bool own = ln < kGroups;
    const GroupHdr gh = p.gh;
    const bool in = (uint32_t)ln < cnt;
    const uint32_t t_sl = p.f;  // 0 past the run
    uint32_t *V = s.slot + h.slot_off;
    uint32_t *W = V + h.slot_cap;
    const uint32_t *B = V + 2ull * h.slot_cap;
    uint32_t *back = s.back + h.back_off;
    // lbit: list work for this lane's group (0 for a dense group, which only
    // counts); a slot's back entries are its listed bits'
    const uint32_t dm = DN ? (uint32_t)__builtin_amdgcn_readfirstlane((int)h.dmask) : 0u, sm = ~dm;  // (uniform)
    const uint32_t bit = own ? 1u << ln : 0u, lbit = bit & sm, below = (bit - 1u) & sm;
    uint64_t *m = s.mem + gh.off;

    // the deletes' slot records; this group's last c members
    const uint32_t cw = in ? W[t_sl] : 0u, cbo = in ? B[t_sl] : 0u;
    const uint32_t S = gh.size, c = lbit ? min(cnt, S) : 0u;
    uint64_t e[K];
#pragma unroll
    for (int i = 0; i < K; ++i) e[i] = (uint32_t)i < c ? m[S - c + i] : 0ull;

    // member positions of this group's deletes
    uint32_t pos[K], wq[K], sq[K];
    uint32_t kj = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        wq[q] = rl32(cw, q);  // 0 for q >= cnt
        sq[q] = rl32(t_sl, q);
        pos[q] = (wq[q] & lbit) ? back[rl32(cbo, q) + __popc(wq[q] & below)] : 0u;
        kj += (wq[q] & bit) ? 1u : 0u;
    }
    // a listed group: kj <= c, the tail [S2, S) is prefetched
    const uint32_t S2 = S - kj;
    // slot records of the tail [S2, S), read together with the positions
    // above: the survivors among them are the members that move
    uint32_t eb[K], ew[K], es[K];  // es: the tail members' slots
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const bool cand = (uint32_t)i < c && (uint32_t)i + kj >= c;
        es[i] = (uint32_t)e[i] - h.sbase;
        eb[i] = cand ? B[es[i]] : 0u;
        ew[i] = cand ? W[es[i]] : 0u;
    }

    // survivors: tail members whose slot is not one of this group's deletes
    bool surv[K];
    uint32_t srank[K], ns = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        bool del = false;
#pragma unroll
        for (int q = 0; q < K; ++q) del |= (wq[q] & lbit) && es[i] == sq[q];
        surv[i] = (uint32_t)i < c && (uint32_t)i + kj >= c && !del;
        srank[i] = ns;
        ns += surv[i] ? 1u : 0u;
    }
    // the k-th hole (deletes in slot order, position below S2) takes the
    // k-th survivor
    bool hole[K], bad = false;
    uint64_t mv[K];
    uint32_t mb[K], mw[K];  // the moved members' back offsets and weights
    uint32_t hr = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const bool h0 = (wq[q] & lbit) && pos[q] < S2;
        uint64_t x = 0;
        uint32_t xb = 0, xw = 0;
        bool found = false;
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (surv[i] && srank[i] == hr) {
                x = e[i];
                xb = eb[i];
                xw = ew[i];
                found = true;
            }
        bad |= h0 && !found;  // no survivor left: the store is inconsistent
        hole[q] = h0 && found;
        mv[q] = x;
        mb[q] = xb;
        mw[q] = xw;
        hr += h0 ? 1u : 0u;
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
        if (hole[q]) m[pos[q]] = mv[q];
#pragma unroll
    for (int q = 0; q < K; ++q)
        if (hole[q]) back[mb[q] + __popc(mw[q] & below)] = pos[q];
    if (bad) atomicOr(&s.ctl->error, (unsigned)ERR_OVERRUN);
    if (kj) s.grp[(size_t)u * kGroups + ln].size = S2;
    uint64_t fsub = 0;
    if (ln == 0) {
        s.hdr[u].nlive = h.nlive - cnt;
        s.hdr[u].dend = max(h.dend, rl32(t_sl, (int)cnt - 1) + 1u);  // a run's slots ascend
    }
    if (in) {
        W[t_sl] = 0u;
        if (s.fw) {
            uint32_t *FR = V + 3ull * h.slot_cap;
            fsub = FR[t_sl];
            FR[t_sl] = 0u;
        }
    }
    const uint64_t F = s.fw ? s.fnode[u].F - dpp_sum64(fsub) : 0ull;
    alias_rebuild(s, u, own ? S2 : 0u, gh.off, s.nbr ? make_hmeta(h.hash_off, h.hash_cap) : 0ull, F, dm, ln);
}

// One bit group j of a long delete run, by one wave with lanes across the
// deletes: mark the group's deleted members and count them (k_j), compact the
// tail [S2, size) in place, then fill the holes (deletes in slot order with a
// member below S2) with the compacted survivors in order.  Groups of one node
// touch disjoint members and back entries and only read the slot arrays, so
// the 32 tasks of a run run concurrently; tombstones wait for the last one.
template <bool DN>
__device__ void delete_group(const DevStore &s, const uint32_t *__restrict__ slots, uint32_t cnt, int32_t u, int j,
                             int ln)
{
    const NodeHdr h = s.hdr[u];
    const uint32_t *W = s.slot + h.slot_off + h.slot_cap;
    const uint32_t *B = W + h.slot_cap;
    uint32_t *back = s.back + h.back_off;
    const GroupHdr gh = s.grp[(size_t)u * kGroups + j];
    const uint32_t dm = DN ? h.dmask : 0u;
    const uint32_t bit = 1u << j, below = (bit - 1u) & ~dm;
    uint64_t *m = s.mem + gh.off;
    const uint64_t lt = (1ull << ln) - 1ull;
    uint32_t kj = 0;
    if ((dm >> j) & 1u) {  // a dense group only counts
        for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {
            const uint32_t t = c0 + (uint32_t)ln;
            const uint32_t w = t < cnt ? W[slots[t]] : 0u;
            kj += (uint32_t)__popcll(__ballot((w & bit) != 0u));
        }
        if (kj && ln == 0) s.grp[(size_t)u * kGroups + j].size = gh.size - kj;
        return;
    }
    for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {  // mark
        const uint32_t t = c0 + (uint32_t)ln;
        const uint32_t sl = t < cnt ? slots[t] : 0u;
        const uint32_t w = t < cnt ? W[sl] : 0u, bo = t < cnt ? B[sl] : 0u;
        const bool has = (w & bit) != 0u;
        if (has) reinterpret_cast<uint32_t *>(m + back[bo + __popc(w & below)])[0] = sl | kMark;
        kj += (uint32_t)__popcll(__ballot(has));
    }
    __threadfence_block();  // marks by any lane before the tail is read
    const uint32_t S = gh.size, S2 = S - kj;
    uint32_t ns = 0;
    for (uint32_t p0 = S2; p0 < S; p0 += 64) {  // survivors of the tail, in order, to [S2, S2 + ns)
        const uint32_t p = p0 + (uint32_t)ln;
        const uint64_t e = p < S ? m[p] : (uint64_t)kMark;
        const bool surv = !((uint32_t)e & kMark);
        const uint64_t b = __ballot(surv);
        if (surv) m[S2 + ns + (uint32_t)__popcll(b & lt)] = e;  // never above p
        ns += (uint32_t)__popcll(b);
    }
    __threadfence_block();  // compacted survivors before they are read
    uint32_t hr = 0;
    bool bad = false;
    for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {  // holes take the survivors
        const uint32_t t = c0 + (uint32_t)ln;
        const uint32_t sl = t < cnt ? slots[t] : 0u;
        const uint32_t w = t < cnt ? W[sl] : 0u, bo = t < cnt ? B[sl] : 0u;
        const bool has = (w & bit) != 0u;
        const uint32_t pos = has ? back[bo + __popc(w & below)] : 0u;
        const bool hole = has && pos < S2;
        const uint64_t hb = __ballot(hole);
        const uint32_t rank = hr + (uint32_t)__popcll(hb & lt);
        const bool ok = hole && rank < ns;
        bad |= hole && !ok;
        const uint64_t mv = ok ? m[S2 + rank] : 0ull;
        if (ok) m[pos] = mv;
        const uint32_t ms = (uint32_t)mv - h.sbase;
        const uint32_t mb = ok ? B[ms] : 0u, mw = ok ? W[ms] : 0u;
        if (ok) back[mb + __popc(mw & below)] = pos;
        hr += (uint32_t)__popcll(hb);
    }
    if (__any(bad) && ln == 0) atomicOr(&s.ctl->error, (unsigned)ERR_OVERRUN);
    if (kj && ln == 0) s.grp[(size_t)u * kGroups + j].size = S2;
}

// After the last group task of a long run: tombstones and the node's alias
// table.
template <bool DN>
__device__ void delete_finish(const DevStore &s, const uint32_t *__restrict__ slots, uint32_t cnt, int32_t u, int ln)
{
    const NodeHdr h = s.hdr[u];
    uint32_t *V = s.slot + h.slot_off;
    uint32_t *W = V + h.slot_cap;
    uint32_t *FR = V + 3ull * h.slot_cap;  // float weights only
    uint64_t fsub = 0;
    for (uint32_t t = ln; t < cnt; t += 64) {
        const uint32_t sl = slots[t];
        W[sl] = 0u;
        if (s.fw) {
            fsub += FR[sl];
            FR[sl] = 0u;
        }
    }
    if (ln == 0) {
        s.hdr[u].nlive = h.nlive - cnt;
        s.hdr[u].dend = max(h.dend, slots[cnt - 1] + 1u);  // a run's slots ascend
    }
    GroupHdr gh = {0, 0, 0};
    if (ln < kGroups) gh = s.grp[(size_t)u * kGroups + ln];
    const uint64_t F = s.fw ? s.fnode[u].F - dpp_sum64(fsub) : 0ull;
    alias_rebuild(s, u, ln < kGroups ? gh.size : 0u, gh.off, s.nbr ? make_hmeta(h.hash_off, h.hash_cap) : 0ull, F,
                  DN ? h.dmask : 0u, ln);
}

// ------------------------------------------------------------------------
// neighbour multisets, kept apart from the run kernels: one thread per arc
// or delete, after the apply kernel has placed (and, when it grew, rehashed)
// every node's table.  A hub's thousands of arcs then probe in parallel
// instead of one 64-arc chunk at a time inside its run.
// ------------------------------------------------------------------------
__global__ void k_hash_insert(DevStore s, const DevEdge *__restrict__ E, int64_t k)
{
    if (s.ctl->ustatus) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    const DevEdge x = E[i];
    if (x.u < 1 || x.u > s.n) return;
    const NodeHdr h = s.hdr[x.u];
    if (h.hash_cap && hash_add(s.hash + h.hash_off, h.hash_cap, (uint32_t)x.v, 1u, x.w))
        atomicAdd(&s.hdr[x.u].hash_used, 1u);
}

__global__ void k_hash_delete(DevStore s, const uint32_t *__restrict__ us, const uint32_t *__restrict__ slots,
                              int64_t k)
{
    if (s.ctl->ustatus) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int32_t u = (int32_t)us[i];
    if (u < 1 || u > s.n) return;
    const NodeHdr h = s.hdr[u];
    // before the apply kernel tombstones the slot: its weight is still there
    if (h.hash_cap)
        hash_dec(s.hash + h.hash_off, h.hash_cap, s.slot[h.slot_off + slots[i]], s.slot[h.slot_off + h.slot_cap + slots[i]]);
}

// one wave per long delete run: tombstones, fractions and alias table
template <bool DN>
__global__ __launch_bounds__(256) void k_delete_finish(DevStore s, const uint32_t *__restrict__ slots, Runs r,
                                                       const uint32_t *__restrict__ big)
{
    if (s.ctl->ustatus) return;
    const int ln = lane_id();
    const uint32_t nbig = s.ctl->nbig;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t bi = wave; bi < nbig; bi += nwaves) {
        const uint32_t ri = big[bi];
        const int32_t u = (int32_t)r.node[ri];
        if (u >= 1 && u <= s.n) delete_finish<DN>(s, slots + r.start[ri], r.count[ri], u, ln);
    }
}

__global__ void k_delete_big_list(Runs r, uint32_t *__restrict__ big, Ctl *ctl)
{
    const uint32_t nruns = *r.nruns;
    const uint32_t ri = blockIdx.x * blockDim.x + threadIdx.x;
    if (ri < nruns && r.count[ri] > kBigDelete) big[atomicAdd(&ctl->nbig, 1u)] = ri;
}

// Tasks: first one per (long run, bit group), then every other run; a wave
// takes tasks by grid stride, so long runs' group tasks start first.  Long
// runs are finished by k_delete_finish after this kernel: the kernel boundary
// orders every group task's writes before the finish, where a device-scope
// fence inside the kernel would write back and invalidate an XCD's L2.
template <bool DN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_delete_apply(DevStore s, const uint32_t *__restrict__ slots, Runs r,
                                                      const uint32_t *__restrict__ big)
{
    if (s.ctl->ustatus) return;  // a captured batch that does not apply
    const int ln = lane_id();
    const uint32_t nruns = *r.nruns;
    const uint32_t nbig = s.ctl->nbig;
    const uint32_t ngt = nbig * kGroups;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    uint32_t task = wave;
    for (; task < ngt; task += nwaves) {
        {
            const uint32_t bi = task / kGroups, ri = big[bi];
            const int32_t u = (int32_t)r.node[ri];
            if (u < 1 || u > s.n) continue;
            delete_group<DN>(s, slots + r.start[ri], r.count[ri], u, (int)(task % kGroups), ln);
        }
    }
    // the other runs, continuing this wave's stride (long runs were done above)
    auto meta = [&](uint32_t t) {
        RunMeta m = run_meta(s, r, t - ngt, nruns);
        if (m.cnt > kBigDelete) m.u = 0;
        return m;
    };
    // one run ahead only: a second stage's registers would cost this
    // kernel a wave per SIMD of occupancy (measured slower)
    RunMeta m0 = meta(task), m1 = meta(task + nwaves);
    RunPre p0 = run_pre<true>(s, nullptr, slots, m0, ln);
    for (; task - ngt < nruns; task += nwaves) {
        const RunMeta m2 = meta(task + 2 * nwaves);
        const RunPre p1 = run_pre<true>(s, nullptr, slots, m1, ln);
        if (m0.u) {
            if (m0.cnt == 1)
                delete_small<1, DN>(s, m0.cnt, m0.u, p0, ln);
            else if (m0.cnt == 2)
                delete_small<2, DN>(s, m0.cnt, m0.u, p0, ln);
            else if (m0.cnt <= 4)
                delete_small<4, DN>(s, m0.cnt, m0.u, p0, ln);
            else if (m0.cnt <= (uint32_t)kDB)
                delete_small<kDB, DN>(s, m0.cnt, m0.u, p0, ln);
            else
                delete_run<DN>(s, slots + m0.start, m0.cnt, m0.u, p0, ln);
        }
        m0 = m1;
        m1 = m2;
        p0 = p1;
    }
}

// ------------------------------------------------------------------------
// compaction: tombstoned slots are reclaimed in place.  A listed node's live
// slots are renumbered by rank (the k-th live slot becomes slot k) and their
// records packed to the front of the node's slot arrays; its listed groups'
// member lists and its back entries are rebuilt from the slots in slot order
// (the k-th member of group j is the k-th live slot with bit j).  Masses and
// the alias table stay as they are.  The neighbour table is rebuilt from the
// live slots (its dead keys go).  The reference tombstones and never reuses a
// slot (include/operation.cuh:209-210, Bingo.cuh:199-218).
// ------------------------------------------------------------------------
__global__ void k_compact_reset(Ctl *ctl)
{
    if (threadIdx.x == 0) {
        ctl->ncomp[0] = ctl->ncomp[1] = ctl->ncomp[2] = ctl->ncomp[3] = 0;
        ctl->comp_freed = 0;
        ctl->comp_bigslots = ctl->comp_nseg = 0;
    }
}

// A team of NT threads working on one node: one wave (NT = 64) or one
// workgroup (NT = 1024, scan partials in LDS).
template <int NT>
struct Team {
    uint32_t *part;  // NT / 64 words of LDS (NT > 64)
    int t;
    __device__ __forceinline__ void sync() const
    {
        if (NT > 64) __syncthreads();
        else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    // exclusive team prefix sum of v (threads in t order) and its total
    __device__ __forceinline__ uint32_t excl(uint32_t v, uint32_t &total) const
    {
        const uint32_t inc = dpp_incl_add(v);
        if (NT == 64) {
            total = rl32(inc, 63);
            return inc - v;
        }
        const int w = t >> 6;
        if ((t & 63) == 63) part[w] = inc;
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            const uint32_t x = part[i];
            before += i < w ? x : 0u;
            tot += x;
        }
        __syncthreads();  // part is reused by the next scan
        total = tot;
        return before + inc - v;
    }
};

// Selection in tiles of 2048 nodes per workgroup (8 per thread, strided so
// header reads coalesce): list positions come from one atomic per list per
// tile, the totals from one atomic per workgroup (a single contended 64-bit
// atomic costs ~12 ns on this part; one per wave made this kernel 6.5 ms on
// 2^25 nodes).
constexpr uint32_t kSelPer = 8;
__global__ __launch_bounds__(256) void k_compact_select(DevStore s, uint32_t thr16, uint32_t *__restrict__ small_list,
                                                        uint32_t *__restrict__ big_list,
                                                        uint32_t *__restrict__ shift_list)
{
    __shared__ uint32_t part[4];
    __shared__ uint32_t base[4];
    const Team<256> tm{part, (int)threadIdx.x};
    const int64_t n1 = (int64_t)s.n + 1;
    const int64_t tile = 256 * kSelPer;
    uint64_t freed = 0, bslots = 0, bseg = 0;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < n1; t0 += (int64_t)gridDim.x * tile) {
        // this thread's picks (bit k: node t0 + k 256 + tid): rebuilt on the
        // wave path (ps) or in segments (pb); a dead prefix dropped by a wave
        // (pz) or a workgroup (py)
        uint32_t ps = 0, pb = 0, pz = 0, py = 0;
#pragma unroll
        for (uint32_t k = 0; k < kSelPer; ++k) {
            const int64_t u = t0 + k * 256 + threadIdx.x;
            if (u < 1 || u >= n1) continue;
            const NodeHdr h = s.hdr[u];
            const uint32_t dead = h.nslots - h.nlive;
            if (!(dead >= 1u && (uint64_t)dead * 65536u >= (uint64_t)thr16 * h.nslots)) continue;
            freed += dead;
            const bool big = h.nslots > kBigCompact;
            if (h.dend == dead) {
                (big ? py : pz) |= 1u << k;
            } else if (big) {
                pb |= 1u << k;
                bslots += h.nslots;
                bseg += (h.nslots + kCompactSeg - 1) / kCompactSeg;
            } else {
                ps |= 1u << k;
            }
        }
        // the wave-path lists in node order (one scan per k, both lists in
        // one word), so that a run of list entries is a run of neighbouring
        // headers and regions; the workgroup-path ones in thread order
        uint32_t es[kSelPer], tk[kSelPer], ts = 0, tz = 0, tbz;
#pragma unroll
        for (uint32_t k = 0; k < kSelPer; ++k) {
            es[k] = tm.excl(((ps >> k) & 1u) | ((pz >> k) & 1u) << 16, tk[k]);
            ts += tk[k] & 0xFFFFu;
            tz += tk[k] >> 16;
        }
        const uint32_t ebz = tm.excl((uint32_t)__popc(pb) | (uint32_t)__popc(py) << 16, tbz);
        if (threadIdx.x == 0) {
            base[0] = ts ? atomicAdd(&s.ctl->ncomp[0], ts) : 0u;
            base[1] = (tbz & 0xFFFFu) ? atomicAdd(&s.ctl->ncomp[1], tbz & 0xFFFFu) : 0u;
            base[2] = tz ? atomicAdd(&s.ctl->ncomp[2], tz) : 0u;
            base[3] = (tbz >> 16) ? atomicAdd(&s.ctl->ncomp[3], tbz >> 16) : 0u;
        }
        __syncthreads();
        uint32_t os = base[0], oz = base[2], ob = base[1] + (ebz & 0xFFFFu), oy = base[3] + (ebz >> 16);
#pragma unroll
        for (uint32_t k = 0; k < kSelPer; ++k) {
            const uint32_t u = (uint32_t)(t0 + k * 256 + threadIdx.x);
            if (ps & (1u << k)) small_list[os + (es[k] & 0xFFFFu)] = u;
            if (pz & (1u << k)) shift_list[oz + (es[k] >> 16)] = u;
            os += tk[k] & 0xFFFFu;
            oz += tk[k] >> 16;
            if (pb & (1u << k)) big_list[ob++] = u;
            if (py & (1u << k)) shift_list[n1 - 1 - oy++] = u;  // from the list's end
        }
        __syncthreads();  // base is the next tile's
    }
    // workgroup totals from 32-bit sums of 32-, 16- and 16-bit pieces (exact)
    const uint64_t vals[3] = {freed, bslots, bseg};
    unsigned long long *dst[3] = {&s.ctl->comp_freed, &s.ctl->comp_bigslots, &s.ctl->comp_nseg};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const uint64_t x = vals[q];
        uint32_t hi, lo16, hi16;
        tm.excl((uint32_t)(x >> 32), hi);
        tm.excl((uint32_t)(x & 0xFFFFu), lo16);
        tm.excl((uint32_t)((x >> 16) & 0xFFFFu), hi16);
        const uint64_t sum = ((uint64_t)hi << 32) + ((uint64_t)hi16 << 16) + lo16;
        if (threadIdx.x == 0 && sum) atomicAdd(dst[q], (unsigned long long)sum);
    }
}

// One chunk of a node's slots, [c0, min(c0 + 64, end)), by one wave: lane l
// holds slot c0 + l as read (ChunkIn; a run of chunks is read before any of
// them is rebuilt).
struct ChunkIn {
    uint32_t v, w, f;
    bool in;
};
// Every lane reads (a lane past end reads slot c0 again, c0 < end), so that
// a run's reads are a fixed number of instructions with no branches between
// them.
__device__ __forceinline__ ChunkIn chunk_load(const DevStore &s, const NodeHdr &h, uint32_t c0, uint32_t end, int ln)
{
    const uint32_t *V = s.slot + h.slot_off;
    const uint32_t i = c0 + (uint32_t)ln;
    ChunkIn x;
    x.in = i < end;
    const uint32_t k = x.in ? i : c0;
    x.v = V[k];
    x.w = V[h.slot_cap + k];
    x.f = s.fw ? V[3ull * h.slot_cap + k] : 0u;
    return x;
}

// A chunk's members of each listed group, in slot order from position gc_j;
// each slot's back entries (one per listed bit, in bit order) at bdst[b0..]
// (lw: the lane's listed bits, ent: its member word).
template <typename BT>
__device__ __forceinline__ void chunk_members(const DevStore &s, uint32_t lw, uint64_t ent, uint32_t b0, uint64_t go,
                                              uint32_t &gc, BT *bdst, int ln)
{
    const uint64_t lt = (1ull << ln) - 1ull;
    for (uint32_t any = dpp_or32(lw); any; any &= any - 1u) {
        const int j = __ffs(any) - 1;
        const bool has = (lw >> j) & 1u;
        const uint64_t m = __ballot(has);
        const uint32_t base = rl32(gc, j);
        const uint64_t gj = rl64(go, j);
        if (has) {
            const uint32_t pos = base + (uint32_t)__popcll(m & lt);
            s.mem[gj + pos] = ent;
            bdst[b0 + (uint32_t)__popc(lw & ((1u << j) - 1u))] = pos;
        }
        if (ln == j) gc += (uint32_t)__popcll(m);
    }
}

constexpr uint32_t kBackTab = 1024;
// The chunk rebuilt: live slots take the ids nid, nid + 1, ... and back
// offsets nb, ... in slot order; the members of each listed group are
// written in slot order from position gc_j of the group's list (lane j < 32
// holds gc_j and the list's offset go_j in the member pool), each with its
// back entry.  Neither the old member lists nor the old back entries are
// read: the lists follow from the slots.  STAGE (big nodes, whose segments
// other waves move concurrently): each live record goes to rec[id] as {v, w,
// fraction, back offset}; otherwise it moves down to slot id in place.  In
// place is safe because a record's id is at most its position: the chunk's
// writes wait for its reads (they store what was read; vector memory
// operations complete in issue order), and chunks read later lie above every
// position written so far.
// Back entries are gathered in LDS (bt: kBackTab words per wave) and
// written in order, four 64-byte lines per store, rather than scattered
// across the chunk's back range once per group; a chunk with more than
// kBackTab of them writes them directly.
template <bool STAGE>
__device__ __forceinline__ void rebuild_chunk(const DevStore &s, const NodeHdr &h, const ChunkIn &x, uint32_t c0,
                                              uint64_t go, uint32_t &gc, uint32_t &nid, uint32_t &nb, uint4 *rec,
                                              uint32_t *bt, int ln)
{
    uint32_t *V = s.slot + h.slot_off;
    uint32_t *back = s.back + h.back_off;
    const uint32_t i = c0 + (uint32_t)ln;
    const bool live = x.in && (x.w || x.f);
    const uint32_t lw = live ? x.w & ~h.dmask : 0u;  // listed bits: one member and one back entry each
    const uint64_t lt = (1ull << ln) - 1ull;
    const uint64_t lv = __ballot(live);
    const uint32_t id = nid + (uint32_t)__popcll(lv & lt);
    const uint32_t pc = (uint32_t)__popc(lw), pinc = dpp_incl_add(pc);
    const uint32_t nbk = nb + pinc - pc;
    if (STAGE) {
        if (live) rec[id] = make_uint4(x.v, x.w, x.f, nbk);
    } else if (live) {
        if (id != i) {
            V[id] = x.v;
            V[h.slot_cap + id] = x.w;
            if (s.fw) V[3ull * h.slot_cap + id] = x.f;
        }
        V[2ull * h.slot_cap + id] = nbk;
    }
    const uint64_t ent = ((uint64_t)x.v << 32) | id;
    const uint32_t P = rl32(pinc, 63);  // the chunk's back entries
    if (P <= kBackTab) {  // entry nb + q goes to bt[q]
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the previous chunk's table reads first
        chunk_members(s, lw, ent, nbk - nb, go, gc, bt, ln);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        for (uint32_t q = (uint32_t)ln; q < P; q += 64) back[nb + q] = bt[q];
    } else {
        chunk_members(s, lw, ent, nbk, go, gc, back, ln);
    }
    nid += (uint32_t)__popcll(lv);
    nb += P;
}

// a compacted node's header (its neighbour table, if any, is rebuilt by
// k_compact_hash)
__device__ __forceinline__ void compact_finish(const DevStore &s, int32_t u, const NodeHdr &h, uint32_t nid,
                                               uint32_t nb, int ln)
{
    if (ln == 0) {
        NodeHdr nh = h;
        nh.nslots = nid;
        nh.nlive = nid;
        nh.back_used = nb;
        nh.sbase = 0;  // member words were rewritten with the new slot numbers
        nh.dend = 0;
        s.hdr[u] = nh;
    }
}

// lane l's header as seen by every lane
__device__ __forceinline__ NodeHdr hdr_of_lane(const NodeHdr &x, int l)
{
    static_assert(sizeof(NodeHdr) == 64, "NodeHdr is 16 words");
    uint32_t a[16], b[16];
    __builtin_memcpy(a, &x, 64);
#pragma unroll
    for (int q = 0; q < 16; ++q) b[q] = rl32(a[q], l);
    NodeHdr r;
    __builtin_memcpy(&r, b, 64);
    return r;
}

// lane j (and j + 32) holds the member-pool offset of group j's list
__device__ __forceinline__ uint64_t group_offsets(const DevStore &s, int32_t u, int ln)
{
    return s.grp[(size_t)u * kGroups + (ln & (kGroups - 1))].off;
}

// the value of x once its load has completed (a use that the compiler must
// wait for here, so that later uses carry no pending load)
__device__ __forceinline__ void settle(uint32_t &x) { asm volatile("" : "+v"(x)); }

// listed nodes of at most kBigCompact slots, one wave each.  Batches of 64
// consecutive list entries go round robin over the waves, and a batch's 64
// headers are read at once, one per lane.  Its chunks of slots are then
// taken kChunkRun at a time: all their reads (and the group offsets of each
// node they start) are issued before any of their writes.  Vector memory
// operations complete in issue order and a wait for a read also waits for
// every write before it, so the wave waits once per run, not per chunk.
// Eight waves per SIMD: the grid (8 workgroups per CU) is resident at once.
constexpr int kChunkRun = 4;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_compact_small(
    DevStore s, const uint32_t *__restrict__ list)
{
    __shared__ uint32_t btab[4][kBackTab];
    const int ln = lane_id();
    uint32_t *bt = btab[threadIdx.x >> 6];
    const uint64_t cnt = s.ctl->ncomp[0];
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint64_t step = (uint64_t)((gridDim.x * blockDim.x) >> 6) * 64;
    for (uint64_t k0 = (uint64_t)wave * 64; k0 < cnt; k0 += step) {
        uint32_t uk = k0 + ln < cnt ? list[k0 + ln] : 0u;
        NodeHdr hk{};
        if (k0 + ln < cnt) hk = s.hdr[uk];
        {
            uint32_t a[16];
            __builtin_memcpy(a, &hk, 64);
            settle(uk);
#pragma unroll
            for (int q = 0; q < 16; ++q) settle(a[q]);
            __builtin_memcpy(&hk, a, 64);
        }
        const int nq = (int)min<uint64_t>(64, cnt - k0);
        int q = 0;       // the next chunk to read: node q of the batch, from slot c0
        uint32_t c0 = 0;
        uint64_t go = 0;  // the node being rebuilt: group offsets, members per group, ids and back entries so far
        uint32_t gc = 0, nid = 0, nb = 0;
        while (q < nq) {
            ChunkIn X[kChunkRun];
            uint64_t G[kChunkRun];
            int Q[kChunkRun];
            uint32_t C[kChunkRun];
#pragma unroll
            for (int r = 0; r < kChunkRun; ++r) {  // past the batch's end: node 0's first chunk, unused
                const bool ok = q < nq;
                const int qr = ok ? q : 0;
                Q[r] = q;
                C[r] = ok ? c0 : 0u;
                const NodeHdr h = hdr_of_lane(hk, qr);
                X[r] = chunk_load(s, h, C[r], h.nslots, ln);
                G[r] = group_offsets(s, (int32_t)rl32(uk, qr), ln);
                if (ok && c0 + 64 < h.nslots) {
                    c0 += 64;
                } else if (ok) {
                    ++q;
                    c0 = 0;
                }
            }
#pragma unroll
            for (int r = 0; r < kChunkRun; ++r) {
                if (Q[r] >= nq) break;
                const NodeHdr h = hdr_of_lane(hk, Q[r]);
                if (C[r] == 0) {
                    go = G[r];
                    gc = nid = nb = 0;
                }
                rebuild_chunk<false>(s, h, X[r], C[r], go, gc, nid, nb, nullptr, bt, ln);
                if (C[r] + 64 >= h.nslots) compact_finish(s, (int32_t)rl32(uk, Q[r]), h, nid, nb, ln);
            }
        }
    }
}

// neighbour tables of compacted nodes (stores with them), one wave per node,
// from the live slots: counts and weight sums are unchanged, keys whose
// count fell to zero go.  list[0..ncomp[c]), or with from_end the last
// ncomp[c] entries of a list of n1.
__global__ __launch_bounds__(256) void k_compact_hash(DevStore s, const uint32_t *__restrict__ list, int c,
                                                      int from_end, uint32_t n1)
{
    const int ln = lane_id();
    const Team<64> tm{nullptr, ln};
    const uint32_t cnt = s.ctl->ncomp[c];
    if (from_end) list += n1 - cnt;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = wave; k < cnt; k += nwaves) {
        const int32_t u = (int32_t)list[k];
        const NodeHdr h = s.hdr[u];
        if (!h.hash_cap) continue;
        uint64_t *tab = s.hash + h.hash_off;
        const uint32_t *V = s.slot + h.slot_off;
        for (uint32_t q = (uint32_t)ln; q < 2u * h.hash_cap; q += 64) tab[q] = 0ull;
        __threadfence_block();
        tm.sync();
        uint32_t fresh = 0, hused;
        for (uint32_t i = (uint32_t)ln; i < h.nslots; i += 64) fresh += hash_add(tab, h.hash_cap, V[i], 1u, V[h.slot_cap + i]);
        tm.excl(fresh, hused);
        if (ln == 0) s.hdr[u].hash_used = hused;
    }
}

// ---- nodes whose dead slots are their first d ones (a sliding window's
// expiries): the live records move down by d (back offsets lowered by D,
// the first live slot's) and the back entries down by D; member words keep
// their place and their slot fields, the node's sbase grows by d.  Moving
// down in place is safe as in rebuild_chunk: a run of chunks is read before
// any of it is written, and later runs read above every position written.
constexpr int kShiftRun = 4;

// A node's n live records moved from index + d to index (back offsets
// lowered by D) and its nb back entries from index + D to index, by NT
// threads (t: this thread), in runs: each run's records and back entries are
// read together before any of them is written.
template <int NT>
__device__ __forceinline__ void shift_node(const DevStore &s, const NodeHdr &h, uint32_t d, uint32_t D, uint32_t n,
                                           uint32_t nb, int t)
{
    uint32_t *V = s.slot + h.slot_off;
    uint32_t *back = s.back + h.back_off;
    constexpr uint32_t RS = NT * kShiftRun, RB = 2 * NT * kShiftRun;  // records, back entries per run
    for (uint32_t k = 0; k * RS < n || k * RB < nb; ++k) {
        const uint32_t c0 = k * RS, b0 = k * RB;
        uint32_t v[kShiftRun], w[kShiftRun], b[kShiftRun], f[kShiftRun], x[2 * kShiftRun];
        if (c0 < n) {
#pragma unroll
            for (int r = 0; r < kShiftRun; ++r) {
                const uint32_t i = c0 + (uint32_t)(r * NT + t), q = (i < n ? i : c0) + d;
                v[r] = V[q];
                w[r] = V[h.slot_cap + q];
                b[r] = V[2ull * h.slot_cap + q];
                f[r] = s.fw ? V[3ull * h.slot_cap + q] : 0u;
            }
        }
        if (b0 < nb) {
#pragma unroll
            for (int r = 0; r < 2 * kShiftRun; ++r) {
                const uint32_t i = b0 + (uint32_t)(r * NT + t);
                x[r] = back[(i < nb ? i : b0) + D];
            }
        }
        if (NT > 64) __syncthreads();  // every wave's reads before any wave's writes
        if (c0 < n) {
#pragma unroll
            for (int r = 0; r < kShiftRun; ++r) {
                const uint32_t i = c0 + (uint32_t)(r * NT + t);
                if (i < n) {
                    V[i] = v[r];
                    V[h.slot_cap + i] = w[r];
                    V[2ull * h.slot_cap + i] = b[r] - D;
                    if (s.fw) V[3ull * h.slot_cap + i] = f[r];
                }
            }
        }
        if (b0 < nb) {
#pragma unroll
            for (int r = 0; r < 2 * kShiftRun; ++r) {
                const uint32_t i = b0 + (uint32_t)(r * NT + t);
                if (i < nb) back[i] = x[r];
            }
        }
    }
}

// When the shift would take sbase past s.sbase_max, the node's member slot
// fields (slot + sbase, each at least sbase + d: its slot was live) are
// lowered by sbase + d instead, in place and in order, and sbase returns to 0
template <int NT>
__device__ __forceinline__ bool shift_renorm(const DevStore &s, int32_t u, const NodeHdr &h, uint32_t d, int t)
{
    if ((uint64_t)h.sbase + d <= s.sbase_max) return false;
    const uint64_t sub = (uint64_t)h.sbase + d;
    for (int j = 0; j < kGroups; ++j) {
        if ((h.dmask >> j) & 1u) continue;  // a dense group keeps no list
        const GroupHdr g = s.grp[(size_t)u * kGroups + j];
        for (uint32_t i = (uint32_t)t; i < g.size; i += NT) s.mem[g.off + i] -= sub;
    }
    return true;
}

// the node's header after the shift
__device__ __forceinline__ void shift_header(const DevStore &s, int32_t u, NodeHdr h, uint32_t d, uint32_t D,
                                             bool renorm)
{
    h.nslots -= d;
    h.back_used -= D;
    h.sbase = renorm ? 0u : h.sbase + d;
    h.dend = 0;
    s.hdr[u] = h;
}

// nodes of at most kBigCompact slots (shift_list[0..ncomp[2])), one wave
// each; batches of 64 list entries with their headers read at once
__global__ __launch_bounds__(256) void k_compact_shift(DevStore s, const uint32_t *__restrict__ list)
{
    const int ln = lane_id();
    const uint64_t cnt = s.ctl->ncomp[2];
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint64_t step = (uint64_t)((gridDim.x * blockDim.x) >> 6) * 64;
    for (uint64_t k0 = (uint64_t)wave * 64; k0 < cnt; k0 += step) {
        uint32_t uk = k0 + ln < cnt ? list[k0 + ln] : 0u;
        NodeHdr hk{};
        uint32_t Dk = 0;
        if (k0 + ln < cnt) {
            hk = s.hdr[uk];
            const uint32_t d = hk.nslots - hk.nlive;
            Dk = hk.nlive ? s.slot[hk.slot_off + 2ull * hk.slot_cap + d] : hk.back_used;
        }
        const int nq = (int)min<uint64_t>(64, cnt - k0);
        for (int q = 0; q < nq; ++q) {
            const NodeHdr h = hdr_of_lane(hk, q);
            const uint32_t d = h.nslots - h.nlive, D = rl32(Dk, q);
            const int32_t u = (int32_t)rl32(uk, q);
            shift_node<64>(s, h, d, D, h.nlive, h.back_used - D, ln);
            const bool renorm = shift_renorm<64>(s, u, h, d, ln);
            if (ln == 0) shift_header(s, u, h, d, D, renorm);
        }
    }
}

// larger nodes (the last ncomp[3] entries of shift_list), a workgroup of
// 1024 threads each: a hub's move is one workgroup's chain of round trips,
// so it takes as many loads in flight as one workgroup holds
__global__ __launch_bounds__(1024) void k_compact_shift_big(DevStore s, const uint32_t *__restrict__ list, uint32_t n1)
{
    const uint32_t cnt = s.ctl->ncomp[3];
    for (uint32_t k = blockIdx.x; k < cnt; k += gridDim.x) {
        const int32_t u = (int32_t)list[n1 - cnt + k];
        const NodeHdr h = s.hdr[u];
        const uint32_t d = h.nslots - h.nlive;
        const uint32_t D = h.nlive ? s.slot[h.slot_off + 2ull * h.slot_cap + d] : h.back_used;
        shift_node<1024>(s, h, d, D, h.nlive, h.back_used - D, (int)threadIdx.x);
        const bool renorm = shift_renorm<1024>(s, u, h, d, (int)threadIdx.x);
        if (threadIdx.x == 0) shift_header(s, u, h, d, D, renorm);
    }
}

// ---- big nodes: segments of kCompactSeg slots over the whole grid ----
// Renumbering needs each segment's live count, back entries and per-group
// member counts before it; the in-place moves of one node's records would
// collide across waves, so the live records are staged out and written back
// by a second kernel.  Members and back entries go straight to their final
// places, which no other segment writes.

// the big node a segment belongs to (largest i with segbase[i] <= seg)
__device__ __forceinline__ uint32_t seg_node(const uint32_t *segbase, uint32_t nbig, uint32_t seg)
{
    uint32_t lo = 0, hi = nbig - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (segbase[mid] <= seg) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// one workgroup: segbase = exclusive prefix of each big node's segment count
__global__ __launch_bounds__(1024) void k_compact_segbase(DevStore s, const uint32_t *__restrict__ big, uint32_t nbig,
                                                          CompactBig cb)
{
    __shared__ uint32_t part[16];
    const Team<1024> tm{part, (int)threadIdx.x};
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < nbig; c0 += 1024) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t ns = i < nbig ? s.hdr[big[i]].nslots : 0u;
        uint32_t tot;
        const uint32_t ex = tm.excl((ns + kCompactSeg - 1) / kCompactSeg, tot);
        if (i < nbig) cb.segbase[i] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) cb.segbase[nbig] = run;
}

// per segment (one wave each): live slots, back entries and members of each
// listed group
__global__ __launch_bounds__(256) void k_compact_segcount(DevStore s, const uint32_t *__restrict__ big, uint32_t nbig,
                                                          uint64_t nseg, CompactBig cb)
{
    const int ln = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint64_t sg = wave; sg < nseg; sg += nwaves) {
        const uint32_t i = seg_node(cb.segbase, nbig, (uint32_t)sg);
        const NodeHdr h = s.hdr[big[i]];
        const uint32_t *V = s.slot + h.slot_off;
        const uint32_t *W = V + h.slot_cap, *F = V + 3ull * h.slot_cap;
        const uint32_t a = ((uint32_t)sg - cb.segbase[i]) * kCompactSeg, b = min(a + kCompactSeg, h.nslots);
        uint32_t live = 0, pc = 0, gc = 0;
        for (uint32_t c0 = a; c0 < b; c0 += 64) {
            const uint32_t j = c0 + (uint32_t)ln;
            const bool in = j < b;
            const uint32_t w = in ? W[j] : 0u, f = (s.fw && in) ? F[j] : 0u;
            const bool lv = in && (w || f);
            const uint32_t lw = lv ? w & ~h.dmask : 0u;
            live += lv ? 1u : 0u;
            pc += (uint32_t)__popc(lw);
            for (uint32_t any = dpp_or32(lw); any; any &= any - 1u) {
                const int g = __ffs(any) - 1;
                const uint64_t m = __ballot((lw >> g) & 1u);
                if (ln == g) gc += (uint32_t)__popcll(m);
            }
        }
        live = dpp_sum32(live);
        pc = dpp_sum32(pc);
        if (ln < kGroups) cb.seg_gc[sg * kGroups + ln] = gc;
        if (ln == 0) {
            cb.seg_live[sg] = live;
            cb.seg_pc[sg] = pc;
        }
    }
}

// per big node (one wave each): in place of its segments' counts, the first
// new id, back offset and per-group member position of each; and its totals
__global__ __launch_bounds__(256) void k_compact_segscan(uint32_t nbig, CompactBig cb)
{
    const int ln = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wave; i < nbig; i += nwaves) {
        const uint32_t s0 = cb.segbase[i], s1 = cb.segbase[i + 1];
        uint32_t rl = 0, rp = 0, rg = 0;  // rg: lane j's running member count of group j
        for (uint32_t c0 = s0; c0 < s1; c0 += 64) {
            const uint32_t sg = c0 + (uint32_t)ln;
            const bool in = sg < s1;
            const uint32_t l = in ? cb.seg_live[sg] : 0u, p = in ? cb.seg_pc[sg] : 0u;
            const uint32_t il = dpp_incl_add(l), ip = dpp_incl_add(p);
            if (in) {
                cb.seg_live[sg] = rl + il - l;
                cb.seg_pc[sg] = rp + ip - p;
            }
            rl += rl32(il, 63);
            rp += rl32(ip, 63);
            uint32_t x[kGroups];
#pragma unroll
            for (int g = 0; g < kGroups; ++g) x[g] = in ? cb.seg_gc[(uint64_t)sg * kGroups + g] : 0u;
#pragma unroll
            for (int g = 0; g < kGroups; ++g) {
                const uint32_t ig = dpp_incl_add(x[g]);
                if (in) cb.seg_gc[(uint64_t)sg * kGroups + g] = rl32(rg, g) + ig - x[g];
                const uint32_t tg = rl32(ig, 63);
                if (ln == g) rg += tg;
            }
        }
        if (ln == 0) {
            cb.nlive[i] = rl;
            cb.nback[i] = rp;
        }
    }
}

// one workgroup: the big nodes' record staging offsets (prefix of their live counts)
__global__ __launch_bounds__(1024) void k_compact_nodescan(uint32_t nbig, CompactBig cb)
{
    __shared__ uint32_t part[16];
    const Team<1024> tm{part, (int)threadIdx.x};
    uint64_t trec = 0;
    for (uint32_t c0 = 0; c0 < nbig; c0 += 1024) {
        const uint32_t i = c0 + threadIdx.x;
        const bool in = i < nbig;
        // 32-bit scans of the high and low 16 bits (1024 values each)
        const uint32_t nl = in ? cb.nlive[i] : 0u;
        uint32_t a, b;
        const uint32_t lh = tm.excl(nl >> 16, a), ll = tm.excl(nl & 0xFFFFu, b);
        if (in) cb.trec[i] = trec + ((uint64_t)lh << 16) + ll;
        trec += ((uint64_t)a << 16) + b;
    }
}

// per segment (one wave each): stage the live records, write the members
// and back entries in their final places
__global__ __launch_bounds__(256) void k_compact_stage(DevStore s, const uint32_t *__restrict__ big, uint32_t nbig,
                                                       uint64_t nseg, CompactBig cb)
{
    __shared__ uint32_t btab[4][kBackTab];
    const int ln = lane_id();
    uint32_t *bt = btab[threadIdx.x >> 6];
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint64_t sg = wave; sg < nseg; sg += nwaves) {
        const uint32_t i = seg_node(cb.segbase, nbig, (uint32_t)sg);
        const int32_t u = (int32_t)big[i];
        const NodeHdr h = s.hdr[u];
        const uint64_t go = group_offsets(s, u, ln);
        uint32_t gc = ln < kGroups ? cb.seg_gc[sg * kGroups + ln] : 0u;
        uint32_t nid = cb.seg_live[sg], nb = cb.seg_pc[sg];
        uint4 *rec = cb.rec + cb.trec[i];
        const uint32_t a = ((uint32_t)sg - cb.segbase[i]) * kCompactSeg, b = min(a + kCompactSeg, h.nslots);
        for (uint32_t c0 = a; c0 < b; c0 += kChunkRun * 64) {  // runs of chunks, as in k_compact_small
            ChunkIn X[kChunkRun];
#pragma unroll
            for (int r = 0; r < kChunkRun; ++r) X[r] = chunk_load(s, h, c0 + r * 64 < b ? c0 + r * 64 : c0, b, ln);
#pragma unroll
            for (int r = 0; r < kChunkRun; ++r)
                if (c0 + r * 64 < b) rebuild_chunk<true>(s, h, X[r], c0 + r * 64, go, gc, nid, nb, rec, bt, ln);
        }
    }
}

// write the staged records to the front of the node's slot arrays
__global__ __launch_bounds__(256) void k_compact_unstage(DevStore s, const uint32_t *__restrict__ big, uint32_t nbig,
                                                         uint64_t nseg, CompactBig cb)
{
    for (uint64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
        const uint32_t i = seg_node(cb.segbase, nbig, (uint32_t)sg);
        const NodeHdr h = s.hdr[big[i]];
        uint32_t *V = s.slot + h.slot_off;
        uint32_t *W = V + h.slot_cap, *B = V + 2ull * h.slot_cap, *F = V + 3ull * h.slot_cap;
        const bool last = (uint32_t)sg + 1 == cb.segbase[i + 1];
        const uint32_t id0 = cb.seg_live[sg], id1 = last ? cb.nlive[i] : cb.seg_live[sg + 1];
        const uint4 *rec = cb.rec + cb.trec[i];
        for (uint32_t id = id0 + threadIdx.x; id < id1; id += 256) {
            const uint4 r = rec[id];
            V[id] = r.x;
            W[id] = r.y;
            if (s.fw) F[id] = r.z;
            B[id] = r.w;
        }
    }
}

// headers of the big nodes, and their neighbour tables from the live slots
__global__ __launch_bounds__(256) void k_compact_bigfinish(DevStore s, const uint32_t *__restrict__ big, uint32_t nbig,
                                                           CompactBig cb)
{
    __shared__ uint32_t part[4];
    const Team<256> tm{part, (int)threadIdx.x};
    for (uint32_t i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int32_t u = (int32_t)big[i];
        const NodeHdr h = s.hdr[u];
        const uint32_t nid = cb.nlive[i];
        uint32_t hused = h.hash_used;
        if (s.nbr && h.hash_cap) {
            uint64_t *tab = s.hash + h.hash_off;
            const uint32_t *V = s.slot + h.slot_off, *W = V + h.slot_cap;
            for (uint32_t q = threadIdx.x; q < 2u * h.hash_cap; q += 256) tab[q] = 0ull;
            __syncthreads();
            uint32_t fresh = 0;
            for (uint32_t j = threadIdx.x; j < nid; j += 256) fresh += hash_add(tab, h.hash_cap, V[j], 1u, W[j]);
            tm.excl(fresh, hused);
        }
        if (threadIdx.x == 0) {
            NodeHdr nh = h;
            nh.nslots = nid;
            nh.nlive = nid;
            nh.back_used = cb.nback[i];
            nh.hash_used = hused;
            nh.sbase = 0;
            nh.dend = 0;
            s.hdr[u] = nh;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// walks: one lane per walker, state in registers
// ------------------------------------------------------------------------
// One draw on node `cur` (NodeData::sample, include/Bingo.cuh:220-267):
// slot i = top 5 bits of r.x; the 64-bit coin (r.y:r.z) against T_i picks
// group i or its partner; member floor(r.w * size / 2^32) of that group.
// A dense group j (no member list) returns -2 - j, for dense_draw; -1 is a
// node without live arcs.
// `hm` receives the slot's meta word without the partner index: cur's
// neighbour-table locator (kHmetaMask bits) and the 17-bit field above it
// (mass code or float threshold ceiling).  Split in
// two so that callers can issue the slot read together with other reads.
// The slot is read as two 16-byte loads.  The empty asm takes both as
// operands, so the compiler can neither narrow them to the fields a caller
// happens to use nor turn them into a load of T followed by a dependent load
// of the chosen locator: random gathers here are bound by the number of
// requests, and two independent requests per slot is the fewest.
// Split in two so that a caller can do other work between issuing the loads
// and pinning them (the pin waits for them).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void slot_load(const AliasEnt *__restrict__ alias, int32_t cur, uint4 r, u32x4 &a, u32x4 &b)
{
    const u32x4 *p = reinterpret_cast<const u32x4 *>(alias + (size_t)cur * kGroups + (r.x >> 27));
    a = p[0];
    b = p[1];
}
__device__ __forceinline__ void slot_pin(u32x4 a, u32x4 b, uint4 &q0, uint4 &q1)
{
    asm volatile("" : "+v"(a), "+v"(b));
    q0 = make_uint4(a.x, a.y, a.z, a.w);
    q1 = make_uint4(b.x, b.y, b.z, b.w);
}
__device__ __forceinline__ void alias_slot(const AliasEnt *__restrict__ alias, int32_t cur, uint4 r, uint4 &q0,
                                           uint4 &q1)
{
    u32x4 a, b;
    slot_load(alias, cur, r, a, b);
    slot_pin(a, b, q0, q1);
}
__device__ __forceinline__ int32_t slot_member(const uint64_t *__restrict__ mem, uint4 r, uint4 q0, uint4 q1,
                                               uint64_t &hm)
{
    const uint64_t T = ((uint64_t)q0.y << 32) | q0.x;
    const uint64_t coin = ((uint64_t)r.y << 32) | r.z;
    const bool first = ((T >> 63) != 0) == (coin < T);  // the draw's outcome is the slot's first locator
    const uint64_t loc = first ? (((uint64_t)q0.w << 32) | q0.z) : (((uint64_t)q1.y << 32) | q1.x);
    hm = (((uint64_t)q1.w << 32) | q1.z) >> 5;  // locator | the 17-bit field << 42
    const uint32_t sz = (uint32_t)(loc >> kOffBits);
    if (sz == 0u) return (loc & kDenseLoc) ? -2 - (int32_t)(loc & 31u) : -1;
    const uint32_t pos = __umulhi(r.w, sz);
    return (int32_t)reinterpret_cast<const uint32_t *>(mem + (loc & kOffMask36) + pos)[1];
}
__device__ __forceinline__ int32_t sample_step(const AliasEnt *__restrict__ alias, const uint64_t *__restrict__ mem,
                                               int32_t cur, uint4 r, uint64_t &hm)
{
    uint4 q0, q1;
    alias_slot(alias, cur, r, q0, q1);
    return slot_member(mem, r, q0, q1, hm);
}

// One draw on `cur` that needs no neighbour-table locator (integer weights,
// apps other than node2vec): the slot's first 16 bytes (T and the likelier
// outcome's locator) in one load, the second locator only when the less
// likely outcome comes up.  Random gathers here are bound by the lane
// addresses the memory pipeline handles, so one 16-byte load instead of two
// per draw is the saving.
__device__ __forceinline__ int32_t sample_lite(const AliasEnt *__restrict__ alias, const uint64_t *__restrict__ mem,
                                               int32_t cur, uint4 r)
{
    const u32x4 *p = reinterpret_cast<const u32x4 *>(alias + (size_t)cur * kGroups + (r.x >> 27));
    u32x4 a = p[0];
    asm volatile("" : "+v"(a));
    const uint64_t T = ((uint64_t)a.y << 32) | a.x;
    const uint64_t coin = ((uint64_t)r.y << 32) | r.z;
    uint64_t loc = ((uint64_t)a.w << 32) | a.z;
    if (((T >> 63) != 0) != (coin < T)) {
        const uint2 b = *reinterpret_cast<const uint2 *>(p + 1);
        loc = ((uint64_t)b.y << 32) | b.x;
    }
    const uint32_t sz = (uint32_t)(loc >> kOffBits);
    if (sz == 0u) return (loc & kDenseLoc) ? -2 - (int32_t)(loc & 31u) : -1;
    const uint32_t pos = __umulhi(r.w, sz);
    return (int32_t)reinterpret_cast<const uint32_t *>(mem + (loc & kOffMask36) + pos)[1];
}

// One draw on `cur` with float weights (NodeData::sample's wtype::Float
// branch, include/Bingo.cuh:254-264).  With probability F / (S + F) -- the
// 64-bit coin of the fractional stream (counter word 3 = kFracTag | attempt)
// below the node's threshold, or no integer mass -- the fractional part: a
// uniform slot accepted when a 32-bit uniform is below its fraction, round
// `it` drawing from counter word 3 = kFracTag | it << 16 | attempt, up to
// kFracIters rounds.  Otherwise, or when every round rejects, the integer
// draw from `r`.  The alias slot and the threshold are read together.  FW
// false compiles to sample_step alone.
template <bool FW>
__device__ __forceinline__ int32_t draw(const DevStore &s, int32_t cur, uint4 r, uint64_t &hm, uint32_t w0, uint32_t w1,
                                        uint32_t step, uint32_t att, uint32_t k0, uint32_t k1)
{
    if (!FW) return sample_step(s.alias, s.mem, cur, r, hm);
    // the fractional stream's first block is computed while the slot loads
    // (pinned with the slot, so it is not sunk past the wait)
    u32x4 sa, sb;
    slot_load(s.alias, cur, r, sa, sb);
    const uint4 b0 = philox(w0, w1, step, kFracTag | att, k0, k1);
    u32x4 pb = {b0.x, b0.y, b0.z, b0.w};
    asm volatile("" : "+v"(sa), "+v"(sb), "+v"(pb));
    const uint4 q0 = make_uint4(sa.x, sa.y, sa.z, sa.w), q1 = make_uint4(sb.x, sb.y, sb.z, sb.w);
    uint4 b = make_uint4(pb.x, pb.y, pb.z, pb.w);
    const uint32_t th = q1.w >> 15;  // ceil(thr / 2^47), saturated; 0 iff thr = 0
    if (th) {
        const uint64_t coin = ((uint64_t)b.x << 32) | b.y;
        const uint32_t c = (uint32_t)(coin >> 47);
        // (th - 1) 2^47 < thr <= th 2^47 (no upper bound when saturated):
        // the top 17 bits decide unless they fall on the boundary
        bool frac;
        if (c + 1u < th)
            frac = true;
        else if (c >= th && th < kThrHiMax)
            frac = false;
        else {
            const uint64_t thr = s.fnode[cur].thr;
            frac = thr == ~0ull || coin < thr;
        }
        if (frac) {
            const NodeHdr h = s.hdr[cur];
            const uint32_t *V = s.slot + h.slot_off, *FR = V + 3ull * h.slot_cap;
            for (uint32_t it = 0; it < kFracIters; ++it) {
                if (it) b = philox(w0, w1, step, kFracTag | (it << 16) | att, k0, k1);
                const uint32_t sl = __umulhi(b.z, h.nslots);
                if (b.w < FR[sl]) {
                    hm = make_hmeta(h.hash_off, h.hash_cap);
                    return (int32_t)V[sl];
                }
            }
        }
    }
    return slot_member(s.mem, r, q0, q1, hm);
}

// A draw from dense group j of node `cur`, which keeps no member list: the
// reference's rejection for groups classed dense (include/Bingo.cuh:244-252),
// a uniform slot until its weight has bit j (tombstones have no bits).
// Trials come four per Philox block (counter word 3 = kDenseTag | it << 16 |
// attempt), words x, y, z, w in order, slot floor(word nslots / 2^32); after
// kDenseBlocks blocks without a hit (probability below (1 - count /
// nslots)^4096), the first slot with the bit after the last trial's slot,
// cyclically.
// `trials` counts the slot weights read.
__device__ __forceinline__ int32_t dense_draw(const DevStore &s, int32_t cur, int j, uint32_t w0, uint32_t w1,
                                           uint32_t step, uint32_t att, uint32_t k0, uint32_t k1, uint32_t &trials)
{
    const NodeHdr h = s.hdr[cur];
    const uint32_t n = h.nslots, bit = 1u << j;
    if (!n) return -1;
    const uint32_t *V = s.slot + h.slot_off, *W = V + h.slot_cap;
    uint32_t sl = 0;
    for (uint32_t it = 0; it < kDenseBlocks; ++it) {
        const uint4 b = philox(w0, w1, step, kDenseTag | (it << 16) | att, k0, k1);
        const uint32_t x[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sl = __umulhi(x[q], n);
            ++trials;
            if (W[sl] & bit) return (int32_t)V[sl];
        }
    }
    for (uint32_t t = 1; t <= n; ++t) {
        ++trials;
        const uint32_t s2 = sl + t >= n ? sl + t - n : sl + t;
        if (W[s2] & bit) return (int32_t)V[s2];
    }
    return -1;
}

// DN: the store has dense groups (DevStore::rate16); the other instances
// carry no rejection code (its registers would cost them occupancy).
template <int APP, bool FW, bool DN>
__global__ __launch_bounds__(256) void k_walk(DevStore s, WalkArgs a)
{
    const int64_t wid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    uint32_t steps = 0, draws = 0, probes = 0, folds = 0, dtr = 0;
    if (wid < a.nwalkers) {
        const int32_t start = a.starts ? a.starts[wid] : (int32_t)(wid % s.n) + 1;
        int32_t L = a.lens ? a.lens[wid] : a.max_len;
        if (APP == 3 && L > 2) L = 2;
        int32_t *row = a.paths + wid * (int64_t)a.max_len;
        int32_t j = 1;
        if (start < 1 || start > s.n || L < 1 || L > a.max_len) {
            j = 0;
            atomicOr(&s.ctl->werr, (unsigned)ERR_INVAL);
        } else {
            const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
            const uint32_t w0 = (uint32_t)wid, w1 = (uint32_t)((uint64_t)wid >> 32);
            row[0] = start;
            int32_t cur = start, prev = -1;
            uint64_t hcur = 0ull, hprev = 0ull;  // neighbour-table locators of cur and prev
            const uint32_t tlo = min(a.thr_neighbor, a.thr_other), thi = max(a.thr_neighbor, a.thr_other);
            for (; j < L; ++j) {
                uint4 r;
                int32_t nx;
                if (APP == 1 && j >= 2) {
                    // folded return region (integer weights only): see fold_step
                    const bool fold = !FW && a.fold_r;
                    FoldState fs = {0, 0, 0, false};
                    for (uint32_t at = 0;; ++at) {
                        r = philox(w0, w1, (uint32_t)j, at, k0, k1);
                        nx = draw<FW>(s, cur, r, hcur, w0, w1, (uint32_t)j, at, k0, k1);
                        ++draws;
                        if (fold && fold_step(s, cur, prev, hcur, a.fold_r, a.fold_cap, fs, folds, w0, w1,
                                              (uint32_t)j, at, k0, k1)) {
                            nx = prev;
                            break;
                        }
                        if (DN && nx < -1) nx = dense_draw(s, cur, -2 - nx, w0, w1, (uint32_t)j, at, k0, k1, dtr);
                        if (nx < 0 || (int64_t)at + 1 >= a.max_attempts) break;
                        const uint32_t aux = r.x & kAuxMask;
                        bool acc;
                        if (nx == prev) acc = aux < a.thr_return;
                        else if (aux < tlo) acc = true;
                        else if (aux >= thi) acc = false;
                        else {
                            ++probes;
                            acc = aux < (is_neighbor(s.hash, hprev, (uint32_t)nx) ? a.thr_neighbor : a.thr_other);
                        }
                        if (acc) break;
                    }
                } else {
                    r = philox(w0, w1, (uint32_t)j, 0u, k0, k1);
                    if (!FW && APP != 1)
                        nx = sample_lite(s.alias, s.mem, cur, r);
                    else
                        nx = draw<FW>(s, cur, r, hcur, w0, w1, (uint32_t)j, 0u, k0, k1);
                    if (DN && nx < -1) nx = dense_draw(s, cur, -2 - nx, w0, w1, (uint32_t)j, 0u, k0, k1, dtr);
                    ++draws;
                }
                if (nx < 0) break;
                row[j] = nx;
                ++steps;
                prev = cur;
                cur = nx;
                hprev = hcur;
                if (APP == 2 && (r.x & kAuxMask) < a.thr_stop) {
                    ++j;
                    break;
                }
            }
        }
        for (; j < a.max_len; ++j) row[j] = -1;
    }
    const uint32_t tot = wave_sum32(steps), tdraw = wave_sum32(draws), tprobe = wave_sum32(probes),
                   tfold = wave_sum32(folds), tdtr = DN ? wave_sum32(dtr) : 0u;
    if (lane_id() == 0) {
        if (tot) {
            atomicAdd(a.steps, (unsigned long long)tot);
            atomicAdd(&s.ctl->steps, (unsigned long long)tot);
        }
        if (tdraw) atomicAdd(&s.ctl->attempts, (unsigned long long)tdraw);
        if (tprobe) atomicAdd(&s.ctl->probes, (unsigned long long)tprobe);
        if (tfold) atomicAdd(&s.ctl->folds, (unsigned long long)tfold);
        if (tdtr) atomicAdd(&s.ctl->dtrials, (unsigned long long)tdtr);
    }
}

// ------------------------------------------------------------------------
// persistent walks (batched form): each lane carries WPL walker slots; a
// slot whose walker finished is refilled from a global cursor with one
// atomic per wave (ballot of empty slots, popcount ranks), so walkers of
// uneven length (PPR, per-walker caps, dead ends, node2vec rejections) never
// leave lanes idle.  Each loop iteration makes one draw per active slot; the
// WPL slots of a lane are independent chains, so their gathers overlap.
// With D > 0 the path is staged D vertices at a time in LDS (odd row pitch,
// no bank conflicts on the per-step writes) and written out in 16-byte
// vectors.  Walker i's path depends only on (seed, i, start, len), so any
// schedule gives the same paths as k_walk and the oracle.
// ------------------------------------------------------------------------
// Path rows are written once and not read back by the walk.  Whole lines
// written by a wave together go out as non-temporal stores (kept from
// displacing the store's lines in L2); a lane's own 16-byte pieces of its
// row stay ordinary stores, so that L2 merges them into lines.
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_row4(int32_t *p, int32_t a, int32_t b, int32_t c, int32_t d)
{
    const i32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<i32x4 *>(p));
}
__device__ __forceinline__ void store4(int32_t *p, int32_t a, int32_t b, int32_t c, int32_t d)
{
    *reinterpret_cast<int4 *>(p) = make_int4(a, b, c, d);
}

template <int WPL, int D>
struct PathOut {
    int32_t *stage;  // this lane's WPL rows of D + 1 ints (LDS)
    __device__ __forceinline__ void put(const WalkArgs &a, int k, int64_t wid, int32_t j, int32_t v) const
    {
        if (D == 0) {
            a.paths[wid * (int64_t)a.max_len + j] = v;
        } else {
            int32_t *st = stage + k * (D + 1);
            st[j % D] = v;
            if (j % D == D - 1) flush(a, k, wid, j - (D - 1), D);
        }
    }
    __device__ __forceinline__ void flush(const WalkArgs &a, int k, int64_t wid, int32_t j0, int cnt) const
    {
        const int32_t *st = stage + k * (D + 1);
        int32_t *dst = a.paths + wid * (int64_t)a.max_len + j0;
#pragma unroll
        for (int c = 0; c < D / 4; ++c)
            if (4 * c < cnt) store4(dst + 4 * c, st[4 * c], st[4 * c + 1], st[4 * c + 2], st[4 * c + 3]);
    }
    // write out the partial stage (its tail -1) and return the first position
    // not yet written (a multiple of D when D > 0)
    __device__ __forceinline__ int32_t close(const WalkArgs &a, int k, int64_t wid, int32_t j) const
    {
        if (D == 0) return j;
        const int32_t r = j % D;
        if (r) {
            int32_t *st = stage + k * (D + 1);
            for (int t = r; t < D; ++t) st[t] = -1;
            flush(a, k, wid, j - r, D);
            j += D - r;
        }
        return j;
    }
    // pad positions j..max_len-1 with -1 (this lane alone)
    __device__ __forceinline__ void finish(const WalkArgs &a, int k, int64_t wid, int32_t j) const
    {
        int32_t *row = a.paths + wid * (int64_t)a.max_len;
        if (D == 0) {
            for (; j < a.max_len; ++j) row[j] = -1;
            return;
        }
        const int32_t r = j % D;
        if (r) {
            int32_t *st = stage + k * (D + 1);
            for (int t = r; t < D; ++t) st[t] = -1;
            flush(a, k, wid, j - r, D);
            j += D - r;
        }
        int4 *dst = reinterpret_cast<int4 *>(row + j);
        for (int32_t c = 0; c < (a.max_len - j) / 4; ++c) dst[c] = make_int4(-1, -1, -1, -1);
    }
};

template <int APP, int WPL, int D, bool FW, bool DN>
__global__ __launch_bounds__(512) void k_walk_persist(DevStore s, WalkArgs a)
{
    extern __shared__ int32_t lds_stage[];
    const int ln = lane_id();
    const uint64_t lt = (1ull << ln) - 1ull;
    PathOut<WPL, D> out{lds_stage + (size_t)threadIdx.x * WPL * (D + 1)};
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint32_t tlo = min(a.thr_neighbor, a.thr_other), thi = max(a.thr_neighbor, a.thr_other);
    int64_t wid[WPL];
    int32_t cur[WPL], prev[WPL], jj[WPL], len[WPL];
    uint32_t att[WPL];
    uint64_t hp[WPL];    // neighbour-table locator of each slot's previous node
    FoldState fst[WPL];  // node2vec: each slot's fold state for its current step
    int64_t padw[WPL];   // walker whose row tail the wave pads at the end of this iteration (-1: none)
    int32_t padj[WPL];   // ... from this position on
    const bool fold = !FW && APP == 1 && a.fold_r;
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
        wid[k] = -1;
        cur[k] = prev[k] = jj[k] = len[k] = 0;
        att[k] = 0;
        hp[k] = 0;
        fst[k] = FoldState{0, 0, 0, false};
        padw[k] = -1;
        padj[k] = 0;
    }
    uint32_t steps = 0, draws = 0, probes = 0, folds = 0, dtr = 0;
    bool more = true;
    // walker ids are claimed from the global cursor a chunk at a time into a
    // wave-local range [lp, le) and handed to empty slots from there: one
    // atomic per chunk instead of one per refill (short walkers refill every
    // few iterations, and the cursor is one contended address).  Chunks stay
    // small enough that every wave of the grid gets about 8 of them.
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t chunk = (uint32_t)min<uint64_t>(4u * 64u * WPL, (uint64_t)a.nwalkers / (8u * nwaves));
    uint64_t lp = 0, le = 0;
    for (;;) {
        if (more) {
            uint64_t emp[WPL];
            uint32_t need = 0;
#pragma unroll
            for (int k = 0; k < WPL; ++k) {
                emp[k] = __ballot(wid[k] < 0);
                need += (uint32_t)__popcll(emp[k]);
            }
            if (need) {
                // ranks below n0 take the rest of the local range, the others
                // a fresh claim of at least `need` ids
                const uint64_t r0 = lp, n0 = le - lp;
                unsigned long long b = 0;
                if (n0 < need) {
                    const uint32_t claim = max(chunk, need);
                    if (ln == 0) b = atomicAdd(a.cursor, (unsigned long long)claim);
                    b = __shfl(b, 0, 64);
                    lp = b + (need - n0);
                    le = b + claim;
                } else {
                    lp += need;
                }
                uint32_t off = 0;
#pragma unroll
                for (int k = 0; k < WPL; ++k) {
                    if (wid[k] < 0) {
                        const uint64_t rk = off + (uint32_t)__popcll(emp[k] & lt);
                        const uint64_t c = rk < n0 ? r0 + rk : b + (rk - n0);
                        if (c < (uint64_t)a.nwalkers) {
                            const int64_t w = (int64_t)c;
                            const int32_t st = a.starts ? a.starts[w] : (int32_t)(w % s.n) + 1;
                            int32_t L = a.lens ? a.lens[w] : a.max_len;
                            if (APP == 3 && L > 2) L = 2;
                            if (st < 1 || st > s.n || L < 1 || L > a.max_len) {
                                atomicOr(&s.ctl->werr, (unsigned)ERR_INVAL);
                                out.finish(a, k, w, 0);
                            } else {
                                wid[k] = w;
                                cur[k] = st;
                                prev[k] = -1;
                                jj[k] = 1;
                                len[k] = L;
                                att[k] = 0;
                                out.put(a, k, w, 0, st);
                                if (L == 1) {
                                    out.finish(a, k, w, 1);
                                    wid[k] = -1;
                                }
                            }
                        }
                    }
                    off += (uint32_t)__popcll(emp[k]);
                }
                // nothing valid left locally; the cursor is past the end too
                if (lp >= (uint64_t)a.nwalkers) more = false;
            }
        }
        bool mine = false;
#pragma unroll
        for (int k = 0; k < WPL; ++k) mine |= wid[k] >= 0;
        if (!__any(mine)) {
            if (!more) break;
            continue;
        }
#pragma unroll
        for (int k = 0; k < WPL; ++k) {
            if (wid[k] < 0) continue;
            const int64_t w = wid[k];
            const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)((uint64_t)w >> 32);
            const uint4 r = philox(w0, w1, (uint32_t)jj[k], att[k], k0, k1);
            uint64_t hcur = 0;
            int32_t nx = (!FW && APP != 1) ? sample_lite(s.alias, s.mem, cur[k], r)
                                            : draw<FW>(s, cur[k], r, hcur, w0, w1, (uint32_t)jj[k], att[k], k0, k1);
            ++draws;
            bool taken = false;  // the fold coin took the return arc
            if (fold && jj[k] >= 2) {
                if (att[k] == 0) fst[k].known = false;
                if (fold_step(s, cur[k], prev[k], hcur, a.fold_r, a.fold_cap, fst[k], folds, w0, w1, (uint32_t)jj[k],
                              att[k], k0, k1)) {
                    nx = prev[k];
                    taken = true;
                }
            }
            if (DN && nx < -1) nx = dense_draw(s, cur[k], -2 - nx, w0, w1, (uint32_t)jj[k], att[k], k0, k1, dtr);
            if (nx < 0) {
                padj[k] = out.close(a, k, w, jj[k]);
                padw[k] = w;
                wid[k] = -1;
                continue;
            }
            const uint32_t aux = r.x & kAuxMask;
            if (APP == 1 && !taken && jj[k] >= 2 && (int64_t)att[k] + 1 < a.max_attempts) {
                bool acc;
                if (nx == prev[k]) acc = aux < a.thr_return;
                else if (aux < tlo) acc = true;
                else if (aux >= thi) acc = false;
                else {
                    ++probes;
                    acc = aux < (is_neighbor(s.hash, hp[k], (uint32_t)nx) ? a.thr_neighbor : a.thr_other);
                }
                if (!acc) {
                    ++att[k];
                    continue;
                }
            }
            out.put(a, k, w, jj[k], nx);
            ++steps;
            prev[k] = cur[k];
            cur[k] = nx;
            ++jj[k];
            att[k] = 0;
            if (APP == 1) {
                hp[k] = hcur;
            }
            if ((APP == 2 && aux < a.thr_stop) || jj[k] >= len[k]) {
                padj[k] = out.close(a, k, w, jj[k]);
                padw[k] = w;
                wid[k] = -1;
            }
        }
        // Row tails of the walkers that finished in this iteration, padded
        // with -1.  Few rows: by the whole wave one row at a time, consecutive
        // lanes writing consecutive 16-byte (D > 0: rows are 16-byte aligned
        // from position j, a multiple of D) or 4-byte pieces.  Many rows (a
        // wave's worth of one-step walkers): each lane its own row, which
        // then takes fewer store instructions.
        constexpr int32_t per = D > 0 ? 4 : 1;
#pragma unroll
        for (int k = 0; k < WPL; ++k) {
            const uint64_t fin = __ballot(padw[k] >= 0);
            const int32_t rows = __popcll(fin);
            if ((int64_t)rows * ((a.max_len + 64 * per - 1) / (64 * per)) > (a.max_len + per - 1) / per) {
                if (padw[k] >= 0) {
                    int32_t *row = a.paths + padw[k] * (int64_t)a.max_len;
                    if (D > 0) {
                        for (int32_t t = padj[k]; t < a.max_len; t += 4) store4(row + t, -1, -1, -1, -1);
                    } else {
                        for (int32_t t = padj[k]; t < a.max_len; ++t) row[t] = -1;
                    }
                }
                padw[k] = -1;
                continue;
            }
            for (uint64_t m = fin; m; m &= m - 1) {
                const int src = __ffsll((unsigned long long)m) - 1;
                const uint32_t wlo = __shfl((uint32_t)padw[k], src, 64);
                const uint32_t whi = __shfl((uint32_t)((uint64_t)padw[k] >> 32), src, 64);
                const int32_t j0 = __shfl(padj[k], src, 64);
                int32_t *row = a.paths + (int64_t)(((uint64_t)whi << 32) | wlo) * a.max_len;
                if (D > 0) {
                    for (int32_t t = j0 + 4 * ln; t < a.max_len; t += 256)
                        store_row4(row + t, -1, -1, -1, -1);
                } else {
                    for (int32_t t = j0 + ln; t < a.max_len; t += 64) row[t] = -1;
                }
            }
            padw[k] = -1;
        }
    }
    const uint32_t tot = wave_sum32(steps), tdraw = wave_sum32(draws), tprobe = wave_sum32(probes),
                   tfold = wave_sum32(folds), tdtr = DN ? wave_sum32(dtr) : 0u;
    if (ln == 0) {
        if (tot) {
            atomicAdd(a.steps, (unsigned long long)tot);
            atomicAdd(&s.ctl->steps, (unsigned long long)tot);
        }
        if (tdraw) atomicAdd(&s.ctl->attempts, (unsigned long long)tdraw);
        if (tprobe) atomicAdd(&s.ctl->probes, (unsigned long long)tprobe);
        if (tfold) atomicAdd(&s.ctl->folds, (unsigned long long)tfold);
        if (tdtr) atomicAdd(&s.ctl->dtrials, (unsigned long long)tdtr);
    }
}

// ------------------------------------------------------------------------
// serving: one persistent launch runs the walkers of many requests.  The
// walkers of all admitted requests form one index space (request r owns
// [base_r, base_r + nwalkers_r)); lanes claim indices with the same
// wave-aggregated refill as k_walk_persist, map them to (request, local
// walker) by binary search over the request table held in LDS, and run the
// request's own app, parameters and seed.  Requests of different apps,
// sizes, seeds and parameters share every wave, and a request's paths do
// not depend on what it was batched with.  CLS is the launch's class:
// kServeOther (DeepWalk, PPR, sampling only) and kServeN2V (node2vec only)
// keep only their own state in registers; kServeMixed runs any app.
// ------------------------------------------------------------------------
constexpr int kServeOther = 0, kServeN2V = 1, kServeMixed = 2;
// walker slots per lane: one where node2vec state is carried only
__host__ __device__ constexpr int serve_wpl(int cls) { return cls == kServeOther ? 2 : 1; }
constexpr int kServeD = 8;  // path positions staged per slot in LDS

// A served walker's row: staged kServeD positions at a time in LDS and
// written as 16-byte stores when the request's rows allow it (max_len a
// multiple of kServeD, 16-byte aligned paths), else position by position.
struct ServeRow {
    int32_t *st;  // this slot's kServeD + 1 staging ints
    __device__ __forceinline__ void flush(int32_t *row, int32_t j0) const
    {
#pragma unroll
        for (int c = 0; c < kServeD / 4; ++c) store4(row + j0 + 4 * c, st[4 * c], st[4 * c + 1], st[4 * c + 2], st[4 * c + 3]);
    }
    __device__ __forceinline__ void put(bool stg, int32_t *row, int32_t j, int32_t v) const
    {
        if (!stg) {
            row[j] = v;
            return;
        }
        st[j % kServeD] = v;
        if (j % kServeD == kServeD - 1) flush(row, j - (kServeD - 1));
    }
    __device__ __forceinline__ void finish(bool stg, int32_t *row, int32_t j, int32_t max_len) const
    {
        if (!stg) {
            for (; j < max_len; ++j) row[j] = -1;
            return;
        }
        const int32_t r = j % kServeD;
        if (r) {
            for (int t = r; t < kServeD; ++t) st[t] = -1;
            flush(row, j - r);
            j += kServeD - r;
        }
        for (; j < max_len; j += 4) store4(row + j, -1, -1, -1, -1);
    }
};

template <bool FW, int CLS, bool DN>
__global__ __launch_bounds__(256) void k_walk_serve(DevStore s, const ServeReq *__restrict__ reqs, int nreq,
                                                    unsigned long long total, unsigned long long *cursor)
{
    constexpr int WPL = serve_wpl(CLS);
    constexpr bool MIX = CLS == kServeMixed;
    __shared__ ServeReq rq[kMaxServeReqs];
    __shared__ int32_t stage[256 * WPL * (kServeD + 1)];  // odd pitch: no bank conflicts
    // Steps per request are counted per workgroup: the LDS copy of an entry
    // holds this workgroup's count in place of the counter's address (which
    // stays in reqs), added to the global counter once at the end.  Lanes
    // counting straight into the global counters (adjacent words, so one L2
    // channel) serialised ~10^6 atomics per launch.
    auto count = [&](int i) { return &rq[i].count; };
    for (int i = threadIdx.x; i < nreq; i += blockDim.x) {
        rq[i] = reqs[i];
        rq[i].count = 0ull;
    }
    __syncthreads();
    const int ln = lane_id();
    const uint64_t lt = (1ull << ln) - 1ull;
    ServeRow out[WPL];
    bool stg[WPL];
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
        out[k].st = stage + ((size_t)threadIdx.x * WPL + k) * (kServeD + 1);
        stg[k] = false;
    }
    int32_t rid[WPL], cur[WPL], prev[WPL], jj[WPL], len[WPL];
    int64_t wid[WPL];
    uint32_t att[WPL], acc[WPL];
    int32_t acc_req[WPL];
    uint64_t hp[WPL];
    FoldState fst[WPL];
    uint32_t lookups = 0, dtr = 0;  // return-weight lookups, dense-group trials (not reported by the served form)
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
        rid[k] = acc_req[k] = -1;
        cur[k] = prev[k] = jj[k] = len[k] = 0;
        wid[k] = 0;
        att[k] = acc[k] = 0;
        hp[k] = 0;
        fst[k] = FoldState{0, 0, 0, false};
    }
    bool more = true;
    // indices are claimed from the launch cursor a chunk at a time into a
    // wave-local range [lp, le), as in k_walk_persist: walkers that finish
    // at different iterations (node2vec rejections, PPR stops) would
    // otherwise cost one atomic on the one cursor address per wave and
    // iteration
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t chunk = (uint32_t)min<uint64_t>(4u * 64u * WPL, (uint64_t)total / (8u * nwaves));
    uint64_t lp = 0, le = 0;
    for (;;) {
        if (more) {
            uint64_t emp[WPL];
            uint32_t need = 0;
#pragma unroll
            for (int k = 0; k < WPL; ++k) {
                emp[k] = __ballot(rid[k] < 0);
                need += (uint32_t)__popcll(emp[k]);
            }
            if (need) {
                // ranks below n0 take the rest of the local range, the others
                // a fresh claim of at least `need` indices
                const uint64_t r0 = lp, n0 = le - lp;
                unsigned long long b = 0;
                if (n0 < need) {
                    const uint32_t claim = max(chunk, need);
                    if (ln == 0) b = atomicAdd(cursor, (unsigned long long)claim);
                    b = __shfl(b, 0, 64);
                    lp = b + (need - n0);
                    le = b + claim;
                } else {
                    lp += need;
                }
                uint32_t off = 0;
#pragma unroll
                for (int k = 0; k < WPL; ++k) {
                    if (rid[k] < 0) {
                        const uint64_t rk = off + (uint32_t)__popcll(emp[k] & lt);
                        const unsigned long long c = rk < n0 ? r0 + rk : b + (rk - n0);
                        if (c < total) {
                            int lo = 0, hi = nreq - 1;  // last request with base <= c
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (rq[mid].base <= c) lo = mid; else hi = mid - 1;
                            }
                            const ServeReq &q = rq[lo];
                            const int64_t w = q.first + (int64_t)(c - q.base);
                            if (acc_req[k] != lo) {
                                if (acc_req[k] >= 0 && acc[k]) atomicAdd(count(acc_req[k]), (unsigned long long)acc[k]);
                                acc_req[k] = lo;
                                acc[k] = 0;
                            }
                            const int32_t st = q.starts ? q.starts[w] : (int32_t)(w % s.n) + 1;
                            int32_t L = q.lens ? q.lens[w] : q.max_len;
                            if (q.app == 3 && L > 2) L = 2;
                            int32_t *row = q.paths + w * (int64_t)q.max_len;
                            if (st < 1 || st > s.n || L < 1 || L > q.max_len || (!MIX && (q.app == 1) != (CLS == kServeN2V))) {
                                atomicOr(&s.ctl->werr, (unsigned)ERR_INVAL);
                                for (int32_t t = 0; t < q.max_len; ++t) row[t] = -1;
                            } else {
                                stg[k] = q.max_len % kServeD == 0 && ((uintptr_t)q.paths & 15u) == 0;
                                out[k].put(stg[k], row, 0, st);
                                rid[k] = lo;
                                wid[k] = w;
                                cur[k] = st;
                                prev[k] = -1;
                                jj[k] = 1;
                                len[k] = L;
                                att[k] = 0;
                                if (L == 1) {
                                    out[k].finish(stg[k], row, 1, q.max_len);
                                    rid[k] = -1;
                                }
                            }
                        }
                    }
                    off += (uint32_t)__popcll(emp[k]);
                }
                // nothing valid left locally; the cursor is past the end too
                if (lp >= (uint64_t)total) more = false;
            }
        }
        bool mine = false;
#pragma unroll
        for (int k = 0; k < WPL; ++k) mine |= rid[k] >= 0;
        if (!__any(mine)) {
            if (!more) break;
            continue;
        }
#pragma unroll
        for (int k = 0; k < WPL; ++k) {
            if (rid[k] < 0) continue;
            const ServeReq &q = rq[rid[k]];
            const int64_t w = wid[k];
            int32_t *row = q.paths + w * (int64_t)q.max_len;
            const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)((uint64_t)w >> 32);
            const uint32_t k0 = (uint32_t)q.seed, k1 = (uint32_t)(q.seed >> 32);
            const uint4 r = philox(w0, w1, (uint32_t)jj[k], att[k], k0, k1);
            const bool n2v = MIX ? q.app == 1 : CLS == kServeN2V;
            const bool fold = !FW && n2v && q.fold_r && jj[k] >= 2;
            uint64_t hcur = 0;
            int32_t nx = (!FW && !n2v) ? sample_lite(s.alias, s.mem, cur[k], r)
                                              : draw<FW>(s, cur[k], r, hcur, w0, w1, (uint32_t)jj[k], att[k], k0, k1);
            bool taken = false;  // the fold coin took the return arc
            if (fold) {
                if (att[k] == 0) fst[k].known = false;
                if (fold_step(s, cur[k], prev[k], hcur, q.fold_r, q.fold_cap, fst[k], lookups, w0, w1, (uint32_t)jj[k],
                              att[k], k0, k1)) {
                    nx = prev[k];
                    taken = true;
                }
            }
            if (DN && nx < -1) nx = dense_draw(s, cur[k], -2 - nx, w0, w1, (uint32_t)jj[k], att[k], k0, k1, dtr);
            bool done = nx < 0;
            if (!done) {
                const uint32_t aux = r.x & kAuxMask;
                if (n2v && !taken && jj[k] >= 2 && (int64_t)att[k] + 1 < q.max_attempts) {
                    const uint32_t tlo = min(q.thr_neighbor, q.thr_other), thi = max(q.thr_neighbor, q.thr_other);
                    bool ok;
                    if (nx == prev[k]) ok = aux < q.thr_return;
                    else if (aux < tlo) ok = true;
                    else if (aux >= thi) ok = false;
                    else ok = aux < (is_neighbor(s.hash, hp[k], (uint32_t)nx) ? q.thr_neighbor : q.thr_other);
                    if (!ok) {
                        ++att[k];
                        continue;
                    }
                }
                out[k].put(stg[k], row, jj[k], nx);
                ++acc[k];
                prev[k] = cur[k];
                cur[k] = nx;
                hp[k] = hcur;
                ++jj[k];
                att[k] = 0;
                done = (q.app == 2 && aux < q.thr_stop) || jj[k] >= len[k];
            }
            if (done) {
                out[k].finish(stg[k], row, jj[k], q.max_len);
                rid[k] = -1;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < WPL; ++k)
        if (acc_req[k] >= 0 && acc[k]) atomicAdd(count(acc_req[k]), (unsigned long long)acc[k]);
    __syncthreads();  // every wave leaves the loop above
    for (int i = threadIdx.x; i < nreq; i += blockDim.x) {
        const unsigned long long c = *count(i);
        if (c) atomicAdd(reqs[i].steps, c);
    }
}

// ------------------------------------------------------------------------
// export and point queries
// ------------------------------------------------------------------------
__global__ void k_export_a(DevStore s, uint32_t *nslots, uint64_t *mass, uint32_t *bsize, uint64_t *T, uint8_t *alias,
                           uint64_t *fmass, uint64_t *fthr, uint32_t *dmask)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t total = ((int64_t)s.n + 1) * kGroups;
    if (i >= total) return;
    const int j = (int)(i & (kGroups - 1));
    const int64_t u = i >> 5;
    const AliasEnt e = s.alias[i];
    const uint32_t sz = s.grp[i].size;
    bsize[i] = sz;
    T[i] = e.T;
    alias[i] = (uint8_t)(e.meta & 31u);
    if (j == 0) {
        nslots[u] = s.hdr[u].nslots;
        dmask[u] = s.hdr[u].dmask;
        uint64_t acc = 0;
        for (int g = 0; g < kGroups; ++g) acc += (uint64_t)s.grp[u * kGroups + g].size << g;
        mass[u] = acc;
        if (fmass) {
            const FNode f = s.fnode[u];
            fmass[u] = f.F;
            fthr[u] = f.thr;
        }
    }
}

__global__ __launch_bounds__(256) void k_export_b(DevStore s, const uint64_t *slot_out_off, const uint64_t *mem_out_off,
                                                  uint32_t *slot_v, uint32_t *slot_w, uint64_t *members,
                                                  uint32_t *backpos, uint32_t *slot_f)
{
    const int ln = lane_id();
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (wave > s.n) return;
    const int64_t u = wave;
    const NodeHdr h = s.hdr[u];
    const uint32_t *V = s.slot + h.slot_off, *W = V + h.slot_cap, *B = V + 2ull * h.slot_cap;
    const uint64_t so = slot_out_off[u];
    for (uint32_t i = ln; i < h.nslots; i += 64) {
        slot_v[so + i] = V[i];
        slot_w[so + i] = W[i];
        if (slot_f) slot_f[so + i] = V[3ull * h.slot_cap + i];
    }
    uint64_t mo = mem_out_off[u];
    for (int j = 0; j < kGroups; ++j) {  // member lists (dense groups keep none)
        if ((h.dmask >> j) & 1u) continue;
        const GroupHdr g = s.grp[u * kGroups + j];
        for (uint32_t i = ln; i < g.size; i += 64) members[mo + i] = s.mem[g.off + i] - h.sbase;  // slot fields
        mo += g.size;
    }
    uint64_t bo = mem_out_off[u];
    for (uint32_t c = 0; c < h.nslots; c += 64) {
        const uint32_t i = c + ln;
        const uint32_t w = i < h.nslots ? W[i] & ~h.dmask : 0u;
        const uint32_t pc = (uint32_t)__popc(w);
        const uint32_t incl = wave_incl32(pc);
        for (uint32_t k = 0; k < pc; ++k) backpos[bo + incl - pc + k] = s.back[h.back_off + B[i] + k];
        bo += __shfl(incl, 63, 64);
    }
}

__global__ void k_edge_count(DevStore s, const int32_t *us, const int32_t *vs, int64_t k, uint32_t *counts)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int32_t u = us[i], v = vs[i];
    uint32_t c = 0;
    if (u >= 1 && u <= s.n && v >= 1 && v <= s.n) {
        const NodeHdr h = s.hdr[u];
        if (s.nbr) {
            if (h.hash_cap) {
                const uint64_t *tab = s.hash + h.hash_off;
                uint32_t p = hslot((uint32_t)v, h.hash_cap);
                for (uint32_t probe = 0; probe < h.hash_cap; ++probe) {
                    const uint64_t e = tab[2ull * p];
                    if ((uint32_t)e == (uint32_t)v) {
                        c = (uint32_t)(e >> 32);
                        break;
                    }
                    if ((uint32_t)e == 0u) break;
                    p = (p + 1u) & (h.hash_cap - 1u);
                }
            }
        } else {
            const uint32_t *V = s.slot + h.slot_off, *W = V + h.slot_cap, *FR = V + 3ull * h.slot_cap;
            for (uint32_t t = 0; t < h.nslots; ++t)
                c += ((W[t] != 0u || (s.fw && FR[t] != 0u)) && V[t] == (uint32_t)v);
        }
    }
    counts[i] = c;
}

// plain streaming copy: 16-byte vectors, grid-stride (tail bytes by lane 0)
__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n16,
                                                uint8_t *__restrict__ dtail, const uint8_t *__restrict__ stail,
                                                uint32_t ntail)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

// Random-gather probe (the ceiling a walk's dependent gathers are judged
// against): each thread makes `iters` rounds of 8 independent W-byte reads
// at hashed positions in [base, base + units * W).  W = 32 reads like an
// alias slot (two 16-byte loads), W = 8 like a member.
__device__ __forceinline__ uint64_t probe_mix(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

template <int W>
__global__ __launch_bounds__(256) void k_gather_probe(const uint8_t *__restrict__ base, uint64_t units, uint32_t iters,
                                                      unsigned long long *sink)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint64_t acc = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t i = probe_mix(t * 0x9E3779B97F4A7C15ull + (uint64_t)it * 8 + k) % units;
            if (W == 8) {
                v[k] = reinterpret_cast<const uint64_t *>(base)[i];
            } else {
                const uint4 *p = reinterpret_cast<const uint4 *>(base + i * 32);
                const uint4 a = p[0], b = p[1];
                v[k] = (uint64_t)a.x + a.y + a.z + a.w + b.x + b.y + b.z + b.w;
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    if (acc == 0x5eed5eed5eedull) *sink = acc;  // keeps the loads; practically never taken
}

// Random read-modify-write probe (the ceiling for the update kernels'
// scattered partial-line writes): rounds of 8 independent 8-byte words at
// hashed positions in [base, base + units * 8), each read and stored back
// unchanged, so the store it runs over keeps its contents.  A random 8-byte
// store is one 32-byte write request to memory on this part.
__global__ __launch_bounds__(256) void k_rmw_probe(uint64_t *__restrict__ base, uint64_t units, uint32_t iters)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint32_t it = 0; it < iters; ++it) {
        uint64_t i[8], v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            i[k] = probe_mix(t * 0x9E3779B97F4A7C15ull + (uint64_t)it * 8 + k) % units;
            v[k] = base[i[k]];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            asm volatile("" : "+v"(v[k]));  // the store is not provably the loaded value: kept
            base[i[k]] = v[k];
        }
    }
}

// ------------------------------------------------------------------------
// deferred region moves of an insert batch (see CopyTask): phase 0 copies
// regions and zeroes new neighbour tables, phase 1 re-inserts the old tables'
// live keys.  A workgroup per chunk; the chunk's task is found by binary
// search over the tasks' chunk bases, held in LDS.
// ------------------------------------------------------------------------
template <int PHASE>
__global__ __launch_bounds__(256) void k_copy_tasks(DevStore s)
{
    __shared__ uint32_t cb[kMaxCopyTasks];
    const uint32_t nt = min((uint32_t)(s.ctl->taskctr >> 32), kMaxCopyTasks);
    if (nt == 0 || s.ctl->ustatus) return;
    for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) cb[i] = s.tasks[i].cbase;
    __syncthreads();
    const uint32_t nchunks = cb[nt - 1] + s.tasks[nt - 1].nch;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint32_t lo = 0, hi = nt - 1;  // last task with cbase <= c
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (cb[mid] <= c) lo = mid; else hi = mid - 1;
        }
        const CopyTask t = s.tasks[lo];
        const uint32_t k = c - t.cbase;
        if (PHASE == 0) {
            const uint64_t total = t.kind == TASK_COPY ? t.n : 16u * t.n;
            const uint64_t b0 = (uint64_t)k * kCopyChunk, b1 = min(b0 + kCopyChunk, total);
            if (t.kind == TASK_COPY && ((t.dst | t.src | t.n) & 7u) == 0u) {
                uint64_t *d = reinterpret_cast<uint64_t *>(t.dst);
                const uint64_t *q = reinterpret_cast<const uint64_t *>(t.src);
                for (uint64_t i = b0 / 8u + threadIdx.x; i < b1 / 8u; i += blockDim.x) d[i] = q[i];
            } else if (t.kind == TASK_COPY) {
                uint32_t *d = reinterpret_cast<uint32_t *>(t.dst);
                const uint32_t *q = reinterpret_cast<const uint32_t *>(t.src);
                for (uint64_t i = b0 / 4u + threadIdx.x; i < b1 / 4u; i += blockDim.x) d[i] = q[i];
            } else {
                uint64_t *d = reinterpret_cast<uint64_t *>(t.dst);
                for (uint64_t i = b0 / 8u + threadIdx.x; i < b1 / 8u; i += blockDim.x) d[i] = 0ull;
            }
        } else if (t.kind == TASK_REHASH) {
            // this chunk's share of the old table's entries
            const uint64_t e0 = (uint64_t)t.old_cap * k / t.nch, e1 = (uint64_t)t.old_cap * (k + 1) / t.nch;
            const uint64_t *old = reinterpret_cast<const uint64_t *>(t.src);
            uint32_t fresh = 0;
            for (uint64_t i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
                const uint64_t x = old[2u * i], w = old[2u * i + 1u];
                if ((uint32_t)x && (x >> 32))
                    fresh += hash_add(reinterpret_cast<uint64_t *>(t.dst), (uint32_t)t.n, (uint32_t)x,
                                      (uint32_t)(x >> 32), w);
            }
            fresh = wave_sum32(fresh);
            if (lane_id() == 0 && fresh) atomicAdd(&s.hdr[t.u].hash_used, fresh);
        }
    }
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
static inline unsigned run_blocks(int64_t k)
{
    // one wave per run, 4 waves per block, capped at ~8 resident waves per SIMD
    const int64_t b = (k + 3) / 4;
    return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

void launch_reset_ctl(Ctl *ctl, hipStream_t st) { hipLaunchKernelGGL(k_reset_ctl, dim3(1), dim3(64), 0, st, ctl); }

uint64_t launch_gather_probe(void *base, uint64_t units, int width, unsigned long long *sink, hipStream_t st)
{
    const unsigned blocks = 4096, threads = 256;
    const uint32_t iters = 64;
    if (width == 0)
        hipLaunchKernelGGL(k_rmw_probe, dim3(blocks), dim3(threads), 0, st,
                           static_cast<uint64_t *>(base), units, iters);
    else if (width == 32)
        hipLaunchKernelGGL(k_gather_probe<32>, dim3(blocks), dim3(threads), 0, st,
                           static_cast<const uint8_t *>(base), units, iters, sink);
    else
        hipLaunchKernelGGL(k_gather_probe<8>, dim3(blocks), dim3(threads), 0, st, static_cast<const uint8_t *>(base),
                           units, iters, sink);
    return (uint64_t)blocks * threads * iters * 8;
}

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const size_t n16 = bytes / 16, done = n16 * 16;
    const unsigned blocks = (unsigned)std::min<size_t>(std::max<size_t>((n16 + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, st, static_cast<uint4 *>(dst),
                       static_cast<const uint4 *>(src), n16, static_cast<uint8_t *>(dst) + done,
                       static_cast<const uint8_t *>(src) + done, (uint32_t)(bytes - done));
}

void launch_copy_tasks(const DevStore &s, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_tasks<0>, dim3(2048), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_copy_tasks<1>, dim3(2048), dim3(256), 0, st, s);
}

void launch_reset_walk(Ctl *ctl, hipStream_t st) { hipLaunchKernelGGL(k_reset_walk, dim3(1), dim3(64), 0, st, ctl); }

void launch_init_nodes(DevStore s, hipStream_t st)
{
    const int64_t total = ((int64_t)s.n + 1) * kGroups;
    hipLaunchKernelGGL(k_init_nodes, dim3(blocks_for(total, 256)), dim3(256), 0, st, s);
}

void launch_insert_keys(const DevEdge *e, int64_t k, const unsigned long long *dk, int32_t n, int fw, uint32_t *keys,
                        uint32_t *vals, Ctl *ctl, hipStream_t st)
{
    hipLaunchKernelGGL(k_insert_keys, dim3(blocks_for(k, 256)), dim3(256), 0, st, e, k, dk, n, fw, keys, vals, ctl);
}

void launch_gather_edges(const DevEdge *e, const uint32_t *order, int64_t k, const unsigned long long *dk, DevEdge *out,
                         uint32_t *frac, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_edges, dim3(blocks_for(k, 256)), dim3(256), 0, st, e, order, k, dk, out, frac);
}

__global__ void k_set_batch(Ctl *ctl, int64_t kins, int64_t kdel)
{
    if (threadIdx.x == 0) {
        ctl->bk[0] = (unsigned long long)kins;
        ctl->bk[1] = (unsigned long long)kdel;
    }
}

void launch_set_batch(Ctl *ctl, int64_t kins, int64_t kdel, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_batch, dim3(1), dim3(64), 0, st, ctl, kins, kdel);
}

__global__ void k_delete_guard(Ctl *ctl)
{
    if (threadIdx.x == 0 && (ctl->error & (ERR_INVAL | ERR_OVERRUN))) atomicOr(&ctl->ustatus, (unsigned)UST_INVAL);
}

void launch_delete_guard(Ctl *ctl, hipStream_t st) { hipLaunchKernelGGL(k_delete_guard, dim3(1), dim3(64), 0, st, ctl); }

__global__ void k_clear_ustatus(Ctl *ctl)
{
    if (threadIdx.x == 0) ctl->ustatus = 0u;
}

void launch_clear_ustatus(Ctl *ctl, hipStream_t st) { hipLaunchKernelGGL(k_clear_ustatus, dim3(1), dim3(64), 0, st, ctl); }

// the merged free lists back into s.fl (captured batches keep one list buffer)
__global__ __launch_bounds__(256) void k_freelist_copyback(DevStore s)
{
    if (s.ctl->ustatus) return;
    const uint64_t n = s.ctl->fl_total;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s.fl[i] = s.fl2[i];
}

void launch_freelist_copyback(DevStore s, hipStream_t st)
{
    hipLaunchKernelGGL(k_freelist_copyback, dim3(1024), dim3(256), 0, st, s);
}

size_t insert_wcls_bytes(int64_t k) { return (size_t)run_blocks(k) * 2 * kPC * sizeof(uint32_t); }

__global__ void k_alloc_reset(FreeMeta *fm)
{
    const int t = threadIdx.x;
    if (t < kPC) {
        fm->take[t] = 0u;
        fm->freed[t] = 0u;
    }
}

void launch_alloc_reset(FreeMeta *fm, hipStream_t st) { hipLaunchKernelGGL(k_alloc_reset, dim3(1), dim3(kPC), 0, st, fm); }

void launch_alloc_resolve(DevStore s, int fw, hipStream_t st, int captured, uint64_t flcap, uint64_t pfcap)
{
    hipLaunchKernelGGL(k_alloc_resolve, dim3(1), dim3(kPC), 0, st, s, fw, captured, flcap, pfcap);
}

void launch_freelist_merge(DevStore s, hipStream_t st)
{
    hipLaunchKernelGGL(k_freelist_merge, dim3(512), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_freelist_commit, dim3(1), dim3(kPC), 0, st, s);
}

void launch_insert_plan(DevStore s, const DevEdge *sorted, Runs r, int64_t k, hipStream_t st)
{
    if (s.rate16)
        hipLaunchKernelGGL(k_insert_plan<true>, dim3(run_blocks(k)), dim3(256), 0, st, s, sorted, r);
    else
        hipLaunchKernelGGL(k_insert_plan<false>, dim3(run_blocks(k)), dim3(256), 0, st, s, sorted, r);
}

void launch_insert_apply(DevStore s, const DevEdge *sorted, const uint32_t *frac, Runs r, int64_t k, hipStream_t st)
{
    // same grid as the plan: workgroup b takes the index ranges workgroup b reserved
    if (s.rate16)
        hipLaunchKernelGGL(k_insert_apply<true>, dim3(run_blocks(k)), dim3(256), 0, st, s, sorted, frac, r);
    else
        hipLaunchKernelGGL(k_insert_apply<false>, dim3(run_blocks(k)), dim3(256), 0, st, s, sorted, frac, r);
    hipLaunchKernelGGL(k_commit_tops, dim3(1), dim3(64), 0, st, s.ctl);
}

void launch_hash_insert(DevStore s, const DevEdge *sorted, int64_t k, hipStream_t st)
{
    hipLaunchKernelGGL(k_hash_insert, dim3(blocks_for(k, 256)), dim3(256), 0, st, s, sorted, k);
}

void launch_hash_delete(DevStore s, const uint32_t *us, const uint32_t *slots, int64_t k, hipStream_t st)
{
    hipLaunchKernelGGL(k_hash_delete, dim3(blocks_for(k, 256)), dim3(256), 0, st, s, us, slots, k);
}

void launch_delete_keys(const int32_t *dels, int64_t k, const unsigned long long *dk, int32_t n, Ctl *ctl,
                        unsigned long long *keys, hipStream_t st)
{
    hipLaunchKernelGGL(k_delete_keys, dim3(blocks_for(k, 256)), dim3(256), 0, st, dels, k, dk, n, ctl, keys);
}

void launch_delete_check(DevStore s, const unsigned long long *keys, int64_t k, int pads, uint32_t *us, uint32_t *slots,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_delete_check, dim3(blocks_for(k, 256)), dim3(256), 0, st, s, keys, k, pads, us, slots);
}

void launch_delete_big_list(Runs r, int64_t k, uint32_t *big, Ctl *ctl, hipStream_t st)
{
    hipLaunchKernelGGL(k_delete_big_list, dim3(blocks_for(k, 256)), dim3(256), 0, st, r, big, ctl);
}

void launch_delete_apply(DevStore s, const uint32_t *slots, Runs r, const uint32_t *big, int64_t k, hipStream_t st)
{
    const int64_t nbig_max = k / kBigDelete + 1;  // waves loop over the list: at most 1024 blocks
    const dim3 fin(std::min<unsigned>(blocks_for(nbig_max * 64, 256), 1024u));
    if (s.rate16) {
        hipLaunchKernelGGL(k_delete_apply<true>, dim3(run_blocks(k)), dim3(256), 0, st, s, slots, r, big);
        hipLaunchKernelGGL(k_delete_finish<true>, fin, dim3(256), 0, st, s, slots, r, big);
    } else {
        hipLaunchKernelGGL(k_delete_apply<false>, dim3(run_blocks(k)), dim3(256), 0, st, s, slots, r, big);
        hipLaunchKernelGGL(k_delete_finish<false>, fin, dim3(256), 0, st, s, slots, r, big);
    }
}

template <bool FW, bool DN>
static void launch_walk_fd(DevStore s, int app, const WalkArgs &a, dim3 grid, dim3 blk, hipStream_t st)
{
    switch (app) {
    case 0: hipLaunchKernelGGL((k_walk<0, FW, DN>), grid, blk, 0, st, s, a); break;
    case 1: hipLaunchKernelGGL((k_walk<1, FW, DN>), grid, blk, 0, st, s, a); break;
    case 2: hipLaunchKernelGGL((k_walk<2, FW, DN>), grid, blk, 0, st, s, a); break;
    default: hipLaunchKernelGGL((k_walk<3, FW, DN>), grid, blk, 0, st, s, a); break;
    }
}

void launch_walk(DevStore s, int app, const WalkArgs &a, int block, hipStream_t st)
{
    if (block != 64 && block != 128 && block != 256) block = 256;
    const dim3 grid(blocks_for(a.nwalkers, block)), blk(block);
    if (s.fw)
        s.rate16 ? launch_walk_fd<true, true>(s, app, a, grid, blk, st) : launch_walk_fd<true, false>(s, app, a, grid, blk, st);
    else
        s.rate16 ? launch_walk_fd<false, true>(s, app, a, grid, blk, st)
                 : launch_walk_fd<false, false>(s, app, a, grid, blk, st);
}

int serve_walkers_per_lane(int cls) { return serve_wpl(cls); }

template <bool FW, bool DN>
static const void *serve_fn(int cls)
{
    return cls == kServeN2V     ? reinterpret_cast<const void *>(&k_walk_serve<FW, kServeN2V, DN>)
           : cls == kServeOther ? reinterpret_cast<const void *>(&k_walk_serve<FW, kServeOther, DN>)
                                : reinterpret_cast<const void *>(&k_walk_serve<FW, kServeMixed, DN>);
}

int serve_max_blocks_per_cu(int fw, int cls, int dense)
{
    int nb = 0;
    const void *fn = fw ? (dense ? serve_fn<true, true>(cls) : serve_fn<true, false>(cls))
                        : (dense ? serve_fn<false, true>(cls) : serve_fn<false, false>(cls));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb > 0 ? nb : 1;
}

template <bool FW, bool DN>
static void launch_serve_cls(DevStore s, const ServeReq *reqs, int nreq, unsigned long long total,
                             unsigned long long *cursor, int cls, int blocks, hipStream_t st)
{
    const dim3 g(blocks), b(256);
    if (cls == kServeN2V)
        hipLaunchKernelGGL((k_walk_serve<FW, kServeN2V, DN>), g, b, 0, st, s, reqs, nreq, total, cursor);
    else if (cls == kServeOther)
        hipLaunchKernelGGL((k_walk_serve<FW, kServeOther, DN>), g, b, 0, st, s, reqs, nreq, total, cursor);
    else
        hipLaunchKernelGGL((k_walk_serve<FW, kServeMixed, DN>), g, b, 0, st, s, reqs, nreq, total, cursor);
}

void launch_walk_serve(DevStore s, const ServeReq *reqs, int nreq, unsigned long long total,
                       unsigned long long *cursor, int cls, int blocks, hipStream_t st)
{
    if (s.fw)
        s.rate16 ? launch_serve_cls<true, true>(s, reqs, nreq, total, cursor, cls, blocks, st)
                 : launch_serve_cls<true, false>(s, reqs, nreq, total, cursor, cls, blocks, st);
    else
        s.rate16 ? launch_serve_cls<false, true>(s, reqs, nreq, total, cursor, cls, blocks, st)
                 : launch_serve_cls<false, false>(s, reqs, nreq, total, cursor, cls, blocks, st);
}

size_t persist_lds_bytes(int block, int wpl, int depth)
{
    return depth ? (size_t)block * wpl * (depth + 1) * sizeof(int32_t) : 0;
}

template <int APP, bool FW, bool DN>
static const void *persist_fn(int wpl, int depth)
{
#define BINGO_PF(W, D) reinterpret_cast<const void *>(&k_walk_persist<APP, W, D, FW, DN>)
    switch (wpl * 100 + depth) {
    case 100: return BINGO_PF(1, 0);
    case 104: return BINGO_PF(1, 4);
    case 116: return BINGO_PF(1, 16);
    case 200: return BINGO_PF(2, 0);
    case 204: return BINGO_PF(2, 4);
    case 216: return BINGO_PF(2, 16);
    case 400: return BINGO_PF(4, 0);
    case 404: return BINGO_PF(4, 4);
    case 416: return BINGO_PF(4, 16);
    default: return nullptr;
    }
#undef BINGO_PF
}

template <bool FW, bool DN>
static const void *persist_kernel_fd(int app, int wpl, int depth)
{
    switch (app) {
    case 0: return persist_fn<0, FW, DN>(wpl, depth);
    case 1: return persist_fn<1, FW, DN>(wpl, depth);
    case 2: return persist_fn<2, FW, DN>(wpl, depth);
    default: return persist_fn<3, FW, DN>(wpl, depth);
    }
}

static const void *persist_kernel(int app, int wpl, int depth, int fw, int dense)
{
    if (fw) return dense ? persist_kernel_fd<true, true>(app, wpl, depth) : persist_kernel_fd<true, false>(app, wpl, depth);
    return dense ? persist_kernel_fd<false, true>(app, wpl, depth) : persist_kernel_fd<false, false>(app, wpl, depth);
}

int persist_max_blocks_per_cu(int app, int wpl, int depth, int block, int fw, int dense)
{
    const void *fn = persist_kernel(app, wpl, depth, fw, dense);
    if (!fn) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, persist_lds_bytes(block, wpl, depth)) !=
        hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

void launch_walk_persist(DevStore s, int app, const WalkArgs &a, const PersistCfg &c, hipStream_t st)
{
    const void *fn = persist_kernel(app, c.wpl, c.depth, s.fw, s.rate16 != 0);
    if (!fn) return;
    void *args[] = {&s, const_cast<WalkArgs *>(&a)};
    (void)hipLaunchKernel(fn, dim3(c.blocks), dim3(c.block), args, persist_lds_bytes(c.block, c.wpl, c.depth), st);
}

void launch_export_a(DevStore s, uint32_t *nslots, uint64_t *mass, uint32_t *bsize, uint64_t *T, uint8_t *alias,
                     uint64_t *fmass, uint64_t *fthr, uint32_t *dmask, hipStream_t st)
{
    const int64_t total = ((int64_t)s.n + 1) * kGroups;
    hipLaunchKernelGGL(k_export_a, dim3(blocks_for(total, 256)), dim3(256), 0, st, s, nslots, mass, bsize, T, alias,
                       fmass, fthr, dmask);
}

void launch_export_b(DevStore s, const uint64_t *slot_out_off, const uint64_t *mem_out_off, uint32_t *slot_v,
                     uint32_t *slot_w, uint64_t *members, uint32_t *backpos, uint32_t *slot_f, hipStream_t st)
{
    const int64_t waves = (int64_t)s.n + 1;
    hipLaunchKernelGGL(k_export_b, dim3(blocks_for(waves * 64, 256)), dim3(256), 0, st, s, slot_out_off, mem_out_off,
                       slot_v, slot_w, members, backpos, slot_f);
}

void launch_compact_select(DevStore s, uint32_t thr16, uint32_t *small_list, uint32_t *big_list, uint32_t *shift_list,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_compact_reset, dim3(1), dim3(64), 0, st, s.ctl);
    const int64_t n1 = (int64_t)s.n + 1;
    int dev = 0, ncu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const unsigned gb = (unsigned)std::min<int64_t>((n1 + 256 * kSelPer - 1) / (256 * kSelPer), (int64_t)ncu * 8);
    hipLaunchKernelGGL(k_compact_select, dim3(gb), dim3(256), 0, st, s, thr16, small_list, big_list, shift_list);
}

void launch_compact_apply(DevStore s, const uint32_t *small_list, const uint32_t *big_list, const uint32_t *shift_list,
                          uint32_t nsmall, uint32_t nbig, uint32_t nzs, uint32_t nzb, uint64_t nseg,
                          const CompactBig &cb, hipStream_t st)
{
    int dev = 0, ncu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t n1 = (uint32_t)s.n + 1;
    if (nzs) {
        const unsigned zb = (unsigned)std::min<int64_t>((nzs + 255) / 256, (int64_t)ncu * 8);
        hipLaunchKernelGGL(k_compact_shift, dim3(zb), dim3(256), 0, st, s, shift_list);
        if (s.nbr) hipLaunchKernelGGL(k_compact_hash, dim3(zb), dim3(256), 0, st, s, shift_list, 2, 0, n1);
    }
    if (nzb) {
        hipLaunchKernelGGL(k_compact_shift_big, dim3(std::min<uint32_t>(nzb, (uint32_t)ncu * 2)), dim3(1024), 0, st, s,
                           shift_list, n1);
        if (s.nbr)
            hipLaunchKernelGGL(k_compact_hash, dim3((nzb + 3) / 4), dim3(256), 0, st, s, shift_list, 3, 1, n1);
    }
    if (nsmall) {
        const unsigned sb = (unsigned)std::min<int64_t>((nsmall + 255) / 256, (int64_t)ncu * 8);
        hipLaunchKernelGGL(k_compact_small, dim3(sb), dim3(256), 0, st, s, small_list);
        if (s.nbr) hipLaunchKernelGGL(k_compact_hash, dim3(sb), dim3(256), 0, st, s, small_list, 0, 0, n1);
    }
    if (!nbig) return;
    const unsigned gw = (unsigned)std::min<uint64_t>((nseg + 3) / 4, (uint64_t)ncu * 8);  // a wave per segment
    const unsigned gb = (unsigned)std::min<uint64_t>(nseg, (uint64_t)ncu * 8);
    hipLaunchKernelGGL(k_compact_segbase, dim3(1), dim3(1024), 0, st, s, big_list, nbig, cb);
    hipLaunchKernelGGL(k_compact_segcount, dim3(gw), dim3(256), 0, st, s, big_list, nbig, nseg, cb);
    hipLaunchKernelGGL(k_compact_segscan, dim3((nbig + 3) / 4), dim3(256), 0, st, nbig, cb);
    hipLaunchKernelGGL(k_compact_nodescan, dim3(1), dim3(1024), 0, st, nbig, cb);
    hipLaunchKernelGGL(k_compact_stage, dim3(gw), dim3(256), 0, st, s, big_list, nbig, nseg, cb);
    hipLaunchKernelGGL(k_compact_unstage, dim3(gb), dim3(256), 0, st, s, big_list, nbig, nseg, cb);
    hipLaunchKernelGGL(k_compact_bigfinish, dim3((unsigned)std::min<uint32_t>(nbig, (uint32_t)ncu * 4)), dim3(256), 0,
                       st, s, big_list, nbig, cb);
}

void launch_edge_count(DevStore s, const int32_t *us, const int32_t *vs, int64_t k, uint32_t *counts, hipStream_t st)
{
    hipLaunchKernelGGL(k_edge_count, dim3(blocks_for(k, 256)), dim3(256), 0, st, s, us, vs, k, counts);
}

// ------------------------------------------------------------------------
// rocPRIM wrappers
// ------------------------------------------------------------------------
// Update batches sort node ids (23-30 significant bits; deletes also the
// slot).  rocPRIM's default takes a block merge sort up to 2^20 items, which
// for a 10^6-arc batch costs ~20 launches and 200-250 us; past 2^16 items
// the onesweep radix sort (a few passes of 8-bit digits) is used instead.
using BatchSortCfg = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                rocprim::default_config, (size_t)1 << 16>;

size_t sort_pairs_u32(void *tmp, size_t tmp_bytes, uint32_t *k0, uint32_t *k1, uint32_t *v0, uint32_t *v1, int64_t k,
                      int bits, hipStream_t st, bool *second)
{
    size_t b = tmp_bytes;
    rocprim::double_buffer<uint32_t> keys(k0, k1), vals(v0, v1);
    (void)rocprim::radix_sort_pairs<BatchSortCfg>(tmp, b, keys, vals, (size_t)k, 0, (unsigned)bits, st);
    if (second) *second = keys.current() == k1;
    return b;
}

size_t sort_keys_u64(void *tmp, size_t tmp_bytes, const unsigned long long *kin, unsigned long long *kout, int64_t k,
                     int bits, hipStream_t st)
{
    size_t b = tmp_bytes;
    (void)rocprim::radix_sort_keys<BatchSortCfg>(tmp, b, kin, kout, (size_t)k, 0, (unsigned)bits, st);
    return b;
}

size_t run_length_encode_u32(void *tmp, size_t tmp_bytes, const uint32_t *keys, int64_t k, uint32_t *uniq,
                             uint32_t *counts, uint32_t *nruns, hipStream_t st)
{
    size_t b = tmp_bytes;
    (void)rocprim::run_length_encode(tmp, b, keys, (size_t)k, uniq, counts, nruns, st);
    return b;
}

size_t exclusive_scan_u32(void *tmp, size_t tmp_bytes, const uint32_t *in, uint32_t *out, int64_t k, hipStream_t st)
{
    size_t b = tmp_bytes;
    (void)rocprim::exclusive_scan(tmp, b, in, out, 0u, (size_t)k, rocprim::plus<uint32_t>(), st);
    return b;
}

size_t merge_sort_pairs_u32(void *tmp, size_t tmp_bytes, const uint32_t *kin, uint32_t *kout, const uint32_t *vin,
                            uint32_t *vout, int64_t k, hipStream_t st)
{
    size_t b = tmp_bytes;
    (void)rocprim::merge_sort(tmp, b, kin, kout, vin, vout, (size_t)k, rocprim::less<uint32_t>(), st);
    return b;
}

size_t merge_sort_keys_u64(void *tmp, size_t tmp_bytes, const unsigned long long *kin, unsigned long long *kout,
                           int64_t k, hipStream_t st)
{
    size_t b = tmp_bytes;
    (void)rocprim::merge_sort(tmp, b, kin, kout, (size_t)k, rocprim::less<unsigned long long>(), st);
    return b;
}

}  // namespace bingo

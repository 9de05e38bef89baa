"""The batched workload at full size: R-MAT 2^23 nodes / 2^27 undirected
edges (2^28 arcs) with Zipf weights (s = 1.2, at most 2^24), 10^7 PPR
walkers (stop probability 0.15, lengths uniform in [5, 200]) replayed from a
captured HIP graph with the autotuned schedule, an update batch of
log-uniform size in between.  Strided blocks of walkers are checked bit for
bit against the oracle (their own Philox streams) before and after the
update, and the whole store's slot counts, masses and bucket sizes against
numpy."""
import ctypes as C

import numpy as np
import pytest

import bingo_amd as B

pytestmark = pytest.mark.gpu

SCALE, EDGES, NW, MAXLEN, TP = 23, 1 << 27, 10_000_000, 200, 0.15


def _rows(g, d_paths, blocks, per):
    """Rows [b, b + per) of the device path matrix for every b in blocks."""
    out = np.empty((len(blocks) * per, MAXLEN), np.int32)
    for i, b in enumerate(blocks):
        dst = out[i * per:(i + 1) * per]
        B._check(B.lib().bingo_memcpy_d2h(g._h, dst.ctypes.data, C.c_void_p(d_paths.ptr.value + int(b) * MAXLEN * 4),
                                          dst.nbytes))
    return out


def _check_sample(g, o, orc, d_paths, lens, rng, what):
    per = 64
    blocks = np.sort(rng.choice(NW // per, 400, replace=False)) * per
    ids = (blocks[:, None] + np.arange(per)[None, :]).ravel()
    got = _rows(g, d_paths, blocks, per)
    want, _ = o.walk_ids(B.PPR, orc.walk_params(seed=77, tp=TP, max_len=MAXLEN), ids,
                         starts=(ids % g.n + 1).astype(np.int32), lens=lens[ids], nthreads=16)
    bad = np.flatnonzero((got != want).any(axis=1))
    assert not len(bad), f"{what}: walkers {ids[bad[:5]].tolist()} differ"


def test_batched_workload_full_scale(orc):
    n = 1 << SCALE
    print("[c3] generating", flush=True)
    base = B.gen_rmat(SCALE, EDGES, B.W_ZIPF24, seed=2026)
    g = B.Graph(neighbor_index=False)
    d_base = B.DeviceArray(g, base)
    g.load_device(n, d_base, len(base))
    d_base.free()
    rng = np.random.default_rng(3)
    lens = rng.integers(5, MAXLEN + 1, size=NW, dtype=np.int32)
    d_lens = B.DeviceArray(g, lens)
    d_paths = B.DeviceArray(g, nbytes=NW * MAXLEN * 4)
    d_steps = B.DeviceArray(g, np.zeros(1, np.uint64))
    wp = B.walk_params(seed=77, tp=TP, max_len=MAXLEN)
    print("[c3] tuning and capturing", flush=True)
    chosen = g.tune(B.PPR, wp, sample=1 << 20)
    wgr = g.walk_graph(B.PPR, wp, NW, d_paths, d_steps, d_lens=d_lens)
    wgr.launch()
    g.sync()
    s1 = int(d_steps.to_host(np.uint64, (1,))[0])
    # at most (1 - 0.85^(L-1)) / 0.15 < 6.67 steps per walker on average
    # (fewer from start nodes without arcs); the sampled rows are exact below
    assert 2 * NW < s1 < 6.67 * NW, s1

    print("[c3] oracle load and sample", flush=True)
    o = orc.OracleGraph(n, False)
    assert o.load(base) == 0
    _check_sample(g, o, orc, d_paths, lens, rng, f"first replay ({chosen})")

    print("[c3] update batch and replay", flush=True)
    k = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e6)))))
    ug = B.UpdateGen(n, SCALE, base, EDGES, B.W_ZIPF24, seed=2026)
    ins, dels = ug.inserts(k), ug.deletes(k)
    g.insert(ins)
    g.delete(dels)
    assert o.insert(ins) == 0 and o.delete(dels) == 0
    wgr.launch()
    g.sync()
    _check_sample(g, o, orc, d_paths, lens, rng, "replay after the update")
    wgr.close()

    print("[c3] whole-store properties", flush=True)
    d = g.export_nodes()
    u = np.concatenate([base["u"], ins["u"]]).astype(np.int64)
    w = np.concatenate([base["w"], ins["w"]]).astype(np.int64)
    assert np.array_equal(d["nslots"].astype(np.int64), np.bincount(u, minlength=n + 1))
    # deleted arcs: slot ordinals in arrival order per node
    order = np.argsort(u, kind="stable")
    first = np.searchsorted(u[order], u[order], side="left")
    slot = np.empty_like(u)
    slot[order] = np.arange(len(u)) - first
    key = (u << 32) | slot
    dkey = (dels["u"].astype(np.int64) << 32) | dels["slot"].astype(np.int64)
    srt = np.argsort(key)
    live = np.ones(len(u), bool)
    live[srt[np.searchsorted(key[srt], dkey)]] = False
    lu, lw = u[live], w[live]
    assert np.array_equal(d["mass"], np.bincount(lu, weights=lw.astype(np.float64), minlength=n + 1).astype(np.uint64))
    for j in range(25):  # Zipf weights are at most 2^24
        cnt = np.bincount(lu[(lw >> j) & 1 == 1], minlength=n + 1)
        assert np.array_equal(d["bsize"][:, j], cnt.astype(np.uint32)), j
    assert not d["bsize"][:, 25:].any()
    o.close()
    g.close()

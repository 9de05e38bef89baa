"""Scale workload checks: R-MAT 2^25 nodes / 2^30 arcs (log-uniform u32
weights) loaded into a store fitted to the graph, one sliding-window step
(10^7 inserts, then the 10^7 oldest arcs expired), then size-independent
properties of the whole store against numpy, alias tables of sampled nodes
against the oracle's construction, walk steps out of probe nodes checked
against their live arcs, and the hub's transition law by chi-square; then a
compaction of every node with a tombstone (each one's oldest slots: a dead
prefix), after which slot counts equal live counts and masses, buckets and
the hub's draws are unchanged.

Needs about 215 GB of HBM, 60 GB of host memory and about 90 s."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

import bingo_amd as B

pytestmark = pytest.mark.gpu


def _per_node_tables(g):
    """nslots, mass, bucket sizes, T and alias of every node (no slot arrays)."""
    n1 = g.n + 1
    out = dict(nslots=np.zeros(n1, np.uint32), mass=np.zeros(n1, np.uint64), bsize=np.zeros((n1, 32), np.uint32),
               T=np.zeros((n1, 32), np.uint64), alias=np.zeros((n1, 32), np.uint8))
    d = B.Dump()
    for k, a in out.items():
        setattr(d, k, a.ctypes.data)
    B._check(B.lib().bingo_export(g._h, C.byref(d)))
    out["total_members"] = d.total_members
    return out


def test_scale_store_window_and_walks(orc):
    scale, m, k = 25, 1 << 30, 10_000_000
    n = 1 << scale
    print("[c5] generating", flush=True)
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=2026, undirected=False)
    g = B.Graph(neighbor_index=False, fit_load=True, pool_slack=0.03, grow_in_place=True)
    d_base = B.DeviceArray(g, base)
    g.load_device(n, d_base, m)
    d_base.free()
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=2026)
    ins, dels = ug.inserts(k), ug.expire(k)
    del ug
    print("[c5] window step", flush=True)
    g.insert(ins)
    g.delete(dels)
    # the expired arcs are exactly the first k of the load
    assert (dels["u"] == base["u"][:k]).all()
    st = _per_node_tables(g)

    print("[c5] per-node slot counts and masses", flush=True)
    u_all = np.concatenate([base["u"], ins["u"]])
    assert np.array_equal(st["nslots"], np.bincount(u_all, minlength=n + 1).astype(np.uint32))
    del u_all
    lu = np.concatenate([base["u"][k:], ins["u"]])
    lv = np.concatenate([base["v"][k:], ins["v"]])
    lw = np.concatenate([base["w"][k:], ins["w"]])
    del base
    lo = np.bincount(lu, weights=(lw & 0xFFFF).astype(np.float64), minlength=n + 1).astype(np.uint64)
    hi = np.bincount(lu, weights=(lw >> 16).astype(np.float64), minlength=n + 1).astype(np.uint64)
    assert np.array_equal(st["mass"], (hi << np.uint64(16)) + lo)
    pc = np.bitwise_count(lw)
    assert st["total_members"] == int(pc.sum(dtype=np.int64))
    members = np.bincount(lu, weights=pc.astype(np.float64), minlength=n + 1).astype(np.int64)
    bs = st["bsize"].astype(np.int64)
    assert np.array_equal(bs.sum(axis=1), members)
    weighted = np.zeros(n + 1, np.uint64)
    for j in range(32):
        weighted += bs[:, j].astype(np.uint64) << np.uint64(j)
    assert np.array_equal(weighted, st["mass"])
    for j in (0, 9, 31):  # three buckets exactly
        sel = (lw >> np.uint32(j)) & np.uint32(1) == 1
        assert np.array_equal(bs[:, j], np.bincount(lu[sel], minlength=n + 1)), j

    print("[c5] alias tables of sampled nodes", flush=True)
    rng = np.random.default_rng(5)
    sample = np.concatenate([[1], rng.integers(1, n + 1, 3000)])
    for u in sample:
        mass_j = bs[u].astype(np.uint64) << np.arange(32, dtype=np.uint64)
        if not mass_j.any():
            continue
        T, a = orc.alias_build(mass_j)
        assert np.array_equal(st["T"][u], T) and np.array_equal(st["alias"][u], a), u

    print("[c5] walks out of probe nodes follow live arcs", flush=True)
    nw = 1 << 20
    starts = (np.arange(nw, dtype=np.int64) * (n // nw) % n + 1).astype(np.int32)
    paths, steps = g.walk(B.DEEPWALK, B.walk_params(seed=9, max_len=80), starts=starts)
    assert steps > nw
    src, dst = paths[:, :-1].ravel(), paths[:, 1:].ravel()
    ok = dst > 0
    src, dst = src[ok], dst[ok]
    probes = np.unique(rng.choice(src, 2000))
    in_probe = np.isin(lu, probes, kind="table")
    arc_keys = np.unique((lu[in_probe].astype(np.int64) << 32) | lv[in_probe].astype(np.int64))
    step_sel = np.isin(src, probes, kind="table")
    step_keys = (src[step_sel].astype(np.int64) << 32) | dst[step_sel].astype(np.int64)
    assert step_sel.sum() > 10_000
    assert np.isin(step_keys, arc_keys).all()

    print("[c5] hub chi-square", flush=True)
    hub = int(np.argmax(st["nslots"]))
    N = 2_000_000
    p2, _ = g.walk(B.SAMPLING, B.walk_params(seed=11, max_len=2), starts=np.full(N, hub, np.int32))
    hv = lu == hub
    tgt = lv[hv].astype(np.int64)
    bins = 64
    exp = np.bincount((tgt * 2654435761) % bins, weights=lw[hv].astype(np.float64), minlength=bins)
    got = np.bincount((p2[:, 1].astype(np.int64) * 2654435761) % bins, minlength=bins)
    assert got.sum() == N
    assert stats.chisquare(got, exp / exp.sum() * N).pvalue > 0.01

    print("[c5] compaction", flush=True)
    nodes, freed = g.compact(0.0)
    assert freed == k and nodes == len(np.unique(dels["u"]))
    st2 = _per_node_tables(g)
    assert np.array_equal(st2["nslots"], np.bincount(lu, minlength=n + 1).astype(np.uint32))
    for f in ("mass", "bsize", "T", "alias"):
        assert np.array_equal(st2[f], st[f]), f
    # every node's dead slots are its oldest, a prefix: compaction drops them
    # without touching member lists, so the hub's draws are the same
    p3, _ = g.walk(B.SAMPLING, B.walk_params(seed=11, max_len=2), starts=np.full(N, hub, np.int32))
    assert np.array_equal(p3, p2)

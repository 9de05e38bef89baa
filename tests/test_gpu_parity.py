"""GPU parity: the HIP store, its batched updates and its walks against the
CPU oracle, bit for bit, on the same seeded inputs (calls go through the C
ABI in include/bingo.h)."""
import random

import numpy as np
import pytest

import bingo_amd as B
from tests.brute import BruteGraph, check_against_brute

pytestmark = pytest.mark.gpu

APPS = [B.DEEPWALK, B.NODE2VEC, B.PPR, B.SAMPLING]


def _pair(orc, n, edges, nbr=True, fit=False):
    print(f"[parity] load n={n} m={len(edges)} nbr={nbr} fit={fit}", flush=True)
    g = B.Graph(neighbor_index=nbr, fit_load=fit)
    g.load(n, edges)
    o = orc.OracleGraph(n, nbr)
    assert o.load(edges) == 0
    return g, o


def _same_state(g, o, what=""):
    a, b = g.export(), o.export()
    for k in b:
        assert a[k].shape == b[k].shape, (what, k, a[k].shape, b[k].shape)
        if not np.array_equal(a[k], b[k]):
            bad = np.argwhere(a[k] != b[k])[:5]
            raise AssertionError(f"{what}: {k} differs at {bad.tolist()}")


def _params(orc, seed, max_len, p=0.5, q=2.0, tp=0.0125, max_attempts=4096):
    return (B.walk_params(seed=seed, p=p, q=q, tp=tp, max_len=max_len, max_attempts=max_attempts),
            orc.walk_params(seed=seed, p=p, q=q, tp=tp, max_len=max_len, max_attempts=max_attempts))


def _same_walks(g, o, orc, app, seed, max_len, starts=None, lens=None, nwalkers=None, **kw):
    gp, op = _params(orc, seed, max_len, **kw)
    pa, sa = g.walk(app, gp, starts=starts, lens=lens, nwalkers=nwalkers)
    pb, sb = o.walk(app, op, starts=starts, lens=lens, nwalkers=nwalkers)
    if not np.array_equal(pa, pb):
        bad = np.argwhere((pa != pb).any(axis=1))[:5, 0]
        raise AssertionError(f"app {app} seed {seed}: walkers {bad.tolist()} differ\n{pa[bad[0]]}\n{pb[bad[0]]}")
    assert sa == sb
    return pa, sa


def _rand_edges(rng, n, k, hub=None, wbits=32):
    e = np.zeros(k, B.EDGE_DT)
    e["u"] = [hub if (hub and rng.random() < 0.4) else rng.randint(1, n) for _ in range(k)]
    e["v"] = [rng.randint(1, n) for _ in range(k)]
    e["w"] = [max(1, rng.getrandbits(rng.randint(1, wbits))) for _ in range(k)]
    return e


def test_load_and_walks_small(orc):
    rng = random.Random(1)
    n = 300
    e = _rand_edges(rng, n, 4000, hub=7)
    g, o = _pair(orc, n, e)
    _same_state(g, o, "load")
    bg = BruteGraph(n)
    bg.insert(e)
    check_against_brute(g.export(), bg)
    starts = np.array([rng.randint(1, n) for _ in range(3000)], np.int32)
    lens = np.array([rng.randint(1, 40) for _ in range(3000)], np.int32)
    for app in APPS:
        _same_walks(g, o, orc, app, seed=11 + app, max_len=40, starts=starts, lens=lens)
        _same_walks(g, o, orc, app, seed=5, max_len=17, nwalkers=n)


@pytest.mark.parametrize("nbr,fit", [(True, False), (False, False), (True, True)])
def test_update_rounds_match_oracle(orc, nbr, fit):
    scale = 10
    n = 1 << scale
    base = B.gen_rmat(scale, 6000, B.W_LOGUNI32, seed=3, undirected=False)
    g, o = _pair(orc, n, base, nbr, fit)
    ug = B.UpdateGen(n, scale, base, 6000, B.W_LOGUNI32, seed=3)
    bg = BruteGraph(n)
    bg.insert(base)
    for rnd, (ki, kd) in enumerate([(1, 1), (7, 3), (1000, 900), (5000, 5000), (0, 2000), (3000, 0)]):
        print(f"[parity] round {rnd} inserts={ki} deletes={kd}", flush=True)
        ins, dels = ug.inserts(ki), ug.deletes(kd)
        g.insert(ins)
        assert o.insert(ins) == 0
        g.delete(dels)
        assert o.delete(dels) == 0
        bg.insert(ins)
        bg.delete(dels)
        _same_state(g, o, f"round {rnd}")
        app = [B.DEEPWALK, B.PPR, B.SAMPLING][rnd % 3] if not nbr else APPS[rnd % 4]
        _same_walks(g, o, orc, app, seed=100 + rnd, max_len=30, nwalkers=2 * n)
    check_against_brute(g.export(), bg)


def test_pools_grown_in_place_match_oracle(orc):
    """Pools that grow by mapping more HBM behind a reserved address range
    (grow_in_place, no pool slack: every batch grows some pool) keep the
    store equal to the oracle's, and a captured walk graph replays across
    the growth without being captured again (the pool addresses stay)."""
    rng = random.Random(12)
    n = 800
    e = _rand_edges(rng, n, 3000, hub=11)
    print(f"[parity] load n={n} m={len(e)} grow_in_place", flush=True)
    g = B.Graph(neighbor_index=True, pool_slack=0.0, grow_in_place=True)
    g.load(n, e)
    o = orc.OracleGraph(n, True)
    assert o.load(e) == 0
    _same_state(g, o, "load")
    nw, L = 1500, 16
    dp = B.DeviceArray(g, nbytes=nw * L * 4)
    ds = B.DeviceArray(g, np.zeros(1, np.uint64))
    wg = g.walk_graph(B.NODE2VEC, B.walk_params(seed=8, max_len=L), nw, dp, ds)
    for rnd in range(6):
        ins = _rand_edges(rng, n, 2000 * (rnd + 1), hub=11)
        g.insert(ins)
        assert o.insert(ins) == 0
        d = o.export()
        so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
        dels = []
        for u in rng.sample(range(1, n + 1), 60):
            live = np.flatnonzero(d["slot_w"][so[u]:so[u] + int(d["nslots"][u])] != 0)
            dels += [(u, int(x)) for x in live[: rng.randint(0, 3)]]
        dels = np.array(dels, B.DEL_DT)
        g.delete(dels)
        assert o.delete(dels) == 0
        _same_state(g, o, f"round {rnd}")
        wg.launch()
        g.sync()
        want, _ = o.walk(B.NODE2VEC, orc.walk_params(seed=8, max_len=L), nwalkers=nw)
        assert np.array_equal(dp.to_host(np.int32, (nw, L)), want), rnd
    wg.close()
    for a in (dp, ds):
        a.free()
    g.close()


def test_fitted_store_under_sliding_window(orc):
    """A store fitted to the graph at load (regions sized to their contents
    plus 1/16) holds the same state as the oracle, is smaller than the
    power-of-two store of the same graph, and stays exact while a sliding
    window inserts new arcs and expires the oldest ones (regions outgrow
    their headroom and move)."""
    scale, m = 12, 40000
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=8, undirected=False)
    g, o = _pair(orc, n, base, nbr=False, fit=True)
    _same_state(g, o, "fitted load")
    with B.Graph(neighbor_index=False) as gp:
        gp.load(n, base)
        assert g.stats()["bytes_reserved"] < gp.stats()["bytes_reserved"]
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=8)
    for rnd, k in enumerate([100, 5000, 20000, 20000]):
        print(f"[parity] window round {rnd} k={k}", flush=True)
        ins, dels = ug.inserts(k), ug.expire(k)
        g.insert(ins)
        assert o.insert(ins) == 0
        g.delete(dels)
        assert o.delete(dels) == 0
        _same_state(g, o, f"window round {rnd}")
        _same_walks(g, o, orc, B.DEEPWALK, seed=40 + rnd, max_len=25, nwalkers=n)


def test_many_long_delete_runs(orc):
    """Twenty nodes with hundreds of deletes each in one batch (the per-group
    tasks of long runs run side by side), among short runs of other nodes,
    with and without neighbour tables."""
    scale = 12
    n = 1 << scale
    base = B.gen_rmat(scale, 150_000, B.W_LOGUNI32, seed=21, undirected=False)
    for nbr in (True, False):
        g, o = _pair(orc, n, base, nbr)
        d = o.export()
        so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
        rng = np.random.default_rng(4)
        hubs = np.argsort(d["nslots"])[::-1][:20]
        dels = []
        for u in range(1, n + 1):
            ns = int(d["nslots"][u])
            if not ns:
                continue
            frac = 0.35 if u in hubs else 0.05
            pick = np.flatnonzero(rng.random(ns) < frac)
            dels += [(u, int(s)) for s in pick]
        dels = np.array(dels, B.DEL_DT)
        dels = dels[rng.permutation(len(dels))]
        counts = np.bincount(dels["u"], minlength=n + 1)
        assert (counts > 128).sum() >= 10 and ((counts > 0) & (counts <= 8)).sum() > 100
        g.delete(dels)
        assert o.delete(dels) == 0
        _same_state(g, o, f"long runs nbr={nbr}")
        app = B.NODE2VEC if nbr else B.DEEPWALK
        _same_walks(g, o, orc, app, seed=77, max_len=20, nwalkers=2 * n)
        g.close()
        o.close()


def test_hub_node_long_runs(orc):
    """One node with tens of thousands of arcs: multi-chunk ballots, bucket
    growth and neighbour-table rehash inside one batch, thousands of deletes
    of one node in one batch."""
    rng = random.Random(7)
    n = 5000
    e = _rand_edges(rng, n, 20000, hub=1)
    g, o = _pair(orc, n, e)
    for rnd in range(3):
        ins = _rand_edges(rng, n, 30000, hub=1)
        g.insert(ins)
        assert o.insert(ins) == 0
        d = o.export()
        so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
        hub_live = [s for s in range(int(d["nslots"][1])) if d["slot_w"][so[1] + s]]
        pick = rng.sample(hub_live, len(hub_live) // 3)
        other = []
        for u in rng.sample(range(2, n + 1), 500):
            lw = [s for s in range(int(d["nslots"][u])) if d["slot_w"][so[u] + s]]
            if lw:
                other.append((u, rng.choice(lw)))
        dels = np.array([(1, s) for s in pick] + other, B.DEL_DT)
        dels = dels[np.random.RandomState(rnd).permutation(len(dels))]  # batch order is irrelevant
        g.delete(dels)
        assert o.delete(dels) == 0
        _same_state(g, o, f"hub round {rnd}")
        starts = np.ones(4096, np.int32)
        for app in APPS:
            _same_walks(g, o, orc, app, seed=rnd * 10 + app, max_len=12, starts=starts)


def test_config1_plain_small(orc):
    """The plain small workload: undirected R-MAT (0.57, 0.19, 0.19, 0.05),
    2^17 nodes, 2^21 edges, u32 weights uniform in [1, 2^16); DeepWalk one
    walker per node, length 80; then 10^4 inserts + 10^4 deletes."""
    scale, m = 17, 1 << 21
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_UNIFORM16, seed=2026)
    g, o = _pair(orc, n, base, nbr=True)
    _same_state(g, o, "C1 load")
    pa, steps = _same_walks(g, o, orc, B.DEEPWALK, seed=1, max_len=80, nwalkers=n)
    assert steps > n * 40
    ug = B.UpdateGen(n, scale, base, m, B.W_UNIFORM16, seed=2026)
    ins, dels = ug.inserts(10_000), ug.deletes(10_000)
    g.insert(ins)
    o.insert(ins)
    g.delete(dels)
    o.delete(dels)
    _same_state(g, o, "C1 after updates")
    _same_walks(g, o, orc, B.DEEPWALK, seed=2, max_len=80, nwalkers=n)
    _same_walks(g, o, orc, B.NODE2VEC, seed=3, max_len=80, nwalkers=n // 4)


def test_invalid_batches_rejected_without_change(orc):
    rng = random.Random(4)
    n = 50
    e = _rand_edges(rng, n, 400)
    g, o = _pair(orc, n, e)
    bad_ins = [
        np.array([(0, 1, 1)], B.EDGE_DT),
        np.array([(1, n + 1, 1)], B.EDGE_DT),
        np.array([(1, 2, 0)], B.EDGE_DT),
        np.concatenate([_rand_edges(rng, n, 10), np.array([(-3, 2, 5)], B.EDGE_DT)]),
    ]
    for b in bad_ins:
        with pytest.raises(B.BingoError) as ei:
            g.insert(b)
        assert ei.value.code == -1
    d = o.export()
    bad_dels = [
        np.array([(1, int(d["nslots"][1]))], B.DEL_DT),
        np.array([(n + 1, 0)], B.DEL_DT),
        np.array([(2, 0), (2, 0)], B.DEL_DT),
        np.array([(3, -1)], B.DEL_DT),
    ]
    for b in bad_dels:
        with pytest.raises(B.BingoError) as ei:
            g.delete(b)
        assert ei.value.code == -1
    _same_state(g, o, "after rejected batches")
    g.delete(np.array([(2, 0)], B.DEL_DT))
    o.delete(np.array([(2, 0)], B.DEL_DT))
    with pytest.raises(B.BingoError):
        g.delete(np.array([(2, 0)], B.DEL_DT))  # already dead
    _same_state(g, o, "after dead-slot delete")
    with pytest.raises(B.BingoError):
        g.walk(B.DEEPWALK, B.walk_params(max_len=5), starts=np.array([0], np.int32))
    with pytest.raises(B.BingoError):
        g.walk(B.DEEPWALK, B.walk_params(max_len=5), starts=np.array([1], np.int32), lens=np.array([6], np.int32))


def test_edge_count_matches_oracle(orc):
    rng = random.Random(8)
    n = 80
    e = _rand_edges(rng, n, 3000)
    g, o = _pair(orc, n, e)
    dels = np.array(list({(int(x["u"]), rng.randint(0, 5)) for x in e[:200]}), B.DEL_DT)
    d = o.export()
    dels = dels[dels["slot"] < d["nslots"][dels["u"]]]
    g.delete(dels)
    o.delete(dels)
    us = np.array([rng.randint(0, n + 1) for _ in range(5000)], np.int32)
    vs = np.array([rng.randint(0, n + 1) for _ in range(5000)], np.int32)
    got = g.edge_count(us, vs)
    want = [o.edge_count(int(u), int(v)) for u, v in zip(us, vs)]
    assert got.tolist() == want


def test_device_entry_points(orc):
    rng = random.Random(9)
    n = 500
    e = _rand_edges(rng, n, 8000)
    o = orc.OracleGraph(n)
    o.load(e)
    g = B.Graph()
    de = B.DeviceArray(g, e)
    g.load_device(n, de, len(e))
    ins = _rand_edges(rng, n, 3000)
    di = B.DeviceArray(g, ins)
    g.insert_device(di, len(ins))
    o.insert(ins)
    nw, L = 2000, 25
    dp = B.DeviceArray(g, nbytes=nw * L * 4)
    ds = B.DeviceArray(g, np.zeros(1, np.uint64))
    gp, op = _params(orc, 77, L)
    g.walk_device(B.NODE2VEC, gp, nw, dp, ds)
    g.sync()
    pa = dp.to_host(np.int32, (nw, L))
    sa = int(ds.to_host(np.uint64, (1,))[0])
    pb, sb = o.walk(B.NODE2VEC, op, nwalkers=nw)
    assert np.array_equal(pa, pb) and sa == sb
    st = g.stats()
    assert st["arcs"] == len(e) + len(ins) and st["walk_ms"] > 0
    for a in (de, di, dp, ds):
        a.free()


def test_gpu_transition_law_chi_square(orc):
    """Empirical one-step frequencies of 1.2e7 GPU draws (three walk seeds of
    4e6 draws, pooled) from a 40-arc node with log-uniform weights against
    w / sum(w), bins with more than 20 expected draws (a seeded sample: the
    test's false-failure rate is the 1% of its threshold)."""
    from scipy import stats

    rng = random.Random(12)
    w = [max(1, rng.getrandbits(rng.randint(1, 32))) for _ in range(40)]
    n = 41
    e = np.array([(1, i + 2, x) for i, x in enumerate(w)], B.EDGE_DT)
    g = B.Graph()
    g.load(n, e)
    N = 4_000_000
    cnt = np.zeros(40)
    for seed in (1, 2, 3):
        paths, steps = g.walk(B.SAMPLING, B.walk_params(seed=seed, max_len=2), starts=np.ones(N, np.int32))
        assert steps == N
        cnt += np.bincount(paths[:, 1], minlength=n + 1)[2:]
    exp = np.array(w, float) / sum(w)
    keep = exp * cnt.sum() > 20
    assert stats.chisquare(cnt[keep], exp[keep] / exp[keep].sum() * cnt[keep].sum()).pvalue > 0.01


@pytest.mark.parametrize("knobs", [
    dict(form=B.FORM_PLAIN, block=64),
    dict(form=B.FORM_PLAIN, block=512),
    dict(form=B.FORM_PERSISTENT, block=256, walkers_per_lane=1, lds_depth=0),
    dict(form=B.FORM_PERSISTENT, block=128, walkers_per_lane=2, lds_depth=4),
    dict(form=B.FORM_PERSISTENT, block=512, walkers_per_lane=4, lds_depth=16),
    dict(form=B.FORM_PERSISTENT, block=64, walkers_per_lane=4, lds_depth=0, blocks_per_cu=1),
])
def test_walk_forms_give_identical_paths(orc, knobs):
    """Every scheduling form and knob setting yields the oracle's paths:
    uneven lengths, PPR stops and dead ends included."""
    rng = random.Random(21)
    scale = 11
    n = 1 << scale
    base = B.gen_rmat(scale, 9000, B.W_ZIPF24, seed=21, undirected=False)  # directed: dead ends
    g, o = _pair(orc, n, base)
    nw = 20000
    starts = np.array([rng.randint(1, n) for _ in range(nw)], np.int32)
    lens = np.array([rng.randint(1, 48) for _ in range(nw)], np.int32)
    for app in APPS:
        gp = B.walk_params(seed=31 + app, tp=0.15, max_len=48, **knobs)
        op = orc.walk_params(seed=31 + app, tp=0.15, max_len=48)
        pa, sa = g.walk(app, gp, starts=starts, lens=lens)
        pb, sb = o.walk(app, op, starts=starts, lens=lens)
        assert np.array_equal(pa, pb), (app, knobs)
        assert sa == sb
        pa, sa = g.walk(app, gp, nwalkers=n)  # default starts, max_len everywhere
        pb, sb = o.walk(app, op, nwalkers=n)
        assert np.array_equal(pa, pb) and sa == sb, (app, knobs)


def test_tuner_picks_a_valid_schedule(orc):
    scale = 12
    n = 1 << scale
    base = B.gen_rmat(scale, 30000, B.W_LOGUNI32, seed=5)
    g, o = _pair(orc, n, base)
    wp = B.walk_params(seed=3, max_len=40)
    chosen = g.tune(B.NODE2VEC, wp, sample=n)
    assert chosen["form"] in (B.FORM_PLAIN, B.FORM_PERSISTENT)
    pa, sa = g.walk(B.NODE2VEC, wp, nwalkers=n)  # FORM_AUTO now uses the tuned knobs
    pb, sb = o.walk(B.NODE2VEC, orc.walk_params(seed=3, max_len=40), nwalkers=n)
    assert np.array_equal(pa, pb) and sa == sb
    with pytest.raises(B.BingoError):
        g.walk(B.DEEPWALK, B.walk_params(max_len=40, form=B.FORM_PERSISTENT, lds_depth=16), nwalkers=4)


@pytest.mark.parametrize("form", [B.FORM_PLAIN, B.FORM_PERSISTENT])
def test_walk_graph_replays_match_eager(orc, form):
    """A captured walk replayed several times, with eager walks, syncs and an
    update batch that grows the pools (forcing a re-capture) in between:
    every replay equals the eager walk and the oracle."""
    scale = 11
    n = 1 << scale
    base = B.gen_rmat(scale, 6000, B.W_LOGUNI32, seed=2)
    g, o = _pair(orc, n, base)
    nw, L = 5000, 30
    dp = B.DeviceArray(g, nbytes=nw * L * 4)
    ds = B.DeviceArray(g, np.zeros(1, np.uint64))
    wp = B.walk_params(seed=4, max_len=L, form=form, block=256, walkers_per_lane=2, lds_depth=0)
    op = orc.walk_params(seed=4, max_len=L)
    wgr = g.walk_graph(B.NODE2VEC, wp, nw, dp, ds)
    ug = B.UpdateGen(n, scale, base, 6000, B.W_LOGUNI32, seed=2)
    total = 0
    for rep in range(4):
        if rep == 2:
            ins = ug.inserts(40000)  # several times the loaded arcs: pools must grow
            g.insert(ins)
            assert o.insert(ins) == 0
        print(f"[parity] walk graph replay {rep}", flush=True)
        wgr.launch()
        g.sync()
        pa = dp.to_host(np.int32, (nw, L))
        eager, se = g.walk(B.NODE2VEC, wp, nwalkers=nw)
        pb, sb = o.walk(B.NODE2VEC, op, nwalkers=nw)
        assert np.array_equal(pa, eager) and np.array_equal(pa, pb), rep
        total += sb
    assert int(ds.to_host(np.uint64, (1,))[0]) == total
    wgr.close()


@pytest.mark.parametrize("P,Q", [(0.5, 2.0), (0.01, 1.0), (0.9, 1.0), (0.1, 0.25), (0.25, 4.0), (2.0, 0.5),
                                 (1e-12, 1.0)])
def test_node2vec_parameters_every_form(orc, P, Q):
    """node2vec across the return/in-out parameters -- folded return regions
    large (p = 0.01), small (p = 0.9) and beside an envelope above 1 (q < 1),
    the plain envelope (p >= q) and a fold factor too large to fold
    (p = 1e-12) -- on a graph of multi-arcs and symmetric pairs (returns are
    common and carry several arcs' weight), in the plain, persistent and
    served forms, against the oracle."""
    rng = random.Random(int(P * 1000) + int(Q * 10))
    n = 500
    e = _rand_edges(rng, n, 6000, hub=3)
    back = e.copy()
    back["u"], back["v"] = e["v"], e["u"]
    back["w"] = [max(1, rng.getrandbits(rng.randint(1, 6))) for _ in range(len(e))]
    e = np.concatenate([e, back, e[:500]])  # reverse arcs and repeated arcs
    g, o = _pair(orc, n, e)
    starts = np.array([rng.randint(1, n) for _ in range(4000)], np.int32)
    op = orc.walk_params(seed=21, p=P, q=Q, max_len=32)
    want, steps = o.walk(B.NODE2VEC, op, starts=starts)
    for form, knobs in ((B.FORM_PLAIN, {}), (B.FORM_PERSISTENT, dict(block=128)),
                        (B.FORM_PERSISTENT, dict(block=256, walkers_per_lane=4, lds_depth=16))):
        gp = B.walk_params(seed=21, p=P, q=Q, max_len=32, form=form)
        for k, v in knobs.items():
            setattr(gp, k, v)
        pa, sa = g.walk(B.NODE2VEC, gp, starts=starts)
        assert np.array_equal(pa, want) and sa == steps, (P, Q, form, knobs)
    srv = B.Server(g, arena_bytes=1 << 24)
    t = srv.submit_walk(B.NODE2VEC, B.walk_params(seed=21, p=P, q=Q, max_len=32), starts=starts)
    out, st, _ = srv.wait(t)
    srv.close()
    assert np.array_equal(out, want) and st == steps, (P, Q, "served")
    g.close()
    o.close()


def test_gather_probe_reports_rates():
    g = B.Graph(neighbor_index=False)
    g.load(1 << 12, B.gen_rmat(12, 50_000, B.W_LOGUNI32, seed=1))
    for width in (8, 32):
        n, ms = g.gather_probe(width)
        assert n > 0 and ms > 0
    before = g.export()
    n, ms = g.gather_probe(0)  # read-modify-writes leave the store as it was
    assert n > 0 and ms > 0
    after = g.export()
    assert before.keys() == after.keys()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    with pytest.raises(B.BingoError):
        g.gather_probe(16)
    g.close()

"""GPU parity with float weights (the reference's wtype::Float): the HIP
store's state -- integer parts through the bit groups, per-slot fractions,
fraction masses, fractional-part thresholds -- and every walk form against
the CPU oracle, bit for bit, through load, insert batches and delete batches
of every run length (short, staged, per-group long runs)."""
import numpy as np
import pytest

import bingo_amd as B
from tests.floatw import fbits, float_edges, live_slots

pytestmark = pytest.mark.gpu

APPS = [B.DEEPWALK, B.NODE2VEC, B.PPR, B.SAMPLING]


def _pair(orc, n, edges, nbr=True, fit=False):
    print(f"[float] load n={n} m={len(edges)} nbr={nbr} fit={fit}", flush=True)
    g = B.Graph(neighbor_index=nbr, fit_load=fit, float_weights=True)
    g.load(n, edges)
    o = orc.OracleGraph(n, nbr, float_weights=True)
    assert o.load(edges) == 0
    return g, o


def _same_state(g, o, what=""):
    a, b = g.export(), o.export()
    assert {"slot_f", "fmass", "fthr"} <= set(a) and {"slot_f", "fmass", "fthr"} <= set(b)
    for k in b:
        assert a[k].shape == b[k].shape, (what, k, a[k].shape, b[k].shape)
        if not np.array_equal(a[k], b[k]):
            bad = np.argwhere(a[k] != b[k])[:5]
            raise AssertionError(f"{what}: {k} differs at {bad.tolist()}")


def _same_walks(g, o, orc, app, seed, max_len, starts=None, nwalkers=None, form=B.FORM_AUTO, **kw):
    gp = B.walk_params(seed=seed, max_len=max_len, form=form, **kw)
    op = orc.walk_params(seed=seed, max_len=max_len, float_weights=True, **kw)
    pa, sa = g.walk(app, gp, starts=starts, nwalkers=nwalkers)
    pb, sb = o.walk(app, op, starts=starts, nwalkers=nwalkers)
    if not np.array_equal(pa, pb):
        bad = np.argwhere((pa != pb).any(axis=1))[:5, 0]
        raise AssertionError(f"app {app} seed {seed}: walkers {bad.tolist()} differ\n{pa[bad[0]]}\n{pb[bad[0]]}")
    assert sa == sb
    return pa, sa


def _deletes(d, n, rng, hub, hub_share, other_share):
    """Live slots to delete: a share of the hub's (a long run), a share of
    every other node's (short and staged runs), in shuffled batch order."""
    so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
    dels = []
    for u in range(1, n + 1):
        live = live_slots(d, u, so)
        if not len(live):
            continue
        share = hub_share if u == hub else other_share
        pick = live[rng.random(len(live)) < share]
        dels += [(u, int(s)) for s in pick]
    dels = np.array(dels, B.DEL_DT)
    return dels[rng.permutation(len(dels))]


@pytest.mark.parametrize("nbr,fit", [(True, False), (False, True)])
def test_float_updates_and_walks_match_oracle(orc, nbr, fit):
    rng = np.random.default_rng(31 + nbr)
    n = 400
    hub = 9
    g, o = _pair(orc, n, float_edges(rng, n, 12000, hub=hub), nbr, fit)
    _same_state(g, o, "load")
    for rnd, (ki, hs, os_) in enumerate([(50, 0.02, 0.01), (3000, 0.4, 0.1), (6000, 0.05, 0.3), (0, 0.5, 0.05)]):
        print(f"[float] round {rnd} inserts={ki}", flush=True)
        if ki:
            ins = float_edges(rng, n, ki, hub=hub)
            g.insert(ins)
            assert o.insert(ins) == 0
        dels = _deletes(o.export(), n, rng, hub, hs, os_)
        counts = np.bincount(dels["u"], minlength=n + 1)
        print(f"[float] round {rnd} deletes={len(dels)} hub_run={counts[hub]}", flush=True)
        g.delete(dels)
        assert o.delete(dels) == 0
        _same_state(g, o, f"round {rnd}")
        apps = APPS if nbr else [B.DEEPWALK, B.PPR, B.SAMPLING]
        for app in apps:
            _same_walks(g, o, orc, app, seed=200 + 10 * rnd + app, max_len=24, nwalkers=3 * n)
        starts = np.full(2048, hub, np.int32)
        _same_walks(g, o, orc, apps[rnd % len(apps)], seed=7 + rnd, max_len=8, starts=starts, form=B.FORM_PLAIN)
    g.close()
    o.close()


def test_float_long_delete_runs(orc):
    """Hundreds of deletes on each of several hubs in one batch: the per-group
    tasks subtract the deleted fractions in their last finisher."""
    rng = np.random.default_rng(44)
    n = 2000
    hubs = [3, 17, 400, 1999]
    e = np.concatenate([float_edges(rng, n, 3000, hub=h, hub_frac=0.6) for h in hubs])
    g, o = _pair(orc, n, e)
    d = o.export()
    so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
    dels = []
    for u in range(1, n + 1):
        live = live_slots(d, u, so)
        share = 0.45 if u in hubs else 0.1
        dels += [(u, int(s)) for s in live[rng.random(len(live)) < share]]
    dels = np.array(dels, B.DEL_DT)[rng.permutation(len(dels))]
    counts = np.bincount(dels["u"], minlength=n + 1)
    assert all(counts[h] > 128 for h in hubs)
    g.delete(dels)
    assert o.delete(dels) == 0
    _same_state(g, o, "long runs")
    for app in APPS:
        _same_walks(g, o, orc, app, seed=90 + app, max_len=16, starts=np.repeat(np.array(hubs, np.int32), 512))


def test_float_fraction_only_nodes(orc):
    """Nodes whose arcs all weigh below 1 have no integer mass: every draw
    goes through the fractional part; deleting their last arc leaves a dead
    end, and a node whose integer part is gone keeps walking on fractions."""
    n = 6
    e = np.zeros(6, B.EDGE_DT)
    e["u"] = [1, 1, 1, 2, 2, 3]
    e["v"] = [2, 3, 4, 1, 5, 6]
    e["w"] = [fbits(x) for x in (0.25, 0.5, 0.125, 3.0, 0.75, 0.5)]
    g, o = _pair(orc, n, e)
    _same_state(g, o, "load")
    d = g.export()
    assert int(d["mass"][1]) == 0 and int(d["fthr"][1]) == (1 << 64) - 1
    starts = np.array([1, 2, 3] * 4000, np.int32)
    for app in APPS:
        _same_walks(g, o, orc, app, seed=5 + app, max_len=10, starts=starts)
    dels = np.array([(2, 0), (3, 0)], B.DEL_DT)  # node 2 keeps only 0.75; node 3 has nothing left
    g.delete(dels)
    assert o.delete(dels) == 0
    _same_state(g, o, "after deletes")
    for app in APPS:
        _same_walks(g, o, orc, app, seed=50 + app, max_len=10, starts=starts)
    with pytest.raises(B.BingoError):
        g.delete(np.array([(3, 0)], B.DEL_DT))  # dead already


def test_float_invalid_weights_rejected(orc):
    rng = np.random.default_rng(2)
    n = 40
    g, o = _pair(orc, n, float_edges(rng, n, 500))
    for x in (0.0, -1.0, float("nan"), float("inf"), 4294967296.0, 1e-20, -0.0):
        b = np.array([(1, 2, fbits(x))], B.EDGE_DT)
        with pytest.raises(B.BingoError) as ei:
            g.insert(b)
        assert ei.value.code == -1, x
    _same_state(g, o, "after rejected batches")
    with pytest.raises(B.BingoError):
        g.walk(B.NODE2VEC, B.walk_params(max_len=5, max_attempts=65537), nwalkers=4)
    _same_walks(g, o, orc, B.NODE2VEC, seed=3, max_len=12, nwalkers=400, max_attempts=65536)


def test_float_walk_forms_and_server_agree(orc):
    """Plain, persistent (every knob set), captured-graph and served walks
    give the oracle's paths on a float-weight graph."""
    rng = np.random.default_rng(8)
    scale = 11
    n = 1 << scale
    e = float_edges(rng, n, 30000, hub=5)
    g, o = _pair(orc, n, e)
    for app in APPS:
        want, steps = _same_walks(g, o, orc, app, seed=61 + app, max_len=32, nwalkers=2 * n, form=B.FORM_PLAIN)
        for wpl, depth in ((1, 0), (2, 4), (4, 16)):
            p = B.walk_params(seed=61 + app, max_len=32, form=B.FORM_PERSISTENT, block=256)
            p.walkers_per_lane, p.lds_depth = wpl, depth
            pa, sa = g.walk(app, p, nwalkers=2 * n)
            assert np.array_equal(pa, want) and sa == steps, (app, wpl, depth)
    nw, L = 3000, 20
    dp = B.DeviceArray(g, nbytes=nw * L * 4)
    ds = B.DeviceArray(g, np.zeros(1, np.uint64))
    wg = g.walk_graph(B.PPR, B.walk_params(seed=4, max_len=L), nw, dp, ds)
    want, steps = o.walk(B.PPR, orc.walk_params(seed=4, max_len=L, float_weights=True), nwalkers=nw)
    for rep in range(3):
        wg.launch()
        g.sync()
        assert np.array_equal(dp.to_host(np.int32, (nw, L)), want), rep
    wg.close()
    for a in (dp, ds):
        a.free()
    srv = B.Server(g, arena_bytes=1 << 24, max_launch_walkers=5000)
    reqs = [(app, 900 + 37 * app) for app in APPS] * 2
    tickets = [srv.submit_walk(app, B.walk_params(seed=s, max_len=20), nwalkers=s) for app, s in reqs]
    for (app, s), t in zip(reqs, tickets):
        out, st, _ = srv.wait(t)
        pb, sb = o.walk(app, orc.walk_params(seed=s, max_len=20, float_weights=True), nwalkers=s)
        assert np.array_equal(out, pb) and st == sb, (app, s)
    srv.close()

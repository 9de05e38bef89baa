"""GPU parity with sparse/dense bit groups (the reference's RATE,
include/api.h:25, operation.cuh:72, Bingo.cuh:244-252): the store after
load, insert and delete batches (short, staged and per-group long delete
runs), compaction (wave and segmented paths) and walks of every app in every
form against the oracle built with the same rate, bit for bit; the
rejection fallback; float weights; served and captured walks."""
import random

import numpy as np
import pytest

import bingo_amd as B
from tests.brute import BruteGraph, check_against_brute
from tests.floatw import float_edges

pytestmark = pytest.mark.gpu

APPS = [B.DEEPWALK, B.NODE2VEC, B.PPR, B.SAMPLING]
# (path staging of depth 16 needs max_len a multiple of 16)
FORMS = [dict(form=B.FORM_PLAIN, block=128),
         dict(form=B.FORM_PERSISTENT, block=256, walkers_per_lane=2, lds_depth=16),
         dict(form=B.FORM_PERSISTENT, block=512, walkers_per_lane=4, lds_depth=0)]


def _same_state(g, o, what=""):
    a, b = g.export(), o.export()
    for k in b:
        assert a[k].shape == b[k].shape, (what, k, a[k].shape, b[k].shape)
        if not np.array_equal(a[k], b[k]):
            bad = np.argwhere(a[k] != b[k])[:5]
            raise AssertionError(f"{what}: {k} differs at {bad.tolist()}")


def _same_walks(g, o, orc, app, seed, max_len, knobs, starts=None, lens=None, nwalkers=None, fw=False):
    gp = B.walk_params(seed=seed, tp=0.15, max_len=max_len, **knobs)
    op = orc.walk_params(seed=seed, tp=0.15, max_len=max_len, float_weights=fw)
    pa, sa = g.walk(app, gp, starts=starts, lens=lens, nwalkers=nwalkers)
    pb, sb = o.walk(app, op, starts=starts, lens=lens, nwalkers=nwalkers)
    if not np.array_equal(pa, pb):
        bad = np.argwhere((pa != pb).any(axis=1))[:5, 0]
        raise AssertionError(f"app {app} {knobs}: walkers {bad.tolist()} differ\n{pa[bad[0]]}\n{pb[bad[0]]}")
    assert sa == sb


def _rand_edges(rng, n, k, hub=None, wbits=32, hub_frac=0.4):
    e = np.zeros(k, B.EDGE_DT)
    e["u"] = [hub if (hub and rng.random() < hub_frac) else rng.randint(1, n) for _ in range(k)]
    e["v"] = [rng.randint(1, n) for _ in range(k)]
    e["w"] = [max(1, rng.getrandbits(rng.randint(1, wbits))) for _ in range(k)]
    return e


def _live_dels(rng, o, n, per_node, nodes, hub=None, hub_dels=0):
    d = o.export()
    so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
    out = []
    for u in rng.sample(range(1, n + 1), nodes):
        live = np.flatnonzero(d["slot_w"][so[u]:so[u] + int(d["nslots"][u])] != 0)
        out += [(u, int(x)) for x in rng.sample(list(live), min(len(live), rng.randint(0, per_node)))]
    if hub:
        live = np.flatnonzero(d["slot_w"][so[hub]:so[hub] + int(d["nslots"][hub])] != 0)
        out = [x for x in out if x[0] != hub]
        out += [(hub, int(x)) for x in rng.sample(list(live), min(len(live), hub_dels))]
    return np.array(out, B.DEL_DT)


@pytest.mark.parametrize("rate,nbr", [(0.4, True), (0.1, False), (1e-5, True)])
def test_updates_compaction_and_walks(orc, rate, nbr):
    rng = random.Random(int(rate * 1e5) + 7)
    n, hub = 700, 5
    e = _rand_edges(rng, n, 12000, hub=hub, wbits=rng.choice([16, 32]))
    print(f"[dense] rate={rate} load n={n} m={len(e)}", flush=True)
    g = B.Graph(neighbor_index=nbr, dense_rate=rate, pool_slack=0.0)
    g.load(n, e)
    o = orc.OracleGraph(n, nbr, dense_rate=rate)
    assert o.load(e) == 0
    bg = BruteGraph(n)
    bg.insert(e)
    _same_state(g, o, "load")
    assert g.export()["dmask"].any()
    check_against_brute(g.export(), bg)
    apps = APPS if nbr else [B.DEEPWALK, B.PPR, B.SAMPLING]
    for rnd in range(6):
        print(f"[dense] rate={rate} round {rnd}", flush=True)
        ins = _rand_edges(rng, n, rng.choice([1, 5, 3000, 9000]), hub=hub)
        g.insert(ins)
        assert o.insert(ins) == 0
        bg.insert(ins)
        # short runs everywhere, a long run (per-group tasks) on the hub
        dels = _live_dels(rng, o, n, per_node=rng.choice([1, 3, 20]), nodes=300, hub=hub,
                          hub_dels=rng.choice([0, 60, 600]))
        g.delete(dels)
        assert o.delete(dels) == 0
        bg.delete(dels)
        _same_state(g, o, f"round {rnd}")
        if rnd % 2:
            frac = rng.choice([0.0, 0.05, 0.5])
            got = g.compact(frac)
            assert got == o.compact(frac), (rnd, frac)
            bg = BruteGraph(n)
            d = o.export()
            so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
            for u in range(1, n + 1):
                bg.slots[u] = [[int(v), int(w)] for v, w in zip(d["slot_v"][so[u]:so[u + 1]],
                                                               d["slot_w"][so[u]:so[u + 1]])]
            _same_state(g, o, f"round {rnd} compact")
        app = apps[rnd % len(apps)]
        starts = np.array([hub if rng.random() < 0.3 else rng.randint(1, n) for _ in range(4000)], np.int32)
        lens = np.array([rng.randint(1, 48) for _ in range(4000)], np.int32)
        _same_walks(g, o, orc, app, 40 + rnd, 48, FORMS[rnd % 3], starts=starts, lens=lens)
    for app in apps:
        for knobs in FORMS:
            _same_walks(g, o, orc, app, 90 + app, 32, knobs, nwalkers=2 * n)
    check_against_brute(g.export(), bg)
    g.close()


def test_big_node_segment_compaction(orc):
    """A hub past the wave path's 8192 slots is compacted in segments over
    the grid; with dense groups its back entries are its listed bits only."""
    rng = random.Random(3)
    n, hub = 300, 2
    e = _rand_edges(rng, n, 30000, hub=hub, hub_frac=0.8)
    g = B.Graph(neighbor_index=True, dense_rate=0.4)
    g.load(n, e)
    o = orc.OracleGraph(n, True, dense_rate=0.4)
    assert o.load(e) == 0
    dels = _live_dels(rng, o, n, per_node=4, nodes=100, hub=hub, hub_dels=9000)
    g.delete(dels)
    assert o.delete(dels) == 0
    assert g.compact(0.01) == o.compact(0.01)
    _same_state(g, o, "segment compaction")
    ins = _rand_edges(rng, n, 5000, hub=hub)
    g.insert(ins)
    assert o.insert(ins) == 0
    _same_state(g, o, "insert after compaction")
    for app in APPS:
        _same_walks(g, o, orc, app, 5 + app, 32, FORMS[1], starts=np.full(20000, hub, np.int32))
    g.close()


def test_rejection_fallback(orc):
    """A dense group of one member among 2^18 slots: the draws that pick it
    run out of trials (probability 0.98) and take the member by the scan."""
    n = 4
    first = np.array([(1, 2, 3)], B.EDGE_DT)  # bits 0 and 1 dense
    big = np.zeros(1 << 18, B.EDGE_DT)
    big["u"], big["v"], big["w"] = 1, 3, 2
    g = B.Graph(neighbor_index=True, dense_rate=0.4)
    g.load(n, first)
    g.insert(big)
    o = orc.OracleGraph(n, True, dense_rate=0.4)
    assert o.load(first) == 0 and o.insert(big) == 0
    _same_state(g, o, "fallback store")
    starts = np.ones(1 << 21, np.int32)
    for knobs in (FORMS[0], FORMS[2]):
        gp = B.walk_params(seed=9, max_len=2, **knobs)
        pa, sa = g.walk(B.SAMPLING, gp, starts=starts)
        pb, sb = o.walk(B.SAMPLING, orc.walk_params(seed=9, max_len=2), starts=starts)
        assert np.array_equal(pa, pb) and sa == sb
        assert 0 < int((pa[:, 1] == 2).sum()) < 20
    g.close()


def test_float_weights_with_dense_groups(orc):
    rng = np.random.default_rng(4)
    n, hub = 500, 9
    e = float_edges(rng, n, 8000, hub=hub)
    g = B.Graph(neighbor_index=True, float_weights=True, dense_rate=0.4)
    g.load(n, e)
    o = orc.OracleGraph(n, True, float_weights=True, dense_rate=0.4)
    assert o.load(e) == 0
    _same_state(g, o, "float load")
    ins = float_edges(rng, n, 3000, hub=hub)
    g.insert(ins)
    assert o.insert(ins) == 0
    dels = _live_dels(random.Random(5), o, n, per_node=3, nodes=200, hub=hub, hub_dels=300)
    g.delete(dels)
    assert o.delete(dels) == 0
    _same_state(g, o, "float updates")
    starts = np.array(rng.integers(1, n + 1, 5000), np.int32)
    for app in APPS:
        for knobs in FORMS:
            _same_walks(g, o, orc, app, 60 + app, 32, knobs, starts=starts, fw=True)
    g.close()


def test_served_and_captured_walks(orc):
    rng = random.Random(8)
    n = 1 << 11
    base = B.gen_rmat(11, 20000, B.W_UNIFORM16, seed=8)  # uniform 16-bit: most groups dense
    g = B.Graph(neighbor_index=True, dense_rate=0.4)
    g.load(n, base)
    o = orc.OracleGraph(n, True, dense_rate=0.4)
    assert o.load(base) == 0
    assert int(g.export()["dmask"].astype(bool).sum()) > n // 4
    starts = np.array([rng.randint(1, n) for _ in range(3000)], np.int32)
    srv = B.Server(g, arena_bytes=1 << 24)
    tickets = [(app, srv.submit_walk(app, B.walk_params(seed=70 + app, max_len=30), starts=starts))
               for app in (B.DEEPWALK, B.NODE2VEC, B.PPR)]
    for app, t in tickets:
        out, st, _ = srv.wait(t)
        want, ws = o.walk(app, orc.walk_params(seed=70 + app, max_len=30), starts=starts)
        assert np.array_equal(out, want) and st == ws, app
    srv.close()
    nw, L = 4000, 30
    dp = B.DeviceArray(g, nbytes=nw * L * 4)
    ds = B.DeviceArray(g, np.zeros(1, np.uint64))
    wg = g.walk_graph(B.NODE2VEC, B.walk_params(seed=4, max_len=L, form=B.FORM_PERSISTENT), nw, dp, ds)
    ug = B.UpdateGen(n, 11, base, 20000, B.W_UNIFORM16, seed=8)
    for rep in range(3):
        if rep == 1:
            ins = ug.inserts(30000)
            g.insert(ins)
            assert o.insert(ins) == 0
        wg.launch()
        g.sync()
        want, _ = o.walk(B.NODE2VEC, orc.walk_params(seed=4, max_len=L), nwalkers=nw)
        assert np.array_equal(dp.to_host(np.int32, (nw, L)), want), rep
    wg.close()
    for a in (dp, ds):
        a.free()
    g.close()


def test_dense_limits():
    g = B.Graph(neighbor_index=True, dense_rate=0.4)
    g.load(3, np.array([(1, 2, 3), (2, 3, 1)], B.EDGE_DT))
    with pytest.raises(B.BingoError):
        g.walk(B.NODE2VEC, B.walk_params(max_len=4, max_attempts=65537), nwalkers=3)
    g.walk(B.NODE2VEC, B.walk_params(max_len=4, max_attempts=65536), nwalkers=3)
    g.close()
    with pytest.raises(B.BingoError):
        B.Graph(dense_rate=-0.5)


@pytest.mark.parametrize("rate", [0.0, 0.4])
def test_load_in_batches(orc, monkeypatch, rate):
    """A load larger than the load batch (2^30 arcs; 7000 here through the
    test hook) is applied as consecutive batches in array order: the store
    equals the oracle's fed the same batches, and with dense groups a node's
    classes come from the first batch that holds its arcs."""
    rng = random.Random(17)
    n = 400
    e = _rand_edges(rng, n, 30000, hub=3, wbits=rng.choice([16, 32]))
    monkeypatch.setenv("BINGO_LOAD_CHUNK", "7000")
    g = B.Graph(neighbor_index=True, dense_rate=rate)
    monkeypatch.delenv("BINGO_LOAD_CHUNK")
    g.load(n, e)
    o = orc.OracleGraph(n, True, dense_rate=rate)
    for at in range(0, len(e), 7000):
        assert o.insert(e[at:at + 7000]) == 0
    _same_state(g, o, "batched load")
    one = orc.OracleGraph(n, True, dense_rate=rate)
    assert one.load(e) == 0
    if rate:  # one batch decides classes from all arcs of a node: a different store
        assert not np.array_equal(one.export()["dmask"], o.export()["dmask"])
    _same_walks(g, o, orc, B.NODE2VEC, 3, 32, FORMS[1], nwalkers=n)
    g.close()

"""GPU parity when insert batches move hub regions too large for one wave:
slot arrays, back indices, member lists of tens of thousands of entries and
neighbour tables of 2^17 entries are moved (and the tables re-hashed) by the
grid-wide task kernels after the apply kernel.  Store, edge multiplicities
and walks must still equal the oracle's, through a further insert batch and
deletes on the moved regions."""
import numpy as np
import pytest

import bingo_amd as B
from tests.floatw import float_edges

pytestmark = pytest.mark.gpu


def _edges(rng, n, k, hub, share, fw=False):
    if fw:
        return float_edges(rng, n, k, hub=hub, hub_frac=share)
    e = np.zeros(k, B.EDGE_DT)
    e["u"] = np.where(rng.random(k) < share, hub, rng.integers(1, n + 1, k))
    e["v"] = rng.integers(1, n + 1, k)
    e["w"] = np.floor(2.0 ** (32 * rng.random(k))).clip(1, 2 ** 32 - 1).astype(np.uint32)
    return e


def _same_state(g, o, what):
    a, b = g.export(), o.export()
    for k in b:
        assert a[k].shape == b[k].shape, (what, k)
        if not np.array_equal(a[k], b[k]):
            bad = np.argwhere(a[k] != b[k])[:5]
            raise AssertionError(f"{what}: {k} differs at {bad.tolist()}")
    return b


def _same_walks(g, o, orc, app, seed, starts, fw):
    pa, sa = g.walk(app, B.walk_params(seed=seed, max_len=20), starts=starts)
    pb, sb = o.walk(app, orc.walk_params(seed=seed, max_len=20, float_weights=fw), starts=starts)
    assert np.array_equal(pa, pb) and sa == sb, (app, seed)


@pytest.mark.parametrize("fit,fw", [(False, False), (True, False), (False, True)])
def test_hub_region_moves_match_oracle(orc, fit, fw):
    rng = np.random.default_rng(71 + fit + 2 * fw)
    n, hub = 20000, 7
    base = _edges(rng, n, 80000, hub, 0.625, fw)  # ~50k arcs on the hub
    print(f"[bigmove] load n={n} m={len(base)} fit={fit} float={fw}", flush=True)
    g = B.Graph(neighbor_index=True, fit_load=fit, float_weights=fw)
    g.load(n, base)
    o = orc.OracleGraph(n, True, float_weights=fw)
    assert o.load(base) == 0
    _same_state(g, o, "load")
    vs = np.arange(1, n + 1, dtype=np.int32)
    hubs = np.full(n, hub, np.int32)
    starts = np.concatenate([np.full(3000, hub, np.int32), rng.integers(1, n + 1, 3000).astype(np.int32)])
    for rnd, (k, share) in enumerate([(70000, 0.86), (20000, 0.5)]):
        ins = _edges(rng, n, k, hub, share, fw)
        print(f"[bigmove] round {rnd}: insert {k} (hub share {share})", flush=True)
        g.insert(ins)
        assert o.insert(ins) == 0
        d = _same_state(g, o, f"insert {rnd}")
        got = g.edge_count(hubs, vs)
        want = np.array([o.edge_count(hub, int(v)) for v in vs], np.uint32)
        assert np.array_equal(got, want), f"round {rnd}: hub multiplicities differ"
        for app in (B.DEEPWALK, B.NODE2VEC):
            _same_walks(g, o, orc, app, 300 + 10 * rnd + app, starts, fw)
        # deletes: a third of the hub's live slots and a few elsewhere
        so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
        dels = []
        for u in [hub] + list(rng.integers(1, n + 1, 200)):
            ns = int(d["nslots"][u])
            w = d["slot_w"][so[u]:so[u] + ns]
            live = np.flatnonzero((w != 0) | (d["slot_f"][so[u]:so[u] + ns] != 0) if fw else w != 0)
            pick = live[rng.random(len(live)) < (0.33 if u == hub else 0.5)]
            dels += [(int(u), int(s)) for s in pick]
        dels = np.array(sorted(set(dels)), B.DEL_DT)[rng.permutation(len(set(dels)))]
        print(f"[bigmove] round {rnd}: delete {len(dels)}", flush=True)
        g.delete(dels)
        assert o.delete(dels) == 0
        _same_state(g, o, f"delete {rnd}")
        got = g.edge_count(hubs, vs)
        want = np.array([o.edge_count(hub, int(v)) for v in vs], np.uint32)
        assert np.array_equal(got, want), f"round {rnd}: hub multiplicities differ after deletes"
        _same_walks(g, o, orc, B.NODE2VEC, 400 + rnd, starts, fw)
    g.close()
    o.close()

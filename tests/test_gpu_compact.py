"""GPU parity of compaction (bingo_compact): the store after compaction
rounds bit-exact against the oracle's, with small nodes (one wave each) and
hubs of more than 8192 slots (segments over the grid), integer and float
weights, with and without neighbour tables; walks and edge counts after
compaction; and a sliding window whose generator renumbers with the store."""
import random

import numpy as np
import pytest

import bingo_amd as B
from tests.brute import BruteGraph, check_against_brute
from tests.floatw import float_edges
from tests.test_gpu_parity import APPS, _same_state, _same_walks

pytestmark = pytest.mark.gpu


def _edges(rng, n, k, hubs=(), wbits=32):
    e = np.zeros(k, B.EDGE_DT)
    u = np.array([rng.randint(1, n) for _ in range(k)], np.int32)
    for h in hubs:
        sel = np.array([rng.random() < 0.3 for _ in range(k)])
        u[sel] = h
    e["u"] = u
    e["v"] = [rng.randint(1, n) for _ in range(k)]
    e["w"] = [max(1, rng.getrandbits(rng.randint(1, wbits))) for _ in range(k)]
    return e


def _dels(rng, g, k):
    ex = g.export()
    so = np.concatenate([[0], np.cumsum(ex["nslots"].astype(np.int64))])
    live = np.flatnonzero((ex["slot_w"] != 0) | (ex["slot_f"] != 0 if "slot_f" in ex else False))
    pick = np.sort(np.array(rng.sample(list(live), min(k, len(live))), np.int64))
    u = np.searchsorted(so, pick, side="right") - 1
    d = np.zeros(len(pick), B.DEL_DT)
    d["u"] = u
    d["slot"] = pick - so[u]
    return d


@pytest.mark.parametrize("nbr", [True, False])
def test_compaction_rounds_match_oracle(orc, nbr):
    rng = random.Random(41 if nbr else 42)
    n = 400
    e = _edges(rng, n, 30000, hubs=(3,))  # node 3: about 9000 slots (the workgroup path)
    print(f"[compact] load nbr={nbr}", flush=True)
    g = B.Graph(neighbor_index=nbr)
    e0 = e.copy()
    g.load(n, e)
    assert np.array_equal(e, e0)  # the host arcs are only read
    o = orc.OracleGraph(n, nbr)
    assert o.load(e) == 0
    for rnd, frac in enumerate([0.0, 0.5, 1 / 16, 0.0, 0.25]):
        d = _dels(rng, o, 4000)
        g.delete(d)
        assert o.delete(d) == 0
        print(f"[compact] round {rnd} frac={frac}", flush=True)
        got = g.compact(frac)
        assert got == o.compact(frac), rnd
        _same_state(g, o, f"compact round {rnd}")
        ins = _edges(rng, n, 3000, hubs=(3, 9))
        g.insert(ins)
        assert o.insert(ins) == 0
        _same_state(g, o, f"insert after compact {rnd}")
        apps = APPS if nbr else [B.DEEPWALK, B.PPR, B.SAMPLING]
        _same_walks(g, o, orc, apps[rnd % len(apps)], seed=300 + rnd, max_len=24, nwalkers=4 * n)
        us = np.array([rng.randint(1, n) for _ in range(2000)], np.int32)
        vs = np.array([rng.randint(1, n) for _ in range(2000)], np.int32)
        assert (g.edge_count(us, vs) == np.array([o.edge_count(int(a), int(b)) for a, b in zip(us, vs)])).all()
    g.close()


def test_compaction_many_bits_match_oracle(orc):
    """Weights with 12 to 32 bits set: 64-slot chunks carry from about 800 to
    2048 back entries, on both sides of the 1024 a wave gathers in LDS before
    writing them in order (above it they are written directly), in small
    nodes and in a hub's segments."""
    rng = random.Random(17)
    n = 300
    e = _edges(rng, n, 32000, hubs=(5,))  # node 5: about 9600 slots (segments)
    e["w"] = [(1 << 32) - 1 - sum(1 << b for b in rng.sample(range(32), rng.randint(0, 20))) for _ in range(len(e))]
    print("[compact] many bits", flush=True)
    g = B.Graph(neighbor_index=False)
    g.load(n, e)
    o = orc.OracleGraph(n, False)
    assert o.load(e) == 0
    for rnd, frac in enumerate([0.0, 0.2]):
        d = _dels(rng, o, 5000)
        g.delete(d)
        assert o.delete(d) == 0
        assert g.compact(frac) == o.compact(frac)
        _same_state(g, o, f"many-bits compact {rnd}")
        _same_walks(g, o, orc, B.DEEPWALK, seed=60 + rnd, max_len=16, nwalkers=2 * n)
    g.close()


@pytest.mark.parametrize("nbr,sbase_max", [(False, None), (True, None), (False, 40)])
def test_dead_prefix_shift_matches_oracle(orc, monkeypatch, nbr, sbase_max):
    """Nodes whose dead slots are their oldest (a sliding window) take the
    shift path: records and back entries move down, member words keep their
    place and slot fields (the node's offset grows); small nodes one wave
    each, a hub of more than 8192 slots a workgroup.  With a small offset
    limit (test hook) the slot fields are lowered in place instead.  State,
    deletes and inserts after it, and walks, against the oracle."""
    rng = random.Random(23 if nbr else 24 if sbase_max is None else 25)
    n = 300
    e = _edges(rng, n, 40000, hubs=(5,))  # node 5: about 12000 slots
    print(f"[compact] dead prefix nbr={nbr} sbase_max={sbase_max}", flush=True)
    if sbase_max is not None:
        monkeypatch.setenv("BINGO_SBASE_MAX", str(sbase_max))
    g = B.Graph(neighbor_index=nbr)
    if sbase_max is not None:
        monkeypatch.delenv("BINGO_SBASE_MAX")
    g.load(n, e)
    o = orc.OracleGraph(n, nbr)
    assert o.load(e) == 0

    def oldest(frac):  # each node's live slots among its first ones
        ex = o.export()
        so = np.concatenate([[0], np.cumsum(ex["nslots"].astype(np.int64))])
        dl = [(u, s) for u in range(1, n + 1) for s in range(int(ex["nslots"][u] * frac * rng.random()))
              if ex["slot_w"][so[u] + s]]
        return np.array(dl, dtype=B.DEL_DT)

    for rnd in range(3):
        d = oldest(0.3)
        g.delete(d)
        assert o.delete(d) == 0
        assert g.compact(0.0) == o.compact(0.0), rnd
        _same_state(g, o, f"shift round {rnd}")
        ins = _edges(rng, n, 4000, hubs=(5, 9))
        g.insert(ins)
        assert o.insert(ins) == 0
        d = _dels(rng, o, 3000)  # any live slots, under the new numbers: the next round mixes both paths
        g.delete(d)
        assert o.delete(d) == 0
        _same_state(g, o, f"after shift round {rnd}")
        _same_walks(g, o, orc, B.NODE2VEC if nbr else B.DEEPWALK, seed=80 + rnd, max_len=20, nwalkers=2 * n)
    g.close()


def test_compaction_float_weights_match_oracle(orc):
    rng = random.Random(5)
    n = 200
    e = float_edges(np.random.default_rng(5), n, 12000, hub=4)
    print("[compact] float weights", flush=True)
    g = B.Graph(neighbor_index=True, float_weights=True)
    g.load(n, e)
    o = orc.OracleGraph(n, True, float_weights=True)
    assert o.load(e) == 0
    for rnd, frac in enumerate([0.0, 0.3]):
        d = _dels(rng, o, 3000)
        g.delete(d)
        assert o.delete(d) == 0
        assert g.compact(frac) == o.compact(frac)
        _same_state(g, o, f"float compact {rnd}")
        for app in (B.DEEPWALK, B.NODE2VEC):
            gp = B.walk_params(seed=8 + rnd, max_len=20)
            op = orc.walk_params(seed=8 + rnd, max_len=20, float_weights=True)
            pa, sa = g.walk(app, gp, nwalkers=3 * n)
            pb, sb = o.walk(app, op, nwalkers=3 * n)
            assert np.array_equal(pa, pb) and sa == sb, (rnd, app)
    g.close()


def test_sliding_window_with_compaction(orc):
    """A sliding window (inserts, then the oldest arcs expire) with a
    compaction after every step: the generator's slot numbers follow the
    store's, state equals the oracle's, and once every node has been
    compacted the slot pool stops growing."""
    scale, m = 13, 1 << 17
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=8, undirected=False)
    print("[compact] sliding window", flush=True)
    g = B.Graph(neighbor_index=False, fit_load=True, pool_slack=0.03)
    g.load(n, base)
    o = orc.OracleGraph(n, False)
    assert o.load(base) == 0
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=8)
    used = []
    for step in range(24):
        ins, dels = ug.inserts(m // 50), ug.expire(m // 50)
        g.insert(ins)
        g.delete(dels)
        assert o.insert(ins) == 0 and o.delete(dels) == 0
        ug.compact(1 / 32)
        assert g.compact(1 / 32) == o.compact(1 / 32), step
        used.append(g.stats()["pool_used"][0])
        if step % 6 == 5:
            _same_state(g, o, f"window step {step}")
            _same_walks(g, o, orc, B.DEEPWALK, seed=step, max_len=40, nwalkers=n)
    # the store's buckets hold exactly its live slots' bits (brute force)
    bg = BruteGraph(n)
    ex = o.export()
    so = np.concatenate([[0], np.cumsum(ex["nslots"].astype(np.int64))])
    for u in range(n + 1):
        bg.slots[u] = [[int(v), int(w)] for v, w in zip(ex["slot_v"][so[u]:so[u + 1]], ex["slot_w"][so[u]:so[u + 1]])]
    check_against_brute(g.export(), bg)
    print("[compact] slot pool top per step:", used, flush=True)
    # live arcs stay constant and moved regions are reused: what growth is
    # left comes from capacities that follow each node's largest live count
    assert used[-1] - used[len(used) // 2] <= 0.5 * (used[len(used) // 2] - used[0]), used
    g.close()

"""Compaction of tombstoned slots, on the CPU: the oracle against a brute-force
restatement (drop the dead slots of the selected nodes, keep the order), the
member lists rebuilt in slot order or, after a dead prefix, kept with slot
numbers lowered, the selection rule, and the update
generator's renumbering against the oracle's store."""
import random

import numpy as np
import pytest

import bingo_amd as B
from brute import BruteGraph, check_against_brute, split_dump
from test_oracle_core import _rand_dels, _rand_edges


def _brute_compact(bg, frac):
    """The rule of bingo_compact on the brute-force graph; returns the nodes
    compacted and the slots reclaimed."""
    thr16 = int(np.floor(frac * 65536.0))
    nodes = freed = 0
    for u in range(1, bg.n + 1):
        ns = len(bg.slots[u])
        dead = sum(1 for v, w in bg.slots[u] if not w)
        if dead and dead * 65536 >= thr16 * ns:
            bg.slots[u] = [[v, w] for v, w in bg.slots[u] if w]
            nodes += 1
            freed += dead
    return nodes, freed


@pytest.mark.parametrize("seed", range(6))
def test_compaction_matches_brute_force(orc, seed):
    rng = random.Random(100 + seed)
    n = rng.choice([3, 20, 60])
    g = orc.OracleGraph(n)
    bg = BruteGraph(n)
    e = _rand_edges(rng, n, rng.randint(50, 400), hub=1)
    assert g.load(e) == 0
    bg.insert(e)
    for step in range(10):
        d = _rand_dels(rng, bg, rng.randint(0, 120))
        assert g.delete(d) == 0
        bg.delete(d)
        frac = rng.choice([0.0, 1 / 16, 0.25, 0.5, 1.0])
        got = g.compact(frac)
        assert got == _brute_compact(bg, frac), (step, frac)
        check_against_brute(g.export(), bg)
        e = _rand_edges(rng, n, rng.randint(0, 120), hub=1)  # new arcs take slots from live(u) on
        assert g.insert(e) == 0
        bg.insert(e)
        check_against_brute(g.export(), bg)
        for u in range(1, n + 1):
            for v in rng.sample(range(1, n + 1), min(n, 5)):
                assert g.edge_count(u, v) == sum(1 for s, vv, w in bg.live_arcs(u) if vv == v)


def test_compaction_rebuilds_lists_in_slot_order(orc):
    """A compacted node whose dead slots are not its first ones gets member
    lists that are its live slots with bit j in slot order, under the new
    slot numbers; a dead prefix keeps the lists' order with slot fields
    lowered by its length; nodes left alone keep their lists entry for
    entry.  The one-draw law of a node moves by at most the multiply-shift's
    2^-32 steps."""
    from brute import transition_law
    rng = random.Random(7)
    n = 6
    g = orc.OracleGraph(n)
    bg = BruteGraph(n)
    e = _rand_edges(rng, n, 500, wmax_bits=8, hub=1)
    g.load(e)
    bg.insert(e)
    d = _rand_dels(rng, bg, 200)
    d = d[d["u"] != 6]  # node 6 keeps every slot, so it is not compacted
    g.delete(d)
    bg.delete(d)
    d0 = g.export()
    before = split_dump(d0, n)
    compacted = {u for u in range(1, n + 1) if any(w == 0 for v, w in bg.slots[u])}
    lead = {u: next((s for s, (v, w) in enumerate(bg.slots[u]) if w), len(bg.slots[u])) for u in compacted}
    prefix = {u for u in compacted if lead[u] == sum(1 for v, w in bg.slots[u] if not w)}
    law0 = {u: transition_law(d0["T"][u], d0["alias"][u], before[u][2])
            for u in range(1, n + 1) if int(d0["mass"][u])}
    assert 6 not in compacted and len(compacted) >= 3
    assert g.compact(0.0)[0] == len(compacted)
    _brute_compact(bg, 0.0)
    d1 = g.export()
    check_against_brute(d1, bg)
    after = split_dump(d1, n)
    for u in range(1, n + 1):
        for j in range(32):
            A0 = [int(x) for x in after[u][2][j]]
            if u in prefix:
                want = [int(x) - lead[u] for x in before[u][2][j]]
            elif u in compacted:
                want = [(v << 32) | s for s, (v, w) in enumerate(bg.slots[u]) if w >> j & 1]
            else:
                want = [int(x) for x in before[u][2][j]]
            assert A0 == want, (u, j)
        if u in law0:  # members move, so each may gain or lose one 2^-32 step
            law1 = transition_law(d1["T"][u], d1["alias"][u], after[u][2])
            assert law1.keys() == law0[u].keys(), u
            assert max(abs(law1[v] - law0[u][v]) for v in law1) < 2.0 ** -26, u
            if u not in compacted or u in prefix:
                assert law1 == law0[u], u


def test_dead_prefix_keeps_lists_and_walks(orc):
    """A sliding window's expiries leave each node a dead prefix: compaction
    drops it and lowers every slot number by its length, member lists keep
    their order, and walks (and the exact one-draw law) are unchanged."""
    from brute import transition_law
    rng = random.Random(11)
    n = 40
    e = _rand_edges(rng, n, 2000, wmax_bits=32, hub=1)
    g = orc.OracleGraph(n)
    bg = BruteGraph(n)
    assert g.load(e) == 0
    bg.insert(e)
    more = _rand_edges(rng, n, 600, wmax_bits=16, hub=1)  # later arcs: deletes below move members
    assert g.insert(more) == 0
    bg.insert(more)
    dels = []
    for u in range(1, n + 1):  # each node's oldest arcs, up to a third of them
        k = rng.randint(0, len(bg.slots[u]) // 3)
        dels += [(u, s) for s in range(k)]
    d = np.array(dels, dtype=orc.DEL_DT)
    assert g.delete(d) == 0
    bg.delete(d)
    d0 = g.export()
    before = split_dump(d0, n)
    lead = [next((s for s, (v, w) in enumerate(bg.slots[u]) if w), len(bg.slots[u])) for u in range(n + 1)]
    wp = orc.walk_params(seed=4, max_len=30)
    pa, sa = g.walk(orc.DEEPWALK, wp, nwalkers=4 * n)
    assert g.compact(0.0) == _brute_compact(bg, 0.0)
    d1 = g.export()
    check_against_brute(d1, bg)
    after = split_dump(d1, n)
    for u in range(1, n + 1):
        for j in range(32):
            assert [int(x) for x in after[u][2][j]] == [int(x) - lead[u] for x in before[u][2][j]], (u, j)
        if int(d0["mass"][u]):
            assert (transition_law(d1["T"][u], d1["alias"][u], after[u][2])
                    == transition_law(d0["T"][u], d0["alias"][u], before[u][2])), u
    pb, sb = g.walk(orc.DEEPWALK, wp, nwalkers=4 * n)
    assert np.array_equal(pa, pb) and sa == sb


def test_compaction_rule_and_bad_fraction(orc):
    g = orc.OracleGraph(2)
    e = np.array([(1, 2, 1)] * 10 + [(2, 1, 3)] * 4, dtype=orc.EDGE_DT)
    g.load(e)
    g.delete(np.array([(1, 0)], dtype=orc.DEL_DT))  # node 1: 1 of 10 dead (0.1)
    g.delete(np.array([(2, 1), (2, 2)], dtype=orc.DEL_DT))  # node 2: 2 of 4 dead (0.5)
    assert g.compact(0.5) == (1, 2)  # node 2 only
    assert g.compact(0.11) == (0, 0)  # 0.1 < 0.11
    assert g.compact(0.1) == (1, 1)  # 6553 * 10 <= 65536
    assert g.compact(0.0) == (0, 0)  # nothing dead left
    with pytest.raises(ValueError):
        g.compact(1.5)
    ex = g.export()
    assert ex["nslots"].tolist() == [0, 9, 2]


@pytest.mark.parametrize("mode", ["expire", "random"])
def test_generator_renumbers_like_the_store(orc, mode):
    """The generator's deletes after compactions name live slots of the
    oracle's store; under expiry they name exactly the oldest arcs."""
    n, m = 1 << 9, 6000
    base = B.gen_rmat(9, m, B.W_LOGUNI32, seed=21, undirected=False)
    ug = B.UpdateGen(n, 9, base, m, B.W_LOGUNI32, seed=21)
    g = orc.OracleGraph(n)
    assert g.load(base) == 0
    arrivals = [base]
    expired = 0
    for step in range(12):
        ins = ug.inserts(300)
        assert g.insert(ins) == 0
        arrivals.append(ins)
        dels = ug.expire(350) if mode == "expire" else ug.deletes(350)
        if mode == "expire":
            arcs = np.concatenate(arrivals)
            want = arcs[expired:expired + len(dels)]
            ex = g.export()
            so = np.concatenate([[0], np.cumsum(ex["nslots"].astype(np.int64))])
            at = so[dels["u"]] + dels["slot"]
            assert (dels["slot"] < ex["nslots"][dels["u"]]).all()
            assert (ex["slot_v"][at] == want["v"]).all() and (ex["slot_w"][at] == want["w"]).all(), step
            expired += len(dels)
        assert g.delete(dels) == 0, step  # a dead or out-of-range slot is rejected
        frac = [0.0, 0.05, 0.3][step % 3]
        ug.compact(frac)
        g.compact(frac)
    assert g.export()["slot_w"].astype(bool).sum() == ug.live()

"""Sparse/dense bit groups (the reference's RATE, include/api.h:25,
operation.cuh:72, Bingo.cuh:244-252) in the oracle, on the CPU: the store
against a brute-force restatement that decides the classes itself from each
node's first batch, the exact one-draw law with dense groups by enumeration,
a chi-square of walks, the rejection stream's fallback, compaction, and the
parameter limits."""
import random
from fractions import Fraction

import numpy as np
import pytest

from brute import BruteGraph, check_against_brute, member_law, split_dump
from test_oracle_compact import _brute_compact
from test_oracle_core import _rand_dels, _rand_edges
from test_oracle_walk import _chi2_p

M64 = 1 << 64


def _classes(batch, bg, rate16, masks):
    """The class rule restated: a node without slots before `batch` gets
    dense groups j with count_j >= 1 and count_j 2^16 >= rate16 len over its
    arcs of the batch; other nodes keep theirs."""
    by = {}
    for e in batch:
        by.setdefault(int(e["u"]), []).append(int(e["w"]))
    for u, ws in by.items():
        if len(bg.slots[u]):
            continue
        m = 0
        for j in range(32):
            c = sum(w >> j & 1 for w in ws)
            if rate16 and c and c * 65536 >= rate16 * len(ws):
                m |= 1 << j
        masks[u] = m


def test_rate_quantisation(orc):
    assert orc.dense_rate16(0.0) == 0 and orc.dense_rate16(1.0) == 0 and orc.dense_rate16(2.5) == 0
    assert orc.dense_rate16(-0.3) == 0
    assert orc.dense_rate16(0.4) == 26214 and orc.dense_rate16(0.1) == 6554
    assert orc.dense_rate16(1e-9) == 1 and orc.dense_rate16(0.999999) == 65535


@pytest.mark.parametrize("seed,rate", [(0, 0.4), (1, 0.1), (2, 0.4), (3, 1e-6), (4, 0.7), (5, 0.25)])
def test_store_matches_brute_force(orc, seed, rate):
    rng = random.Random(500 + seed)
    n = rng.choice([3, 12, 40])
    g = orc.OracleGraph(n, dense_rate=rate)
    bg = BruteGraph(n)
    masks = [0] * (n + 1)
    e = _rand_edges(rng, n, rng.randint(0, 300), wmax_bits=rng.choice([6, 16, 32]), hub=1)
    _classes(e, bg, g.rate16, masks)
    assert g.load(e) == 0
    bg.insert(e)
    d = g.export()
    assert [int(x) for x in d["dmask"]] == masks
    check_against_brute(d, bg)
    for step in range(10):
        r = rng.random()
        if r < 0.45:
            e = _rand_edges(rng, n, rng.randint(0, 120), wmax_bits=rng.choice([6, 32]), hub=1)
            _classes(e, bg, g.rate16, masks)
            assert g.insert(e) == 0
            bg.insert(e)
        elif r < 0.85:
            dl = _rand_dels(rng, bg, rng.randint(0, 150))
            assert g.delete(dl) == 0
            bg.delete(dl)
        else:  # compaction may empty a node: its next batch decides its classes again
            frac = rng.choice([0.0, 0.5, 1.0])
            assert g.compact(frac) == _brute_compact(bg, frac)
        d = g.export()
        assert [int(x) for x in d["dmask"]] == masks, step
        check_against_brute(d, bg)
        for u in range(1, n + 1):
            for v in rng.sample(range(1, n + 1), min(n, 4)):
                assert g.edge_count(u, v) == sum(1 for s, vv, w in bg.live_arcs(u) if vv == v)


def _dense_law(T, alias, groups, dmask, slot_w, slot_v):
    """Exact P(target) of one draw from a node's table, members and slots:
    a listed group by multiply-shift position, a dense group by rejection
    over the slots (each trial picks slot floor(x nslots / 2^32); the law of
    an accepted trial is the trial law restricted to the group's slots).
    Also returns the largest probability that a dense draw exhausts its
    4096 trials."""
    from brute import alias_law
    P = alias_law(T, alias)
    ns = len(slot_w)
    trial = member_law(ns) if ns else []
    out, tail = {}, 0.0
    for j in range(32):
        if P[j] == 0:
            continue
        if dmask >> j & 1:
            hit = [s for s in range(ns) if int(slot_w[s]) >> j & 1]
            A = sum(trial[s] for s in hit)
            tail = max(tail, float((1 - A) ** 4096))
            for s in hit:
                v = int(slot_v[s])
                out[v] = out.get(v, 0) + P[j] * trial[s] / A
        else:
            for pos, pm in enumerate(member_law(len(groups[j]))):
                v = int(groups[j][pos]) >> 32
                out[v] = out.get(v, 0) + P[j] * pm
    return out, tail


@pytest.mark.parametrize("weights,rate", [
    ([1, 2, 3, 4, 5, 6, 7], 0.4),
    ([0xFFFFFFFF, 1, 12345, 1 << 31, 7, 3, 3, 1], 0.1),
    ([5] * 9 + [1 << 20], 0.4),
    ([1, 1 << 20, 3, 3, 3, (1 << 32) - 2, 65535, 255, 17], 1e-6),
])
def test_exact_transition_law(orc, weights, rate):
    n = len(weights) + 1
    g = orc.OracleGraph(n, dense_rate=rate)
    # distinct targets, a deleted arc (a tombstone among the trial slots)
    e = [(1, i + 2, w) for i, w in enumerate(weights)] + [(1, 2, 6)]
    assert g.load(np.array(e, dtype=orc.EDGE_DT)) == 0
    assert g.delete(np.array([(1, len(weights))], dtype=orc.DEL_DT)) == 0
    d = g.export()
    assert int(d["dmask"][1]) != 0
    nodes = split_dump(d, g.n)
    sv, sw, groups = nodes[1]
    law, tail = _dense_law(d["T"][1], d["alias"][1], groups, int(d["dmask"][1]), sw, sv)
    assert tail < 1e-12
    live = [(i + 2, w) for i, w in enumerate(weights)] + []
    S = sum(w for _, w in live)
    for v, w in live:
        want = Fraction(w, S)
        got = law.get(v, 0)
        assert abs(got - want) / want < Fraction(1, 10**6), (v, float(got), float(want))


@pytest.mark.parametrize("seed", [1, 2])
def test_chi_square_walk_steps(orc, seed):
    """One-step walks from a node whose low groups are dense (the rejection
    stream as the walks use it) against w / sum(w)."""
    rng = random.Random(seed)
    weights = [rng.randint(1, 255) for _ in range(40)]
    n = len(weights) + 1
    g = orc.OracleGraph(n, dense_rate=0.4)
    assert g.load(np.array([(1, i + 2, w) for i, w in enumerate(weights)], dtype=orc.EDGE_DT)) == 0
    assert int(g.export()["dmask"][1]) != 0
    N = 200_000
    paths, steps = g.walk(orc.SAMPLING, orc.walk_params(seed=seed, max_len=2), starts=np.ones(N, np.int32))
    assert steps == N
    cnt = np.bincount(paths[:, 1], minlength=n + 1)[2:]
    assert _chi2_p(cnt, np.array(weights, float) / sum(weights)) > 0.01


def test_fallback_after_the_trials(orc):
    """A dense group of one member among 2^18 slots: the 4096 trials miss
    (probability 0.98), and the draw takes the first slot with the bit after
    the last trial, which is that member."""
    n = 4
    g = orc.OracleGraph(n, dense_rate=0.4)
    assert g.load(np.array([(1, 2, 3)], dtype=orc.EDGE_DT)) == 0  # bits 0 and 1 dense
    big = np.zeros(1 << 18, dtype=orc.EDGE_DT)
    big["u"], big["v"], big["w"] = 1, 3, 2  # bit 1 only
    assert g.insert(big) == 0
    d = g.export()
    assert int(d["dmask"][1]) == 3 and int(d["bsize"][1][0]) == 1
    T = d["T"][1]
    assert int(T[0]) > 0  # slot 0 stays in group 0 for a coin of 0
    assert g.sample_words(1, np.array([0, 0, 0, 0], np.uint32)) == 2
    # the walks reach it too: group 0 is drawn with probability 1 / (1 + 2^19 + 2)
    paths, _ = g.walk(orc.SAMPLING, orc.walk_params(seed=9, max_len=2), starts=np.ones(1 << 21, np.int32))
    hits = int((paths[:, 1] == 2).sum())
    assert 0 < hits < 20


def test_dense_attempt_limit(orc):
    g = orc.OracleGraph(3, dense_rate=0.4)
    assert g.load(np.array([(1, 2, 3), (2, 3, 1)], dtype=orc.EDGE_DT)) == 0
    with pytest.raises(ValueError):
        g.walk(orc.NODE2VEC, orc.walk_params(seed=1, max_len=4, max_attempts=65537), nwalkers=3)
    g.walk(orc.NODE2VEC, orc.walk_params(seed=1, max_len=4, max_attempts=65536), nwalkers=3)


def test_rate_off_equals_listed_store(orc):
    """dense_rate 0 and >= 1 keep every group listed: the same store as the
    plain oracle, bit for bit."""
    rng = random.Random(77)
    n = 30
    e = _rand_edges(rng, n, 400, hub=1)
    a, b, c = orc.OracleGraph(n), orc.OracleGraph(n, dense_rate=0.0), orc.OracleGraph(n, dense_rate=1.5)
    for g in (a, b, c):
        assert g.load(e) == 0
    da, db, dc = a.export(), b.export(), c.export()
    for k in da:
        assert np.array_equal(da[k], db[k]) and np.array_equal(da[k], dc[k]), k
    assert not da["dmask"].any()

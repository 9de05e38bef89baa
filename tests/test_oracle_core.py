"""Pins for the CPU oracle: Philox known answers, alias-table exactness,
bucket state against a brute-force model, and the delete convention."""
import json
import os
import random
from fractions import Fraction

import numpy as np
import pytest

from tests.brute import BruteGraph, alias_law, alias_prefix, check_against_brute, split_dump

HERE = os.path.dirname(os.path.abspath(__file__))


def test_philox_known_answers(orc):
    kat = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    for src in ("published", "rocrand"):
        for case in kat[src]:
            assert orc.philox(case["counter"], case["key"]) == case["out"], (src, case)


MASS_CASES = [
    [0] * 32,
    [5] + [0] * 31,
    [0] * 31 + [1 << 62],
    [1] * 32,
    [1 << j for j in range(32)],
    [1] + [0] * 30 + [(1 << 62) - 1],  # extreme ratio
    [(1 << 31) * 3 << j if j % 3 == 0 else 0 for j in range(32)],
]


def _random_masses(rng):
    cnt = [rng.choice([0, 0, 1, 2, 3, 17, 1000, 65535, 1 << 20]) for _ in range(32)]
    return [c << j for j, c in enumerate(cnt)]


@pytest.mark.parametrize("case", range(len(MASS_CASES) + 40))
def test_alias_exact_law(orc, case):
    rng = random.Random(case)
    M = MASS_CASES[case] if case < len(MASS_CASES) else _random_masses(rng)
    T, a = orc.alias_build(np.array(M, dtype=np.uint64))
    S = sum(M)
    # the sequential sweep and its prefix-sum restatement give the same table
    T2, a2 = alias_prefix(M)
    assert [int(x) for x in T] == T2 and [int(x) for x in a] == a2
    P = alias_law(T, a)
    assert sum(P) == 1
    for j in range(32):
        if S == 0:
            assert int(T[j]) == (1 << 64) - 1 and a[j] == j
            continue
        want = Fraction(M[j], S)
        # quantisation of a stay threshold is < 2^-64 per slot
        assert abs(P[j] - want) <= Fraction(32, 1 << 64), (j, float(P[j]), float(want))
        if M[j] == 0:
            assert P[j] == 0
        elif want > Fraction(1, 1 << 40):
            assert abs(P[j] - want) / want <= Fraction(1, 10**6)


def test_alias_sweep_equals_prefix_form(orc):
    """Many heavies, heavies of exactly one slot's capacity (zero excess),
    chains of heavies turning light, single groups: the oracle's sequential
    sweep equals the prefix-sum form on every case."""
    rng = random.Random(99)
    cases = []
    for _ in range(400):
        kind = rng.randrange(4)
        if kind == 0:  # equal-ish masses: many heavies near capacity
            M = [rng.choice([0, 1, 2, 3]) << 20 for _ in range(32)]
        elif kind == 1:  # a few heavies, many small lights
            M = [rng.choice([0, 1, 5]) for _ in range(32)]
            for j in rng.sample(range(32), rng.randint(1, 4)):
                M[j] = rng.randint(1, 1 << 40)
        elif kind == 2:  # sizes << j as in the store
            M = [rng.choice([0, 1, 2, 7, 64, 1000]) << j for j in range(32)]
        else:
            M = [rng.randrange(1 << rng.randint(1, 50)) for _ in range(32)]
        cases.append(M)
    cases.append([1] * 16 + [0] * 16)  # 16 heavies of exactly S / 32 * 2
    cases.append([3] + [1] * 31)
    for M in cases:
        T, a = orc.alias_build(np.array(M, dtype=np.uint64))
        T2, a2 = alias_prefix(M)
        assert [int(x) for x in T] == T2 and [int(x) for x in a] == a2, M
        if sum(M):
            P = alias_law(T, a)
            assert all(abs(P[j] - Fraction(M[j], sum(M))) <= Fraction(32, 1 << 64) for j in range(32))


def _rand_edges(rng, n, k, wmax_bits=32, hub=None):
    e = np.zeros(k, dtype=[("u", np.int32), ("v", np.int32), ("w", np.uint32)])
    for i in range(k):
        u = hub if (hub and rng.random() < 0.5) else rng.randint(1, n)
        e[i] = (u, rng.randint(1, n), rng.randint(1, (1 << rng.randint(1, wmax_bits)) - 1) or 1)
    return e


def _rand_dels(rng, bg, k):
    live = [(u, s) for u in range(1, bg.n + 1) for s, (v, w) in enumerate(bg.slots[u]) if w]
    pick = rng.sample(live, min(k, len(live)))
    d = np.zeros(len(pick), dtype=[("u", np.int32), ("slot", np.int32)])
    for i, (u, s) in enumerate(pick):
        d[i] = (u, s)
    return d


@pytest.mark.parametrize("seed", range(6))
def test_store_matches_brute_force(orc, seed):
    rng = random.Random(seed)
    n = rng.choice([1, 7, 40])
    g = orc.OracleGraph(n)
    bg = BruteGraph(n)
    e = _rand_edges(rng, n, rng.randint(0, 300), hub=1)
    assert g.load(e) == 0
    bg.insert(e)
    check_against_brute(g.export(), bg)
    for _ in range(8):
        if rng.random() < 0.5:
            e = _rand_edges(rng, n, rng.randint(0, 120), hub=1)
            assert g.insert(e) == 0
            bg.insert(e)
        else:
            d = _rand_dels(rng, bg, rng.randint(0, 150))
            assert g.delete(d) == 0
            bg.delete(d)
        check_against_brute(g.export(), bg)
        for u in range(1, n + 1):
            for v in range(1, n + 1):
                want = sum(1 for s, vv, w in bg.live_arcs(u) if vv == v)
                assert g.edge_count(u, v) == want


def test_delete_convention_fills_holes_from_tail(orc):
    """Restates the batch-delete convention on the exported lists: positions
    below the new size that were not deleted stay put; the k-th hole (holes
    in ascending slot order) receives the k-th surviving tail member."""
    rng = random.Random(11)
    n = 3
    g = orc.OracleGraph(n)
    bg = BruteGraph(n)
    e = _rand_edges(rng, n, 400, wmax_bits=6, hub=1)
    g.load(e)
    bg.insert(e)
    for _ in range(5):
        before = split_dump(g.export(), n)
        d = _rand_dels(rng, bg, 60)
        dset = {(int(x["u"]), int(x["slot"])) for x in d}
        g.delete(d)
        bg.delete(d)
        after = split_dump(g.export(), n)
        for u in range(1, n + 1):
            for j in range(32):
                B = [int(x) for x in before[u][2][j]]
                A = [int(x) for x in after[u][2][j]]
                dele = [(m & 0xFFFFFFFF) for m in B if (u, m & 0xFFFFFFFF) in dset]
                S2 = len(B) - len(dele)
                assert len(A) == S2
                pos = {m & 0xFFFFFFFF: p for p, m in enumerate(B)}
                holes = [pos[s] for s in sorted(dele) if pos[s] < S2]
                surv = [m for m in B[S2:] if (u, m & 0xFFFFFFFF) not in dset]
                want = list(B[:S2])
                for h, m in zip(holes, surv):
                    want[h] = m
                assert A == want


def test_invalid_updates_leave_state_unchanged(orc):
    g = orc.OracleGraph(5)
    bg = BruteGraph(5)
    e = np.array([(1, 2, 3), (1, 3, 1), (2, 1, 7)], dtype=orc.EDGE_DT)
    g.load(e)
    bg.insert(e)
    snap = g.export()
    bad_inserts = [
        np.array([(0, 1, 1)], dtype=orc.EDGE_DT),
        np.array([(1, 6, 1)], dtype=orc.EDGE_DT),
        np.array([(1, 2, 0)], dtype=orc.EDGE_DT),
    ]
    for b in bad_inserts:
        assert g.insert(b) == -1
    bad_dels = [
        np.array([(1, 2)], dtype=orc.DEL_DT),  # slot out of range
        np.array([(1, 0), (1, 0)], dtype=orc.DEL_DT),  # duplicate
        np.array([(6, 0)], dtype=orc.DEL_DT),
    ]
    for b in bad_dels:
        assert g.delete(b) == -1
    assert g.delete(np.array([(1, 0)], dtype=orc.DEL_DT)) == 0
    assert g.delete(np.array([(1, 0)], dtype=orc.DEL_DT)) == -1  # already dead
    bg.delete(np.array([(1, 0)], dtype=orc.DEL_DT))
    check_against_brute(g.export(), bg)
    for k in ("T", "alias"):
        assert (snap[k][2:] == g.export()[k][2:]).all()

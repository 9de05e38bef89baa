"""Pins for the oracle's samplers: exact one-step law by enumeration,
chi-square of walks against w/sum(w), node2vec's second-order law and the
PPR length law, plus the generators' distributions."""
import random
from fractions import Fraction

import numpy as np
import pytest
from scipy import stats

from tests.brute import split_dump, transition_law


def _chi2_p(counts, probs):
    counts = np.asarray(counts, float)
    exp = np.asarray(probs, float) * counts.sum()
    keep = exp > 0
    assert counts[~keep].sum() == 0
    return stats.chisquare(counts[keep], exp[keep]).pvalue


def _star(orc, weights, extra=()):
    n = len(weights) + 1
    e = [(1, i + 2, w) for i, w in enumerate(weights)] + list(extra)
    g = orc.OracleGraph(n)
    assert g.load(np.array(e, dtype=orc.EDGE_DT)) == 0
    return g


@pytest.mark.parametrize("weights", [
    [1, 2, 3, 4],
    [0xFFFFFFFF, 1, 12345, 1 << 31, 7],
    [5] * 9,
    [1, 1 << 20, 3, 3, 3, (1 << 32) - 2],
])
def test_exact_transition_law(orc, weights):
    g = _star(orc, weights)
    d = g.export()
    nodes = split_dump(d, g.n)
    law = transition_law(d["T"][1], d["alias"][1], nodes[1][2])
    S = sum(weights)
    for i, w in enumerate(weights):
        want = Fraction(w, S)
        got = law.get(i + 2, 0)
        assert abs(got - want) / want < Fraction(1, 10**6), (i, float(got), float(want))


def test_sample_words_follows_table(orc):
    g = _star(orc, [3, 5, 6, 1 << 30])
    d = g.export()
    groups = split_dump(d, g.n)[1][2]
    T, a = d["T"][1], d["alias"][1]
    rng = random.Random(3)
    for _ in range(2000):
        x = [rng.getrandbits(32) for _ in range(4)]
        i = x[0] >> 27
        coin = (x[1] << 32) | x[2]
        grp = i if coin < int(T[i]) else int(a[i])
        mem = groups[grp]
        want = int(mem[(x[3] * len(mem)) >> 32]) >> 32
        assert g.sample_words(1, np.array(x, np.uint32)) == want


@pytest.mark.parametrize("seed", [1, 2])
def test_chi_square_one_step(orc, seed):
    rng = random.Random(seed)
    weights = [rng.randint(1, (1 << rng.randint(1, 20)) - 1) for _ in range(25)]
    g = _star(orc, weights)
    N = 300_000
    p = orc.walk_params(seed=seed, max_len=2)
    paths, steps = g.walk(orc.SAMPLING, p, starts=np.ones(N, np.int32))
    assert steps == N
    cnt = np.bincount(paths[:, 1], minlength=g.n + 1)[2:]
    probs = np.array(weights, float) / sum(weights)
    assert _chi2_p(cnt, probs) > 0.01


def test_deepwalk_dead_end_and_padding(orc):
    # 1 -> 2 -> 3, node 3 has no out-arcs
    g = orc.OracleGraph(3)
    g.load(np.array([(1, 2, 5), (2, 3, 9)], dtype=orc.EDGE_DT))
    p = orc.walk_params(seed=4, max_len=6)
    paths, steps = g.walk(orc.DEEPWALK, p, starts=np.array([1, 3], np.int32))
    assert paths[0].tolist() == [1, 2, 3, -1, -1, -1]
    assert paths[1].tolist() == [3, -1, -1, -1, -1, -1]
    assert steps == 2
    paths, _ = g.walk(orc.DEEPWALK, p, starts=np.array([1], np.int32), lens=np.array([2], np.int32))
    assert paths[0].tolist() == [1, 2, -1, -1, -1, -1]


def test_walks_are_keyed_by_walker(orc):
    rng = random.Random(5)
    n = 30
    e = np.array([(rng.randint(1, n), rng.randint(1, n), rng.randint(1, 1000)) for _ in range(300)],
                 dtype=orc.EDGE_DT)
    g = orc.OracleGraph(n)
    g.load(e)
    p = orc.walk_params(seed=77, max_len=20)
    starts = np.array([rng.randint(1, n) for _ in range(64)], np.int32)
    a, _ = g.walk(orc.NODE2VEC, p, starts=starts, nthreads=1)
    b, _ = g.walk(orc.NODE2VEC, p, starts=starts, nthreads=8)
    assert (a == b).all()
    # a walker's path depends only on (seed, walker index, start)
    c, _ = g.walk(orc.NODE2VEC, p, starts=starts[:10])
    assert (a[:10] == c).all()


@pytest.mark.parametrize("P,Q,fold", [(0.5, 2.0, True), (0.25, 1.0, True), (0.1, 4.0, True), (2.0, 0.5, True),
                                       (1.0, 1.0, True), (0.5, 2.0, False)])
def test_node2vec_second_order_law(orc, P, Q, fold):
    """The second-order law w * alpha on a graph with a triangle (1,2,3),
    pendants 4,5 on 2 and a doubled return arc 2->1 (weight sums), with the
    return arc folded out of the envelope (1/p > max(1, 1/q)) or not."""
    und = [(1, 2, 3), (2, 3, 5), (1, 3, 2), (2, 4, 7), (2, 5, 1), (4, 5, 2)]
    e = []
    for u, v, w in und:
        e += [(u, v, w), (v, u, w)]
    e.append((2, 1, 4))  # a second arc 2 -> 1: the return mass is 3 + 4
    g = orc.OracleGraph(5)
    g.load(np.array(e, dtype=orc.EDGE_DT))
    p = orc.walk_params(seed=9, p=P, q=Q, max_len=3, max_attempts=1 << 20, float_weights=not fold)
    assert (p.fold_r > 0) == (fold and 1 / P > max(1.0, 1 / Q))
    N = 400_000
    paths, _ = g.walk(orc.NODE2VEC, p, starts=np.ones(N, np.int32))
    sel = paths[:, 1] == 2  # condition on 1 -> 2
    third = paths[sel, 2]
    nbr2 = {1: 3 + 4, 3: 5, 4: 7, 5: 1}
    nbr1 = {2, 3}
    alpha = {x: (1 / P if x == 1 else (1.0 if x in nbr1 else 1 / Q)) for x in nbr2}
    xs = sorted(nbr2)
    probs = np.array([nbr2[x] * alpha[x] for x in xs])
    probs /= probs.sum()
    cnt = [(third == x).sum() for x in xs]
    assert sum(cnt) == sel.sum()
    assert _chi2_p(cnt, probs) > 0.01


def test_fold_threshold_matches_ratio(orc):
    """The folded region's threshold against the exact ratio w_ret R /
    (S 2^27 + w_ret R), including the cut to 60 bits."""
    from fractions import Fraction
    rng = np.random.default_rng(6)
    cases = [(1, 1, 1 << 27), (7, 3, 1 << 27), ((1 << 59) - 1, (1 << 59) - 1, 99 << 27), (1 << 40, 1, 1),
             (5, 0, 1 << 27), (1, 1, (1 << 63) - 1)]
    cases += [(int(rng.integers(1, 1 << 59)), 0, int(rng.integers(1, 1 << 40))) for _ in range(50)]
    cases += [(S, int(rng.integers(0, S + 1)), int(rng.integers(1, 1 << 40)))
              for S in (int(x) for x in rng.integers(1, 1 << 59, 2000))]
    cases += [(S, int(rng.integers(0, S + 1)), int(rng.integers(1, 1 << 30)))
              for S in (int(x) for x in rng.integers(1, 1 << 20, 1000))]
    for S, w, R in cases:
        t = orc.fold_threshold(S, w, R)
        if w == 0:
            assert t == 0
            continue
        exact = Fraction(w * R, (S << 27) + w * R)
        assert abs(exact - Fraction(t, 1 << 64)) < Fraction(1, 1 << 58) * exact + Fraction(1, 1 << 63), (S, w, R)


def test_walk_thresholds_fold_rules(orc):
    """Folding applies when 1/p exceeds env = max(1, 1/q) by a factor below
    2^32 on integer weights; otherwise the plain envelope max(1/p, 1, 1/q)."""
    one = 1 << 27
    p = orc.walk_params(p=0.5, q=2.0)
    assert (p.thr_return, p.thr_neighbor, p.thr_other, p.fold_r) == (one, one, one // 2, one)
    p = orc.walk_params(p=0.25, q=0.5)  # env 2, 1/p = 4: factor 1
    assert (p.thr_return, p.thr_neighbor, p.thr_other, p.fold_r) == (one, one // 2, one, one)
    p = orc.walk_params(p=2.0, q=0.5)  # 1/p below env: plain, lim 2
    assert (p.thr_return, p.thr_neighbor, p.thr_other, p.fold_r) == (one // 4, one // 2, one, 0)
    p = orc.walk_params(p=0.5, q=2.0, float_weights=True)  # plain, lim 2
    assert (p.thr_return, p.thr_neighbor, p.thr_other, p.fold_r) == (one, one // 2, one // 4, 0)
    p = orc.walk_params(p=1e-12, q=1.0)  # factor ~1e12 >= 2^32: plain
    assert p.fold_r == 0 and p.thr_return == one and p.thr_neighbor == 0
    p = orc.walk_params(p=1.0 / (1 + 2.0 ** 31), q=1.0)  # factor 2^31: folded
    assert p.fold_r == 1 << 58


def test_ppr_length_law(orc):
    n = 4
    e = np.array([(u, v, 1) for u in range(1, n + 1) for v in range(1, n + 1)], dtype=orc.EDGE_DT)
    g = orc.OracleGraph(n)
    g.load(e)
    tp, L = 0.15, 12
    p = orc.walk_params(seed=13, tp=tp, max_len=L)
    N = 200_000
    paths, steps = g.walk(orc.PPR, p, nwalkers=N)
    k = (paths >= 0).sum(axis=1) - 1  # steps per walker
    assert k.sum() == steps
    t = p.thr_stop / (1 << 27)
    probs = [(1 - t) ** (j - 1) * t for j in range(1, L - 1)] + [(1 - t) ** (L - 2)]
    cnt = np.bincount(k, minlength=L)[1:]
    assert cnt.sum() == N
    assert _chi2_p(cnt, probs) > 0.01


def test_rmat_quadrants_and_weights(orc):
    N = 200_000
    e = orc.gen_rmat(1, N, orc.W_LOGUNI32, seed=5, undirected=False)
    q = np.zeros(4)
    for (uu, vv), k in zip([(1, 1), (1, 2), (2, 1), (2, 2)], range(4)):
        q[k] = ((e["u"] == uu) & (e["v"] == vv)).sum()
    assert _chi2_p(q, [0.57, 0.19, 0.19, 0.05]) > 0.01
    bl = np.floor(np.log2(e["w"].astype(np.float64))).astype(int) + 1
    assert e["w"].min() >= 1
    assert _chi2_p(np.bincount(bl, minlength=33)[1:], [1 / 32] * 32) > 0.01
    w16 = orc.gen_rmat(4, 50_000, orc.W_UNIFORM16, seed=6)["w"]
    assert w16.min() >= 1 and w16.max() < (1 << 16)
    wz = orc.gen_rmat(4, 50_000, orc.W_ZIPF24, seed=6)["w"]
    assert wz.min() >= 1 and wz.max() <= (1 << 24)
    # P(1) = (0.5^-.2 - 1.5^-.2) / (0.5^-.2 - (2^24+.5)^-.2) = 0.2037
    assert abs((wz == 1).mean() - 0.2037) < 0.01
    assert abs((wz == 2).mean() / (wz == 1).mean() - 0.3952) < 0.03
    und = orc.gen_rmat(6, 1000, orc.W_CONST1, seed=1)
    assert (und["u"][0::2] == und["v"][1::2]).all() and (und["v"][0::2] == und["u"][1::2]).all()
    assert und["u"].min() >= 1 and und["u"].max() <= 64


def test_update_generator(orc):
    base = orc.gen_rmat(8, 2000, orc.W_UNIFORM16, seed=3)
    n = 256
    ug = orc.UpdateGen(n, 8, base, 2000, orc.W_UNIFORM16, seed=3)
    deg = np.bincount(base["u"], minlength=n + 1)
    assert ug.live() == len(base)
    ins = ug.inserts(500)
    assert ug.live() == len(base) + 500
    dels = ug.deletes(1500)
    assert len({(int(d["u"]), int(d["slot"])) for d in dels}) == 1500
    # every delete names a slot that exists at that point
    deg2 = deg + np.bincount(ins["u"], minlength=n + 1)
    assert (dels["slot"] < deg2[dels["u"]]).all()
    g = orc.OracleGraph(n)
    assert g.load(base) == 0 and g.insert(ins) == 0 and g.delete(dels) == 0

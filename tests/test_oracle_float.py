"""Float weights in the oracle (the reference's wtype::Float): splitting a
float into integer part and 32-bit fraction, the fractional-part threshold
against the exact ratio, store state after updates against a brute-force
recomputation, and the one-step law of a draw — integer part through the
alias table, fractional part by rejection — against w / sum(w)."""
from fractions import Fraction

import numpy as np
from scipy import stats

from tests.brute import BruteGraph, check_against_brute, member_law, split_dump, transition_law
from tests.floatw import fbits, fexact, float_edges

M32, M64 = 1 << 32, 1 << 64


def test_split_float(orc):
    for x in [1.0, 2.5, 0.25, 3.01, 1e6 + 0.5, 4294967040.0, 2.0 ** -9, 7.75, 0.01, 123.99]:
        b = fbits(x)
        q = fexact(b)
        ip, fr = orc.split_float(b)
        assert ip == int(q) and fr == int((q - int(q)) * M32), x
    for x in [0.0, -1.0, -0.5, float("nan"), float("inf"), 4294967296.0, 1e-20]:
        assert orc.split_float(fbits(x)) is None, x


def test_frac_threshold_matches_ratio(orc):
    rng = np.random.default_rng(3)
    cases = [(1, 1), (1, M32 - 1), (3, 1), (1 << 59, (1 << 60) - 1), ((1 << 60) - 1, 1), (12345, 1 << 40)]
    cases += [(int(rng.integers(1, 1 << 60)), int(rng.integers(1, 1 << 60))) for _ in range(2000)]
    cases += [(int(rng.integers(1, 1000)), int(rng.integers(1, 1 << 36))) for _ in range(2000)]
    cases += [(int(rng.integers(1, 1 << 60)), int(rng.integers(1, 1 << 8))) for _ in range(500)]
    for S, F in cases:
        t = orc.frac_threshold(S, F)
        exact = Fraction(F, (S << 32) + F)
        assert 0 <= exact - Fraction(t, M64) < Fraction(1, M64), (S, F)
    assert orc.frac_threshold(5, 0) == 0
    assert orc.frac_threshold(0, 7) == M64 - 1


def _split(e, orc):
    ip = np.zeros(len(e), np.int64)
    fr = np.zeros(len(e), np.int64)
    for i, b in enumerate(e["w"]):
        ip[i], fr[i] = orc.split_float(int(b))
    return ip, fr


def test_float_store_after_updates(orc):
    rng = np.random.default_rng(11)
    n = 200
    base = float_edges(rng, n, 3000, hub=7)
    for nbr in (True, False):
        o = orc.OracleGraph(n, nbr, float_weights=True)
        assert o.load(base) == 0
        ins = float_edges(rng, n, 800, hub=7)
        assert o.insert(ins) == 0
        arcs = np.concatenate([base, ins])
        ip, fr = _split(arcs, orc)
        # delete a third of node 7's arcs and 300 others
        d0 = o.export()
        so = np.concatenate([[0], np.cumsum(d0["nslots"].astype(np.int64))])
        picks = [(7, s) for s in rng.choice(int(d0["nslots"][7]), int(d0["nslots"][7]) // 3, replace=False)]
        for u in rng.integers(1, n + 1, 300):
            if u != 7 and d0["nslots"][u]:
                picks.append((int(u), int(rng.integers(0, d0["nslots"][u]))))
        dels = np.array(sorted(set(picks)), orc.DEL_DT)
        dels = dels[rng.permutation(len(dels))]
        assert o.delete(dels) == 0
        d = o.export()
        # integer parts: buckets, back positions, masses as for integer weights
        bg = BruteGraph(n)
        bi = arcs.copy()
        bi["w"] = ip
        bg.insert(bi)
        bg.delete(dels)
        check_against_brute(d, bg)
        # fractions: per slot, per node sum, threshold
        slot = np.zeros(len(arcs), np.int64)
        seen = np.zeros(n + 1, np.int64)
        for i, u in enumerate(arcs["u"]):
            slot[i] = seen[u]
            seen[u] += 1
        dead = {(int(a["u"]), int(a["slot"])) for a in dels}
        want_f = np.zeros(len(d["slot_f"]), np.int64)
        fmass = np.zeros(n + 1, np.int64)
        for i in range(len(arcs)):
            u = int(arcs["u"][i])
            if (u, int(slot[i])) in dead:
                continue
            want_f[so[u] + slot[i]] = fr[i]
            fmass[u] += fr[i]
        assert np.array_equal(d["slot_f"].astype(np.int64), want_f)
        assert np.array_equal(d["fmass"].astype(np.int64), fmass)
        for u in range(n + 1):
            S, F = int(d["mass"][u]), int(fmass[u])
            assert int(d["fthr"][u]) == orc.frac_threshold(S, F)
        # a deleted slot is deleted even when its integer part was 0
        live_frac_only = [(int(arcs["u"][i]), int(slot[i])) for i in range(len(arcs))
                          if ip[i] == 0 and (int(arcs["u"][i]), int(slot[i])) not in dead]
        assert live_frac_only
        assert o.delete(np.array(live_frac_only[:1], orc.DEL_DT)) == 0
        assert o.delete(np.array(live_frac_only[:1], orc.DEL_DT)) == -1  # now dead: rejected
        o.close()


def draw_law(d, u, n):
    """Exact law of one draw on node u from the exported state."""
    nodes = split_dump(d, n)
    sv, sw, groups = nodes[u]
    so = int(np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])[u])
    ns = int(d["nslots"][u])
    fr = [int(x) for x in d["slot_f"][so:so + ns]]
    S, F, thr = int(d["mass"][u]), int(d["fmass"][u]), int(d["fthr"][u])
    p_int = transition_law(d["T"][u], d["alias"][u], groups) if S else {}
    if not F:
        return p_int
    pf = Fraction(1) if not S else Fraction(thr, M64)
    ml = member_law(ns)
    acc = sum(ml[s] * Fraction(fr[s], M32) for s in range(ns))
    miss = float(1 - acc) ** 4096  # every round rejects: the integer draw
    law = {}
    for s in range(ns):
        if fr[s]:
            v = int(sv[s])
            law[v] = law.get(v, 0) + pf * ml[s] * Fraction(fr[s], M32) / acc * (1 - Fraction(miss))
    for v, p in p_int.items():
        law[v] = law.get(v, 0) + (1 - pf + pf * Fraction(miss)) * p
    return law


def test_float_one_step_law(orc):
    rng = np.random.default_rng(5)
    n = 60
    e = float_edges(rng, n, 1500, hub=3)
    o = orc.OracleGraph(n, False, float_weights=True)
    assert o.load(e) == 0
    d = o.export()
    ip, fr = _split(e, orc)
    for u in [3] + [int(x) for x in rng.integers(1, n + 1, 12)]:
        sel = e["u"] == u
        if not sel.any():
            continue
        want = {}
        for v, a, b in zip(e["v"][sel], ip[sel], fr[sel]):
            want[int(v)] = want.get(int(v), 0) + Fraction(int(a)) + Fraction(int(b), M32)
        tot = sum(want.values())
        got = draw_law(d, u, n)
        assert set(got) == set(want), u
        for v in want:
            assert abs(float(got[v] / (want[v] / tot)) - 1.0) < 1e-6, (u, v)
    # the drawn frequencies follow the law (fractional rejection included)
    N = 200_000
    paths, _ = o.walk(orc.SAMPLING, orc.walk_params(seed=9, max_len=2), starts=np.full(N, 3, np.int32))
    law = draw_law(d, 3, n)
    vs = sorted(law)
    cnt = np.bincount(paths[:, 1], minlength=n + 1)[vs]
    exp = np.array([float(law[v]) for v in vs]) * N
    assert stats.chisquare(cnt, exp / exp.sum() * cnt.sum()).pvalue > 0.01
    o.close()


def test_float_fraction_only_node(orc):
    """A node whose arcs all weigh less than 1 is walked through the
    fractional part alone."""
    e = np.zeros(3, np.dtype([("u", np.int32), ("v", np.int32), ("w", np.uint32)]))
    e["u"] = 1
    e["v"] = [2, 3, 4]
    e["w"] = [fbits(0.25), fbits(0.5), fbits(0.125)]
    o = orc.OracleGraph(4, False, float_weights=True)
    assert o.load(e) == 0
    d = o.export()
    assert int(d["mass"][1]) == 0 and int(d["fthr"][1]) == M64 - 1
    N = 120_000
    paths, steps = o.walk(orc.SAMPLING, orc.walk_params(seed=4, max_len=2), starts=np.ones(N, np.int32))
    assert steps == N
    cnt = np.bincount(paths[:, 1], minlength=5)[2:5]
    assert stats.chisquare(cnt, np.array([2, 4, 1]) / 7 * N).pvalue > 0.01
    o.close()

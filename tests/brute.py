"""Brute-force restatement of the store's observable state, used to pin the
oracle (and through it the HIP library).  Everything here is recomputed from
the list of live slots only: which arcs exist, with which weights."""
from __future__ import annotations

from fractions import Fraction

import numpy as np

M64 = 1 << 64
M32 = 1 << 32


class BruteGraph:
    def __init__(self, n):
        self.n = n
        self.slots = [[] for _ in range(n + 1)]  # per node: [v, w] (w = 0 -> deleted)

    def insert(self, edges):
        for e in edges:  # batch order inside a node == slot order
            self.slots[int(e["u"])].append([int(e["v"]), int(e["w"])])

    def delete(self, dels):
        for d in dels:
            self.slots[int(d["u"])][int(d["slot"])][1] = 0

    def live_arcs(self, u):
        return [(s, v, w) for s, (v, w) in enumerate(self.slots[u]) if w]


def dense_of(d, n):
    """Dense-group masks of a dump, [n + 1] (zeros if it has none)."""
    dm = d.get("dmask")
    return np.zeros(n + 1, np.uint32) if dm is None else np.asarray(dm, np.uint32)


def split_dump(d, n):
    """Per-node views of a flat dump: slot arrays and per-group member lists
    (empty for dense groups, which keep no list)."""
    so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
    dm = dense_of(d, n)
    dense = ((dm[:, None] >> np.arange(32, dtype=np.uint32)[None, :]) & 1).astype(bool)
    msz = np.where(dense, 0, d["bsize"].astype(np.int64))
    mo = np.concatenate([[0], np.cumsum(msz.reshape(-1))])
    nodes = []
    for u in range(n + 1):
        sv = d["slot_v"][so[u]:so[u + 1]]
        sw = d["slot_w"][so[u]:so[u + 1]]
        groups = []
        for j in range(32):
            k = u * 32 + j
            groups.append(d["members"][mo[k]:mo[k + 1]])
        nodes.append((sv, sw, groups))
    return nodes


def check_against_brute(d, bg: BruteGraph):
    """Bucket j of node u must hold exactly the live slots with bit j set
    (as (v << 32) | slot), every live slot's back positions must point at its
    own member entries, masses are count_j * 2^j summed.  A dense group
    (dmask bit) holds no list: its size is the count of those slots."""
    n = bg.n
    nodes = split_dump(d, n)
    dm = dense_of(d, n)
    bp = 0
    for u in range(n + 1):
        sv, sw, groups = nodes[u]
        ref = bg.slots[u]
        assert len(sv) == len(ref), (u, len(sv), len(ref))
        assert [int(x) for x in sv] == [v for v, _ in ref], u
        assert [int(x) for x in sw] == [w for _, w in ref], u
        mass = 0
        for j in range(32):
            want = sorted((v << 32) | s for s, (v, w) in enumerate(ref) if w >> j & 1)
            if int(dm[u]) >> j & 1:
                assert len(groups[j]) == 0 and int(np.ravel(d["bsize"])[u * 32 + j]) == len(want), (u, j)
            else:
                got = sorted(int(x) for x in groups[j])
                assert got == want, (u, j)
            mass += len(want) << j
        assert int(d["mass"][u]) == mass, u
        for s, (v, w) in enumerate(ref):
            for j in range(32):
                if w >> j & 1 and not int(dm[u]) >> j & 1:
                    pos = int(d["backpos"][bp])
                    bp += 1
                    assert int(groups[j][pos]) == (v << 32) | s, (u, s, j)
    assert bp == len(d["backpos"])


def alias_law(T, alias):
    """Exact probability of each group implied by a 32-slot table used as
    'slot i uniform, stay iff coin64 < T_i else go to alias_i'."""
    P = [Fraction(0)] * 32
    for i in range(32):
        stay = Fraction(int(T[i]), M64)
        P[i] += stay / 32
        P[int(alias[i])] += (1 - stay) / 32
    return P


def alias_prefix(M):
    """The sweep alias table restated through prefix sums (the form the GPU
    computes lane-parallel): light i (s_i < S) takes the first heavy whose
    cumulative excess reaches the light's cumulative deficit before it; a
    heavy whose cumulative excess E_k falls inside a light's deficit range
    splits: tau = S - (end of that range - E_k), alias = the next heavy."""
    S = sum(int(m) for m in M)
    s = [32 * int(m) for m in M]
    light = [j for j in range(32) if s[j] < S]
    heavy = [j for j in range(32) if s[j] >= S]
    tau, ali = [S] * 32, list(range(32))
    A, acc = {}, 0
    for j in light:
        A[j] = acc
        acc += S - s[j]
    E, acc = [], 0
    for h in heavy:
        acc += s[h] - S
        E.append(acc)
    for i in light:
        k = next(k for k in range(len(heavy)) if A[i] <= E[k])
        ali[i], tau[i] = heavy[k], s[i]
    for k, h in enumerate(heavy):
        cont = [i for i in light if A[i] <= E[k] < A[i] + S - s[i]]
        if cont and k + 1 < len(heavy):
            end = A[cont[0]] + S - s[cont[0]]
            tau[h], ali[h] = S - (end - E[k]), heavy[k + 1]
    T = [M64 - 1 if tau[j] == S else (tau[j] << 64) // S for j in range(32)]
    return T, ali


def member_law(size):
    """Exact probability of each position under floor(x * size / 2^32)."""
    cuts = [-(-(p * M32) // size) for p in range(size + 1)]  # ceil(p*2^32/size)
    return [Fraction(cuts[p + 1] - cuts[p], M32) for p in range(size)]


def transition_law(T, alias, groups):
    """Exact P(target) of one draw on a node from its table and members."""
    P = alias_law(T, alias)
    out = {}
    for j in range(32):
        if P[j] == 0:
            continue
        members = groups[j]
        assert len(members) > 0
        for pos, pm in enumerate(member_law(len(members))):
            v = int(members[pos]) >> 32
            out[v] = out.get(v, 0) + P[j] * pm
    return out

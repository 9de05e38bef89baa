"""Full-size checks on the node2vec workload graph (R-MAT 2^22 nodes, 2^26
undirected edges = 2^27 arcs, log-uniform u32 weights) after a 10^6 + 10^6
update batch: size-independent properties of the whole store, node2vec walks
of a walker sample bit-exact against the oracle, and a hub node's transition
law by chi-square."""
import numpy as np
import pytest
from scipy import stats

import bingo_amd as B

pytestmark = pytest.mark.gpu


def _slots_in_order(u):
    """Per-arc slot ordinal: the k-th arc of a node (array order) is slot k."""
    order = np.argsort(u, kind="stable")
    su = u[order]
    first = np.searchsorted(su, su, side="left")
    slot = np.empty_like(u, dtype=np.int64)
    slot[order] = np.arange(len(u), dtype=np.int64) - first
    return slot


def test_config2_full_scale(orc):
    scale, m = 22, 1 << 26
    n = 1 << scale
    print("[scale] generating", flush=True)
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=2026)
    g = B.Graph(neighbor_index=True)
    g.load(n, base)
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=2026)
    ins, dels = ug.inserts(1_000_000), ug.deletes(1_000_000)
    g.insert(ins)
    g.delete(dels)
    print("[scale] exporting per-node state", flush=True)
    d = g.export_nodes()

    arcs = np.concatenate([base, ins])
    u = arcs["u"].astype(np.int64)
    w = arcs["w"].astype(np.int64)
    slot = _slots_in_order(u)
    key = (u << 32) | slot
    dkey = (dels["u"].astype(np.int64) << 32) | dels["slot"].astype(np.int64)
    order = np.argsort(key)
    pos = order[np.searchsorted(key[order], dkey)]
    assert (key[pos] == dkey).all()
    live = np.ones(len(arcs), bool)
    live[pos] = False
    assert d["nslots"].astype(np.int64).tolist() == np.bincount(u, minlength=n + 1).tolist()
    lu, lw = u[live], w[live]
    lo = np.bincount(lu, weights=(lw & 0xFFFF).astype(np.float64), minlength=n + 1).astype(np.uint64)
    hi = np.bincount(lu, weights=(lw >> 16).astype(np.float64), minlength=n + 1).astype(np.uint64)
    assert np.array_equal(d["mass"], (hi << np.uint64(16)) + lo)
    for j in range(32):
        cnt = np.bincount(lu[(lw >> j) & 1 == 1], minlength=n + 1)
        assert np.array_equal(d["bsize"][:, j], cnt.astype(np.uint32)), j
    assert d["total_members"] == int(np.unpackbits(lw.astype(">u4").view(np.uint8)).sum())

    print("[scale] node2vec sample against the oracle", flush=True)
    nw = 20000
    starts = (np.arange(nw, dtype=np.int64) * (n // nw) % n + 1).astype(np.int32)
    pa, sa = g.walk(B.NODE2VEC, B.walk_params(seed=5, max_len=80), starts=starts)
    o = orc.OracleGraph(n)
    assert o.load(base) == 0 and o.insert(ins) == 0 and o.delete(dels) == 0
    pb, sb = o.walk(B.NODE2VEC, orc.walk_params(seed=5, max_len=80), starts=starts, nthreads=16)
    o.close()
    assert np.array_equal(pa, pb) and sa == sb

    print("[scale] hub chi-square", flush=True)
    hub = int(np.argmax(d["nslots"]))
    N = 2_000_000
    paths, _ = g.walk(B.SAMPLING, B.walk_params(seed=11, max_len=2), starts=np.full(N, hub, np.int32))
    bins = 64
    hv = lu == hub
    tgt = arcs["v"][live][hv].astype(np.int64)
    exp = np.bincount((tgt * 2654435761) % bins, weights=lw[hv].astype(np.float64), minlength=bins)
    got = np.bincount((paths[:, 1].astype(np.int64) * 2654435761) % bins, minlength=bins)
    assert got.sum() == N
    assert stats.chisquare(got, exp / exp.sum() * N).pvalue > 0.01

"""Captured update batches (bingo_update_graph_*): insert and delete batches of
uneven sizes replayed from HIP graphs, one per power-of-two size, against the
same batches through bingo_insert/bingo_delete and the oracle, bit for bit,
with eager walks, captured walks and syncs between the replays; invalid
batches and batches past the committed room apply nothing and are reported
by the next sync (ref: insert() and deleteE() wait for the device mid-batch,
include/operation.cuh:144-251)."""
import random

import numpy as np
import pytest

import bingo_amd as B
from tests.floatw import float_edges

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    for k in b:
        assert a[k].shape == b[k].shape, (what, k)
        if not np.array_equal(a[k], b[k]):
            bad = np.argwhere(a[k] != b[k])[:5]
            raise AssertionError(f"{what}: {k} differs at {bad.tolist()}")


def _live_dels(rng, d, n, k):
    so = np.concatenate([[0], np.cumsum(d["nslots"].astype(np.int64))])
    live = np.flatnonzero(d["slot_w"] != 0) if "slot_f" not in d else np.flatnonzero((d["slot_w"] != 0) |
                                                                                      (d["slot_f"] != 0))
    pick = np.array(rng.sample(list(live), min(k, len(live))), np.int64)
    u = np.searchsorted(so, pick, side="right") - 1
    out = np.zeros(len(pick), B.DEL_DT)
    out["u"], out["slot"] = u, pick - so[u]
    return out


@pytest.mark.parametrize("mode", ["n2v", "dense", "float", "deepwalk"])
def test_replays_match_eager_and_oracle(orc, mode):
    rng = random.Random(hash(mode) & 0xFFFF)
    scale = 11
    n = 1 << scale
    fw = mode == "float"
    nbr = mode != "deepwalk"
    rate = 0.4 if mode == "dense" else 0.0
    if fw:
        base = float_edges(np.random.default_rng(3), n, 30000, hub=5)
    else:
        base = B.gen_rmat(scale, 15000, B.W_LOGUNI32, seed=3)
    kw = dict(neighbor_index=nbr, float_weights=fw, dense_rate=rate)
    g, e = B.Graph(**kw), B.Graph(**kw)  # captured, eager
    g.load(n, base)
    e.load(n, base)
    o = orc.OracleGraph(n, nbr, float_weights=fw, dense_rate=rate)
    assert o.load(base) == 0
    ug_gen = B.UpdateGen(n, scale, base if not fw else B.gen_rmat(scale, 15000, B.W_LOGUNI32, seed=3), 15000,
                         B.W_LOGUNI32, seed=3)
    MAXI, MAXD = 40000, 20000
    d_ins = B.DeviceArray(g, nbytes=MAXI * 12)
    d_dels = B.DeviceArray(g, nbytes=MAXD * 8)
    ug = g.update_graph(d_ins, MAXI, d_dels, MAXD)
    L, nw = 32, 3000
    dp = B.DeviceArray(g, nbytes=nw * L * 4)
    ds = B.DeviceArray(g, np.zeros(1, np.uint64))
    app = B.NODE2VEC if nbr else B.DEEPWALK
    wg = g.walk_graph(app, B.walk_params(seed=6, max_len=L, form=B.FORM_PERSISTENT), nw, dp, ds)
    sizes = [(1, 0), (0, 1), (700, 300), (5000, 4000), (33, 1500), (40000, 20000), (2048, 1025), (0, 0), (9000, 10)]
    for rnd, (ki, kd) in enumerate(sizes):
        print(f"[update graph] {mode} replay {rnd}: {ki} inserts, {kd} deletes", flush=True)
        ins = ug_gen.inserts(ki) if ki else np.zeros(0, B.EDGE_DT)
        if fw:
            ins = float_edges(np.random.default_rng(rnd), n, ki, hub=5) if ki else ins
        dels = _live_dels(rng, o.export(), n, kd) if kd else np.zeros(0, B.DEL_DT)
        # the oracle applies inserts then deletes; deletes name slots live before the batch
        d_ins.write(ins)
        d_dels.write(dels)
        ug.launch(len(ins), len(dels))
        if rnd % 3 == 1:
            wg.launch()  # a captured walk right behind the captured batch
        e.insert(ins)
        e.delete(dels)
        assert o.insert(ins) == 0 and o.delete(dels) == 0
        g.sync()
        a = g.export()
        _same(a, e.export(), f"{mode} replay {rnd} vs eager")
        _same(a, o.export(), f"{mode} replay {rnd} vs oracle")
        want, _ = o.walk(app, orc.walk_params(seed=6, max_len=L, float_weights=fw), nwalkers=nw)
        got, _ = g.walk(app, B.walk_params(seed=6, max_len=L), nwalkers=nw)
        assert np.array_equal(got, want), rnd
        if rnd % 3 == 1:
            assert np.array_equal(dp.to_host(np.int32, (nw, L)), want), rnd
    wg.close()
    ug.close()
    g.close()
    e.close()


def test_failed_batches_apply_nothing(orc):
    """An invalid batch, and then a batch past the committed room, apply
    nothing -- nor do the captured batches behind them -- and the next sync
    reports each; replays work again afterwards."""
    scale = 10
    n = 1 << scale
    base = B.gen_rmat(scale, 4000, B.W_LOGUNI32, seed=9)
    g = B.Graph(neighbor_index=True, pool_slack=0.0)
    g.load(n, base)
    o = orc.OracleGraph(n, True)
    assert o.load(base) == 0
    gen = B.UpdateGen(n, scale, base, 4000, B.W_LOGUNI32, seed=9)
    MAXI = 1 << 17
    d_ins = B.DeviceArray(g, nbytes=MAXI * 12)
    d_dels = B.DeviceArray(g, nbytes=64 * 8)
    ug = g.update_graph(d_ins, MAXI, d_dels, 64, room_bytes=4 << 20)  # 4 MiB of room
    ok = gen.inserts(50)
    d_ins.write(ok)
    ug.launch(50, 0)
    assert o.insert(ok) == 0
    g.sync()
    _same(g.export(), o.export(), "small batch")
    bad = gen.inserts(10)
    bad["v"][3] = n + 5  # invalid arc
    d_ins.write(bad)
    ug.launch(10, 0)
    later = gen.inserts(20)
    d_ins.write(later)  # behind the invalid batch: not applied either
    ug.launch(20, 0)
    with pytest.raises(B.BingoError) as ei:
        g.sync()
    assert ei.value.code == B.EINVAL
    _same(g.export(), o.export(), "after invalid")
    dead = np.array([(int(base["u"][0]), 0)], B.DEL_DT)
    d_dels.write(np.array([(0, 0)], B.DEL_DT))  # node 0: invalid delete
    ug.launch(0, 1)
    with pytest.raises(B.BingoError):
        g.sync()
    _same(g.export(), o.export(), "after invalid delete")
    big = gen.inserts(MAXI)  # ~17 MB of new slots and members: past the room
    d_ins.write(big)
    ug.launch(MAXI, 0)
    with pytest.raises(B.BingoError) as ei:
        g.sync()
    assert ei.value.code == B.ENOMEM
    _same(g.export(), o.export(), "after no room")
    d_ins.write(later)  # replays work again
    ug.launch(20, 0)
    d_dels.write(dead)
    ug.launch(0, 1)
    assert o.insert(later) == 0 and o.delete(dead) == 0
    g.sync()
    _same(g.export(), o.export(), "after recovery")
    ug.close()
    g.close()

"""Failure handling on the GPU, forced from the host: a store that would pass
its HBM budget refuses the batch and keeps its state (pools grown by copy
and pools mapped in place), pools whose address reservation is refused grow
by copy next to mapped ones, and walkers with invalid device-side starts are
reported by bingo_sync."""

import numpy as np
import pytest

import bingo_amd as B
from tests.test_gpu_parity import _same_state

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def _graph_and_oracle(orc, n, base, **kw):
    g = B.Graph(**kw)
    g.load(n, base)
    o = orc.OracleGraph(n, kw.get("neighbor_index", True))
    assert o.load(base) == 0
    return g, o


def test_budget_refusal_keeps_state_copied_pools(orc):
    scale, m = 12, 30000
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=3)
    probe = B.Graph()
    probe.load(n, base)
    need = probe.stats()["bytes_reserved"]
    probe.close()
    print(f"[errors] copy pools, budget {need} bytes", flush=True)
    g, o = _graph_and_oracle(orc, n, base, hbm_budget=need, pool_slack=0.0)
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=3)
    dels = ug.deletes(5000)  # drawn before the refused arcs join the generator's view
    with pytest.raises(B.BingoError) as e:
        g.insert(ug.inserts(200000))
    assert e.value.code == -2  # BINGO_ENOMEM
    _same_state(g, o, "after a refused insert")
    g.delete(dels)  # deletes never grow a pool
    assert o.delete(dels) == 0
    _same_state(g, o, "after deletes")
    g.close()


def test_budget_refusal_keeps_state_mapped_pools(orc):
    scale, m = 12, 30000
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=4)
    budget = 5 * GIB  # four pools of one 1 GiB chunk each, and node records
    print("[errors] mapped pools, 5 GiB budget", flush=True)
    g, o = _graph_and_oracle(orc, n, base, hbm_budget=budget, grow_in_place=True)
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=4)
    small = ug.inserts(20000)  # fits the mapped chunks: no pool grows, no budget check
    g.insert(small)
    assert o.insert(small) == 0
    _same_state(g, o, "after an insert inside the mapped chunks")
    big = B.gen_rmat(scale, 25_000_000, B.W_LOGUNI32, seed=5, edge_begin=10 ** 8, undirected=False)
    with pytest.raises(B.BingoError) as e:  # the member pool would need more 1 GiB chunks
        g.insert(big)
    assert e.value.code == -2
    _same_state(g, o, "after a refused insert")
    assert g.stats()["bytes_reserved"] <= budget
    g.close()


def test_refused_reservation_copies_that_pool(orc, monkeypatch):
    """The back-index pool's address reservation is refused (test hook): it
    grows by copy while the other pools grow in place; state stays exact."""
    monkeypatch.setenv("BINGO_REFUSE_VMM", "2")
    scale, m = 11, 8000
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=6)
    print("[errors] refused reservation of pool 1", flush=True)
    g, o = _graph_and_oracle(orc, n, base, grow_in_place=True, pool_slack=0.0)
    monkeypatch.delenv("BINGO_REFUSE_VMM")
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=6)
    for rnd, k in enumerate([20000, 60000, 150000]):  # the copied pool grows several times
        ins, dels = ug.inserts(k), ug.deletes(k // 3)
        g.insert(ins)
        g.delete(dels)
        assert o.insert(ins) == 0 and o.delete(dels) == 0
        _same_state(g, o, f"round {rnd}")
    pa, sa = g.walk(B.NODE2VEC, B.walk_params(seed=2, max_len=30), nwalkers=n)
    pb, sb = o.walk(B.NODE2VEC, orc.walk_params(seed=2, max_len=30), nwalkers=n)
    assert np.array_equal(pa, pb) and sa == sb
    g.close()


def test_invalid_device_starts_reported_by_sync(orc):
    scale, m = 10, 4000
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=7)
    g, o = _graph_and_oracle(orc, n, base)
    starts = np.arange(1, 1001, dtype=np.int32)
    starts[[3, 500]] = [0, n + 1]  # two walkers outside 1..n
    d_starts = B.DeviceArray(g, starts)
    d_paths = B.DeviceArray(g, nbytes=1000 * 16 * 4)
    d_steps = B.DeviceArray(g, np.zeros(1, np.uint64))
    wp = B.walk_params(seed=4, max_len=16)
    g.walk_device(B.DEEPWALK, wp, 1000, d_paths, d_steps, d_starts=d_starts)
    with pytest.raises(B.BingoError) as e:
        g.sync()
    assert e.value.code == -1  # BINGO_EINVAL
    paths = d_paths.to_host(np.int32, (1000, 16))
    assert (paths[[3, 500]] == -1).all()
    good = np.setdiff1d(np.arange(1000), [3, 500])
    want, _ = o.walk_ids(B.DEEPWALK, orc.walk_params(seed=4, max_len=16), good, starts=starts[good])
    assert np.array_equal(paths[good], want)
    ok = B.DeviceArray(g, np.arange(1, 1001, dtype=np.int32))
    g.walk_device(B.DEEPWALK, wp, 1000, d_paths, d_steps, d_starts=ok)
    g.sync()  # the flag clears when the next walk starts
    g.close()

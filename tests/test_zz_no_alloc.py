"""The allocation-free delete (bingo_delete_noalloc) against bingo_delete:
two handles loaded alike, the same batches through each entry point -- a
batch with long per-node runs, then a smaller one -- leave bit-identical
stores, equal to the oracle's; batches past the limit are refused and
change nothing."""
import numpy as np
import pytest

import bingo_amd as B
from tests.test_gpu_parity import _same_state

pytestmark = pytest.mark.gpu


def _pick(o, rng, k, hub):
    ex = o.export()
    so = np.concatenate([[0], np.cumsum(ex["nslots"].astype(np.int64))])
    live = np.flatnonzero(ex["slot_w"] != 0)
    hub_live = live[(live >= so[hub]) & (live < so[hub + 1])]
    k_hub = min(len(hub_live), k // 4)  # a long run on the hub (per-group delete tasks)
    pick = np.concatenate([rng.choice(hub_live, k_hub, replace=False),
                           rng.choice(np.setdiff1d(live, hub_live), k - k_hub, replace=False)])
    pick.sort()
    u = np.searchsorted(so, pick, side="right") - 1
    d = np.zeros(len(pick), B.DEL_DT)
    d["u"], d["slot"] = u, pick - so[u]
    return d


def test_noalloc_delete_equals_delete(orc):
    scale, m = 12, 60000
    n = 1 << scale
    base = B.gen_rmat(scale, m, B.W_LOGUNI32, seed=31)
    hub = int(np.bincount(base["u"]).argmax())
    print(f"[noalloc] load n={n} arcs={len(base)}", flush=True)
    g1, g2 = B.Graph(neighbor_index=True), B.Graph(neighbor_index=True)
    g1.load(n, base)
    g2.load(n, base)
    o = orc.OracleGraph(n, True)
    assert o.load(base) == 0
    rng = np.random.default_rng(5)
    for rnd, k in enumerate([40000, 3000]):  # the second call is smaller
        d = _pick(o, rng, k, hub)
        print(f"[noalloc] call {rnd}: {k} deletes", flush=True)
        g1.delete_noalloc(d)
        g2.delete(d)
        assert o.delete(d) == 0
        _same_state(g1, g2, f"call {rnd}: noalloc vs delete")
        _same_state(g1, o, f"call {rnd}: noalloc vs oracle")
    with pytest.raises(B.BingoError):
        g1.delete_noalloc(np.zeros(65537, B.DEL_DT))  # past the limit
    with pytest.raises(B.BingoError):  # invalid entries: rejected, nothing changed
        g1.delete_noalloc(np.array([(1, 10 ** 6)], B.DEL_DT))
    _same_state(g1, o, "after refused calls")
    g1.close()
    g2.close()

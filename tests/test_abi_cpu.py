"""CPU-only checks of the C ABI library: it loads, exports every symbol that
include/bingo.h declares, its host-side generators agree with the oracle's,
and handle creation fails cleanly without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bingo_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(bingo_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = B.lib()
    declared = _declared()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(B.SIGNATURES), declared ^ set(B.SIGNATURES)
    assert b"gfx950" in L.bingo_version()


@pytest.mark.parametrize("wkind", [B.W_CONST1, B.W_UNIFORM16, B.W_LOGUNI32, B.W_ZIPF24])
def test_rmat_generator_matches_oracle(orc, wkind):
    for und in (True, False):
        a = B.gen_rmat(12, 5000, wkind, seed=42, undirected=und, edge_begin=777)
        b = orc.gen_rmat(12, 5000, wkind, seed=42, undirected=und, edge_begin=777)
        assert (a.view(np.uint32) == b.view(np.uint32)).all()


def test_update_generator_matches_oracle(orc):
    base = B.gen_rmat(10, 3000, B.W_LOGUNI32, seed=9)
    n = 1 << 10
    ga = B.UpdateGen(n, 10, base, 3000, B.W_LOGUNI32, seed=9)
    gb = orc.UpdateGen(n, 10, base, 3000, B.W_LOGUNI32, seed=9)
    for k in (1, 100, 2500):
        ia, ib = ga.inserts(k), gb.inserts(k)
        assert (ia.view(np.uint32) == ib.view(np.uint32)).all()
        da, db = ga.deletes(k), gb.deletes(k)
        assert (da.view(np.int32) == db.view(np.int32)).all()
        assert ga.live() == gb.live()


def test_update_generator_expiry_is_arrival_order():
    """Sliding-window deletes name the oldest live arcs: load order, then the
    inserts, each by its slot ordinal within its node."""
    n, m = 1 << 8, 1000
    base = B.gen_rmat(8, m, B.W_LOGUNI32, seed=3, undirected=False)
    ug = B.UpdateGen(n, 8, base, m, B.W_LOGUNI32, seed=3)
    ins = ug.inserts(50)
    d = np.concatenate([ug.expire(600), ug.expire(420)])
    arcs = np.concatenate([base, ins])
    slot = np.zeros(len(arcs), np.int64)
    seen = np.zeros(n + 1, np.int64)
    for i, u in enumerate(arcs["u"]):
        slot[i] = seen[u]
        seen[u] += 1
    assert (d["u"] == arcs["u"][:1020]).all() and (d["slot"] == slot[:1020]).all()
    assert ug.live() == 30
    with pytest.raises(B.BingoError):
        ug.deletes(1)  # random deletes do not mix with a sliding window
    with pytest.raises(B.BingoError):
        ug.expire(31)


def test_generator_rejects_bad_arguments():
    out = np.zeros(4, B.EDGE_DT)
    assert B.lib().bingo_gen_rmat(0, 0, 2, 0.57, 0.19, 0.19, 1, 1, 1, out.ctypes.data) == -1
    assert B.lib().bingo_gen_rmat(5, 0, 2, 0.9, 0.19, 0.19, 1, 1, 1, out.ctypes.data) == -1


def test_create_without_gpu_fails_cleanly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = C.c_void_p()
    rc = B.lib().bingo_create(None, C.byref(h))
    assert rc == -3 and not h.value
    assert b"device" in B.lib().bingo_last_error().lower()

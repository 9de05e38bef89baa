"""The reference's driver surface: edge-list loading/weighting (CPU) and the
instance() update-and-walk rounds (GPU), plus the CLI binary."""
import os
import random
import subprocess

import numpy as np
import pytest

import bingo_amd as B


def _write_edgelist(path, pairs):
    with open(path, "w") as f:
        for u, v in pairs:
            f.write(f"{u} {v}\n")


def test_read_edgelist_weights_like_the_reference_loader(tmp_path):
    """ids shifted to 1-based; weight of u->v = outdeg(v) + 1 (graph.cuh:22-54)."""
    rng = random.Random(3)
    pairs = [(rng.randint(0, 40), rng.randint(0, 40)) for _ in range(500)]
    p = tmp_path / "g.txt"
    _write_edgelist(p, pairs)
    n, e = B.read_edgelist(str(p))
    assert n == max(max(u, v) for u, v in pairs) + 1
    deg = np.zeros(n + 1, np.int64)
    for u, _ in pairs:
        deg[u + 1] += 1
    assert e["u"].tolist() == [u + 1 for u, _ in pairs]
    assert e["v"].tolist() == [v + 1 for _, v in pairs]
    assert e["w"].tolist() == [int(deg[v + 1]) + 1 for _, v in pairs]


def test_read_edgelist_rejects_bad_input(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1 2\n-4 5\n")
    with pytest.raises(B.BingoError):
        B.read_edgelist(str(p))
    with pytest.raises(B.BingoError):
        B.read_edgelist(str(tmp_path / "missing.txt"))


@pytest.mark.gpu
@pytest.mark.parametrize("app", [B.DEEPWALK, B.NODE2VEC, B.PPR, B.SAMPLING])
@pytest.mark.parametrize("utype", [B.UPD_INSERT, B.UPD_DELETE, B.UPD_MIXED])
def test_instance_rounds(tmp_path, app, utype):
    und = B.gen_rmat(10, 6000, B.W_CONST1, seed=4)
    p = tmp_path / "rmat10.txt"
    _write_edgelist(p, zip((und["u"] - 1).tolist(), (und["v"] - 1).tolist()))
    print(f"[apps] instance app={app} utype={utype}", flush=True)
    r = B.instance(app, utype, str(p), rounds=3, batch=2000, seed=9, walk_len=20, verbose=False)
    m = len(und)
    k = 2000 if utype != B.UPD_MIXED else 1000
    want = m + 3 * k * (utype != B.UPD_DELETE) - 3 * k * (utype != B.UPD_INSERT)
    assert r["arcs"] == want
    assert r["n"] == int(max(und["u"].max(), und["v"].max()))
    assert r["walk_steps"] > 0 and r["walk_s"] > 0 and r["eval_s"] >= r["walk_s"]


@pytest.mark.gpu
def test_cli_prints_reference_lines(tmp_path):
    und = B.gen_rmat(9, 3000, B.W_CONST1, seed=5)
    p = tmp_path / "small.txt"
    _write_edgelist(p, zip((und["u"] - 1).tolist(), (und["v"] - 1).tolist()))
    cli = os.path.join(os.path.dirname(B.LIB_PATH), "bingo_cli")
    r = subprocess.run([cli, "n", "-g", str(p), "-r", "2", "-b", "500"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for line in ("Graph dataset:", "Edge number: 6000.", "Random walk in", "Evaluation time:", "Memory Consumption:"):
        assert line in r.stderr, r.stderr


def test_instance_update_stream_is_deterministic():
    """The driver's update stream (host only): duplicates of existing arcs at
    the source's next slot, deletes of distinct live arcs, seeded."""
    und = B.gen_rmat(8, 800, B.W_CONST1, seed=6)
    n = 256
    a, b = B.InstanceGen(und, n, B.UPD_MIXED, 200, seed=3), B.InstanceGen(und, n, B.UPD_MIXED, 200, seed=3)
    deg = np.bincount(und["u"], minlength=n + 1)
    seen = set()
    for _ in range(3):
        (ia, da), (ib, db) = a.round(), b.round()
        assert len(ia) == len(da) == 100
        assert (ia.view(np.uint32) == ib.view(np.uint32)).all() and (da.view(np.int32) == db.view(np.int32)).all()
        pairs = set(zip(und["u"].tolist(), und["v"].tolist()))
        assert all((int(u), int(v)) in pairs for u, v in zip(ia["u"], ia["v"]))
        for u in ia["u"]:
            deg[u] += 1
        assert (da["slot"] < deg[da["u"]]).all()
        for u, sl in zip(da["u"].tolist(), da["slot"].tolist()):
            assert (u, sl) not in seen
            seen.add((u, sl))


@pytest.mark.gpu
@pytest.mark.parametrize("app,utype", [(B.NODE2VEC, B.UPD_MIXED), (B.DEEPWALK, B.UPD_INSERT), (B.PPR, B.UPD_DELETE)])
def test_instance_state_matches_oracle(tmp_path, orc, app, utype):
    """instance()'s store after its rounds equals the oracle fed the same
    edge list and the driver's update stream; walks on it agree."""
    und = B.gen_rmat(10, 6000, B.W_CONST1, seed=4)
    p = tmp_path / "rmat10.txt"
    _write_edgelist(p, zip((und["u"] - 1).tolist(), (und["v"] - 1).tolist()))
    print(f"[apps] instance vs oracle app={app} utype={utype}", flush=True)
    r = B.instance(app, utype, str(p), rounds=3, batch=2000, seed=9, walk_len=20, verbose=False, keep_graph=True)
    g = r["graph"]
    n, base = B.read_edgelist(str(p))
    o = orc.OracleGraph(n, app == B.NODE2VEC)
    assert o.load(base) == 0
    gen = B.InstanceGen(base, n, utype, 2000, seed=9)
    for _ in range(3):
        ins, dels = gen.round()
        assert o.insert(ins) == 0 and o.delete(dels) == 0
    a, b = g.export(), o.export()
    for k in b:
        assert np.array_equal(a[k], b[k]), k
    pa, sa = g.walk(app, B.walk_params(seed=5, max_len=20), nwalkers=n)
    pb, sb = o.walk(app, orc.walk_params(seed=5, max_len=20), nwalkers=n)
    assert np.array_equal(pa, pb) and sa == sb
    g.close()

"""Serving loop: walk queries and update batches submitted together are
batched into shared launches, yet every walk returns exactly what it returns
alone on the same graph state (and what the oracle returns)."""
import random

import numpy as np
import pytest

import bingo_amd as B

pytestmark = pytest.mark.gpu


def _requests(n, scale, base, m, rng):
    ug = B.UpdateGen(n, scale, base, m, B.W_LOGUNI32, seed=77)
    reqs = []
    for i in range(40):
        if i % 9 == 8:
            reqs.append(("update", ug.inserts(rng.randint(10, 3000)), ug.deletes(rng.randint(10, 2000))))
            continue
        app = [B.DEEPWALK, B.NODE2VEC, B.PPR, B.SAMPLING][i % 4]
        L = rng.choice([8, 20, 33])
        p = B.walk_params(seed=1000 + i, max_len=L, tp=0.1, p=rng.choice([0.5, 2.0]), q=rng.choice([0.5, 2.0]))
        nw = rng.choice([1, 7, 500, 4000])
        starts = np.array([rng.randint(1, n) for _ in range(nw)], np.int32) if i % 3 else None
        lens = np.array([rng.randint(1, L) for _ in range(nw)], np.int32) if i % 5 == 1 else None
        reqs.append(("walk", app, p, nw, starts, lens))
    return reqs


def _load(n, base):
    g = B.Graph(neighbor_index=True)
    g.load(n, base)
    return g


@pytest.mark.parametrize("arena", [1 << 30, 600_000])
def test_served_results_equal_alone(orc, arena):
    rng = random.Random(5)
    scale = 12
    n = 1 << scale
    base = B.gen_rmat(scale, 20000, B.W_LOGUNI32, seed=77)
    reqs = _requests(n, scale, base, 20000, rng)

    alone = []  # each request run by itself, in submission order
    g1 = _load(n, base)
    o = orc.OracleGraph(n)
    o.load(base)
    for i, r in enumerate(reqs):
        if r[0] == "update":
            g1.insert(r[1])
            g1.delete(r[2])
            o.insert(r[1])
            o.delete(r[2])
            alone.append(None)
        else:
            _, app, p, nw, starts, lens = r
            paths, steps = g1.walk(app, p, starts=starts, lens=lens, nwalkers=nw)
            if i % 4 == 0:
                op = orc.walk_params(seed=p.seed, max_len=p.max_len, tp=p.tp, p=p.p, q=p.q)
                pb, sb = o.walk(app, op, starts=starts, lens=lens, nwalkers=nw)
                assert np.array_equal(paths, pb) and steps == sb
            alone.append((paths, steps))
    g1.close()

    g2 = _load(n, base)
    srv = B.Server(g2, arena_bytes=arena, max_launch_walkers=6000)
    tickets = []
    for r in reqs:
        if r[0] == "update":
            tickets.append(srv.submit_update(r[1], r[2]))
        else:
            _, app, p, nw, starts, lens = r
            tickets.append(srv.submit_walk(app, p, nwalkers=nw, starts=starts, lens=lens))
    for i, t in enumerate(tickets):
        out, steps, lat = srv.wait(t)
        assert lat >= 0
        if alone[i] is not None:
            assert np.array_equal(out, alone[i][0]), i
            assert steps == alone[i][1], i
    st = srv.stats()
    assert st["walk_requests"] == sum(r[0] == "walk" for r in reqs)
    assert st["update_requests"] == sum(r[0] == "update" for r in reqs)
    assert st["launches"] < st["walk_requests"]  # requests shared launches
    assert 0 <= st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    srv.close()
    g2.close()


def test_server_device_outputs_and_admission(orc):
    scale = 11
    n = 1 << scale
    base = B.gen_rmat(scale, 8000, B.W_UNIFORM16, seed=3)
    g = _load(n, base)
    srv = B.Server(g, arena_bytes=1 << 16, max_launch_walkers=1 << 20)
    p = B.walk_params(seed=9, max_len=16)
    # a host-bound output larger than the arena is refused at submit
    with pytest.raises(B.BingoError) as ei:
        srv.submit_walk(B.DEEPWALK, p, nwalkers=5000)
    assert ei.value.code == -2
    # device-resident outputs need no arena
    nw = 5000
    dp = B.DeviceArray(g, nbytes=nw * 16 * 4)
    t = srv.submit_walk(B.DEEPWALK, p, nwalkers=nw, out=dp)
    _, steps, _ = srv.wait(t)
    got = dp.to_host(np.int32, (nw, 16))
    o = orc.OracleGraph(n)
    o.load(base)
    want, ws = o.walk(B.DEEPWALK, orc.walk_params(seed=9, max_len=16), nwalkers=nw)
    assert np.array_equal(got, want) and steps == ws
    # an invalid update fails its own ticket only
    bad = srv.submit_update(deletes=np.array([(1, 10 ** 6)], B.DEL_DT))
    with pytest.raises(B.BingoError):
        srv.wait(bad)
    t2 = srv.submit_walk(B.PPR, p, nwalkers=100)
    out, _, _ = srv.wait(t2)
    assert out.shape == (100, 16)
    assert srv.stats()["rejected"] == 1
    srv.close()
    dp.free()
    g.close()


@pytest.mark.parametrize("apps", [(B.NODE2VEC,), (B.DEEPWALK, B.PPR, B.SAMPLING)], ids=["node2vec", "no_node2vec"])
def test_single_class_launches_equal_alone(orc, apps):
    """Launches holding only node2vec requests, or none, run kernel variants
    of their own; their paths must equal each request's stand-alone walk and
    the oracle's."""
    rng = random.Random(11)
    scale = 11
    n = 1 << scale
    base = B.gen_rmat(scale, 12000, B.W_LOGUNI32, seed=21)
    g = _load(n, base)
    o = orc.OracleGraph(n)
    o.load(base)
    reqs = []
    for i in range(10):
        app = apps[i % len(apps)]
        p = B.walk_params(seed=500 + i, max_len=rng.choice([8, 24]), tp=0.1, p=rng.choice([0.5, 2.0]),
                          q=rng.choice([0.5, 2.0]))
        nw = rng.choice([3, 300, 2500])
        starts = np.array([rng.randint(1, n) for _ in range(nw)], np.int32) if i % 2 else None
        reqs.append((app, p, nw, starts))
    alone = []
    for app, p, nw, starts in reqs:
        paths, steps = g.walk(app, p, starts=starts, nwalkers=nw)
        op = orc.walk_params(seed=p.seed, max_len=p.max_len, tp=p.tp, p=p.p, q=p.q)
        pb, sb = o.walk(app, op, starts=starts, nwalkers=nw)
        assert np.array_equal(paths, pb) and steps == sb
        alone.append((paths, steps))
    srv = B.Server(g, arena_bytes=1 << 28, max_launch_walkers=4000)
    tickets = [srv.submit_walk(app, p, nwalkers=nw, starts=starts) for app, p, nw, starts in reqs]
    for i, t in enumerate(tickets):
        out, steps, _ = srv.wait(t)
        assert np.array_equal(out, alone[i][0]), i
        assert steps == alone[i][1], i
    assert srv.stats()["walk_requests"] == len(reqs)
    srv.close()
    g.close()


def test_served_update_paths_equal_alone(orc, monkeypatch):
    """Served update batches replay the captured graphs when they fit, and
    take bingo_insert / bingo_delete past the captured size (2^17 arcs a
    side) or past the committed room (4 MiB here, through the server's test
    hook); an invalid batch fails its own ticket only.  Every walk returns
    what it returns alone after the same updates, and the final store equals
    the oracle's."""
    monkeypatch.setenv("BINGO_SERVE_ROOM", str(4 << 20))
    scale = 10
    n = 1 << scale
    base = B.gen_rmat(scale, 4000, B.W_LOGUNI32, seed=21)
    gen = B.UpdateGen(n, scale, base, 4000, B.W_LOGUNI32, seed=21)
    ups = {
        "small": (gen.inserts(300), gen.deletes(200)),
        "past_room": (gen.inserts(100_000), gen.deletes(500)),  # ~13 MB of new slots and members
        "invalid": (np.zeros(0, B.EDGE_DT), np.array([(0, 0)], B.DEL_DT)),  # node 0
        "past_size": (gen.inserts((1 << 17) + 5), gen.deletes(100)),
        "after": (gen.inserts(200), gen.deletes(300)),
    }
    seq = []
    for i, name in enumerate(ups):
        for app in (B.DEEPWALK, B.NODE2VEC):
            seq.append(("walk", app, B.walk_params(seed=50 + 2 * i + app, max_len=20), 500))
        seq.append(("update", name))
    seq.append(("walk", B.NODE2VEC, B.walk_params(seed=99, max_len=20), 500))

    def graph():
        g = B.Graph(neighbor_index=True, pool_slack=0.0)
        g.load(n, base)
        return g

    g1 = graph()  # each request alone, in order
    o = orc.OracleGraph(n, True)
    assert o.load(base) == 0
    alone = []
    for r in seq:
        if r[0] == "walk":
            alone.append(g1.walk(r[1], r[2], nwalkers=r[3]))
            continue
        ins, dels = ups[r[1]]
        if r[1] == "invalid":
            with pytest.raises(B.BingoError):
                g1.delete(dels)
        else:
            g1.insert(ins)
            g1.delete(dels)
            assert o.insert(ins) == 0 and o.delete(dels) == 0
        alone.append(None)
    g1.close()

    g2 = graph()
    srv = B.Server(g2, max_launch_walkers=4000)
    tickets = []
    for r in seq:
        if r[0] == "walk":
            tickets.append(srv.submit_walk(r[1], r[2], nwalkers=r[3]))
        else:
            tickets.append(srv.submit_update(*ups[r[1]]))
    for i, (r, t) in enumerate(zip(seq, tickets)):
        if r[0] == "update" and r[1] == "invalid":
            with pytest.raises(B.BingoError) as ei:
                srv.wait(t)
            assert ei.value.code == B.EINVAL
            continue
        out, steps, _ = srv.wait(t)
        if r[0] == "walk":
            assert np.array_equal(out, alone[i][0]) and steps == alone[i][1], i
    srv.close()
    got, want = g2.export(), o.export()
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    g2.close()

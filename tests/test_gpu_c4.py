"""The served workload at full size: on the node2vec workload graph (R-MAT
2^22 nodes / 2^26 undirected edges, log-uniform u32 weights), a Poisson
stream (2000 requests/s) of walk queries of 10^3..10^6 walkers (DeepWalk and
node2vec mixed, their own seeds) and update batches of 10^3..10^5 arcs goes
through the serving loop.  Every query's paths must equal the same query run
alone, in submission order, on a second copy of the graph that sees the same
updates; the first walkers of every query must equal the oracle's."""
import time

import numpy as np
import pytest

import bingo_amd as B

pytestmark = pytest.mark.gpu

SCALE, EDGES, NREQ, RATE, LEN = 22, 1 << 26, 40, 2000.0, 80


def _stream(n, base, seed):
    rng = np.random.default_rng(seed)
    ug = B.UpdateGen(n, SCALE, base, EDGES, B.W_LOGUNI32, seed=seed)
    reqs = []
    for i in range(NREQ):
        if i % 10 == 5:
            k = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))))
            reqs.append(("update", ug.inserts(k // 2), ug.deletes(k - k // 2)))
        else:
            nw = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e6)))))
            app = B.DEEPWALK if rng.random() < 0.5 else B.NODE2VEC
            reqs.append(("walk", app, rng.integers(1, n + 1, size=nw, dtype=np.int32), int(rng.integers(1, 2**31))))
    return reqs, rng.exponential(1.0 / RATE, size=NREQ)


def test_served_stream_full_scale(orc):
    n = 1 << SCALE
    print("[c4] generating", flush=True)
    base = B.gen_rmat(SCALE, EDGES, B.W_LOGUNI32, seed=2026)
    reqs, gaps = _stream(n, base, seed=11)
    g1, g2 = B.Graph(neighbor_index=True), B.Graph(neighbor_index=True)
    for g in (g1, g2):
        d = B.DeviceArray(g, base)
        g.load_device(n, d, len(base))
        g.sync()
        d.free()

    print("[c4] serving", flush=True)
    outs = [B.DeviceArray(g1, nbytes=len(r[2]) * LEN * 4) if r[0] == "walk" else None for r in reqs]
    srv = B.Server(g1, arena_bytes=1 << 30)
    tickets, due = [], time.perf_counter()
    for r, gap, out in zip(reqs, gaps, outs):
        due += gap
        while time.perf_counter() < due:
            pass
        if r[0] == "walk":
            tickets.append(srv.submit_walk(r[1], B.walk_params(seed=r[3], max_len=LEN), starts=r[2], out=out))
        else:
            tickets.append(srv.submit_update(r[1], r[2]))
    steps = [srv.wait(t)[1] for t in tickets]
    st = srv.stats()
    srv.close()
    print(f"[c4] served {st['walk_requests']} walks, {st['update_requests']} updates in {st['launches']} launches, "
          f"p99 {st['p99_ms']:.1f} ms", flush=True)

    print("[c4] alone runs and the oracle", flush=True)
    o = orc.OracleGraph(n, True)
    assert o.load(base) == 0
    d_paths = B.DeviceArray(g2, nbytes=max(len(r[2]) for r in reqs if r[0] == "walk") * LEN * 4)
    d_steps = B.DeviceArray(g2, np.zeros(1, np.uint64))
    for i, (r, out) in enumerate(zip(reqs, outs)):
        if r[0] == "update":
            g2.insert(r[1])
            g2.delete(r[2])
            assert o.insert(r[1]) == 0 and o.delete(r[2]) == 0
            continue
        nw = len(r[2])
        wp = B.walk_params(seed=r[3], max_len=LEN)
        B._check(B.lib().bingo_memcpy_h2d(g2._h, d_steps.ptr, np.zeros(1, np.uint64).ctypes.data, 8))
        ds = B.DeviceArray(g2, r[2])
        g2.walk_device(r[1], wp, nw, d_paths, d_steps, d_starts=ds)
        g2.sync()
        ds.free()
        alone = d_paths.to_host(np.int32, (nw, LEN))
        served = out.to_host(np.int32, (nw, LEN))
        assert np.array_equal(served, alone), f"request {i} (app {r[1]}, {nw} walkers) differs from its alone run"
        assert steps[i] == int(d_steps.to_host(np.uint64, (1,))[0]), i
        k = min(nw, 500)
        want, _ = o.walk(r[1], orc.walk_params(seed=r[3], max_len=LEN), starts=r[2][:k], nthreads=16)
        assert np.array_equal(served[:k], want), f"request {i}: first walkers differ from the oracle"
        out.free()
    o.close()
    g1.close()
    g2.close()

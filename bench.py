"""Benchmark of the bias-factorized walk engine on one MI355X.

A step is one round of the dynamic-graph workload: one insert batch, one
delete batch, then one walk of every walker, all on HBM-resident inputs.

Default: the served form (`c4`) line, then the headline line of the scale
workload `c5` (R-MAT 2^25 nodes / 2^30 arcs, u32 weights log-uniform in
[1, 2^32), 3.2e7 DeepWalk walkers of length 80, a sliding window of 10^7
inserts + 10^7 expiries of the oldest arcs and a compaction per step).
`value` is whole-job walk steps/s (walk steps / wall time of the K steps,
updates and compaction included; the default K = 25 spans four compaction
cycles, whose cost comes in bursts); the line also carries walk-only steps/s,
updates/s, the walk kernel's roofline, the plain per-launch form on the same
graph, and the CPU oracle (built -march=native) timed on this host's cores:
load, update rounds and walks of a bounded sample, medians of 3.  `c2` is the
node2vec workload (R-MAT 2^22 / 2^26 undirected edges, 10^6 + 10^6 updates).  `c3` is the batched workload (R-MAT 2^23 / 2^27,
Zipf weights, 10^7 PPR walkers with lengths in [5, 200], update batches of
log-uniform size), replayed from a captured HIP graph; `c1` the small one.

    python bench.py [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5] [--no-cpu]
                    [--app deepwalk|node2vec|ppr|sampling] [--float-weights] [--dense-rate R]

`--app` runs a configuration's graph and schedule with another walk app;
`--float-weights` turns its weights into float32 values w / 256 (the
reference's wtype::Float: integer part through the bit groups, fraction by
rejection), for graph, updates and CPU baseline alike.  `--dense-rate R`
stores with the reference's sparse/dense bit groups (RATE): a group of a
node holding at least R of the node's first arcs keeps only its count and is
drawn from by rejection over the node's slots.

`c4` is the served form: a Poisson stream of walk queries and update
batches through the serving loop, with p50/p99 latency, and a load curve
(the same stream at further rates, `--curve-rates`) giving the highest rate
at which the served form and a one-at-a-time plain server keep p99 of
1M-step queries within 10 ms.  `c5` is the scale
workload: 2^30 R-MAT arcs over 2^25 nodes (log-uniform u32 weights) in a
store fitted to the graph at load, 3.2e7 DeepWalk walkers of length 80, and
a sliding window of 10^7 inserts + 10^7 expiries of the oldest arcs per step;
its CPU baseline runs on the first 2^26 arcs of the same stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walk steps/s + edge updates/s on 1 MI355X (HBM GB/s %peak); p99 query latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes (per walk step / draw / probe): node record 8, alias slot
# 8, group header 8, member 4, target 4, path write 4; a neighbour probe 8; a
# node2vec fold lookup 16 (the return arc's neighbour entry; the node's mass
# comes coded in the alias slot, its exact value read about once in 10^4)
B_STEP, B_PROBE, B_FOLD = 36, 8, 16
# a dense-group rejection trial reads one slot weight (8 B per trial: the
# weight's line share and the slot index)
B_DENSE = 8
APP_NAMES = ["deepwalk", "node2vec", "ppr", "sampling"]

CONFIGS = {
    "c2": dict(workload="rmat22_e26_loguni32_node2vec_upd1e6", scale=22, edges=1 << 26, wkind=2, app=1,
               max_len=80, batch=1_000_000, p=0.5, q=2.0, tp=0.0125, walkers=None, lens=None, graph=False),
    "c1": dict(workload="rmat17_e21_uniform16_deepwalk_upd1e4", scale=17, edges=1 << 21, wkind=1, app=0,
               max_len=80, batch=10_000, p=0.5, q=2.0, tp=0.0125, walkers=None, lens=None, graph=False),
    "c3": dict(workload="rmat23_e27_zipf24_ppr015_len5to200_w1e7", scale=23, edges=1 << 27, wkind=3, app=2,
               max_len=200, batch=(1_000, 1_000_000), p=0.5, q=2.0, tp=0.15, walkers=10_000_000, lens=(5, 200),
               graph=True),
    # 2^30 R-MAT arcs (one arc per generated edge), store fitted to the graph
    # at load, sparse/dense groups at the reference's default RATE 0.4;
    # sliding-window churn: 10^7 new arcs in, the 10^7 oldest out; each node
    # fitted with 1/8 room for its arcs, pools as whole allocations with 6 %
    # room beyond the load: the window then runs without any pool growth (and
    # its update kernels run faster than on pools mapped for growth in place)
    "c5": dict(workload="rmat25_a30_loguni32_deepwalk_w32e6_churn1e7", scale=25, edges=1 << 30, wkind=2, app=0,
               max_len=80, batch=10_000_000, p=0.5, q=2.0, tp=0.0125, walkers=32_000_000, lens=None, graph=False,
               directed=True, fit=True, churn="expire", cpu_sample_edges=1 << 26, grow_in_place=False,
               fit_headroom=1 / 8, pool_slack=0.06, compact=1 / 16, dense_rate=0.4),
}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def profile_figure(name, workload):
    """A figure `name[workload] = value` recorded in a rocprofv3 summary
    under profiles/, or None."""
    import glob
    import re

    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(f):
            m = re.match(r"(\w+)\[(.+)\] = (\S+)", line.strip())
            if m and m.group(1) == name and m.group(2) == workload:
                return float(m.group(3))
    return None


def measured_traffic(workload):
    """HBM bytes per walk launch from the rocprofv3 FETCH_SIZE + WRITE_SIZE
    runs summarised under profiles/ (copy-calibrated), or None."""
    return profile_figure("traffic_bytes_per_launch", workload)


def walker_lens(cfg, nw, seed):
    if not cfg["lens"]:
        return None
    lo, hi = cfg["lens"]
    return np.random.default_rng(seed).integers(lo, hi + 1, size=nw, dtype=np.int32)


def batch_sizes(cfg, k, seed):
    b = cfg["batch"]
    if isinstance(b, int):
        return [b] * k
    lo, hi = np.log(b[0]), np.log(b[1])
    return [int(round(np.exp(x))) for x in np.random.default_rng(seed + 1).uniform(lo, hi, size=k)]


def to_float_bits(edges):
    """u32 weights w -> IEEE-754 bits of float32(w / 256): a log-uniform u32
    law becomes a float law over [2^-8, 2^24) whose values mostly have a
    fractional part."""
    e = edges.copy()
    e["w"] = (edges["w"].astype(np.float64) / 256.0).astype(np.float32).view(np.uint32)
    return e


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(B, base, n, cfg, seed, walk_budget_s=4.0, runs=3):
    """The CPU oracle (a C/OpenMP restatement of the same store, updates and
    sampler), built -march=native on this host and run on all of its cores:
    the graph load, `runs` update rounds (insert batch, delete batch and, where
    the configuration compacts, the compaction) and `runs` walks of a bounded
    walker sample, each timed alone; the medians are reported.  The graph is
    the configuration's own or, where the configuration names a sample, the
    first arcs of the same generator stream, with update batches scaled by
    the sample's share of the arcs."""
    from oracle import oracle as orc

    flags = orc.use_native()
    threads = len(os.sched_getaffinity(0))
    graph_note = "the same graph"
    batch = cfg["batch"] if isinstance(cfg["batch"], int) else int(round(np.sqrt(cfg["batch"][0] * cfg["batch"][1])))
    m_full = cfg["edges"] * (1 if cfg.get("directed") else 2)
    if cfg.get("cpu_sample_edges"):
        k = cfg["cpu_sample_edges"]
        base = B.gen_rmat(cfg["scale"], k, cfg["wkind"], seed=seed, undirected=not cfg.get("directed"))
        batch = max(1, int(batch * len(base) / m_full))
        graph_note = (f"the first {len(base)} arcs of the same R-MAT stream and node space (the full graph's "
                      f"CPU store would not fit host memory), update batches scaled to {batch}")
    if cfg.get("float"):
        base = to_float_bits(base)
    t0 = time.perf_counter()
    o = orc.OracleGraph(n, neighbor_index=cfg["app"] == 1, float_weights=bool(cfg.get("float")),
                        dense_rate=cfg.get("dense_rate", 0.0))
    assert o.load(base) == 0
    load_s = time.perf_counter() - t0
    log(f"cpu oracle ({flags}) loaded {len(base)} arcs in {load_s:.1f}s ({threads} threads)")
    ug = B.UpdateGen(n, cfg["scale"], base, len(base) // (1 if cfg.get("directed") else 2), cfg["wkind"], seed=seed)
    upd_s = []
    for r in range(runs):
        ins = ug.inserts(batch)
        if cfg.get("float"):
            ins = to_float_bits(ins)
        dels = ug.expire(batch) if cfg.get("churn") == "expire" else ug.deletes(batch)
        if cfg.get("compact"):
            ug.compact(cfg["compact"])
        t0 = time.perf_counter()
        assert o.insert(ins) == 0 and o.delete(dels) == 0
        if cfg.get("compact"):
            o.compact(cfg["compact"])
        upd_s.append(time.perf_counter() - t0)
    del ug
    wp = orc.walk_params(seed=seed, p=cfg["p"], q=cfg["q"], tp=cfg["tp"], max_len=cfg["max_len"],
                         float_weights=bool(cfg.get("float")))
    nw = 4096
    while True:  # calibrate the walker sample to about walk_budget_s
        starts = (np.arange(nw, dtype=np.int64) * (n // nw) % n + 1).astype(np.int32)
        t0 = time.perf_counter()
        o.walk(cfg["app"], wp, starts=starts, lens=walker_lens(cfg, nw, seed), nthreads=threads)
        dt = time.perf_counter() - t0
        if dt > walk_budget_s / 4 or nw >= n:
            break
        nw = min(n, nw * 4)
    walk_rates = []
    for r in range(runs):
        t0 = time.perf_counter()
        _, steps = o.walk(cfg["app"], wp, starts=starts, lens=walker_lens(cfg, nw, seed), nthreads=threads)
        walk_rates.append(steps / (time.perf_counter() - t0))
    o.close()
    walk = float(np.median(walk_rates))
    upd = 2 * batch / float(np.median(upd_s))
    return {"value": walk, "unit": "walk steps/s", "cores": threads, "kind": "port",
            "updates_per_s": upd, "load_arcs_per_s": len(base) / load_s,
            "cpu_model": cpu_model(), "build": flags, "runs": runs, "statistic": "median",
            "sample": f"{nw} {APP_NAMES[cfg['app']]} walkers (max length {cfg['max_len']}) over {graph_note}; "
                      f"{steps} steps per walk run, update rounds of {batch} inserts + {batch} "
                      f"{'expiries' if cfg.get('churn') == 'expire' else 'deletes'}"
                      f"{' + compaction' if cfg.get('compact') else ''}; {threads} threads"}


def served(B, args):
    """Served form on the c2 graph: a Poisson stream of walk queries (10^3 to
    10^6 walkers, DeepWalk/node2vec mixed, length 80) with update batches
    (10^3 to 10^5 arcs) among them, submitted to the serving loop at
    `--rate` requests/s; outputs stay in HBM.  Reports latency percentiles
    (submit -> completion) and throughput, next to the same requests run one
    at a time as plain per-launch walks, and the latencies a server running
    them that way (one at a time, arrival order) would give under the same
    arrivals."""
    cfg = CONFIGS["c2"]
    scale, m = cfg["scale"], cfg["edges"]
    n = 1 << scale
    rng = np.random.default_rng(args.seed + 7)
    base = B.gen_rmat(scale, m, cfg["wkind"], seed=args.seed)
    ug = B.UpdateGen(n, scale, base, m, cfg["wkind"], seed=args.seed)
    nreq = args.requests
    reqs = []
    for i in range(nreq):
        if rng.random() < args.update_share:
            k = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))))
            reqs.append(("update", ug.inserts(k // 2), ug.deletes(k - k // 2)))
        else:
            nw = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e6)))))
            app = B.DEEPWALK if rng.random() < 0.5 else B.NODE2VEC
            starts = rng.integers(1, n + 1, size=nw, dtype=np.int32)
            reqs.append(("walk", app, starts, int(rng.integers(1, 2 ** 31))))
    gaps = rng.exponential(1.0 / args.rate, size=nreq)

    def fresh_graph():
        g = B.Graph(device=0, neighbor_index=True)
        d_base = B.DeviceArray(g, base)
        g.load_device(n, d_base, len(base))
        g.sync()
        d_base.free()
        return g

    def serve_pass(rate):
        """one pass of the stream through a server at `rate` requests/s (the
        same arrivals, time-scaled), on a freshly loaded graph"""
        g = fresh_graph()
        outs = [B.DeviceArray(g, nbytes=len(r[2]) * 80 * 4) if r[0] == "walk" else None for r in reqs]
        srv = B.Server(g, arena_bytes=1 << 30, max_launch_walkers=args.launch_walkers, streams=args.streams)
        log(f"served: {nreq} requests at {rate} req/s")
        tickets = []
        t0 = time.perf_counter()
        due = t0
        for r, gap, out in zip(reqs, gaps * (args.rate / rate), outs):
            due += gap
            while time.perf_counter() < due:
                pass
            if r[0] == "walk":
                tickets.append(srv.submit_walk(r[1], B.walk_params(seed=r[3], max_len=80), starts=r[2], out=out))
            else:
                tickets.append(srv.submit_update(r[1], r[2]))
        submit_span = time.perf_counter() - t0
        res = {"lat_per_mstep": [], "steps_total": 0, "lat_1m": [], "req_steps": [0] * nreq,
               "submit_span": submit_span}
        for i, (r, t) in enumerate(zip(reqs, tickets)):
            _, steps, lat = srv.wait(t)
            if r[0] == "walk":
                res["req_steps"][i] = steps
                res["steps_total"] += steps
                res["lat_per_mstep"].append(lat / max(steps, 1) * 1e6)
                if 0.5e6 <= steps <= 2e6:  # queries of about one million steps
                    res["lat_1m"].append(lat)
        res["wall"] = time.perf_counter() - t0
        res["st"] = srv.stats()
        srv.close()
        for o in outs:
            if o:
                o.free()
        g.close()
        return res

    # served form
    main = serve_pass(args.rate)
    lat_per_mstep, steps_total, lat_1m = main["lat_per_mstep"], main["steps_total"], main["lat_1m"]
    req_steps, submit_span, wall, st = main["req_steps"], main["submit_span"], main["wall"], main["st"]

    # plain form: the same requests one at a time, one plain launch each
    g = fresh_graph()
    dpaths = B.DeviceArray(g, nbytes=max(len(r[2]) for r in reqs if r[0] == "walk") * 80 * 4)
    dsteps = B.DeviceArray(g, np.zeros(1, np.uint64))
    svc = []  # service time of each request run alone (ms)
    t1 = time.perf_counter()
    for r in reqs:
        ta = time.perf_counter()
        if r[0] == "walk":
            ds = B.DeviceArray(g, r[2])
            g.walk_device(r[1], B.walk_params(seed=r[3], max_len=80, form=B.FORM_PLAIN), len(r[2]), dpaths, dsteps,
                          d_starts=ds)
            g.sync()
            ds.free()
        else:
            g.insert(r[1])
            g.delete(r[2])
        svc.append((time.perf_counter() - ta) * 1e3)
    plain_wall = time.perf_counter() - t1
    # the plain form as a server: one request at a time in arrival order, each
    # taking its measured service time, under the same Poisson arrivals
    def plain_server(rate):
        arrive = np.cumsum(gaps * (args.rate / rate)) * 1e3
        free_at, lat, lat1m = 0.0, [], []
        for i, r in enumerate(reqs):
            free_at = max(free_at, arrive[i]) + svc[i]
            if r[0] == "walk":
                lat.append(free_at - arrive[i])
                if 0.5e6 <= req_steps[i] <= 2e6:
                    lat1m.append(free_at - arrive[i])
        return lat, lat1m

    plain_lat, plain_1m = plain_server(args.rate)
    plain_steps = int(dsteps.to_host(np.uint64, (1,))[0])
    g.close()

    # load curve: the same stream at other rates through the server (a pass
    # each) and through the plain server (from the measured service times);
    # the highest rate at which each keeps p99 of 1M-step queries <= 10 ms
    def p99(x):
        return float(np.percentile(x, 99)) if x else None

    curve = []
    for r in sorted(set(args.curve_rates) | {args.rate}):
        res = main if r == args.rate else serve_pass(r)
        curve.append({"rate_per_s": r, "achieved_rate_per_s": nreq / res["wall"],
                      "served_walk_steps_per_s": res["steps_total"] / res["wall"],
                      "p50_ms_1M_step_queries": float(np.percentile(res["lat_1m"], 50)) if res["lat_1m"] else None,
                      "p99_ms_1M_step_queries": p99(res["lat_1m"]),
                      "plain_server_p99_ms_1M_step_queries": p99(plain_server(r)[1])})
    served_ok = [c["rate_per_s"] for c in curve if c["p99_ms_1M_step_queries"] is not None
                 and c["p99_ms_1M_step_queries"] <= 10.0 and c["achieved_rate_per_s"] >= 0.95 * c["rate_per_s"]]
    plain_ok = 0.0
    for r in np.arange(25.0, 4000.0 + 1e-9, 25.0):  # the plain server's p99 grows with the rate
        q = p99(plain_server(float(r))[1])
        if q is None or q > 10.0:
            break
        plain_ok = float(r)
    log(f"p99 of 1M-step queries <= 10 ms up to {max(served_ok, default=0)} req/s served (of the rates run), "
        f"{plain_ok} req/s plain server")
    lp = np.percentile(lat_per_mstep, [50, 99]) if lat_per_mstep else [None, None]
    nupd = sum(len(r[1]) + len(r[2]) for r in reqs if r[0] == "update")
    out = {
        "metric": METRIC,
        "value": steps_total / wall,
        "unit": "walk steps/s",
        "n_gpus": 1, "steps": nreq, "warmup": 0, "ms_per_step": wall * 1e3 / nreq,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "served_rmat22_e26_loguni32_mixed_poisson", "nodes": n, "arcs": int(len(base)),
                   "requests": nreq, "offered_rate_per_s": args.rate, "walkers_per_query": [1000, 1000000],
                   "update_batch": [1000, 100000], "update_share": args.update_share, "walk_len": 80,
                   "apps": ["deepwalk", "node2vec"],
                   "launch_walkers": args.launch_walkers, "streams": args.streams},
        "achieved_rate_per_s": nreq / wall,
        "submit_span_s": submit_span,
        "p50_ms": st["p50_ms"], "p90_ms": st["p90_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
        "p50_ms_per_1M_steps": float(lp[0]), "p99_ms_per_1M_steps": float(lp[1]),
        "queries_0p5M_to_2M_steps": len(lat_1m),
        "p50_ms_1M_step_queries": float(np.percentile(lat_1m, 50)) if lat_1m else None,
        "p99_ms_1M_step_queries": float(np.percentile(lat_1m, 99)) if lat_1m else None,
        "launches": st["launches"], "walk_requests": st["walk_requests"], "update_requests": st["update_requests"],
        "updates_applied": nupd,
        "plain_form_walk_steps_per_s": plain_steps / plain_wall,
        "served_vs_plain": (steps_total / wall) / (plain_steps / plain_wall),
        "plain_server_p50_ms": float(np.percentile(plain_lat, 50)),
        "plain_server_p99_ms": float(np.percentile(plain_lat, 99)),
        "plain_server_p99_ms_1M_step_queries": float(np.percentile(plain_1m, 99)) if plain_1m else None,
        "load_curve": curve,
        "max_rate_p99_1M_step_le_10ms": {"served": max(served_ok, default=0), "plain_server": plain_ok},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def copy_calibration(B, mib):
    nbytes = mib << 20
    g = B.Graph(device=0)
    src = B.DeviceArray(g, np.arange(nbytes // 8, dtype=np.uint64))
    dst = B.DeviceArray(g, nbytes=nbytes)
    ms = 0.0
    for i in range(3):
        g.event(0)
        B._check(B.lib().bingo_memcpy_d2d(g._h, dst.ptr, src.ptr, nbytes))
        g.event(1)
        ms = g.elapsed_ms(0, 1)
        log(f"copy {i}: {nbytes} bytes in {ms:.3f} ms = {2 * nbytes / ms / 1e6:.1f} GB/s (read+write)")
    ok = bool((dst.to_host(np.uint64, (nbytes // 8,)) == np.arange(nbytes // 8, dtype=np.uint64)).all())
    print(json.dumps({"copy_bytes": nbytes, "kernel": "k_copy16", "ms": ms, "correct": ok}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 25 steps hold four compaction cycles of the scale workload: its nodes
    # reach the 1/16 threshold together, every 6 to 7 window steps
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default=None, choices=sorted(CONFIGS) + ["c4"],
                    help="one workload; default: the served form (c4) line, then the scale workload (c5) line")
    ap.add_argument("--rate", type=float, default=2000.0, help="c4: offered requests per second")
    ap.add_argument("--requests", type=int, default=300, help="c4: number of requests")
    ap.add_argument("--update-share", type=float, default=0.1, metavar="F",
                    help="c4: share of the requests that are update batches")
    ap.add_argument("--curve-rates", type=lambda x: [float(v) for v in x.split(",") if v], default=[500.0, 1000.0, 1500.0],
                    help="c4: further offered rates (req/s, comma-separated) for the load curve")
    ap.add_argument("--streams", type=int, default=2, help="c4: server launches in flight at once")
    ap.add_argument("--launch-walkers", type=int, default=1 << 22,
                    help="c4: the serving loop's walkers per launch (larger requests are chunked)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--app", choices=APP_NAMES, default=None, help="walk app instead of the configuration's")
    ap.add_argument("--float-weights", action="store_true", help="float32 weights w / 256 (wtype::Float)")
    ap.add_argument("--dense-rate", type=float, default=None, metavar="R",
                    help="sparse/dense bit groups at rate R (the reference's RATE); 0: every group listed; "
                         "default: the config's")
    ap.add_argument("--symmetric", action="store_true", help="emit both arcs of every generated edge")
    ap.add_argument("--capture-updates", type=int, choices=(0, 1), default=None,
                    help="update batches replayed from captured HIP graphs (1) or called one by one (0); "
                         "default: the config's (captured where its walks are)")
    ap.add_argument("--pool-slack", type=float, default=None,
                    help="pool capacity beyond the load's need (fraction; default: the workload's)")
    ap.add_argument("--fit-headroom", type=float, default=None,
                    help="fitted stores: per-node room at load (fraction of its arcs; default: the workload's)")
    ap.add_argument("--grow-in-place", type=int, choices=(0, 1), default=None,
                    help="pools grow by mapping more HBM (1) or by copying (0); default: the config's choice")
    ap.add_argument("--compact", type=float, default=None, metavar="FRAC",
                    help="compaction threshold (tombstones per slot) after every step; default: the config's")
    ap.add_argument("--schedule", default="", metavar="BLOCK,WPL,DEPTH",
                    help="fixed persistent walk knobs instead of autotuning (e.g. 512,2,16)")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's plain copy kernel over MIB MiB (traffic-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the store is built and walked on a single MI355X; "
              "multi-GPU sharding is a design note only. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import bingo_amd as B

    if args.copy_calibration:
        copy_calibration(B, args.copy_calibration)
        return
    if args.config == "c4":
        served(B, args)
        return
    if args.config is None:  # the served line first, the headline (scale workload) last
        served(B, args)
        args.config = "c5"

    cfg = dict(CONFIGS[args.config])
    if args.app and APP_NAMES.index(args.app) != cfg["app"]:
        cfg["app"] = APP_NAMES.index(args.app)
        cfg["workload"] += f"_app_{args.app}"
    if args.compact is not None:
        cfg["compact"] = args.compact or None
    if args.float_weights:
        cfg["float"] = True
        cfg["workload"] += "_float32w"
    if args.dense_rate is not None:
        cfg["dense_rate"] = args.dense_rate
    if cfg.get("dense_rate"):
        cfg["workload"] += f"_rate{cfg['dense_rate']:g}"
    if args.symmetric and cfg.get("directed"):
        cfg["directed"] = False
        cfg["workload"] = cfg["workload"].replace("_a30_", "_e30_sym_")
    scale, m = cfg["scale"], cfg["edges"]
    n = 1 << scale
    K, W = args.steps, args.warmup
    nw = cfg["walkers"] or n
    app = cfg["app"]

    t0 = time.perf_counter()
    base = B.gen_rmat(scale, m, cfg["wkind"], seed=args.seed, undirected=not cfg.get("directed"))
    log(f"generated {len(base)} arcs in {time.perf_counter() - t0:.1f}s")
    fit = dict(fit_load=True, fit_headroom=cfg.get("fit_headroom", 1.0 / 16),
               pool_slack=cfg.get("pool_slack", 0.03)) if cfg.get("fit") else {}
    if args.pool_slack is not None:
        fit["pool_slack"] = args.pool_slack
    if args.fit_headroom is not None and fit:
        fit["fit_headroom"] = args.fit_headroom
    gip = cfg.get("grow_in_place") if args.grow_in_place is None else bool(args.grow_in_place)
    g = B.Graph(device=0, neighbor_index=app == B.NODE2VEC, float_weights=bool(cfg.get("float")),
                grow_in_place=gip, dense_rate=cfg.get("dense_rate", 0.0), **fit)
    d_base = B.DeviceArray(g, to_float_bits(base) if cfg.get("float") else base)
    g.load_device(n, d_base, len(base))
    g.sync()
    load_ms = g.stats()["load_ms"]
    d_base.free()
    narcs = len(base)
    log(f"loaded on GPU in {load_ms:.1f} ms; store {g.stats()['bytes_reserved'] / 1e9:.1f} GB")

    sizes = batch_sizes(cfg, W + K, args.seed)
    ug = B.UpdateGen(n, scale, base, m, cfg["wkind"], seed=args.seed)
    if cfg.get("cpu_sample_edges"):
        del base  # the CPU baseline regenerates its own sample
        base = None
    batches = []
    for kb in sizes:
        ins = ug.inserts(kb)
        if cfg.get("float"):
            ins = to_float_bits(ins)
        dels = ug.expire(kb) if cfg.get("churn") == "expire" else ug.deletes(kb)
        if cfg.get("compact"):
            ug.compact(cfg["compact"])  # the store compacts at the same point of the stream
        batches.append((B.DeviceArray(g, ins), B.DeviceArray(g, dels), kb))
    del ug
    log(f"staged {W + K} update batches, sizes {sizes}")
    lens = walker_lens(cfg, nw, args.seed)
    d_lens = B.DeviceArray(g, lens) if lens is not None else None
    d_paths = B.DeviceArray(g, nbytes=nw * cfg["max_len"] * 4)
    d_steps = B.DeviceArray(g, np.zeros(1, np.uint64))
    kw = dict(seed=args.seed, p=cfg["p"], q=cfg["q"], tp=cfg["tp"], max_len=cfg["max_len"])
    if args.schedule:
        blk, wpl, dep = (int(x) for x in args.schedule.split(","))
        wp = B.walk_params(form=B.FORM_PERSISTENT, block=blk, walkers_per_lane=wpl, lds_depth=dep, **kw)
        tuned = dict(form=B.FORM_PERSISTENT, block=blk, walkers_per_lane=wpl, lds_depth=dep)
    else:
        wp = B.walk_params(**kw)
        tuned = g.tune(app, wp, sample=nw)  # at the workload's own walker count
    log(f"walk schedule: {tuned}")
    wgraph = g.walk_graph(app, wp, nw, d_paths, d_steps, d_lens=d_lens) if cfg["graph"] else None
    cap_upd = bool(cfg["graph"]) if args.capture_updates is None else bool(args.capture_updates)
    ugraph = None
    if cap_upd:  # one captured insert + delete per step, inputs copied into its buffers
        kmax = max(sizes)
        u_ins, u_dels = B.DeviceArray(g, nbytes=kmax * 12), B.DeviceArray(g, nbytes=kmax * 8)
        ugraph = g.update_graph(u_ins, kmax, u_dels, kmax)

    def step(i):
        di, dd, kb = batches[i]
        if ugraph:
            g.event(0)
            B._check(B.lib().bingo_memcpy_d2d(g._h, u_ins.ptr, di.ptr, kb * 12))
            B._check(B.lib().bingo_memcpy_d2d(g._h, u_dels.ptr, dd.ptr, kb * 8))
            ugraph.launch(kb, kb)
            g.event(2)
            g.event(1)
        else:
            g.event(0)
            g.insert_device(di, kb)
            g.event(1)
            g.delete_device(dd, kb)
            g.event(2)
        if cfg.get("compact"):
            g.compact(cfg["compact"], counts=False)
        g.event(6)
        if wgraph:
            wgraph.launch()
        else:
            g.walk_device(app, wp, nw, d_paths, d_steps, d_lens=d_lens)
        g.event(3)
        st = g.stats()
        st["compact_ms"] = g.elapsed_ms(2, 6)
        if ugraph:  # the replay's own insert and delete phases
            return st["insert_ms"], st["delete_ms"], g.elapsed_ms(6, 3), st
        return g.elapsed_ms(0, 1), g.elapsed_ms(1, 2), g.elapsed_ms(6, 3), st

    for i in range(W):
        step(i)
        log(f"warmup step {i} done")
    g.sync()
    B._check(B.lib().bingo_memcpy_h2d(g._h, d_steps.ptr, np.zeros(1, np.uint64).ctypes.data, 8))
    ins_ms = del_ms = walk_ms = kern_ms = comp_ms = 0.0
    store_trace = []
    attempts = probes = folds = dtrials = 0
    t0 = time.perf_counter()
    for i in range(K):
        a, b, c, st = step(W + i)
        ins_ms += a
        del_ms += b
        comp_ms += st["compact_ms"]
        store_trace.append(st["bytes_reserved"])
        walk_ms += c
        kern_ms += st["walk_kernel_ms"]
        attempts += st["walk_attempts"]
        probes += st["walk_probes"]
        folds += st["walk_folds"]
        dtrials += st["walk_dense_trials"]
        log(f"step {i}: insert {batches[W + i][2]} in {a:.2f} ms, delete {batches[W + i][2]} in {b:.2f} ms"
            f"{' (+ compaction %.2f ms)' % st['compact_ms'] if cfg.get('compact') else ''}, walk {c:.2f} ms, "
            f"store {st['bytes_reserved'] / 1e9:.1f} GB")
    g.sync()
    wall = time.perf_counter() - t0
    steps = int(d_steps.to_host(np.uint64, (1,))[0])
    st = g.stats()
    nupd = 2 * sum(b[2] for b in batches[W:])

    # the plain per-launch form (one thread per walker) on the same graph
    wplain = B.walk_params(form=B.FORM_PLAIN, **kw)
    d_s2 = B.DeviceArray(g, np.zeros(1, np.uint64))
    plain_ms = []
    for _ in range(2):
        g.event(4)
        g.walk_device(app, wplain, nw, d_paths, d_s2, d_lens=d_lens)
        g.event(5)
        plain_ms.append(g.elapsed_ms(4, 5))
    plain_steps = int(d_s2.to_host(np.uint64, (1,))[0]) // 2
    log(f"plain form walk: {min(plain_ms):.2f} ms for {plain_steps} steps")

    # random-read ceiling on this store: independent 32-byte (alias slot) and
    # 8-byte (member, neighbour probe) reads over the member pool; a walk
    # launch's draws and probes at these rates take ceiling_ms
    rates = {}
    for width in (32, 8, 32, 8):  # second pass counts
        nr, ms = g.gather_probe(width)
        rates[width] = nr / (ms / 1e3)
    # a node2vec or float-weight draw reads the whole 32-byte slot (two
    # 16-byte loads); other draws read its first 16 bytes (one load, priced
    # as an 8-byte read) and rarely the second locator
    slot_rate = rates[32] if (app == 1 or cfg.get("float")) else rates[8]
    ceiling_s = attempts / K * (1 / slot_rate + 1 / rates[8]) + (probes + folds + dtrials) / K / rates[8]
    log(f"random reads: {rates[32] / 1e9:.1f} G/s (32 B), {rates[8] / 1e9:.1f} G/s (8 B); "
        f"walk at {ceiling_s / (kern_ms / K / 1e3):.2f} of that ceiling")
    # the update kernels' ceiling: independent random read-modify-writes of
    # 8-byte words (each a line read and a 32-byte write request)
    for _ in range(2):
        nr, ms = g.gather_probe(0)
        rates[0] = nr / (ms / 1e3)
    log(f"random read-modify-writes: {rates[0] / 1e9:.1f} G/s (8 B)")

    bytes_per_launch = (B_STEP * attempts + B_PROBE * probes + B_FOLD * folds + B_DENSE * dtrials) / K
    kern_s = kern_ms / K / 1e3
    achieved = bytes_per_launch / kern_s / 1e9
    traffic = measured_traffic(cfg["workload"])
    walk_rate = steps / (walk_ms / 1e3)
    plain_rate = plain_steps / (min(plain_ms) / 1e3)
    out = {
        "metric": METRIC,
        "value": steps / wall,
        "unit": "walk steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "nodes": n, "arcs": int(narcs), "walkers": nw,
                   "walk_len": cfg["lens"] or cfg["max_len"], "app": APP_NAMES[app],
                   "update_batch": cfg["batch"], "p": cfg["p"], "q": cfg["q"], "tp": cfg["tp"],
                   "hip_graph": cfg["graph"], "captured_updates": cap_upd, "directed": bool(cfg.get("directed")),
                   "deletes": "oldest arcs (sliding window)" if cfg.get("churn") == "expire" else "uniform live arcs",
                   "store": "fitted at load" if cfg.get("fit") else "power-of-two regions",
                   "pools": "grown in place" if gip else "grown by copy",
                   "pool_slack": fit.get("pool_slack"), "fit_headroom": fit.get("fit_headroom"),
                   "weights": "float32 w/256" if cfg.get("float") else "u32",
                   "dense_rate": cfg.get("dense_rate", 0.0)},
        "walk_steps_per_s": walk_rate,
        "updates_per_s": nupd / ((ins_ms + del_ms + comp_ms) / 1e3),
        "insert_ms": ins_ms / K,
        "delete_ms": del_ms / K,
        "compact_ms": comp_ms / K if cfg.get("compact") else None,
        "store_bytes_first_last": [store_trace[0], store_trace[-1]] if store_trace else None,
        "walk_ms": walk_ms / K,
        "load_ms": load_ms,
        "walk_draws_per_step": attempts / max(steps, 1),
        "walk_fold_lookups_per_step": folds / max(steps, 1),
        "walk_dense_trials_per_step": dtrials / max(steps, 1),
        "walk_schedule": tuned,
        "plain_form_walk_steps_per_s": plain_rate,
        "batched_vs_plain": walk_rate / plain_rate,
        "store_bytes": st["bytes_reserved"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_GBps": traffic / kern_s / 1e9 if traffic else None,
                     "l2_hit_rate": profile_figure("l2_hit_rate", cfg["workload"]),
                     "kernel": "k_walk_persist" if tuned.get("form") == B.FORM_PERSISTENT else "k_walk",
                     "kernel_ms": kern_ms / K, "algorithmic_bytes_per_launch": bytes_per_launch},
        "gather_ceiling": {"reads32_per_s": rates[32], "reads8_per_s": rates[8], "rmw8_per_s": rates[0],
                           "kernel_ms_at_ceiling": ceiling_s * 1e3, "frac": ceiling_s / kern_s},
        "cpu_baseline": None,
    }
    if wgraph:
        wgraph.close()
    if ugraph:
        ugraph.close()
        u_ins.free()
        u_dels.free()
    for di, dd, _ in batches:
        di.free()
        dd.free()
    d_paths.free()
    g.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(B, base, n, cfg, args.seed)
        out["gpu_vs_cpu_walk"] = out["walk_steps_per_s"] / out["cpu_baseline"]["value"]
        out["gpu_vs_cpu_updates"] = out["updates_per_s"] / out["cpu_baseline"]["updates_per_s"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

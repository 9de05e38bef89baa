"""bingo_amd: Python view of the gfx950 bias-factorized random-walk engine.

Everything here calls the C ABI of ``bingo_amd/libbingo.so`` (declared in
``include/bingo.h``) through ctypes.  There is no CPU fallback: if the HIP
library is missing or no GPU is present, the calls raise.

Mapping to the reference (Joeyw02/Bingo):
    Graph.load      build()       include/operation.cuh:33-121
    Graph.insert    insert()      include/operation.cuh:144-197
    Graph.delete    deleteE()     include/operation.cuh:215-251
    Graph.walk      randomWalk()  include/graph.cuh:93-150
    gen_rmat / UpdateGen          the generators of include/graph.cuh:16-92
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbingo.so")

DEEPWALK, NODE2VEC, PPR, SAMPLING = 0, 1, 2, 3
APPS = {"deepwalk": DEEPWALK, "node2vec": NODE2VEC, "ppr": PPR, "sampling": SAMPLING}
W_CONST1, W_UNIFORM16, W_LOGUNI32, W_ZIPF24 = 0, 1, 2, 3

EDGE_DT = np.dtype([("u", np.int32), ("v", np.int32), ("w", np.uint32)])
DEL_DT = np.dtype([("u", np.int32), ("slot", np.int32)])

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ESTATE", -5: "EUNSUPPORTED"}
EINVAL, ENOMEM, EDEVICE, ESTATE, EUNSUPPORTED = -1, -2, -3, -4, -5


class BingoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("neighbor_index", C.c_int32),
        ("hbm_budget", C.c_uint64),
        ("pool_slack", C.c_double),
        ("min_slots", C.c_int32),
        ("fit_load", C.c_int32),
        ("fit_headroom", C.c_double),
        ("float_weights", C.c_int32),
        ("grow_in_place", C.c_int32),
        ("dense_rate", C.c_double),
    ]


class WalkParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("p", C.c_double),
        ("q", C.c_double),
        ("tp", C.c_double),
        ("max_len", C.c_int32),
        ("max_attempts", C.c_int32),
        ("form", C.c_int32),
        ("block", C.c_int32),
        ("walkers_per_lane", C.c_int32),
        ("lds_depth", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("load_ms", C.c_double),
        ("insert_ms", C.c_double),
        ("delete_ms", C.c_double),
        ("walk_ms", C.c_double),
        ("walk_kernel_ms", C.c_double),
        ("walk_steps", C.c_uint64),
        ("walk_attempts", C.c_uint64),
        ("walk_probes", C.c_uint64),
        ("walk_folds", C.c_uint64),
        ("arcs", C.c_uint64),
        ("bytes_reserved", C.c_uint64),
        ("pool_used", C.c_uint64 * 4),
        ("compact_ms", C.c_double),
        ("walk_dense_trials", C.c_uint64),
    ]


class Dump(C.Structure):
    _fields_ = [
        ("total_slots", C.c_uint64),
        ("total_members", C.c_uint64),
        ("nslots", C.c_void_p),
        ("mass", C.c_void_p),
        ("bsize", C.c_void_p),
        ("T", C.c_void_p),
        ("alias", C.c_void_p),
        ("slot_v", C.c_void_p),
        ("slot_w", C.c_void_p),
        ("members", C.c_void_p),
        ("backpos", C.c_void_p),
        ("slot_f", C.c_void_p),
        ("fmass", C.c_void_p),
        ("fthr", C.c_void_p),
        ("dmask", C.c_void_p),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("arena_bytes", C.c_uint64),
        ("max_launch_walkers", C.c_int64),
        ("blocks_per_cu", C.c_int32),
        ("streams", C.c_int32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("requests", C.c_uint64),
        ("walk_requests", C.c_uint64),
        ("update_requests", C.c_uint64),
        ("rejected", C.c_uint64),
        ("walk_steps", C.c_uint64),
        ("launches", C.c_uint64),
        ("p50_ms", C.c_double),
        ("p90_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("busy_s", C.c_double),
    ]


class InstanceOpts(C.Structure):
    _fields_ = [
        ("rounds", C.c_int32),
        ("batch", C.c_int64),
        ("seed", C.c_uint64),
        ("walk_len", C.c_int32),
        ("device", C.c_int32),
        ("print", C.c_int32),
        ("keep_graph", C.c_int32),
    ]


class InstanceResult(C.Structure):
    _fields_ = [
        ("walk_s", C.c_double),
        ("eval_s", C.c_double),
        ("memory_gb", C.c_double),
        ("walk_steps", C.c_uint64),
        ("arcs", C.c_int64),
        ("n", C.c_int32),
        ("graph", C.c_void_p),
    ]


# every symbol include/bingo.h declares, with its ctypes signature
_vp, _i32, _i64, _u64, _int = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_int
SIGNATURES = {
    "bingo_config_default": (None, [C.POINTER(Config)]),
    "bingo_walk_params_default": (None, [C.POINTER(WalkParams)]),
    "bingo_create": (_int, [C.POINTER(Config), C.POINTER(_vp)]),
    "bingo_destroy": (None, [_vp]),
    "bingo_last_error": (C.c_char_p, []),
    "bingo_version": (C.c_char_p, []),
    "bingo_load": (_int, [_vp, _i32, _vp, _i64]),
    "bingo_load_device": (_int, [_vp, _i32, _vp, _i64]),
    "bingo_insert": (_int, [_vp, _vp, _i64]),
    "bingo_insert_device": (_int, [_vp, _vp, _i64]),
    "bingo_delete": (_int, [_vp, _vp, _i64]),
    "bingo_delete_device": (_int, [_vp, _vp, _i64]),
    "bingo_delete_noalloc": (_int, [_vp, _vp, _i64]),
    "bingo_walk": (_int, [_vp, _int, C.POINTER(WalkParams), _vp, _vp, _i64, _vp, C.POINTER(_u64)]),
    "bingo_walk_device": (_int, [_vp, _int, C.POINTER(WalkParams), _vp, _vp, _i64, _vp, _vp]),
    "bingo_walk_tune": (_int, [_vp, _int, C.POINTER(WalkParams), _i64, C.POINTER(WalkParams)]),
    "bingo_walk_graph_create": (_int, [_vp, _int, C.POINTER(WalkParams), _vp, _vp, _i64, _vp, _vp, C.POINTER(_vp)]),
    "bingo_walk_graph_launch": (_int, [_vp]),
    "bingo_walk_graph_destroy": (None, [_vp]),
    "bingo_update_graph_create": (_int, [_vp, _vp, _i64, _vp, _i64, _u64, C.POINTER(_vp)]),
    "bingo_update_graph_launch": (_int, [_vp, _i64, _i64]),
    "bingo_update_graph_destroy": (None, [_vp]),
    "bingo_num_nodes": (_i32, [_vp]),
    "bingo_stats_get": (_int, [_vp, C.POINTER(Stats)]),
    "bingo_sync": (_int, [_vp]),
    "bingo_stream": (_vp, [_vp]),
    "bingo_edge_count": (_int, [_vp, _vp, _vp, _i64, _vp]),
    "bingo_export": (_int, [_vp, C.POINTER(Dump)]),
    "bingo_dev_alloc": (_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "bingo_dev_free": (_int, [_vp, _vp]),
    "bingo_memcpy_h2d": (_int, [_vp, _vp, _vp, C.c_size_t]),
    "bingo_memcpy_d2h": (_int, [_vp, _vp, _vp, C.c_size_t]),
    "bingo_memcpy_d2d": (_int, [_vp, _vp, _vp, C.c_size_t]),
    "bingo_gather_probe": (_int, [_vp, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "bingo_event_record": (_int, [_vp, _int]),
    "bingo_event_elapsed": (_int, [_vp, _int, _int, C.POINTER(C.c_float)]),
    "bingo_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "bingo_server_create": (_int, [_vp, C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "bingo_server_submit_walk": (_int, [_vp, _int, C.POINTER(WalkParams), _vp, _vp, _i64, _vp, _int, C.POINTER(_u64)]),
    "bingo_server_submit_update": (_int, [_vp, _vp, _i64, _vp, _i64, C.POINTER(_u64)]),
    "bingo_server_wait": (_int, [_vp, _u64, C.POINTER(_u64), C.POINTER(C.c_double)]),
    "bingo_server_stats_get": (_int, [_vp, C.POINTER(ServerStats)]),
    "bingo_server_destroy": (None, [_vp]),
    "bingo_instance_opts_default": (None, [C.POINTER(InstanceOpts)]),
    "bingo_read_edgelist": (_int, [C.c_char_p, C.POINTER(_i32), _vp, _i64, C.POINTER(_i64)]),
    "bingo_instance": (_int, [_int, _int, C.c_char_p, C.POINTER(InstanceOpts), C.POINTER(InstanceResult)]),
    "bingo_gen_rmat": (_int, [_int, _i64, _i64, C.c_double, C.c_double, C.c_double, _int, _u64, _int, _vp]),
    "bingo_updgen_create": (_int, [_i32, _int, _vp, _i64, _i64, C.c_double, C.c_double, C.c_double, _int, _u64,
                                   C.POINTER(_vp)]),
    "bingo_updgen_destroy": (None, [_vp]),
    "bingo_updgen_inserts": (_int, [_vp, _i64, _vp]),
    "bingo_updgen_deletes": (_int, [_vp, _i64, _vp]),
    "bingo_updgen_expire": (_int, [_vp, _i64, _vp]),
    "bingo_updgen_live": (_i64, [_vp]),
    "bingo_instgen_create": (_int, [_vp, _i64, _i32, _int, _i64, _u64, C.POINTER(_vp)]),
    "bingo_instgen_round": (_int, [_vp, _vp, C.POINTER(_i64), _vp, C.POINTER(_i64)]),
    "bingo_instgen_live": (_i64, [_vp]),
    "bingo_instgen_destroy": (None, [_vp]),
    "bingo_updgen_compact": (_int, [_vp, C.c_double]),
    "bingo_compact": (_int, [_vp, C.c_double, C.POINTER(_u64), C.POINTER(_u64)]),
}

_lib = None


def lib():
    """Load libbingo.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m bingo_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise BingoError(rc, lib().bingo_last_error().decode())
    return rc


FORM_AUTO, FORM_PLAIN, FORM_PERSISTENT = 0, 1, 2


def float_bits(w):
    """float32 weights -> the uint32 bit patterns a float_weights graph takes as bingo_edge.w"""
    return np.ascontiguousarray(w, np.float32).view(np.uint32)


def walk_params(seed=1, p=0.5, q=2.0, tp=0.0125, max_len=80, max_attempts=4096, form=FORM_AUTO, block=0,
                walkers_per_lane=0, lds_depth=-1, blocks_per_cu=0) -> WalkParams:
    wp = WalkParams()
    lib().bingo_walk_params_default(C.byref(wp))
    wp.seed, wp.p, wp.q, wp.tp = seed, p, q, tp
    wp.max_len, wp.max_attempts, wp.form, wp.block = max_len, max_attempts, form, block
    wp.walkers_per_lane, wp.lds_depth, wp.blocks_per_cu = walkers_per_lane, lds_depth, blocks_per_cu
    return wp


def gen_rmat(scale, nedges, wkind, seed, undirected=True, edge_begin=0, abc=(0.57, 0.19, 0.19)):
    out = np.zeros(nedges * (2 if undirected else 1), EDGE_DT)
    _check(lib().bingo_gen_rmat(scale, edge_begin, nedges, *abc, wkind, seed, int(undirected), _ptr(out)))
    return out


def read_edgelist(path):
    """Parse and weight a 0-based "u v" edge list the way the reference's loader does."""
    n, m = C.c_int32(0), C.c_int64(0)
    _check(lib().bingo_read_edgelist(path.encode(), C.byref(n), None, 0, C.byref(m)))
    out = np.zeros(m.value, EDGE_DT)
    _check(lib().bingo_read_edgelist(path.encode(), C.byref(n), _ptr(out), m.value, C.byref(m)))
    return n.value, out


def instance(app, utype, dataset, rounds=10, batch=100000, seed=1, walk_len=80, device=0, verbose=True,
             keep_graph=False):
    """The reference's instance(app, utype, dataset) (include/apps.h:4).
    keep_graph: the result's "graph" is a Graph holding the state after the rounds."""
    o = InstanceOpts()
    lib().bingo_instance_opts_default(C.byref(o))
    o.rounds, o.batch, o.seed, o.walk_len, o.device, o.print = rounds, batch, seed, walk_len, device, int(verbose)
    o.keep_graph = int(keep_graph)
    r = InstanceResult()
    _check(lib().bingo_instance(app, utype, dataset.encode(), C.byref(o), C.byref(r)))
    out = {k: getattr(r, k) for k, _ in InstanceResult._fields_}
    out["graph"] = Graph.adopt(r.graph, r.n) if r.graph else None
    return out


class InstanceGen:
    """The instance() driver's update stream (host only)."""

    def __init__(self, base_edges, n, utype, batch, seed):
        base = np.ascontiguousarray(base_edges, EDGE_DT)
        h = C.c_void_p()
        _check(lib().bingo_instgen_create(_ptr(base), len(base), n, utype, batch, seed, C.byref(h)))
        self._h = h
        self.per = batch // 2 if utype == UPD_MIXED else batch

    def round(self):
        ins, dels = np.zeros(self.per, EDGE_DT), np.zeros(self.per, DEL_DT)
        ni, nd = C.c_int64(0), C.c_int64(0)
        _check(lib().bingo_instgen_round(self._h, _ptr(ins), C.byref(ni), _ptr(dels), C.byref(nd)))
        return ins[:ni.value], dels[:nd.value]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().bingo_instgen_destroy(self._h)
            self._h = None


UPD_INSERT, UPD_DELETE, UPD_MIXED = 0, 1, 2


class UpdateGen:
    """Seeded insert/delete stream over a loaded arc list (slots in load order)."""

    def __init__(self, n, scale, base_edges, edge_begin, wkind, seed, abc=(0.57, 0.19, 0.19)):
        base = np.ascontiguousarray(base_edges, EDGE_DT)
        h = C.c_void_p()
        _check(lib().bingo_updgen_create(n, scale, _ptr(base), len(base), edge_begin, *abc, wkind, seed, C.byref(h)))
        self._h = h

    def inserts(self, k):
        out = np.zeros(k, EDGE_DT)
        _check(lib().bingo_updgen_inserts(self._h, k, _ptr(out)))
        return out

    def deletes(self, k):
        out = np.zeros(k, DEL_DT)
        _check(lib().bingo_updgen_deletes(self._h, k, _ptr(out)))
        return out

    def expire(self, k):
        """The k oldest live arcs (sliding-window deletes)."""
        out = np.zeros(k, DEL_DT)
        _check(lib().bingo_updgen_expire(self._h, k, _ptr(out)))
        return out

    def live(self):
        return lib().bingo_updgen_live(self._h)

    def compact(self, min_dead_frac):
        """Renumber slots as Graph.compact(min_dead_frac) renumbers the store's."""
        _check(lib().bingo_updgen_compact(self._h, min_dead_frac))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().bingo_updgen_destroy(self._h)
            self._h = None


def _need(arr, nbytes, what):
    """A device buffer the kernels will touch must hold nbytes: an undersized
    one would be written out of bounds on the GPU."""
    if arr is None:
        return
    if arr.ptr is None or arr.nbytes < nbytes:
        raise BingoError(-1, f"{what}: device buffer of {arr.nbytes} bytes, the call touches {nbytes}")


class DeviceArray:
    """A device allocation owned through a graph handle (HBM-resident inputs)."""

    def __init__(self, graph: "Graph", host: np.ndarray | None = None, nbytes: int | None = None):
        self.g = graph
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        p = C.c_void_p()
        _check(lib().bingo_dev_alloc(graph._h, self.nbytes, C.byref(p)))
        self.ptr = p
        if host is not None and self.nbytes:
            _check(lib().bingo_memcpy_h2d(graph._h, self.ptr, _ptr(np.ascontiguousarray(host)), self.nbytes))

    def write(self, host: np.ndarray):
        """Copy host data to the start of the buffer (ordered on the graph's stream)."""
        h = np.ascontiguousarray(host)
        if h.nbytes > self.nbytes:
            raise BingoError(-1, f"write of {h.nbytes} bytes into a device buffer of {self.nbytes}")
        if h.nbytes:
            _check(lib().bingo_memcpy_h2d(self.g._h, self.ptr, _ptr(h), h.nbytes))

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().bingo_memcpy_d2h(self.g._h, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().bingo_dev_free(self.g._h, self.ptr)
            self.ptr = None


class WalkGraph:
    """A captured walk launch (reset + walk kernel) replayed with hipGraphLaunch."""

    def __init__(self, graph, app, params, nwalkers, d_paths, d_steps, d_starts=None, d_lens=None):
        h = C.c_void_p()
        _check(lib().bingo_walk_graph_create(graph._h, app, C.byref(params), d_starts.ptr if d_starts else None,
                                             d_lens.ptr if d_lens else None, nwalkers, d_paths.ptr, d_steps.ptr,
                                             C.byref(h)))
        self._h = h
        self.g = graph  # the store must outlive its captured walks

    def launch(self):
        _check(lib().bingo_walk_graph_launch(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().bingo_walk_graph_destroy(self._h)
            self._h = None

    __del__ = close


class UpdateGraph:
    """Captured update batches (insert then delete) replayed with hipGraphLaunch:
    launch(nins, ndels) applies d_ins[:nins] and d_dels[:ndels] as they are
    when the launch runs, without waiting for the host."""

    def __init__(self, graph, d_ins, max_ins, d_dels, max_dels, room_bytes=0):
        if max_ins:
            _need(d_ins, max_ins * EDGE_DT.itemsize, "update graph inserts")
        if max_dels:
            _need(d_dels, max_dels * DEL_DT.itemsize, "update graph deletes")
        h = C.c_void_p()
        _check(lib().bingo_update_graph_create(graph._h, d_ins.ptr if d_ins else None, max_ins,
                                               d_dels.ptr if d_dels else None, max_dels, room_bytes, C.byref(h)))
        self._h = h
        self.g = graph

    def launch(self, nins, ndels):
        _check(lib().bingo_update_graph_launch(self._h, nins, ndels))

    def close(self):
        if getattr(self, "_h", None):
            lib().bingo_update_graph_destroy(self._h)
            self._h = None

    __del__ = close


class Server:
    """Serving loop on a loaded graph: submit walks and update batches from any
    thread, collect results with wait().  The graph must not be used directly
    while the server runs."""

    def __init__(self, graph, arena_bytes=4 << 30, max_launch_walkers=1 << 22, blocks_per_cu=0, streams=2):
        cfg = ServerConfig()
        lib().bingo_server_config_default(C.byref(cfg))
        cfg.arena_bytes, cfg.max_launch_walkers, cfg.blocks_per_cu = arena_bytes, max_launch_walkers, blocks_per_cu
        cfg.streams = streams
        h = C.c_void_p()
        _check(lib().bingo_server_create(graph._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self.g = graph
        self._keep = {}  # host output buffers of pending requests

    def submit_walk(self, app, params: WalkParams, nwalkers=None, starts=None, lens=None, out=None):
        """out: a numpy int32 [nwalkers][max_len] array (filled before
        completion) or a DeviceArray (written in HBM).  Returns a ticket."""
        if starts is not None:
            starts = np.ascontiguousarray(starts, np.int32)
            nwalkers = len(starts)
        if lens is not None:
            lens = np.ascontiguousarray(lens, np.int32)
        on_dev = isinstance(out, DeviceArray)
        if on_dev:
            _need(out, nwalkers * params.max_len * 4, "served walk paths")
        if out is None:
            out = np.empty((nwalkers, params.max_len), np.int32)
        ptr = out.ptr if on_dev else _ptr(out)
        t = C.c_uint64(0)
        _check(lib().bingo_server_submit_walk(self._h, app, C.byref(params), _ptr(starts), _ptr(lens), nwalkers, ptr,
                                              int(on_dev), C.byref(t)))
        self._keep[t.value] = out
        return t.value

    def submit_update(self, inserts=None, deletes=None):
        ins = np.ascontiguousarray(inserts if inserts is not None else np.zeros(0, EDGE_DT), EDGE_DT)
        dels = np.ascontiguousarray(deletes if deletes is not None else np.zeros(0, DEL_DT), DEL_DT)
        t = C.c_uint64(0)
        _check(lib().bingo_server_submit_update(self._h, _ptr(ins), len(ins), _ptr(dels), len(dels), C.byref(t)))
        return t.value

    def wait(self, ticket):
        """Returns (output, steps, latency_ms); raises on a failed request."""
        steps, lat = C.c_uint64(0), C.c_double(0)
        rc = lib().bingo_server_wait(self._h, ticket, C.byref(steps), C.byref(lat))
        out = self._keep.pop(ticket, None)
        _check(rc)
        return out, steps.value, lat.value

    def stats(self):
        s = ServerStats()
        _check(lib().bingo_server_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().bingo_server_destroy(self._h)
            self._h = None

    __del__ = close


def _walk_buffers(params, nwalkers, d_paths, d_steps, d_starts, d_lens):
    _need(d_paths, nwalkers * params.max_len * 4, "walk paths")
    _need(d_steps, 8, "walk step counter")
    _need(d_starts, nwalkers * 4, "walk starts")
    _need(d_lens, nwalkers * 4, "walk lengths")


class Graph:
    """One HBM-resident dynamic graph on one GPU (ids 1..n).

    float_weights=True: edge weights are the IEEE-754 bits of float32 values
    (the reference's wtype::Float), see float_bits().  dense_rate: the
    reference's RATE (0: every group keeps a member list)."""

    def __init__(self, device=0, neighbor_index=True, hbm_budget=0, pool_slack=0.25, min_slots=16, fit_load=False,
                 fit_headroom=1.0 / 16, float_weights=False, grow_in_place=None, dense_rate=0.0):
        cfg = Config()
        lib().bingo_config_default(C.byref(cfg))
        cfg.device, cfg.neighbor_index, cfg.hbm_budget = device, int(neighbor_index), hbm_budget
        cfg.pool_slack, cfg.min_slots = pool_slack, min_slots
        cfg.fit_load, cfg.fit_headroom = int(fit_load), fit_headroom
        if grow_in_place is not None:
            cfg.grow_in_place = int(grow_in_place)
        cfg.float_weights = int(float_weights)
        cfg.dense_rate = float(dense_rate)
        h = C.c_void_p()
        _check(lib().bingo_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.n = 0
        self.float_weights = bool(float_weights)

    def close(self):
        if getattr(self, "_h", None):
            lib().bingo_destroy(self._h)
            self._h = None

    __del__ = close

    @classmethod
    def adopt(cls, handle, n, float_weights=False):
        """A Graph owning an existing loaded handle (e.g. instance(keep_graph=True))."""
        g = cls.__new__(cls)
        g._h = C.c_void_p(handle)
        g.n = n
        g.float_weights = float_weights
        return g

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- updates --
    def load(self, n, edges):
        e = np.ascontiguousarray(edges, EDGE_DT)
        _check(lib().bingo_load(self._h, n, _ptr(e), len(e)))
        self.n = n

    def load_device(self, n, d_edges: DeviceArray, m):
        _need(d_edges, m * EDGE_DT.itemsize, "load edges")
        _check(lib().bingo_load_device(self._h, n, d_edges.ptr, m))
        self.n = n

    def insert(self, edges):
        e = np.ascontiguousarray(edges, EDGE_DT)
        _check(lib().bingo_insert(self._h, _ptr(e), len(e)))

    def insert_device(self, d_edges: DeviceArray, k):
        _need(d_edges, k * EDGE_DT.itemsize, "insert edges")
        _check(lib().bingo_insert_device(self._h, d_edges.ptr, k))

    def delete(self, dels):
        d = np.ascontiguousarray(dels, DEL_DT)
        _check(lib().bingo_delete(self._h, _ptr(d), len(d)))

    def delete_noalloc(self, dels):
        """bingo_delete with every batch buffer in the library's static device
        scratch (at most BINGO_NOALLOC_MAX_DELETES deletes)."""
        d = np.ascontiguousarray(dels, DEL_DT)
        _check(lib().bingo_delete_noalloc(self._h, _ptr(d), len(d)))

    def delete_device(self, d_dels: DeviceArray, k):
        _need(d_dels, k * DEL_DT.itemsize, "deletes")
        _check(lib().bingo_delete_device(self._h, d_dels.ptr, k))

    def compact(self, min_dead_frac, counts=True):
        """Reclaim tombstoned slots of nodes with dead >= 1 and dead / nslots
        >= min_dead_frac (16-bit fixed point); live slots are renumbered by
        rank.  Returns (nodes, slots) reclaimed, or None with counts=False
        (then the call does not wait for the device)."""
        if not counts:
            _check(lib().bingo_compact(self._h, min_dead_frac, None, None))
            return None
        nodes, slots = C.c_uint64(0), C.c_uint64(0)
        _check(lib().bingo_compact(self._h, min_dead_frac, C.byref(nodes), C.byref(slots)))
        return nodes.value, slots.value

    # -- walks --
    def walk(self, app, params: WalkParams, starts=None, lens=None, nwalkers=None):
        if starts is not None:
            starts = np.ascontiguousarray(starts, np.int32)
            nwalkers = len(starts)
        if lens is not None:
            lens = np.ascontiguousarray(lens, np.int32)
        if nwalkers is None:
            nwalkers = self.n
        paths = np.empty((nwalkers, params.max_len), np.int32)
        steps = C.c_uint64(0)
        _check(lib().bingo_walk(self._h, app, C.byref(params), _ptr(starts), _ptr(lens), nwalkers, _ptr(paths),
                                C.byref(steps)))
        return paths, steps.value

    def walk_device(self, app, params: WalkParams, nwalkers, d_paths: DeviceArray, d_steps: DeviceArray,
                    d_starts: DeviceArray | None = None, d_lens: DeviceArray | None = None):
        _walk_buffers(params, nwalkers, d_paths, d_steps, d_starts, d_lens)
        _check(lib().bingo_walk_device(self._h, app, C.byref(params), d_starts.ptr if d_starts else None,
                                       d_lens.ptr if d_lens else None, nwalkers, d_paths.ptr, d_steps.ptr))

    def walk_graph(self, app, params: WalkParams, nwalkers, d_paths: DeviceArray, d_steps: DeviceArray,
                   d_starts: DeviceArray | None = None, d_lens: DeviceArray | None = None):
        """Capture one walk launch as a HIP graph; call .launch() to replay it."""
        _walk_buffers(params, nwalkers, d_paths, d_steps, d_starts, d_lens)
        return WalkGraph(self, app, params, nwalkers, d_paths, d_steps, d_starts, d_lens)

    def update_graph(self, d_ins: DeviceArray | None, max_ins, d_dels: DeviceArray | None, max_dels, room_bytes=0):
        """Capture insert + delete batches of up to max_ins / max_dels entries as
        HIP graphs; .launch(nins, ndels) replays them."""
        return UpdateGraph(self, d_ins, max_ins, d_dels, max_dels, room_bytes)

    def tune(self, app, params: WalkParams, sample):
        """Autotune the walk schedule for `app`; returns the chosen knobs."""
        out = WalkParams()
        _check(lib().bingo_walk_tune(self._h, app, C.byref(params), sample, C.byref(out)))
        return dict(form=out.form, block=out.block, walkers_per_lane=out.walkers_per_lane, lds_depth=out.lds_depth)

    # -- inspection --
    def sync(self):
        _check(lib().bingo_sync(self._h))

    def stats(self):
        s = Stats()
        _check(lib().bingo_stats_get(self._h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in Stats._fields_ if k != "pool_used"}
        d["pool_used"] = list(s.pool_used)
        return d

    def event(self, slot):
        _check(lib().bingo_event_record(self._h, slot))

    def elapsed_ms(self, a, b):
        ms = C.c_float(0)
        _check(lib().bingo_event_elapsed(self._h, a, b, C.byref(ms)))
        return ms.value

    def gather_probe(self, width):
        """(reads, kernel ms) of independent random `width`-byte reads over the member pool;
        width 0: (read-modify-writes, kernel ms) of 8-byte words stored back unchanged"""
        n, ms = C.c_uint64(0), C.c_float(0)
        _check(lib().bingo_gather_probe(self._h, width, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def edge_count(self, us, vs):
        us = np.ascontiguousarray(us, np.int32)
        vs = np.ascontiguousarray(vs, np.int32)
        out = np.zeros(len(us), np.uint32)
        _check(lib().bingo_edge_count(self._h, _ptr(us), _ptr(vs), len(us), _ptr(out)))
        return out

    def export_nodes(self):
        """Per-node slot counts, masses and bucket sizes only (cheap at any size)."""
        d = Dump()
        n1 = self.n + 1
        out = dict(nslots=np.zeros(n1, np.uint32), mass=np.zeros(n1, np.uint64), bsize=np.zeros((n1, 32), np.uint32))
        for k, a in out.items():
            setattr(d, k, a.ctypes.data)
        _check(lib().bingo_export(self._h, C.byref(d)))
        out["total_slots"], out["total_members"] = d.total_slots, d.total_members
        return out

    def export(self):
        d = Dump()
        _check(lib().bingo_export(self._h, C.byref(d)))
        n1 = self.n + 1
        out = dict(
            nslots=np.zeros(n1, np.uint32),
            mass=np.zeros(n1, np.uint64),
            bsize=np.zeros((n1, 32), np.uint32),
            T=np.zeros((n1, 32), np.uint64),
            alias=np.zeros((n1, 32), np.uint8),
            slot_v=np.zeros(d.total_slots, np.uint32),
            slot_w=np.zeros(d.total_slots, np.uint32),
            members=np.zeros(d.total_members, np.uint64),
            backpos=np.zeros(d.total_members, np.uint32),
            dmask=np.zeros(n1, np.uint32),
        )
        if self.float_weights:
            out.update(slot_f=np.zeros(d.total_slots, np.uint32), fmass=np.zeros(n1, np.uint64),
                       fthr=np.zeros(n1, np.uint64))
        for k, a in out.items():
            setattr(d, k, a.ctypes.data)
        _check(lib().bingo_export(self._h, C.byref(d)))
        return out

"""TEST INFRASTRUCTURE: ctypes view of the CPU oracle (oracle/liboracle.so).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module.  It is the checker, never the thing measured or
shipped: the product (``bingo_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EDGE_DT = np.dtype([("u", np.int32), ("v", np.int32), ("w", np.uint32)])
DEL_DT = np.dtype([("u", np.int32), ("slot", np.int32)])

DEEPWALK, NODE2VEC, PPR, SAMPLING = 0, 1, 2, 3
W_CONST1, W_UNIFORM16, W_LOGUNI32, W_ZIPF24 = 0, 1, 2, 3


class WalkParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("thr_return", C.c_uint32),
        ("thr_neighbor", C.c_uint32),
        ("thr_other", C.c_uint32),
        ("thr_stop", C.c_uint32),
        ("max_len", C.c_int32),
        ("max_attempts", C.c_int32),
        ("fold_r", C.c_uint64),
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


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


NATIVE_PATH = os.path.join(HERE, "_native", "liboracle.so")
NATIVE_FLAGS = "-O3 -march=native -mtune=native -fopenmp"


def use_native() -> str:
    """Build the oracle for this machine's CPU (-march=native, into _native/)
    and load that build instead of the portable one: the CPU baseline is
    timed with it.  Must run before the first lib() call; returns the flags."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the portable oracle is loaded already")
    subprocess.run(["make", "-s", "-C", HERE, "native"], check=True)
    LIB_PATH = NATIVE_PATH
    return NATIVE_FLAGS


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i32, C.c_int]
        L.orc_create_ex.restype = vp
        L.orc_create_ex.argtypes = [i32, C.c_int, C.c_int]
        L.orc_create_rate.restype = vp
        L.orc_create_rate.argtypes = [i32, C.c_int, C.c_int, C.c_uint32]
        L.orc_frac_threshold.restype = u64
        L.orc_frac_threshold.argtypes = [u64, u64]
        L.orc_split_float.restype = C.c_int
        L.orc_split_float.argtypes = [C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_destroy.argtypes = [vp]
        L.orc_insert.argtypes = [vp, vp, i64]
        L.orc_delete.argtypes = [vp, vp, i64]
        L.orc_walk.argtypes = [vp, C.c_int, C.POINTER(WalkParams), vp, vp, i64, vp, C.POINTER(u64), C.c_int]
        L.orc_walk_ids.argtypes = [vp, C.c_int, C.POINTER(WalkParams), vp, vp, vp, i64, vp, C.POINTER(u64), C.c_int]
        L.orc_walk_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(WalkParams)]
        L.orc_walk_thresholds_ex.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(WalkParams)]
        L.orc_fold_threshold.restype = C.c_uint64
        L.orc_fold_threshold.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_export.argtypes = [vp, C.POINTER(Dump)]
        L.orc_compact.restype = i64
        L.orc_compact.argtypes = [vp, C.c_double, C.POINTER(u64)]
        L.orc_edge_count.restype = C.c_uint32
        L.orc_edge_count.argtypes = [vp, i32, i32]
        L.orc_alias_build.argtypes = [vp, vp, vp]
        L.orc_philox.argtypes = [vp, vp, vp]
        L.orc_sample_words.restype = i32
        L.orc_sample_words.argtypes = [vp, i32, vp]
        L.orc_rmat_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, vp]
        L.orc_gen_rmat.argtypes = [C.c_int, i64, i64, vp, C.c_int, u64, C.c_int, vp, C.c_int]
        L.orc_gen_weight.restype = C.c_uint32
        L.orc_gen_weight.argtypes = [C.c_int, u64, u64]
        L.orc_updgen_create.restype = vp
        L.orc_updgen_create.argtypes = [i32, C.c_int, vp, i64, i64, vp, C.c_int, u64]
        L.orc_updgen_destroy.argtypes = [vp]
        L.orc_updgen_inserts.argtypes = [vp, i64, vp]
        L.orc_updgen_deletes.argtypes = [vp, i64, vp]
        L.orc_updgen_live.restype = i64
        L.orc_updgen_live.argtypes = [vp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def walk_params(seed=1, p=0.5, q=2.0, tp=0.0125, max_len=80, max_attempts=4096, float_weights=False) -> WalkParams:
    """float_weights: parameters for a float-weight graph (node2vec keeps the
    plain envelope there; folding needs integer weight sums)"""
    wp = WalkParams()
    lib().orc_walk_thresholds_ex(p, q, tp, 0 if float_weights else 1, C.byref(wp))
    wp.seed = seed
    wp.max_len = max_len
    wp.max_attempts = max_attempts
    return wp


def fold_threshold(S, wret, fold_r):
    return int(lib().orc_fold_threshold(S, wret, fold_r))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_ptr(c), _ptr(k), _ptr(out))
    return [int(x) for x in out]


def alias_build(mass):
    m = np.ascontiguousarray(mass, dtype=np.uint64)
    assert m.shape == (32,)
    T = np.zeros(32, np.uint64)
    a = np.zeros(32, np.uint8)
    lib().orc_alias_build(_ptr(m), _ptr(T), _ptr(a))
    return T, a


def rmat_thresholds(a=0.57, b=0.19, c=0.19):
    thr = np.zeros(3, np.uint32)
    lib().orc_rmat_thresholds(a, b, c, _ptr(thr))
    return thr


def gen_rmat(scale, nedges, wkind, seed, undirected=True, edge_begin=0, abc=(0.57, 0.19, 0.19), nthreads=8):
    thr = rmat_thresholds(*abc)
    out = np.zeros(nedges * (2 if undirected else 1), EDGE_DT)
    lib().orc_gen_rmat(scale, edge_begin, nedges, _ptr(thr), wkind, seed, int(undirected), _ptr(out), nthreads)
    return out


class UpdateGen:
    def __init__(self, n, scale, base_edges, edge_begin, wkind, seed, abc=(0.57, 0.19, 0.19)):
        thr = rmat_thresholds(*abc)
        base = np.ascontiguousarray(base_edges, EDGE_DT)
        self._h = lib().orc_updgen_create(n, scale, _ptr(base), len(base), edge_begin, _ptr(thr), wkind, seed)

    def inserts(self, k):
        out = np.zeros(k, EDGE_DT)
        lib().orc_updgen_inserts(self._h, k, _ptr(out))
        return out

    def deletes(self, k):
        out = np.zeros(k, DEL_DT)
        rc = lib().orc_updgen_deletes(self._h, k, _ptr(out))
        if rc:
            raise ValueError("not enough live arcs")
        return out

    def live(self):
        return lib().orc_updgen_live(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_updgen_destroy(self._h)
            self._h = None


def frac_threshold(S, F):
    return lib().orc_frac_threshold(S, F)


def split_float(bits):
    """(integer part, fraction x 2^32) of float weight bits, or None."""
    ip, fr = C.c_uint32(0), C.c_uint32(0)
    if not lib().orc_split_float(int(bits), C.byref(ip), C.byref(fr)):
        return None
    return ip.value, fr.value


def dense_rate16(rate):
    """The sparse/dense rate in units of 2^-16, as the product quantises its
    configuration's dense_rate: 0 (off) for rate <= 0 or >= 1, else
    round(rate 2^16) within [1, 65535]."""
    if not (rate > 0.0) or rate >= 1.0:
        return 0
    return min(max(int(rate * 65536.0 + 0.5), 1), 65535)


class OracleGraph:
    """Sequential CPU model of the store (n nodes, ids 1..n).  With
    float_weights, the `w` field of edges carries IEEE-754 float bits;
    dense_rate (ref RATE) classes groups sparse or dense (0: all listed)."""

    def __init__(self, n, neighbor_index=True, float_weights=False, dense_rate=0.0):
        self.n = n
        self.float_weights = float_weights
        self.rate16 = dense_rate16(dense_rate)
        self._h = lib().orc_create_rate(n, int(neighbor_index), int(float_weights), self.rate16)
        if not self._h:
            raise ValueError("bad n")

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    __del__ = close

    def insert(self, edges):
        e = np.ascontiguousarray(edges, EDGE_DT)
        return lib().orc_insert(self._h, _ptr(e), len(e))

    load = insert

    def delete(self, dels):
        d = np.ascontiguousarray(dels, DEL_DT)
        return lib().orc_delete(self._h, _ptr(d), len(d))

    def compact(self, min_dead_frac):
        """Renumber live slots of nodes past the tombstone threshold; returns
        (nodes compacted, slots reclaimed)."""
        freed = C.c_uint64(0)
        nodes = lib().orc_compact(self._h, min_dead_frac, C.byref(freed))
        if nodes < 0:
            raise ValueError("min_dead_frac outside [0, 1]")
        return nodes, freed.value

    def walk(self, app, params: WalkParams, starts=None, lens=None, nwalkers=None, nthreads=8):
        if starts is not None:
            starts = np.ascontiguousarray(starts, np.int32)
            nwalkers = len(starts)
        if lens is not None:
            lens = np.ascontiguousarray(lens, np.int32)
        if nwalkers is None:
            nwalkers = self.n
        paths = np.empty((nwalkers, params.max_len), np.int32)
        steps = C.c_uint64(0)
        rc = lib().orc_walk(self._h, app, C.byref(params), _ptr(starts), _ptr(lens), nwalkers, _ptr(paths),
                            C.byref(steps), nthreads)
        if rc:
            raise ValueError(f"orc_walk rc={rc}")
        return paths, steps.value

    def walk_ids(self, app, params: WalkParams, wids, starts=None, lens=None, nthreads=8):
        """Rows for walkers wids (their own Philox streams); starts/lens are
        the rows' own (None: node (wid % n) + 1, max_len)."""
        wids = np.ascontiguousarray(wids, np.int64)
        if starts is not None:
            starts = np.ascontiguousarray(starts, np.int32)
        if lens is not None:
            lens = np.ascontiguousarray(lens, np.int32)
        paths = np.empty((len(wids), params.max_len), np.int32)
        steps = C.c_uint64(0)
        rc = lib().orc_walk_ids(self._h, app, C.byref(params), _ptr(starts), _ptr(lens), _ptr(wids), len(wids),
                                _ptr(paths), C.byref(steps), nthreads)
        if rc:
            raise ValueError(f"orc_walk_ids rc={rc}")
        return paths, steps.value

    def edge_count(self, u, v):
        return lib().orc_edge_count(self._h, u, v)

    def sample_words(self, u, words):
        w = np.ascontiguousarray(words, np.uint32)
        return lib().orc_sample_words(self._h, u, _ptr(w))

    def export(self):
        d = Dump()
        lib().orc_export(self._h, C.byref(d))
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
        lib().orc_export(self._h, C.byref(d))
        return out

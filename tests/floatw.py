"""Float-weight test inputs (the reference's wtype::Float): float32 bit
patterns of mixed laws, including weights below 1 (fraction only) and the
reference's own generator law d + rd(100) * .01 (include/graph.cuh:52)."""
import struct
from fractions import Fraction

import numpy as np

EDGE_DT = np.dtype([("u", np.int32), ("v", np.int32), ("w", np.uint32)])


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def fexact(bits):
    """Exact value of float32 bits (a Python float holds any float32 exactly)."""
    return Fraction(struct.unpack("<f", struct.pack("<I", bits))[0])


def float_edges(rng, n, k, hub=None, hub_frac=0.3):
    """k arcs on nodes 1..n (a share hub_frac of them from `hub`), weights
    drawn from five laws: integers, d + rd(100) * .01, below 1, large, small
    reals."""
    e = np.zeros(k, EDGE_DT)
    e["u"] = np.where(rng.random(k) < hub_frac, hub, rng.integers(1, n + 1, k)) if hub else rng.integers(1, n + 1, k)
    e["v"] = rng.integers(1, n + 1, k)
    kind = rng.integers(0, 5, k)
    vals = np.where(kind == 0, rng.integers(1, 50, k).astype(np.float64),
           np.where(kind == 1, rng.integers(1, 50, k) + rng.integers(0, 100, k) * 0.01,
           np.where(kind == 2, rng.random(k) * 0.99 + 0.005,
           np.where(kind == 3, rng.random(k) * 2e6 + 1.0,
                    rng.random(k) * 7.0 + 0.125))))
    e["w"] = vals.astype(np.float32).view(np.uint32)
    return e


def live_slots(d, u, so):
    """Live slots of node u in an export (integer part or fraction non-zero)."""
    ns = int(d["nslots"][u])
    w = d["slot_w"][so[u]:so[u] + ns]
    f = d["slot_f"][so[u]:so[u] + ns]
    return np.flatnonzero((w != 0) | (f != 0))

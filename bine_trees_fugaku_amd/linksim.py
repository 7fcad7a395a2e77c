"""Torus link-traffic simulation on the GPU (mirror of simulation/simulator.py,
extended with per-link accounting).

A case is (torus dimensions, collective, algorithm).  Its schedule is a list
of step classes (:func:`classes`); every message of a class carries the share
num/den of the per-rank vector and the class repeats ``rep`` times.  The GPU
routes every message over the torus (dimension-ordered, the shorter way round,
the positive way on a tie) and counts per directed link how many messages
cross it with 32-bit integer atomics.  Link bytes at S bytes per rank are
``count * floor(S * num / den)``: integers, exact at any S.

simulator.py (:125-177) reports per step the peer distance and the weighted
sums ``bw_new``/``bw_ref``/``bw_ideal``; :func:`simpy_stats` derives the same
figures from the GPU counts (the minimal route of a simulator.py message is
exactly |distance| hops).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

COLLS = {"allreduce": 0, "reduce_scatter": 1, "allgather": 2, "bcast": 3, "alltoall": 4, "scatter": 5,
         "gather": 6, "reduce": 7}
ALGOS = {"bine": 0, "recdoub": 1, "ring": 2, "bruck": 3, "simpy": 4}
# Algorithms defined per collective (the rest return BINE_ERR_UNSUPPORTED).
DEFINED = {
    "allreduce": ("bine", "recdoub", "ring", "simpy"),
    "reduce_scatter": ("bine", "recdoub", "ring", "simpy"),
    "allgather": ("bine", "recdoub", "ring"),
    "alltoall": ("bine", "recdoub", "bruck"),
    "bcast": ("bine", "recdoub"),
    "scatter": ("bine", "recdoub"),
    "gather": ("bine", "recdoub"),
    "reduce": ("bine", "recdoub"),
}
NONE = 0xFFFFFFFF


def _case(dims, coll, algo) -> _lib.SimCase:
    c = _lib.SimCase()
    dims = list(dims)
    if not 1 <= len(dims) <= 3:
        raise ValueError("1 to 3 torus dimensions")
    for i, d in enumerate(dims):
        c.dimensions[i] = int(d)
    c.dimensions_num = len(dims)
    c.coll = COLLS[coll] if isinstance(coll, str) else int(coll)
    c.algo = ALGOS[algo] if isinstance(algo, str) else int(algo)
    return c


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def classes(dims, coll, algo) -> dict:
    """Step classes of a case: arrays kind, d, k, num, den, rep (host only)."""
    L = _lib.load()
    c = _case(dims, coll, algo)
    n = ctypes.c_uint32(0)
    _lib.check(L.bine_sim_classes(ctypes.byref(c), ctypes.byref(n), None, None, None, None, None, None),
               "bine_sim_classes")
    out = {k: np.zeros(n.value, np.uint32) for k in ("kind", "d", "k", "num", "den", "rep")}
    _lib.check(L.bine_sim_classes(ctypes.byref(c), ctypes.byref(n), *(_u32p(out[k]) for k in
                                  ("kind", "d", "k", "num", "den", "rep"))), "bine_sim_classes")
    return out


def peers(dims, coll, algo, cls) -> np.ndarray:
    """Destination of every rank's message in class cls (NONE: the rank is silent)."""
    L = _lib.load()
    c = _case(dims, coll, algo)
    out = np.zeros(int(np.prod(dims)), np.uint32)
    _lib.check(L.bine_sim_peers(ctypes.byref(c), int(cls), _u32p(out)), "bine_sim_peers")
    return out


@dataclass
class ClassStats:
    num: np.ndarray
    den: np.ndarray
    rep: np.ndarray
    max_mult: np.ndarray  # messages on the most loaded link
    sum_mult: np.ndarray  # link crossings of all messages (= sum of route lengths)
    max_hops: np.ndarray  # longest route

    def bytes_share(self, nbytes: int) -> np.ndarray:
        """Bytes of one message of each class at nbytes per rank (exact)."""
        return (np.uint64(nbytes) * self.num.astype(np.uint64)) // self.den.astype(np.uint64)


class LinkSim:
    """A set of cases uploaded to one GPU; :meth:`run` counts them all in one pass."""

    def __init__(self, cases, device: int = 0):
        L = _lib.load()
        self.cases = [tuple(c) for c in cases]
        arr = (_lib.SimCase * len(self.cases))(*[_case(*c) for c in self.cases])
        h = ctypes.c_void_p()
        _lib.check(L.bine_sim_create(arr, len(self.cases), int(device), ctypes.byref(h)), "bine_sim_create")
        self._h = h

    def run(self, stream=None) -> None:
        s = None if stream is None else ctypes.c_void_p(getattr(stream, "cuda_stream", stream))
        _lib.check(_lib.load().bine_sim_run(self._h, s), "bine_sim_run")

    def info(self, i: int):
        nc, nl = ctypes.c_uint32(), ctypes.c_uint32()
        _lib.check(_lib.load().bine_sim_case_info(self._h, i, ctypes.byref(nc), ctypes.byref(nl)), "bine_sim_case_info")
        return nc.value, nl.value

    def stats(self, i: int) -> ClassStats:
        nc, _ = self.info(i)
        a = {k: np.zeros(nc, np.uint32) for k in ("num", "den", "rep", "max_mult", "max_hops")}
        sm = np.zeros(nc, np.uint64)
        _lib.check(_lib.load().bine_sim_class_stats(
            self._h, i, _u32p(a["num"]), _u32p(a["den"]), _u32p(a["rep"]), _u32p(a["max_mult"]),
            sm.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _u32p(a["max_hops"])), "bine_sim_class_stats")
        return ClassStats(a["num"], a["den"], a["rep"], a["max_mult"], sm, a["max_hops"])

    def link_counts(self, i: int, cls: int) -> np.ndarray:
        _, nl = self.info(i)
        out = np.zeros(nl, np.uint32)
        _lib.check(_lib.load().bine_sim_link_counts(self._h, i, int(cls), _u32p(out)), "bine_sim_link_counts")
        return out

    def link_bytes(self, i: int, nbytes: int) -> np.ndarray:
        """Bytes carried by every directed link over the whole collective at nbytes per rank."""
        st = self.stats(i)
        share = st.bytes_share(nbytes)
        _, nl = self.info(i)
        tot = np.zeros(nl, np.uint64)
        for c in range(len(share)):
            tot += self.link_counts(i, c).astype(np.uint64) * (share[c] * np.uint64(st.rep[c]))
        return tot

    def close(self) -> None:
        if self._h:
            _lib.check(_lib.load().bine_sim_destroy(self._h), "bine_sim_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simpy_stats(stats: ClassStats, P: int) -> tuple:
    """simulator.py's (sum_of_distances, bw_new, bw_ideal) for ALLREDUCE /
    REDUCESCATTER (simulator.py:160-175) from the GPU counts of a simpy case."""
    dist = stats.sum_mult.astype(np.float64) / P
    if not np.all(stats.sum_mult % P == 0):
        raise ValueError("route lengths differ across ranks")
    frac = 1.0 / 2.0 ** (np.arange(len(dist)) + 1)
    sum_d = bw_new = bw_ideal = 0.0
    for s in range(len(dist)):  # same accumulation order as simulator.py
        sum_d += dist[s]
        bw_ideal += frac[s]
        bw_new += frac[s] * dist[s]
    return sum_d, bw_new, bw_ideal

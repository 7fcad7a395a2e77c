"""Expected link traffic for the link-simulator tests, built only from the oracle.

Three independent sources of messages, each routed by the oracle's
dimension-ordered minimal routing (``oracle.route_bytes``):

* the oracle's own message trace of a restated reference collective
  (allreduce/reduce_scatter/allgather on the torus, alltoall, bcast, reduce,
  scatter, gather) -- the ground truth for every Bine/recursive-doubling/Bruck
  case it covers;
* closed forms for schedules the reference's MPI layer does not contain
  (ring, XOR butterfly, binomial trees);
* simulator.py's peer function restated in the oracle (``oracle.sim_peer``).
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

FAM = {"bine": "BINE", "recdoub": "RECDOUB", "bruck": "BRUCK"}


def per_step_bytes(msgs, P):
    """Aggregate a trace into {step: (src, dst, bytes)} arrays, one entry per (src, dst)."""
    out = {}
    for s in np.unique(msgs["step"]):
        sel = msgs["step"] == s
        key = msgs["src"][sel].astype(np.int64) * P + msgs["dst"][sel]
        u, inv = np.unique(key, return_inverse=True)
        tot = np.zeros(len(u), np.uint64)
        np.add.at(tot, inv, msgs["bytes"][sel])
        out[int(s)] = ((u // P).astype(np.int32), (u % P).astype(np.int32), tot)
    return out


def trace_case(dims, coll, algo, c=2):
    """Run the oracle collective with S = P*c int32 elements per rank (every
    share num/den of the model is then an integer) and return (S, trace)."""
    P = int(np.prod(dims))
    rng = np.random.default_rng(11)
    fam = FAM[algo]
    big = rng.integers(-9, 9, (P, P * c)).astype(np.int32)
    small = rng.integers(-9, 9, (P, c)).astype(np.int32)
    if coll == "allreduce":
        _, m = O.trace(O.allreduce, big, dims, family=fam)
    elif coll == "reduce_scatter":
        _, m = O.trace(O.reduce_scatter, big, dims, family=fam)
    elif coll == "allgather":
        _, m = O.trace(O.allgather, small, dims, family=fam)
    elif coll == "alltoall":
        _, m = O.trace(O.alltoall, big, P, family=fam)
    elif coll == "bcast":
        _, m = O.trace(O.bcast, big, 0)
    elif coll == "reduce":
        _, m = O.trace(O.reduce, big, 0)
    elif coll == "scatter":
        _, m = O.trace(O.scatter, big.reshape(-1)[: P * c], P, 0)
    elif coll == "gather":
        _, m = O.trace(O.gather, small, 0)
    else:
        raise ValueError(coll)
    return P * c * 4, m


# Cases whose messages the oracle's restated reference collectives produce.
TRACED = [("allreduce", "bine"), ("allreduce", "recdoub"), ("reduce_scatter", "bine"),
          ("reduce_scatter", "recdoub"), ("allgather", "bine"), ("allgather", "recdoub"),
          ("alltoall", "bine"), ("alltoall", "bruck"), ("bcast", "bine"), ("reduce", "bine"),
          ("scatter", "bine"), ("gather", "bine")]


def closed_form_dest(dims, coll, algo, cls, L):
    """Destinations of class cls for the schedules the oracle does not trace
    (None when the case is not one of them).  Silent ranks: -1."""
    P = int(np.prod(dims))
    r = np.arange(P, dtype=np.int64)
    if algo == "ring":
        return (r + 1) % P
    if coll == "alltoall" and algo == "recdoub":
        return r ^ (1 << cls)
    if algo == "recdoub" and coll in ("bcast", "scatter"):
        m = 1 << (L - 1 - cls)  # masks from P/2 down to 1
        return np.where((r % (2 * m) == 0) & (r + m < P), r + m, -1)
    if algo == "recdoub" and coll in ("reduce", "gather"):
        m = 1 << cls
        return np.where(r % (2 * m) == m, r - m, -1)
    if algo == "simpy":
        return np.array([O.sim_peer(int(x), cls, dims) for x in r])
    return None


def simpy_to_lib(dims, ids):
    """simulator.py ids -> library row-major ids (simulator.py:11-26)."""
    ids = np.asarray(ids, np.int64)
    if len(dims) == 3:
        x = (ids // dims[1]) % dims[0]
        y = ids % dims[1]
        z = ids // (dims[0] * dims[1])
        return (x * dims[1] + y) * dims[2] + z
    return ids  # 1D and 2D orders agree with the library's

"""ctypes front-end to the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a plain-C restatement of the reference's MPI-layer Bine
collectives (HLC-Lab/bine_trees_fugaku lib/*.cc, simulation/*.py) over P
virtual ranks in host memory.  Only tests, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` use it, always as the checker; the
product library never loads it.

Buffers are numpy arrays shaped ``(P, per_rank_elems)``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_REMAP = os.path.join(HERE, "_ref", "remap_rank")

FAMILY = {"DEFAULT": 0, "BINE": 1, "RECDOUB": 2, "BRUCK": 3, "RING": 4}
DIST = {"INCREASING": 0, "DECREASING": 1}
AR_ALGO = {"L": 0, "B": 1, "REDUCE_BCAST": 2, "B_CONT": 3, "B_COALESCE": 4}
DTYPE = {"int32": 0, "int64": 1, "float32": 2, "float64": 3, "int8": 4}
OP = {"SUM": 0, "VOID": 1}
NP = {"int32": np.int32, "int64": np.int64, "float32": np.float32, "float64": np.float64, "int8": np.int8}

_lock = threading.Lock()
_trace_lock = threading.Lock()  # one traced collective at a time (the trace is process-global)
_lib = None


class OracleError(RuntimeError):
    pass


class Torus(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_uint32 * 3), ("ndims", ctypes.c_uint32), ("size", ctypes.c_uint32),
                ("num_steps", ctypes.c_uint32), ("steps_per_dim", ctypes.c_uint32 * 3)]


def _stale() -> bool:
    if not os.path.exists(LIB_PATH):
        return True
    t = os.path.getmtime(LIB_PATH)
    return any(os.path.getmtime(os.path.join(HERE, f)) > t for f in ("bine_oracle.c", "bine_oracle.h"))


def build() -> None:
    """Compile the oracle with gcc (seconds); also used where only the sources travelled."""
    subprocess.run(["make", "-s", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if _stale():
                build()
            L = ctypes.CDLL(LIB_PATH)
            u32, i32, vp = ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p
            pu32 = ctypes.POINTER(ctypes.c_uint32)
            L.or_torus_init.argtypes = [ctypes.POINTER(Torus), pu32, u32]
            L.or_compute_peers.argtypes = [ctypes.POINTER(Torus), u32, i32, i32, pu32]
            L.or_step_info.argtypes = [ctypes.POINTER(Torus), u32, pu32, pu32]
            L.or_block_step.argtypes = [ctypes.POINTER(Torus), u32, u32, i32, pu32]
            L.or_step_bitmaps.argtypes = [ctypes.POINTER(Torus), u32, u32, i32, u32, vp, vp]
            L.or_get_tree.argtypes = [ctypes.POINTER(Torus), u32, u32, i32, i32, pu32, pu32, pu32, pu32, pu32]
            L.or_cont_remapped_rank.argtypes = [ctypes.POINTER(Torus), u32, u32, i32]
            L.or_cont_remapped_rank.restype = u32
            L.or_allreduce.argtypes = [ctypes.POINTER(Torus), i32, i32, i32, i32, i32, i32, vp, vp]
            L.or_reduce_scatter.argtypes = [ctypes.POINTER(Torus), i32, i32, i32, i32, i32, i32, vp, vp]
            L.or_allgather.argtypes = [ctypes.POINTER(Torus), i32, i32, i32, i32, i32, i32, vp, vp]
            L.or_alltoall.argtypes = [u32, i32, i32, i32, vp, vp]
            L.or_bcast.argtypes = [u32, i32, i32, i32, i32, vp]
            L.or_reduce.argtypes = [u32, i32, i32, i32, i32, i32, vp, vp]
            L.or_scatter.argtypes = [u32, i32, i32, i32, vp, vp]
            L.or_gather.argtypes = [u32, i32, i32, i32, vp, vp]
            L.or_btonb.argtypes = [ctypes.c_int32]
            L.or_btonb.restype = u32
            L.or_nbtob.argtypes = [u32]
            L.or_nbtob.restype = ctypes.c_int32
            L.or_remap_rank.argtypes = [u32, u32]
            L.or_remap_rank.restype = u32
            L.or_rank_negabinary.argtypes = [u32, u32]
            L.or_rank_negabinary.restype = u32
            L.or_get_nu.argtypes = [u32, u32]
            L.or_get_nu.restype = u32
            L.or_remap_distance_doubling.argtypes = [u32]
            L.or_ceil_log2.argtypes = [ctypes.c_ulonglong]
            L.or_reverse.argtypes = [u32]
            L.or_reverse.restype = u32
            L.or_sim_peer.argtypes = [i32, i32, pu32, u32, i32, ctypes.POINTER(ctypes.c_int)]
            L.or_sim_run.argtypes = [pu32, u32, i32, ctypes.POINTER(ctypes.c_int), vp, vp]
            L.or_threads.argtypes = []
            L.or_trace_start.argtypes = []
            L.or_trace_stop.argtypes = []
            L.or_trace_stop.restype = ctypes.c_size_t
            L.or_trace_fetch.argtypes = [vp, vp, vp, vp]
            L.or_route_bytes.argtypes = [pu32, u32, ctypes.c_size_t, vp, vp, vp, vp]
            L.or_route_bytes.restype = u32
            _lib = L
    return _lib


def _u32(seq):
    return (ctypes.c_uint32 * len(seq))(*[int(x) for x in seq])


def torus(dims) -> Torus:
    t = Torus()
    rc = lib().or_torus_init(ctypes.byref(t), _u32(dims), len(dims))
    if rc:
        raise OracleError(f"bad torus {dims}")
    return t


def _check(rc, what):
    if rc:
        raise OracleError(f"{what}: oracle returned {rc}")


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- schedule ---------------------------------------------------------------

def compute_peers(dims, rank, port=0, family="BINE"):
    t = torus(dims)
    out = (ctypes.c_uint32 * max(t.num_steps, 1))()
    lib().or_compute_peers(ctypes.byref(t), rank, port, FAMILY[family], out)
    return [out[i] for i in range(t.num_steps)]


def step_info(dims, port=0):
    t = torus(dims)
    d = (ctypes.c_uint32 * max(t.num_steps, 1))()
    s = (ctypes.c_uint32 * max(t.num_steps, 1))()
    lib().or_step_info(ctypes.byref(t), port, d, s)
    return [(d[i], s[i]) for i in range(t.num_steps)]


def block_step(dims, rank, port=0, family="BINE"):
    t = torus(dims)
    out = (ctypes.c_uint32 * t.size)()
    lib().or_block_step(ctypes.byref(t), rank, port, FAMILY[family], out)
    return np.frombuffer(out, dtype=np.uint32).copy()


def step_bitmaps(dims, rank, step, port=0, family="BINE"):
    t = torus(dims)
    s = np.zeros(t.size, np.uint8)
    r = np.zeros(t.size, np.uint8)
    lib().or_step_bitmaps(ctypes.byref(t), rank, port, FAMILY[family], step, _ptr(s), _ptr(r))
    return s, r


def get_tree(dims, root, port=0, family="BINE", dist="INCREASING"):
    t = torus(dims)
    P = t.size
    arrs = [(ctypes.c_uint32 * P)() for _ in range(5)]
    _check(lib().or_get_tree(ctypes.byref(t), root, port, FAMILY[family], DIST[dist], *arrs), "get_tree")
    names = ("parent", "reached_at_step", "remapped_ranks", "remapped_ranks_max", "subtree_roots")
    return {n: np.frombuffer(a, dtype=np.uint32).copy() for n, a in zip(names, arrs)}


def cont_remapped_rank(dims, rank, port=0, family="BINE"):
    return lib().or_cont_remapped_rank(ctypes.byref(torus(dims)), rank, port, FAMILY[family])


def btonb(x):
    return lib().or_btonb(int(x))


def nbtob(x):
    return lib().or_nbtob(int(x))


def remap_rank(rank, size):
    return lib().or_remap_rank(rank, size)


def rank_negabinary(rank, size):
    return lib().or_rank_negabinary(rank, size)


def get_nu(rank, size):
    return lib().or_get_nu(rank, size)


def remap_distance_doubling(x):
    return lib().or_remap_distance_doubling(x)


def ceil_log2(x):
    return lib().or_ceil_log2(x)


def ref_remap_rank(size, rank):
    """Run the reference's own simulation/remap_rank.c (built under _ref/)."""
    out = subprocess.run([REF_REMAP, str(size), str(rank)], check=True, capture_output=True, text=True).stdout
    return int(out.strip().split("remap_rank:")[1])


# ---- collectives ------------------------------------------------------------

def _dt(a: np.ndarray) -> str:
    return {np.dtype(np.int32): "int32", np.dtype(np.int64): "int64", np.dtype(np.float32): "float32",
            np.dtype(np.float64): "float64", np.dtype(np.int8): "int8"}[a.dtype]


def allreduce(send: np.ndarray, dims, algo="B", family="BINE", ports=1, op="SUM") -> np.ndarray:
    send = np.ascontiguousarray(send)
    recv = np.zeros_like(send)
    t = torus(dims)
    assert send.shape[0] == t.size
    _check(lib().or_allreduce(ctypes.byref(t), ports, FAMILY[family], AR_ALGO[algo], send.shape[1],
                              DTYPE[_dt(send)], OP[op], _ptr(send), _ptr(recv)), "allreduce")
    return recv


def reduce_scatter(send: np.ndarray, dims, family="BINE", dist="INCREASING", op="SUM") -> np.ndarray:
    send = np.ascontiguousarray(send)
    t = torus(dims)
    P = t.size
    rc = send.shape[1] // P
    recv = np.zeros((P, rc), dtype=send.dtype)
    _check(lib().or_reduce_scatter(ctypes.byref(t), 1, FAMILY[family], DIST[dist], rc, DTYPE[_dt(send)], OP[op],
                                   _ptr(send), _ptr(recv)), "reduce_scatter")
    return recv


AG_VARIANT = {"CONT_PERMUTE": 0, "CONT_SEND": 1, "BLOCKS": 2}


def allgather(send: np.ndarray, dims, variant="CONT_PERMUTE", family="BINE", dist="DECREASING", ports=1):
    send = np.ascontiguousarray(send)
    t = torus(dims)
    P = t.size
    recv = np.zeros((P, P * send.shape[1]), dtype=send.dtype)
    _check(lib().or_allgather(ctypes.byref(t), ports, FAMILY[family], DIST[dist], AG_VARIANT[variant],
                              send.shape[1], DTYPE[_dt(send)], _ptr(send), _ptr(recv)), "allgather")
    return recv


def alltoall(send: np.ndarray, P, family="BINE") -> np.ndarray:
    send = np.ascontiguousarray(send)
    recv = np.zeros_like(send)
    _check(lib().or_alltoall(P, FAMILY[family], send.shape[1] // P, DTYPE[_dt(send)], _ptr(send), _ptr(recv)),
           "alltoall")
    return recv


BCAST_VARIANT = {"BINOMIAL_TREE": 0, "SCATTER_ALLGATHER": 2}
REDUCE_VARIANT = {"BINOMIAL_TREE": 0, "REDUCE_SCATTER_GATHER": 1}


def bcast(buf: np.ndarray, root=0, variant="BINOMIAL_TREE") -> np.ndarray:
    buf = np.ascontiguousarray(buf).copy()
    _check(lib().or_bcast(buf.shape[0], BCAST_VARIANT[variant], buf.shape[1], DTYPE[_dt(buf)], root, _ptr(buf)),
           "bcast")
    return buf


def reduce(send: np.ndarray, root=0, variant="BINOMIAL_TREE", op="SUM") -> np.ndarray:
    send = np.ascontiguousarray(send)
    out = np.zeros(send.shape[1], dtype=send.dtype)
    _check(lib().or_reduce(send.shape[0], REDUCE_VARIANT[variant], send.shape[1], DTYPE[_dt(send)], OP[op], root,
                           _ptr(send), _ptr(out)), "reduce")
    return out


def scatter(send_root: np.ndarray, P, root=0) -> np.ndarray:
    send_root = np.ascontiguousarray(send_root)
    count = send_root.size // P
    recv = np.zeros((P, count), dtype=send_root.dtype)
    _check(lib().or_scatter(P, count, DTYPE[_dt(send_root)], root, _ptr(send_root), _ptr(recv)), "scatter")
    return recv


def gather(send: np.ndarray, root=0) -> np.ndarray:
    send = np.ascontiguousarray(send)
    P, count = send.shape
    out = np.zeros(P * count, dtype=send.dtype)
    _check(lib().or_gather(P, count, DTYPE[_dt(send)], root, _ptr(send), _ptr(out)), "gather")
    return out


# ---- simulator.py restatement -------------------------------------------------

JACOBSTHAL = [1, 1, 3, 5, 11, 21, 43, 85, 171, 341]


def sim_peer(sender, step, dims, distances=JACOBSTHAL, allgather=False):
    d = (ctypes.c_int * len(distances))(*distances)
    return lib().or_sim_peer(sender, step, _u32(dims), len(dims), int(allgather), d)


def sim_run(dims, data: np.ndarray, collective="ALLREDUCE", distances=JACOBSTHAL):
    data = np.ascontiguousarray(data, dtype=np.float64).copy()
    stats = np.zeros(5, np.float64)
    d = (ctypes.c_int * len(distances))(*distances)
    code = {"ALLREDUCE": 0, "REDUCESCATTER": 1, "ALLGATHER": 2}[collective]
    _check(lib().or_sim_run(_u32(dims), len(dims), code, d, _ptr(data), _ptr(stats)), "sim_run")
    return data, stats


# ---- message trace and link routing -------------------------------------------

def threads() -> int:
    """OpenMP threads the oracle runs its data movement and reductions on."""
    return int(lib().or_threads())


def trace(fn, *args, **kwargs):
    """Run one oracle collective with the engine's message trace on.

    Returns (result, msgs): msgs holds int32 arrays step/src/dst and uint64
    bytes, one entry per delivered message between distinct ranks, in the
    order the engine delivered them."""
    L = lib()
    with _trace_lock:
        L.or_trace_start()
        try:
            res = fn(*args, **kwargs)
        finally:
            n = L.or_trace_stop()
        msgs = {k: np.zeros(n, np.int32) for k in ("step", "src", "dst")}
        msgs["bytes"] = np.zeros(n, np.uint64)
        if n:
            L.or_trace_fetch(_ptr(msgs["step"]), _ptr(msgs["src"]), _ptr(msgs["dst"]), _ptr(msgs["bytes"]))
    return res, msgs


def route_bytes(dims, src, dst, nbytes):
    """Dimension-ordered minimal routing of messages src[i] -> dst[i] carrying
    nbytes[i]; returns (bytes per directed link, longest route in hops).
    Link id = (node * ndims + d) * 2 + (1 if the hop goes the negative way)."""
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    nbytes = np.ascontiguousarray(np.broadcast_to(np.asarray(nbytes, np.uint64), src.shape))
    P = int(np.prod(dims))
    out = np.zeros(P * len(dims) * 2, np.uint64)
    longest = lib().or_route_bytes(_u32(dims), len(dims), len(src), _ptr(src), _ptr(dst), _ptr(nbytes), _ptr(out))
    return out, int(longest)

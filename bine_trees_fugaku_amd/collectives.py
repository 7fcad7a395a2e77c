"""Host-side mirror of the reference's collective interface (lib/libbine.cc).

``VComm`` replaces ``MPI_Comm`` plus the ``LIBBINE_*`` environment: it holds
the virtual torus, the ports and each collective's (family, algorithm,
distance), and owns the compiled-schedule cache of libbine_hip.so.  Its
methods keep the names and argument meaning of the MPI entry points the
reference interposes (``MPI_Allreduce`` -> ``VComm.allreduce`` ...).  Buffers
are torch tensors on the GPU holding all P virtual ranks: shape ``(P, n)`` (or
any contiguous tensor of P*n elements) where row r is rank r's MPI buffer;
root-only buffers (scatter send, gather/reduce receive) hold one rank.

Every call goes through the C ABI; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

from . import _lib

FAMILY = {"DEFAULT": 0, "BINE": 1, "RECDOUB": 2, "BRUCK": 3, "RING": 4}
DISTANCE = {"INCREASING": 0, "DECREASING": 1}
EXEC = {"AUTO": 0, "STEP": 1, "FUSED": 2}
OP = {"SUM": 0, "VOID": 1}
COLL = {"allreduce": 0, "reduce_scatter": 1, "allgather": 2, "bcast": 3, "alltoall": 4, "scatter": 5, "gather": 6,
        "reduce": 7}
ALGOS = {
    "allreduce": {"L": 0, "B": 1, "REDUCE_BCAST": 2, "B_CONT": 3, "B_COALESCE": 4},
    "allgather": {"VEC_DOUBLING_CONT_PERMUTE": 0, "VEC_DOUBLING_CONT_SEND": 1, "VEC_DOUBLING_BLOCKS": 2,
                  "GATHER_BCAST": 3},
    "reduce_scatter": {"VEC_HALVING_CONT_PERMUTE": 0, "VEC_HALVING_CONT_SEND": 1, "VEC_HALVING_BLOCKS": 2,
                       "REDUCE_SCATTER": 3},
    "bcast": {"BINOMIAL_TREE": 0, "BINOMIAL_TREE_TMPBUF": 1, "SCATTER_ALLGATHER": 2},
    "alltoall": {"LOG": 0},
    "scatter": {"BINOMIAL_TREE_CONT_PERMUTE": 0, "BINOMIAL_TREE_CONT_SEND": 1},
    "gather": {"BINOMIAL_TREE_CONT_PERMUTE": 0, "BINOMIAL_TREE_CONT_SEND": 1},
    "reduce": {"BINOMIAL_TREE": 0, "REDUCE_SCATTER_GATHER": 1},
}


def _torch_dtype_code(t) -> int:
    import torch
    codes = {torch.int32: 0, torch.int64: 1, torch.float32: 2, torch.float64: 3, torch.int8: 4}
    if t.dtype not in codes:
        raise TypeError(f"unsupported element type {t.dtype}")
    return codes[t.dtype]


def make_config(dims: Sequence[int], num_ports: int = 1, exec: str = "AUTO", device: int = 0, from_env: bool = False,
                **overrides) -> _lib.VCommConfig:
    """bine_vcomm_config_t with init_env defaults (lib/libbine.cc:28-73), optional
    LIBBINE_* overlay, then explicit overrides such as ``allreduce_algo="B_CONT"``,
    ``alltoall_family="BRUCK"`` or ``allgather_distance="INCREASING"``."""
    L = _lib.load()
    if not 1 <= len(dims) <= 3:  # LIBBINE_MAX_SUPPORTED_DIMENSIONS (lib/libbine_common.h:19)
        raise ValueError(f"a torus has 1 to 3 dimensions, got {len(dims)}")
    cfg = _lib.VCommConfig()
    size = 1
    for d in dims:
        size *= int(d)
    _lib.check(L.bine_vcomm_config_init(ctypes.byref(cfg), size), "bine_vcomm_config_init")
    for i in range(3):
        cfg.dimensions[i] = int(dims[i]) if i < len(dims) else 1
    cfg.dimensions_num = len(dims)
    cfg.num_ports = num_ports
    cfg.exec = EXEC[exec]
    cfg.device = device
    if from_env:
        _lib.check(L.bine_vcomm_config_from_env(ctypes.byref(cfg)), "bine_vcomm_config_from_env")
    for key, val in overrides.items():
        coll, _, field = key.rpartition("_")
        if coll not in ALGOS or field not in ("algo", "family", "distance"):
            raise ValueError(f"unknown override {key}")
        c = getattr(cfg, coll)
        if field == "algo":
            c.algo = ALGOS[coll][val] if isinstance(val, str) else int(val)
        elif field == "family":
            c.family = FAMILY[val] if isinstance(val, str) else int(val)
        else:
            c.distance = DISTANCE[val] if isinstance(val, str) else int(val)
    return cfg


class VComm:
    """A virtual communicator of P ranks living as HBM partitions of one GPU."""

    def __init__(self, dims: Sequence[int], num_ports: int = 1, exec: str = "AUTO", device: int = 0,
                 from_env: bool = False, config: Optional[_lib.VCommConfig] = None, **overrides):
        """config: a complete bine_vcomm_config_t to use instead of the other arguments."""
        self._lib = _lib.load()
        if config is not None:
            self.config = _lib.VCommConfig()
            ctypes.memmove(ctypes.byref(self.config), ctypes.byref(config), ctypes.sizeof(config))
        else:
            self.config = make_config(dims, num_ports, exec, device, from_env, **overrides)
        self.dims = [self.config.dimensions[i] for i in range(self.config.dimensions_num)]
        self.num_ports = self.config.num_ports
        self.handle = ctypes.c_void_p()
        _lib.check(self._lib.bine_vcomm_create(ctypes.byref(self.config), ctypes.byref(self.handle)),
                   "bine_vcomm_create")
        size = ctypes.c_uint32()
        _lib.check(self._lib.bine_vcomm_size(self.handle, ctypes.byref(size)), "bine_vcomm_size")
        self.size = size.value

    def with_exec(self, exec: str) -> "VComm":
        """A new vcomm with this one's configuration and another execution mode."""
        cfg = _lib.VCommConfig()
        ctypes.memmove(ctypes.byref(cfg), ctypes.byref(self.config), ctypes.sizeof(cfg))
        cfg.exec = EXEC[exec]
        return VComm(self.dims, config=cfg)

    @classmethod
    def from_env(cls, exec: str = "AUTO", device: int = 0) -> "VComm":
        """The torus and algorithms of the LIBBINE_* environment (LIBBINE_DIMENSIONS required)."""
        if not os.environ.get("LIBBINE_DIMENSIONS"):
            raise ValueError("LIBBINE_DIMENSIONS is not set")
        return cls([1], exec=exec, device=device, from_env=True)

    def close(self):
        if self.handle:
            self._lib.bine_vcomm_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_exec(self, exec: str):
        _lib.check(self._lib.bine_vcomm_set_exec(self.handle, EXEC[exec]), "bine_vcomm_set_exec")

    # ---- helpers -----------------------------------------------------------
    @staticmethod
    def _stream(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def _per_rank(self, t, root_only=False) -> int:
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("buffers must be contiguous GPU tensors")
        n = t.numel()
        if root_only:
            return n
        if n % self.size:
            raise ValueError(f"buffer of {n} elements does not split into {self.size} ranks")
        return n // self.size

    @staticmethod
    def _expect(t, n: int, what: str, like=None):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{what} must be a contiguous GPU tensor")
        if like is not None and t.dtype != like.dtype:
            raise ValueError(f"{what} has element type {t.dtype}, expected {like.dtype}")
        if t.numel() != n:
            raise ValueError(f"{what} holds {t.numel()} elements, expected {n}")

    # ---- MPI entry points (lib/libbine.cc) --------------------------------------
    def allreduce(self, sendbuf, recvbuf, op: str = "SUM", stream=None):
        """MPI_Allreduce (lib/libbine.cc:1012).  sendbuf is recvbuf: MPI_IN_PLACE (per-step schedule)."""
        count = self._per_rank(sendbuf)
        self._expect(recvbuf, self.size * count, "recvbuf", sendbuf)
        _lib.check(self._lib.bine_allreduce(sendbuf.data_ptr(), recvbuf.data_ptr(), count, _torch_dtype_code(sendbuf),
                                            OP[op], self.handle, self._stream(stream)), "bine_allreduce")

    def reduce_scatter(self, sendbuf, recvbuf, op: str = "SUM", stream=None):
        """MPI_Reduce_scatter with equal recvcounts (lib/libbine.cc:1064)."""
        rc = self._per_rank(recvbuf)
        self._expect(sendbuf, self.size * self.size * rc, "sendbuf", recvbuf)
        counts = (ctypes.c_int * self.size)(*([rc] * self.size))
        _lib.check(self._lib.bine_reduce_scatter(sendbuf.data_ptr(), recvbuf.data_ptr(), counts,
                                                 _torch_dtype_code(sendbuf), OP[op], self.handle, self._stream(stream)),
                   "bine_reduce_scatter")

    def allgather(self, sendbuf, recvbuf, stream=None):
        """MPI_Allgather (lib/libbine.cc:1168)."""
        c = self._per_rank(sendbuf)
        self._expect(recvbuf, self.size * self.size * c, "recvbuf", sendbuf)
        dt = _torch_dtype_code(sendbuf)
        _lib.check(self._lib.bine_allgather(sendbuf.data_ptr(), c, dt, recvbuf.data_ptr(), c, dt, self.handle,
                                            self._stream(stream)), "bine_allgather")

    def bcast(self, buffer, root: int = 0, stream=None):
        """MPI_Bcast_f (lib/libbine.cc:1252)."""
        c = self._per_rank(buffer)
        _lib.check(self._lib.bine_bcast(buffer.data_ptr(), c, _torch_dtype_code(buffer), root, self.handle,
                                        self._stream(stream)), "bine_bcast")

    def alltoall(self, sendbuf, recvbuf, stream=None):
        """MPI_Alltoall (lib/libbine.cc:1291); count = elements per rank pair."""
        c = self._per_rank(sendbuf) // self.size
        self._expect(sendbuf, self.size * self.size * c, "sendbuf")
        self._expect(recvbuf, self.size * self.size * c, "recvbuf", sendbuf)
        dt = _torch_dtype_code(sendbuf)
        _lib.check(self._lib.bine_alltoall(sendbuf.data_ptr(), c, dt, recvbuf.data_ptr(), c, dt, self.handle,
                                           self._stream(stream)), "bine_alltoall")

    def scatter(self, sendbuf, recvbuf, root: int = 0, stream=None):
        """MPI_Scatter (lib/libbine.cc:1312); sendbuf is the root's P*count elements."""
        c = self._per_rank(recvbuf)
        self._expect(sendbuf, self.size * c, "sendbuf", recvbuf)
        dt = _torch_dtype_code(recvbuf)
        _lib.check(self._lib.bine_scatter(sendbuf.data_ptr(), c, dt, recvbuf.data_ptr(), c, dt, root, self.handle,
                                          self._stream(stream)), "bine_scatter")

    def gather(self, sendbuf, recvbuf, root: int = 0, stream=None):
        """MPI_Gather (lib/libbine.cc:1376); recvbuf is the root's P*count elements."""
        c = self._per_rank(sendbuf)
        self._expect(recvbuf, self.size * c, "recvbuf", sendbuf)
        dt = _torch_dtype_code(sendbuf)
        _lib.check(self._lib.bine_gather(sendbuf.data_ptr(), c, dt, recvbuf.data_ptr(), c, dt, root, self.handle,
                                         self._stream(stream)), "bine_gather")

    def reduce(self, sendbuf, recvbuf, op: str = "SUM", root: int = 0, stream=None):
        """MPI_Reduce (lib/libbine.cc:1440); recvbuf is the root's count elements."""
        c = self._per_rank(sendbuf)
        self._expect(recvbuf, c, "recvbuf", sendbuf)
        _lib.check(self._lib.bine_reduce(sendbuf.data_ptr(), recvbuf.data_ptr(), c, _torch_dtype_code(sendbuf),
                                         OP[op], root, self.handle, self._stream(stream)), "bine_reduce")

    # ---- local reductions (lib/libbine_coll.cc:153-267) ------------------------
    def reduce_local(self, inbuf, inoutbuf, op: str = "SUM", stream=None):
        """reduce_local, 2-argument form: inoutbuf = inbuf + inoutbuf (one partition, any length)."""
        self._expect(inoutbuf, inbuf.numel(), "inoutbuf", inbuf)
        _lib.check(self._lib.bine_reduce_local(inbuf.data_ptr(), inoutbuf.data_ptr(), inbuf.numel(),
                                               _torch_dtype_code(inbuf), OP[op], self._stream(stream)),
                   "bine_reduce_local")

    def reduce_local3(self, a, b, out, op: str = "SUM", stream=None):
        """reduce_local, 3-argument form: out = a + b."""
        self._expect(b, a.numel(), "inbuf_b", a)
        self._expect(out, a.numel(), "outbuf", a)
        _lib.check(self._lib.bine_reduce_local3(a.data_ptr(), b.data_ptr(), out.data_ptr(), a.numel(),
                                                _torch_dtype_code(a), OP[op], self._stream(stream)),
                   "bine_reduce_local3")

    # ---- plan cache --------------------------------------------------------------
    def set_cache_limit(self, max_plans: int):
        _lib.check(self._lib.bine_vcomm_set_cache_limit(self.handle, max_plans), "bine_vcomm_set_cache_limit")

    def trim(self):
        _lib.check(self._lib.bine_vcomm_trim(self.handle), "bine_vcomm_trim")

    def cache_info(self) -> dict:
        n, t = ctypes.c_uint32(), ctypes.c_uint64()
        _lib.check(self._lib.bine_vcomm_cache_info(self.handle, ctypes.byref(n), ctypes.byref(t)),
                   "bine_vcomm_cache_info")
        return {"plans": n.value, "tmp_bytes": t.value}

    # ---- schedule introspection ----------------------------------------------
    def plan_stats(self, coll: str, count: int, dtype, op: str = "SUM", root: int = 0, exec: str = "AUTO") -> dict:
        st = _lib.PlanStats()
        _lib.check(self._lib.bine_plan_stats(self.handle, COLL[coll], count, _dtype_code(dtype), OP[op], root,
                                             EXEC[exec], ctypes.byref(st)), "bine_plan_stats")
        return {f: getattr(st, f) for f, _ in _lib.PlanStats._fields_}

    def autotune(self, coll: str, sendbuf, recvbuf, count: int, op: str = "SUM", root: int = 0, exec: str = "AUTO",
                 iters: int = 3, stream=None):
        dt = _torch_dtype_code(recvbuf)
        _lib.check(self._lib.bine_plan_autotune(self.handle, COLL[coll],
                                                sendbuf.data_ptr() if sendbuf is not None else None,
                                                recvbuf.data_ptr(), count, dt, OP[op], root, EXEC[exec],
                                                self._stream(stream), iters), "bine_plan_autotune")


def _dtype_code(dtype) -> int:
    names = {"int32": 0, "int64": 1, "float32": 2, "float64": 3, "int8": 4}
    if isinstance(dtype, str):
        return names[dtype]
    try:
        import numpy as np
        if isinstance(dtype, type) or isinstance(dtype, np.dtype):
            return names[np.dtype(dtype).name]
    except Exception:
        pass
    s = str(dtype).replace("torch.", "")
    return names[s]


def plan_dump(dims: Sequence[int], coll: str, count: int, dtype, op: str = "SUM", root: int = 0, exec: str = "AUTO",
              num_ports: int = 1, tile_elems: int = 0, in_place: bool = False, two_phase: int = 0, **overrides):
    """Compile a schedule on the host (no GPU) and return its lowered work as
    numpy arrays: (info dict, items structured array, pool structured array,
    launches (n, 2) array).  in_place: the allreduce schedule for sendbuf == recvbuf.
    two_phase: 1 or 2 lowers reductions with several outputs as two launches (see Batch.set_tuning); adding 4
    gives the reduction launch items of two tiles."""
    import numpy as np
    L = _lib.load()
    cfg = make_config(dims, num_ports=num_ports, **overrides)
    info = _lib.PlanDumpInfo()
    args = (ctypes.byref(cfg), COLL[coll], count, _dtype_code(dtype), OP[op], root, 1 if in_place else 0, EXEC[exec],
            tile_elems, two_phase, ctypes.byref(info))
    _lib.check(L.bine_plan_dump(*args, None, None, None), "bine_plan_dump")
    item_dt = np.dtype([("leaf_off", "<u4"), ("out_off", "<u4"), ("lg", "<u2"), ("flags", "<u2"), ("nouts", "<u4"),
                        ("delta", "<u4"), ("n", "<u4"), ("bucket", "<u4"), ("pad", "<u4")])
    pool_dt = np.dtype([("rb", "<u4"), ("off", "<u4")])
    items = np.zeros(info.n_items, item_dt)
    pool = np.zeros(info.n_pool, pool_dt)
    launches = np.zeros((info.n_launches, 2), np.uint32)
    _lib.check(L.bine_plan_dump(*args, items.ctypes.data_as(ctypes.c_void_p), pool.ctypes.data_as(ctypes.c_void_p),
                                launches.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))), "bine_plan_dump")
    meta = {"P": info.P, "extent": list(info.extent), "single": list(info.single), "fused_ok": bool(info.fused_ok)}
    return meta, items, pool, launches


def _batch_sizes(coll: str, P: int, count: int):
    """(sendbuf elements or None, recvbuf elements) of one batch item, laid out
    as for the single call with that MPI count argument."""
    sizes = {
        "allreduce": (P * count, P * count),
        "reduce_scatter": (P * P * count, P * count),
        "allgather": (P * count, P * P * count),
        "bcast": (None, P * count),
        "alltoall": (P * P * count, P * P * count),
        "scatter": (P * count, P * count),
        "gather": (P * count, P * count),
        "reduce": (P * count, count),
    }
    if coll not in sizes:
        raise ValueError(f"unknown collective {coll!r}")
    return sizes[coll]


class Batch:
    """Many collectives on one vcomm compiled into one fused work list.

    ``items`` are dicts with ``coll`` (e.g. "allreduce"), ``sendbuf`` and
    ``recvbuf`` tensors (laid out as for the single call), ``count`` (the MPI
    count argument), optional ``op`` and ``root``.  The buffers are bound at
    creation, as for a captured graph."""

    def __init__(self, comm: VComm, items):
        self.comm = comm
        self._lib = comm._lib
        self._keep = []
        arr = (_lib.BatchItem * len(items))()
        for i, it in enumerate(items):
            recv = it["recvbuf"]
            send = it.get("sendbuf")
            n_send, n_recv = _batch_sizes(it["coll"], comm.size, int(it["count"]))
            VComm._expect(recv, n_recv, f"batch item {i} recvbuf")
            if n_send is not None:
                if send is None:
                    raise ValueError(f"batch item {i} ({it['coll']}) needs a sendbuf")
                VComm._expect(send, n_send, f"batch item {i} sendbuf", recv)
            self._keep.append((send, recv))
            arr[i].coll = COLL[it["coll"]]
            arr[i].sendbuf = send.data_ptr() if send is not None else None
            arr[i].recvbuf = recv.data_ptr()
            arr[i].count = int(it["count"])
            arr[i].datatype = _torch_dtype_code(recv)
            arr[i].op = OP[it.get("op", "SUM")]
            arr[i].root = int(it.get("root", 0))
        self.handle = ctypes.c_void_p()
        _lib.check(self._lib.bine_batch_create(comm.handle, arr, len(items), ctypes.byref(self.handle)),
                   "bine_batch_create")

    def launch(self, stream=None):
        _lib.check(self._lib.bine_batch_launch(self.handle, VComm._stream(stream)), "bine_batch_launch")

    def launch_graph(self, stream=None):
        _lib.check(self._lib.bine_batch_launch_graph(self.handle, VComm._stream(stream)), "bine_batch_launch_graph")

    def autotune(self, iters: int = 3, stream=None):
        _lib.check(self._lib.bine_batch_autotune(self.handle, VComm._stream(stream), iters), "bine_batch_autotune")

    def set_tuning(self, vpt: int, threads: int, persistent: int = 0, nontemporal: int = 0, two_phase: int = 0,
                   double_tiles: int = 0):
        """two_phase: 0 one launch per list; 1 reductions with several outputs write their first output, and a
        second launch copies it to the others (copies in group order); 2 the same, copies tile by tile.
        double_tiles: with two phases, reduction items of two tiles (half as many items)."""
        flags = (1 if nontemporal else 0) | (int(two_phase) << 1) | (8 if double_tiles else 0)
        _lib.check(self._lib.bine_batch_set_tuning(self.handle, vpt, threads, persistent, flags),
                   "bine_batch_set_tuning")

    def tuning(self) -> dict:
        v, t, p, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.bine_batch_get_tuning(self.handle, ctypes.byref(v), ctypes.byref(t), ctypes.byref(p),
                                                   ctypes.byref(f)), "bine_batch_get_tuning")
        return {"vpt": v.value, "threads": t.value, "persistent_wgs": p.value, "nontemporal": f.value & 1,
                "two_phase": (f.value >> 1) & 3, "double_tiles": (f.value >> 3) & 1}

    def time_shapes(self, shapes, iters: int = 3, stream=None):
        """Average ms of one launch at each launch shape (rows of vpt, threads,
        persistent, flags: bit 0 non-temporal, bits 1-2 the two-phase mode, bit 3 reduction items of two tiles);
        the batch keeps its current shape."""
        import numpy as np
        arr = np.ascontiguousarray(np.asarray(shapes, np.int32).reshape(-1, 4))
        ms = np.zeros(len(arr), np.float32)
        _lib.check(self._lib.bine_batch_time_shapes(self.handle, VComm._stream(stream), iters,
                                                    arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(arr),
                                                    ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "bine_batch_time_shapes")
        return ms

    def stats(self) -> dict:
        st = _lib.PlanStats()
        _lib.check(self._lib.bine_batch_stats(self.handle, ctypes.byref(st)), "bine_batch_stats")
        return {f: getattr(st, f) for f, _ in _lib.PlanStats._fields_}

    def close(self):
        if self.handle:
            self._lib.bine_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_candidates(device: int = 0):
    """The autotuner's launch shapes on a device: (n, 4) int32 rows of
    vectors per thread, threads, persistent workgroups, non-temporal."""
    import numpy as np
    L = _lib.load()
    n = ctypes.c_int()
    _lib.check(L.bine_tuning_candidates(device, None, 0, ctypes.byref(n)), "bine_tuning_candidates")
    out = np.zeros((n.value, 4), np.int32)
    _lib.check(L.bine_tuning_candidates(device, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value,
                                        ctypes.byref(n)), "bine_tuning_candidates")
    return out

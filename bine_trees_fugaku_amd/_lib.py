"""ctypes binding of libbine_hip.so (the C ABI declared in include/bine_hip.h).

The library is built in-tree by ``python -m bine_trees_fugaku_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbine_hip.so")

BINE_OK, BINE_ERR_ARG, BINE_ERR_UNSUPPORTED, BINE_ERR_HIP, BINE_ERR_NOMEM, BINE_ERR_NO_DEVICE = range(6)
STATUS_NAMES = {0: "BINE_OK", 1: "BINE_ERR_ARG", 2: "BINE_ERR_UNSUPPORTED", 3: "BINE_ERR_HIP", 4: "BINE_ERR_NOMEM",
                5: "BINE_ERR_NO_DEVICE"}

EXPORTED = [
    "bine_vcomm_config_init", "bine_vcomm_config_from_env", "bine_vcomm_create", "bine_vcomm_destroy",
    "bine_vcomm_size", "bine_vcomm_set_exec", "bine_vcomm_set_cache_limit", "bine_vcomm_trim",
    "bine_vcomm_cache_info", "bine_allreduce", "bine_allreduce_noalloc", "bine_reduce_scatter", "bine_allgather",
    "bine_bcast", "bine_alltoall", "bine_scatter", "bine_gather", "bine_reduce", "bine_compute_peers",
    "bine_compute_step_info", "bine_get_peer", "bine_get_mirroring_port", "bine_block_step", "bine_get_tree", "bine_remap_rank", "bine_reduce_local",
    "bine_reduce_local3",    "bine_plan_stats", "bine_plan_autotune", "bine_plan_dump", "bine_batch_create", "bine_batch_launch", "bine_batch_launch_graph",
    "bine_batch_autotune", "bine_batch_stats", "bine_batch_get_tuning", "bine_batch_set_tuning",
    "bine_batch_destroy", "bine_tuning_candidates", "bine_batch_time_shapes", "bine_sim_classes", "bine_sim_peers", "bine_sim_create", "bine_sim_run", "bine_sim_case_info",
    "bine_sim_class_stats", "bine_sim_link_counts", "bine_sim_destroy", "bine_mlp_create", "bine_mlp_num_params",
    "bine_mlp_get_params", "bine_mlp_set_params", "bine_mlp_forward", "bine_mlp_gradients", "bine_mlp_train_step",
    "bine_mlp_destroy", "bine_last_error",
]


class BineError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str):
        super().__init__(f"{what}: {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class CollConfig(ctypes.Structure):
    _fields_ = [("family", ctypes.c_int32), ("algo", ctypes.c_int32), ("distance", ctypes.c_int32)]


class VCommConfig(ctypes.Structure):
    _fields_ = [("dimensions", ctypes.c_uint32 * 3), ("dimensions_num", ctypes.c_uint32),
                ("num_ports", ctypes.c_uint32), ("allreduce", CollConfig), ("allgather", CollConfig),
                ("reduce_scatter", CollConfig), ("bcast", CollConfig), ("alltoall", CollConfig),
                ("scatter", CollConfig), ("gather", CollConfig), ("reduce", CollConfig),
                ("exec", ctypes.c_int32), ("device", ctypes.c_int32)]


class PlanStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("items", ctypes.c_uint64), ("bytes_read", ctypes.c_uint64),
                ("bytes_written", ctypes.c_uint64), ("max_tree_leaves", ctypes.c_uint64)]


class PlanDumpInfo(ctypes.Structure):
    _fields_ = [("n_items", ctypes.c_uint64), ("n_pool", ctypes.c_uint64), ("n_launches", ctypes.c_uint64),
                ("extent", ctypes.c_int64 * 4), ("single", ctypes.c_int32 * 4), ("P", ctypes.c_uint32),
                ("fused_ok", ctypes.c_int32)]


class SimCase(ctypes.Structure):
    _fields_ = [("dimensions", ctypes.c_uint32 * 3), ("dimensions_num", ctypes.c_uint32), ("coll", ctypes.c_int32),
                ("algo", ctypes.c_int32)]


class BatchItem(ctypes.Structure):
    _fields_ = [("coll", ctypes.c_int32), ("sendbuf", ctypes.c_void_p), ("recvbuf", ctypes.c_void_p),
                ("count", ctypes.c_int32), ("datatype", ctypes.c_int32), ("op", ctypes.c_int32),
                ("root", ctypes.c_int32)]


_lock = threading.Lock()
_lib = None


def _sig(L):
    vp, i32, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32
    pu32 = ctypes.POINTER(ctypes.c_uint32)
    pint = ctypes.POINTER(ctypes.c_int)
    L.bine_vcomm_config_init.argtypes = [ctypes.POINTER(VCommConfig), u32]
    L.bine_vcomm_config_from_env.argtypes = [ctypes.POINTER(VCommConfig)]
    L.bine_vcomm_create.argtypes = [ctypes.POINTER(VCommConfig), ctypes.POINTER(vp)]
    L.bine_vcomm_destroy.argtypes = [vp]
    L.bine_vcomm_size.argtypes = [vp, pu32]
    L.bine_vcomm_set_exec.argtypes = [vp, i32]
    L.bine_vcomm_set_cache_limit.argtypes = [vp, u32]
    L.bine_vcomm_trim.argtypes = [vp]
    L.bine_vcomm_cache_info.argtypes = [vp, pu32, ctypes.POINTER(ctypes.c_uint64)]
    L.bine_allreduce.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    L.bine_allreduce_noalloc.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    L.bine_reduce_scatter.argtypes = [vp, vp, pint, i32, i32, vp, vp]
    L.bine_allgather.argtypes = [vp, i32, i32, vp, i32, i32, vp, vp]
    L.bine_bcast.argtypes = [vp, i32, i32, i32, vp, vp]
    L.bine_alltoall.argtypes = [vp, i32, i32, vp, i32, i32, vp, vp]
    L.bine_scatter.argtypes = [vp, i32, i32, vp, i32, i32, i32, vp, vp]
    L.bine_gather.argtypes = [vp, i32, i32, vp, i32, i32, i32, vp, vp]
    L.bine_reduce.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp]
    L.bine_compute_peers.argtypes = [pu32, u32, u32, i32, i32, pu32]
    L.bine_get_peer.argtypes = [pu32, u32, ctypes.POINTER(ctypes.c_int32), u32, u32, i32, ctypes.POINTER(ctypes.c_int32)]
    L.bine_get_mirroring_port.argtypes = [i32, u32]
    L.bine_compute_step_info.argtypes = [pu32, u32, u32, pu32, pu32]
    L.bine_block_step.argtypes = [pu32, u32, u32, u32, i32, pu32]
    L.bine_get_tree.argtypes = [pu32, u32, u32, u32, i32, i32, pu32, pu32, pu32, pu32, pu32]
    L.bine_remap_rank.argtypes = [u32, u32]
    L.bine_remap_rank.restype = u32
    L.bine_reduce_local.argtypes = [vp, vp, i32, i32, i32, vp]
    L.bine_reduce_local3.argtypes = [vp, vp, vp, i32, i32, i32, vp]
    L.bine_plan_stats.argtypes = [vp, i32, i32, i32, i32, i32, i32, ctypes.POINTER(PlanStats)]
    L.bine_plan_autotune.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, i32, vp, i32]
    L.bine_plan_dump.argtypes = [ctypes.POINTER(VCommConfig), i32, i32, i32, i32, i32, i32, i32, ctypes.c_int64,
                                 i32, ctypes.POINTER(PlanDumpInfo), vp, vp, pu32]
    L.bine_batch_create.argtypes = [vp, ctypes.POINTER(BatchItem), i32, ctypes.POINTER(vp)]
    L.bine_batch_launch.argtypes = [vp, vp]
    L.bine_batch_launch_graph.argtypes = [vp, vp]
    L.bine_batch_autotune.argtypes = [vp, vp, i32]
    L.bine_batch_stats.argtypes = [vp, ctypes.POINTER(PlanStats)]
    L.bine_batch_get_tuning.argtypes = [vp, pint, pint, pint, pint]
    L.bine_batch_set_tuning.argtypes = [vp, i32, i32, i32, i32]
    L.bine_batch_destroy.argtypes = [vp]
    L.bine_tuning_candidates.argtypes = [i32, ctypes.POINTER(ctypes.c_int32), i32, pint]
    L.bine_batch_time_shapes.argtypes = [vp, vp, i32, ctypes.POINTER(ctypes.c_int32), i32,
                                         ctypes.POINTER(ctypes.c_float)]
    psc = ctypes.POINTER(SimCase)
    pu64 = ctypes.POINTER(ctypes.c_uint64)
    L.bine_sim_classes.argtypes = [psc, pu32, pu32, pu32, pu32, pu32, pu32, pu32]
    L.bine_sim_peers.argtypes = [psc, u32, pu32]
    L.bine_sim_create.argtypes = [psc, i32, i32, ctypes.POINTER(vp)]
    L.bine_sim_run.argtypes = [vp, vp]
    L.bine_sim_case_info.argtypes = [vp, i32, pu32, pu32]
    L.bine_sim_class_stats.argtypes = [vp, i32, pu32, pu32, pu32, pu32, pu64, pu32]
    L.bine_sim_link_counts.argtypes = [vp, i32, u32, pu32]
    L.bine_sim_destroy.argtypes = [vp]
    pf = ctypes.POINTER(ctypes.c_float)
    L.bine_mlp_create.argtypes = [ctypes.POINTER(ctypes.c_int32), i32, i32, ctypes.c_uint64, i32, ctypes.POINTER(vp)]
    L.bine_mlp_num_params.argtypes = [vp, pu64]
    L.bine_mlp_get_params.argtypes = [vp, pf]
    L.bine_mlp_set_params.argtypes = [vp, pf]
    L.bine_mlp_forward.argtypes = [vp, vp, i32, vp, vp]
    L.bine_mlp_gradients.argtypes = [vp, vp, vp, vp, i32, vp, vp, vp]
    L.bine_mlp_train_step.argtypes = [vp, vp, vp, vp, i32, ctypes.c_float, vp, vp]
    L.bine_mlp_destroy.argtypes = [vp]
    L.bine_last_error.restype = ctypes.c_char_p
    L.bine_last_error.argtypes = []


def load():
    """Load libbine_hip.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bine_trees_fugaku_amd.build`")
            # PyTorch ships its own HIP runtime (its libraries need libamdhip64.so from their own directory).
            # Loaded first, it also satisfies this library's libamdhip64.so.7, so the process has one runtime.
            # Loaded after this library, torch would bring a second HIP/HSA runtime, which then finds no GPU.
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            L = ctypes.CDLL(LIB_PATH)
            _sig(L)
            _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != BINE_OK:
        raise BineError(rc, what, load().bine_last_error().decode(errors="replace"))

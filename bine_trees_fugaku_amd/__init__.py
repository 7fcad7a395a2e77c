"""bine_trees_fugaku_amd: Bine-tree collectives over P virtual ranks held as
HBM partitions of one AMD Instinct MI355X (gfx950).

The compute path is libbine_hip.so (C ABI in include/bine_hip.h): a host
schedule compiler plus one hand-written HIP gather-reduce kernel.  This
package is the Python mirror of the reference's MPI-style interface.
"""
from .collectives import Batch, VComm, make_config, plan_dump  # noqa: F401
from ._lib import BineError, load  # noqa: F401

__all__ = ["Batch", "VComm", "make_config", "plan_dump", "BineError", "load"]

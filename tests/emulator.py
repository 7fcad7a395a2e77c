"""Numpy model of the gather-reduce kernel, for CPU tests of the host plans.

It executes the work items that libbine_hip.so would upload (obtained with
``plan_dump``) on host arrays: each item reduces its 2^lg leaves with the
balanced pairwise tree ((l0+l1)+(l2+l3))+... -- the association the HIP
kernel implements -- and stores the result to every output.  ITEM_RUN items
are runs of single copies whose source and destination advance by a fixed
(rank, offset) step.  Items of one
launch touch disjoint data (the plan compiler proves it), so running them in
order is equivalent to the GPU's concurrent execution.
"""
from __future__ import annotations

import numpy as np

BUF_SEND, BUF_RECV, BUF_TMP, BUF_TMP2 = 0, 1, 2, 3
ITEM_RUN = 2


def alloc_buffers(meta, dtype, send=None, recv=None):
    """Host buffers shaped (partitions, extent) for the four buffer ids."""
    P = meta["P"]
    bufs = {}
    for b in range(4):
        ext = meta["extent"][b]
        if ext <= 0:
            continue
        parts = 1 if meta["single"][b] else P
        bufs[b] = np.zeros((parts, ext), dtype=dtype)
    if send is not None:
        bufs[BUF_SEND][...] = np.asarray(send, dtype=dtype).reshape(bufs[BUF_SEND].shape)
    if recv is not None:
        bufs[BUF_RECV][...] = np.asarray(recv, dtype=dtype).reshape(bufs[BUF_RECV].shape)
    return bufs


def run(meta, items, pool, launches, bufs, reverse=False):
    """reverse: run each launch's items last to first (a launch's result must not
    depend on the order its items run in)."""
    single = meta["single"]

    def view(rb, off, n):
        buf, rank = int(rb) >> 28, int(rb) & 0x0FFFFFFF
        arr = bufs[buf]
        row = 0 if single[buf] else rank
        return arr[row, off:off + n]

    with np.errstate(over="ignore"):
        for first, cnt in launches:
            launch = items[first:first + cnt]
            for it in (launch[::-1] if reverse else launch):
                n = int(it["n"])
                d = int(it["delta"])
                if int(it["flags"]) & ITEM_RUN:  # k copies advancing by a fixed (rank, offset) step
                    lf, ls = pool[it["leaf_off"]], pool[it["leaf_off"] + 1]
                    of, os_ = pool[it["out_off"]], pool[it["out_off"] + 1]
                    for j in range(int(it["pad"])):
                        src = view(int(lf["rb"]) + j * int(ls["rb"]), int(lf["off"]) + j * int(ls["off"]) + d, n).copy()
                        view(int(of["rb"]) + j * int(os_["rb"]), int(of["off"]) + j * int(os_["off"]) + d, n)[...] = src
                    continue
                nl = 1 << int(it["lg"])
                lv = pool[it["leaf_off"]:it["leaf_off"] + nl]
                vals = [view(e["rb"], int(e["off"]) + d, n).copy() for e in lv]
                while len(vals) > 1:
                    vals = [vals[i] + vals[i + 1] for i in range(0, len(vals), 2)]
                for e in pool[it["out_off"]:it["out_off"] + int(it["nouts"])]:
                    view(e["rb"], int(e["off"]) + d, n)[...] = vals[0]
    return bufs

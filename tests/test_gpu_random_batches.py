"""Seeded random batches on the GPU: each batch holds two to six collectives of different kinds, element types,
counts and roots on one virtual communicator (an in-place allreduce now and then), is launched directly, replayed
as a captured graph, retuned and run in two phases, and every item's output is compared with the oracle bit for
bit after each.

Rooted reductions on a multi-dimensional torus use int32 (exact whatever the tree); copies (bcast, scatter,
gather, allgather, alltoall) have a unique result."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import Batch, VComm  # noqa: E402
from bine_trees_fugaku_amd._lib import BineError  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.random_cases import NP, data, pick  # noqa: E402

pytestmark = pytest.mark.gpu

DIMS = [[8], [16], [32], [64], [4, 4], [2, 2, 2], [4, 4, 4], [2, 8]]
N_BATCHES = 60


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def draw_batch(seed):
    rng = np.random.default_rng(5000 + seed)
    dims = pick(rng, DIMS)
    P = int(np.prod(dims))
    algo = pick(rng, ["B", "B_COALESCE", "B_CONT", "L"])
    exec_mode = pick(rng, ["FUSED", "AUTO", "STEP"])
    items = []
    for _ in range(int(rng.integers(2, 7))):
        coll = pick(rng, ["allreduce", "allreduce", "reduce_scatter", "allgather", "alltoall", "bcast", "reduce",
                          "scatter", "gather"])
        it = {"coll": coll, "root": 0}
        if coll == "allreduce":
            it["dtype"] = pick(rng, ["int8", "int32", "int64", "float32", "float64"])
            it["count"] = P * int(rng.integers(1, 40)) + int(rng.integers(0, P))
            it["send"] = data((P, it["count"]), it["dtype"], rng)
            it["in_place"] = bool(rng.random() < 0.2)
            it["want"] = O.allreduce(it["send"], dims, algo=algo)
        elif coll == "reduce_scatter":
            it["dtype"] = pick(rng, ["int32", "float32", "float64"])
            it["count"] = int(rng.integers(1, 50))
            it["send"] = data((P, P * it["count"]), it["dtype"], rng)
            it["want"] = O.reduce_scatter(it["send"], dims)
        elif coll == "allgather":
            it["dtype"] = pick(rng, ["int8", "int32", "float64"])
            it["count"] = int(rng.integers(1, 50))
            it["send"] = data((P, it["count"]), it["dtype"], rng)
            it["want"] = np.tile(it["send"].reshape(1, -1), (P, 1))
        elif coll == "alltoall":
            it["dtype"] = pick(rng, ["int32", "int64", "float32"])
            it["count"] = int(rng.integers(1, 20))
            s = data((P, P * it["count"]), it["dtype"], rng)
            it["send"] = s
            it["want"] = s.reshape(P, P, it["count"]).transpose(1, 0, 2).reshape(P, P * it["count"])
        else:
            it["root"] = 0 if rng.random() < 0.5 else 2 * int(rng.integers(P // 2))
            it["count"] = int(rng.integers(1, 400))
            if coll == "bcast":
                it["dtype"] = pick(rng, ["int8", "int32", "float64"])
                b = np.zeros((P, it["count"]), NP[it["dtype"]])
                b[it["root"]] = data((it["count"],), it["dtype"], rng)
                it["send"] = b
                it["want"] = np.tile(b[it["root"]].reshape(1, -1), (P, 1))
            elif coll == "reduce":
                it["dtype"] = pick(rng, ["int32", "float32", "float64"]) if len(dims) == 1 else "int32"
                it["send"] = data((P, it["count"]), it["dtype"], rng)
                it["want"] = O.reduce(it["send"], it["root"])
            elif coll == "scatter":
                it["root"] = 0
                it["dtype"] = pick(rng, ["int32", "float32"])
                it["send"] = data((P * it["count"],), it["dtype"], rng)
                it["want"] = it["send"].reshape(P, it["count"])
            else:
                it["dtype"] = pick(rng, ["int32", "float64"])
                it["send"] = data((P, it["count"]), it["dtype"], rng)
                it["want"] = it["send"].reshape(-1)
        items.append(it)
    name = f"{seed}: {dims} {algo} {exec_mode} " + ", ".join(
        f"{i['coll']}/{i['dtype']}/{i['count']}" + ("/in-place" if i.get("in_place") else "") +
        (f"/root={i['root']}" if i["root"] else "") for i in items)
    return dims, P, algo, exec_mode, items, name


def buffers(it, P):
    tdt = getattr(torch, it["dtype"])
    coll, n = it["coll"], it["count"]
    if coll == "allreduce":
        s = dev(it["send"])
        return s, (s if it["in_place"] else torch.zeros_like(s))
    if coll == "reduce_scatter":
        return dev(it["send"]), torch.zeros((P, n), dtype=tdt, device="cuda")
    if coll == "allgather":
        return dev(it["send"]), torch.zeros((P, P * n), dtype=tdt, device="cuda")
    if coll == "alltoall":
        return dev(it["send"]), torch.zeros((P, P * n), dtype=tdt, device="cuda")
    if coll == "bcast":
        return None, dev(it["send"])
    if coll == "reduce":
        return dev(it["send"]), torch.zeros(n, dtype=tdt, device="cuda")
    if coll == "scatter":
        return dev(it["send"]), torch.zeros((P, n), dtype=tdt, device="cuda")
    return dev(it["send"]), torch.zeros(P * n, dtype=tdt, device="cuda")


def check(items, bufs, name, when):
    torch.cuda.synchronize()
    for i, (it, (s, r)) in enumerate(zip(items, bufs)):
        got = r.cpu().numpy()
        assert got.tobytes() == np.ascontiguousarray(it["want"]).tobytes(), f"{name}: item {i} after {when}"


def reset(items, bufs):
    """Inputs of in-place and bcast items are overwritten by a run: restore them before the next."""
    for it, (s, r) in zip(items, bufs):
        if it["coll"] == "bcast" or (it["coll"] == "allreduce" and it["in_place"]):
            r.copy_(dev(it["send"]))


@pytest.mark.parametrize("seed", range(N_BATCHES))
def test_random_batch(gpu, seed):
    dims, P, algo, exec_mode, items, name = draw_batch(seed)
    print(name, flush=True)
    with VComm(dims, exec=exec_mode, allreduce_algo=algo) as c:
        bufs = [buffers(it, P) for it in items]
        spec = [{"coll": it["coll"], "sendbuf": s, "recvbuf": r, "count": it["count"], "root": it["root"]}
                for it, (s, r) in zip(items, bufs)]
        try:
            bt = Batch(c, spec)
        except BineError as e:
            assert e.code == 2, f"{name}: {e}"
            pytest.skip(f"refused as in the reference: {name}")
        try:
            bt.launch()
            check(items, bufs, name, "launch")
            reset(items, bufs)
            bt.launch_graph()
            check(items, bufs, name, "graph replay")
            reset(items, bufs)
            bt.autotune(iters=1)
            reset(items, bufs)
            bt.launch()
            check(items, bufs, name, "retuned launch")
            reset(items, bufs)
            bt.set_tuning(2, 128, 0, 1, two_phase=1 + seed % 2)  # reductions, then copies of their first output
            bt.launch()
            check(items, bufs, name, "two-phase launch")
        finally:
            bt.close()

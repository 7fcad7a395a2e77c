"""CPU tests of the host schedule compiler of libbine_hip.so.

Every schedule is compiled with ``plan_dump`` (no GPU involved), executed by
the numpy model of the gather-reduce kernel (tests/emulator.py) and compared
byte for byte with the CPU oracle, in both execution modes (STEP: one launch
per schedule step; FUSED: one launch).  Only the HIP kernel itself is left for
the GPU tests.
"""
import numpy as np
import pytest

from bine_trees_fugaku_amd import plan_dump
from bine_trees_fugaku_amd._lib import BineError
from oracle import oracle as O
from tests import emulator as E

NP = {"int32": np.int32, "int64": np.int64, "float32": np.float32, "float64": np.float64}


def data(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype.startswith("int"):
        return rng.integers(-1000, 1001, size=shape).astype(NP[dtype])
    return rng.standard_normal(shape).astype(NP[dtype])


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def run_plan(dims, coll, count, dtype, exec, send=None, recv=None, tile=1 << 16, **kw):
    meta, items, pool, launches = plan_dump(dims, coll, count, dtype, exec=exec, tile_elems=tile, **kw)
    bufs = E.alloc_buffers(meta, NP[dtype], send, recv)
    E.run(meta, items, pool, launches, bufs)
    return meta, bufs


def modes(dims, coll, count, dtype, **kw):
    out = ["STEP"]
    meta, *_ = plan_dump(dims, coll, count, dtype, exec="STEP", **kw)
    if meta["fused_ok"]:
        out.append("FUSED")
    return out


AR_TORI = [[8], [2, 2, 2], [4, 4], [4, 4, 4], [2, 8, 2], [16], [6], [10], [6, 6], [2, 6, 2], [4, 6, 4]]


@pytest.mark.parametrize("dims", AR_TORI)
@pytest.mark.parametrize("algo", ["B", "B_COALESCE", "L", "B_CONT"])
@pytest.mark.parametrize("dtype", ["int32", "float32", "float64"])
def test_allreduce(dims, algo, dtype):
    P = int(np.prod(dims))
    if algo == "B_CONT" and P & (P - 1):
        pytest.skip("B_CONT is power-of-two only")
    count = 3 * P + 5
    x = data((P, count), dtype, P)
    ref = O.allreduce(x, dims, algo=algo)
    for ex in modes(dims, "allreduce", count, dtype, allreduce_algo=algo):
        _, bufs = run_plan(dims, "allreduce", count, dtype, ex, send=x, allreduce_algo=algo)
        assert same_bits(bufs[E.BUF_RECV], ref), (dims, algo, ex)


def test_allreduce_fused_available_on_p2():
    for dims in ([2, 2, 2], [4, 4, 4], [8, 8, 8]):
        meta, *_ = plan_dump(dims, "allreduce", 4096 * 8, "float64", exec="STEP")
        assert meta["fused_ok"]


@pytest.mark.parametrize("ports", [2, 3, 6])
@pytest.mark.parametrize("algo", ["B", "L"])
def test_allreduce_multiport(ports, algo):
    dims = [4, 4, 4]
    count = 64 * ports + 11
    x = data((64, count), "float32", ports)
    ref = O.allreduce(x, dims, algo=algo, ports=ports)
    for ex in modes(dims, "allreduce", count, "float32", num_ports=ports, allreduce_algo=algo):
        _, bufs = run_plan(dims, "allreduce", count, "float32", ex, send=x, num_ports=ports, allreduce_algo=algo)
        assert same_bits(bufs[E.BUF_RECV], ref)


def test_allreduce_recdoub():
    dims = [4, 4]
    x = data((16, 100), "float64", 1)
    ref = O.allreduce(x, dims, family="RECDOUB")
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan(dims, "allreduce", 100, "float64", ex, send=x, allreduce_family="RECDOUB")
        assert same_bits(bufs[E.BUF_RECV], ref)


def test_allreduce_void():
    dims = [4, 4]
    x = data((16, 48), "int32", 2)
    ref = O.allreduce(x, dims, op="VOID")
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan(dims, "allreduce", 48, "int32", ex, send=x, op="VOID")
        assert same_bits(bufs[E.BUF_RECV], ref)


def test_allreduce_preconditions():
    with pytest.raises(BineError):  # count < ports * P (lib/libbine.cc:1036)
        plan_dump([4, 4], "allreduce", 10, "int32")
    with pytest.raises(BineError):  # B needs INCREASING distance (lib/libbine.cc:1035)
        plan_dump([4, 4], "allreduce", 64, "int32", allreduce_distance="DECREASING")
    with pytest.raises(BineError):  # B_CONT on non power of two
        plan_dump([6], "allreduce", 64, "int32", allreduce_algo="B_CONT")


@pytest.mark.parametrize("dims", [[8], [4, 4], [2, 2, 2], [4, 4, 4], [16]])
@pytest.mark.parametrize("dtype", ["int32", "float32", "float64"])
def test_reduce_scatter(dims, dtype):
    P = int(np.prod(dims))
    c = 5
    x = data((P, c * P), dtype, P)
    ref = O.reduce_scatter(x, dims)
    for ex in modes(dims, "reduce_scatter", c, dtype):
        _, bufs = run_plan(dims, "reduce_scatter", c, dtype, ex, send=x)
        assert same_bits(bufs[E.BUF_RECV], ref), ex


@pytest.mark.parametrize("dims", [[8], [4, 4], [2, 2, 2], [4, 4, 4], [16]])
@pytest.mark.parametrize("algo", ["VEC_DOUBLING_CONT_PERMUTE", "VEC_DOUBLING_CONT_SEND", "VEC_DOUBLING_BLOCKS"])
def test_allgather(dims, algo):
    if algo == "VEC_DOUBLING_BLOCKS" and len(dims) > 1:
        pytest.skip("1D only")
    P = int(np.prod(dims))
    a = data((P, 7), "int32", P)
    variant = {"VEC_DOUBLING_CONT_PERMUTE": "CONT_PERMUTE", "VEC_DOUBLING_CONT_SEND": "CONT_SEND",
               "VEC_DOUBLING_BLOCKS": "BLOCKS"}[algo]
    ref = O.allgather(a, dims, variant=variant)
    for ex in modes(dims, "allgather", 7, "int32", allgather_algo=algo):
        _, bufs = run_plan(dims, "allgather", 7, "int32", ex, send=a, allgather_algo=algo)
        assert same_bits(bufs[E.BUF_RECV], ref), ex


def test_allgather_multiport():
    a = data((64, 12), "float32", 4)
    ref = O.allgather(a, [4, 4, 4], ports=3)
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan([4, 4, 4], "allgather", 12, "float32", ex, send=a, num_ports=3)
        assert same_bits(bufs[E.BUF_RECV], ref)


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64])
@pytest.mark.parametrize("family", ["BINE", "BRUCK"])
def test_alltoall(P, family):
    c = 3
    x = data((P, c * P), "int64", P)
    ref = O.alltoall(x, P, family=family)
    for ex in modes([P], "alltoall", c, "int64", alltoall_family=family):
        _, bufs = run_plan([P], "alltoall", c, "int64", ex, send=x, alltoall_family=family)
        assert same_bits(bufs[E.BUF_RECV], ref), ex


@pytest.mark.parametrize("P", [6, 10, 12])
def test_alltoall_bruck_non_p2(P):
    x = data((P, 2 * P), "int32", P)
    ref = O.alltoall(x, P, family="BRUCK")
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan([P], "alltoall", 2, "int32", ex, send=x, alltoall_family="BRUCK")
        assert same_bits(bufs[E.BUF_RECV], ref)


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64])
def test_bcast(P):
    for root in sorted({0, 1, P // 2, P - 1}):
        b = np.zeros((P, 9), np.float32)
        b[root] = data((9,), "float32", root)
        ref = O.bcast(b, root)
        for ex in modes([P], "bcast", 9, "float32", root=root):
            _, bufs = run_plan([P], "bcast", 9, "float32", ex, recv=b, root=root)
            assert same_bits(bufs[E.BUF_RECV], ref), (root, ex)
    b = np.zeros((P, 13), np.int32)
    b[0] = data((13,), "int32", 5)
    ref = O.bcast(b, 0, "SCATTER_ALLGATHER")
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan([P], "bcast", 13, "int32", ex, recv=b, bcast_algo="SCATTER_ALLGATHER",
                           bcast_distance="INCREASING")
        assert same_bits(bufs[E.BUF_RECV], ref)


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64])
@pytest.mark.parametrize("dtype", ["int32", "float64"])
def test_reduce(P, dtype):
    for root in sorted({0, 1, P - 1}):
        s = data((P, 9), dtype, root + P)
        ref = O.reduce(s, root)
        for ex in modes([P], "reduce", 9, dtype, root=root):
            _, bufs = run_plan([P], "reduce", 9, dtype, ex, send=s, root=root)
            assert same_bits(bufs[E.BUF_RECV][0], ref), (root, ex)
    s = data((P, 13), dtype, 3)
    ref = O.reduce(s, 0, "REDUCE_SCATTER_GATHER")
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan([P], "reduce", 13, dtype, ex, send=s, reduce_algo="REDUCE_SCATTER_GATHER")
        assert same_bits(bufs[E.BUF_RECV][0], ref)


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64])
def test_scatter_gather(P):
    sr = data((P * 4,), "int32", P)
    ref = O.scatter(sr, P, 0)
    for ex in ("STEP", "FUSED"):
        _, bufs = run_plan([P], "scatter", 4, "int32", ex, send=sr, root=0)
        assert same_bits(bufs[E.BUF_RECV], ref)
    s = data((P, 4), "int32", P + 1)
    for root in range(0, P, 2):
        ref = O.gather(s, root)
        for ex in ("STEP", "FUSED"):
            _, bufs = run_plan([P], "gather", 4, "int32", ex, send=s, root=root)
            assert same_bits(bufs[E.BUF_RECV][0], ref), (root, ex)


def test_unsupported_configurations_are_errors():
    with pytest.raises(BineError):
        plan_dump([4], "gather", 4, "int32", root=1)  # odd root: wrong data in the reference
    with pytest.raises(BineError):
        plan_dump([6], "gather", 4, "int32")  # non power of two: reference deadlocks
    with pytest.raises(BineError):
        plan_dump([8], "scatter", 4, "int32", root=2)
    with pytest.raises(BineError):
        plan_dump([4, 4], "reduce_scatter", 4, "int32", num_ports=2)  # recvbuf + offset_port quirk


def test_fused_bytes_are_compulsory():
    """The fused allreduce reads every input once and writes every output once."""
    P, count = 64, 64 * 32
    meta, items, pool, launches = plan_dump([4, 4, 4], "allreduce", count, "float32", exec="FUSED")
    reads = sum(int(it["n"]) << int(it["lg"]) for it in items)
    writes = sum(int(it["n"]) * int(it["nouts"]) for it in items)
    assert reads == P * count and writes == P * count
    assert len(launches) == 1


@pytest.mark.parametrize("coll,dims,count,dtype", [
    ("allgather", [4, 4, 4], 1001, "float32"), ("allreduce", [4, 4], 16 * 37 + 3, "float32"),
    ("alltoall", [8], 17, "int32"), ("scatter", [8], 1001, "float64"), ("allreduce", [4, 4], 16 * 37 + 3, "int8"),
    ("allreduce", [4, 4], 16 * 40 + 4, "int8")])
def test_vector_items_are_dword_aligned(coll, dims, count, dtype):
    """A tile is lowered as 16-byte vectors (ITEM_VEC) only when every leaf and
    output offset it touches is dword-aligned and its length is a whole number
    of vectors.  Leaves and outputs need not share a residue modulo the vector
    (the kernels load unaligned and realign stores), so odd-count copies and
    reductions still vectorize; int8 blocks off a dword boundary do not."""
    size = {"int8": 1, "int32": 4, "float32": 4, "float64": 8}[dtype]
    vec = 16 // size
    for ex in modes(dims, coll, count, dtype):
        meta, items, pool, launches = plan_dump(dims, coll, count, dtype, exec=ex)
        vitems = items[(items["flags"] & 1) != 0]
        if dtype != "int8":
            assert len(vitems) > 0, ex
        for it in vitems:
            assert it["n"] % vec == 0
            if it["flags"] & 2:  # copy run: first entry and per-copy step
                offs = [pool[it["leaf_off"]]["off"], pool[it["leaf_off"] + 1]["off"],
                        pool[it["out_off"]]["off"], pool[it["out_off"] + 1]["off"]]
            else:
                ents = np.concatenate([pool[it["leaf_off"]: it["leaf_off"] + (1 << it["lg"])],
                                       pool[it["out_off"]: it["out_off"] + it["nouts"]]])
                offs = [int(e) + int(it["delta"]) for e in ents["off"]]
            assert all((int(o) * size) % 4 == 0 for o in offs), (ex, it)


@pytest.mark.parametrize("coll,count", [("scatter", 64 * 5 + 12), ("gather", 64 * 6 + 4), ("scatter", 64 * 2),
                                        ("gather", 64 * 3 + 5)])
def test_long_copies_lower_to_tile_pairs(coll, count):
    """A single copy of at least two tiles is lowered as copy runs of two consecutive tiles (first entry plus a
    step of one tile, same rank); odd tile counts and ragged tails stay ordinary items.  The emulated result
    equals the oracle."""
    P, dims, tile = 8, [8], 64
    x = data((P, count), "int32", count)
    meta, items, pool, launches = plan_dump(dims, coll, count, "int32", exec="FUSED", tile_elems=tile)
    runs = items[(items["flags"] & 2) != 0]
    pairs = [it for it in runs if pool[it["leaf_off"] + 1]["rb"] == 0 and pool[it["out_off"] + 1]["rb"] == 0]
    assert len(pairs) > 0
    for it in pairs:  # two consecutive whole tiles of one block
        assert it["n"] == tile and it["pad"] == 2 and it["flags"] & 1
        assert pool[it["leaf_off"] + 1]["off"] == tile and pool[it["out_off"] + 1]["off"] == tile
    if coll == "scatter":
        send = x.reshape(-1)
        _, bufs = run_plan(dims, coll, count, "int32", "FUSED", send=send, tile=tile, root=0)
        assert same_bits(bufs[E.BUF_RECV], O.scatter(send, P, 0))
    else:
        _, bufs = run_plan(dims, coll, count, "int32", "FUSED", send=x, tile=tile, root=0)
        assert same_bits(bufs[E.BUF_RECV][0], O.gather(x, 0))


# ---- MPI_IN_PLACE allreduce ------------------------------------------------------------

IN_PLACE_TORI = [[8], [2, 2, 2], [2, 8, 2], [8, 8, 8], [4, 4], [16], [6], [10], [6, 6], [2, 6, 2]]


@pytest.mark.parametrize("dims", IN_PLACE_TORI)
@pytest.mark.parametrize("algo", ["B", "B_COALESCE", "L", "B_CONT"])
@pytest.mark.parametrize("ports", [1, 3])
def test_allreduce_in_place(dims, algo, ports):
    """sendbuf == recvbuf: the in-place schedule reads RECV only and gives the oracle's bits
    whatever order a launch's items run in (odd step counts included: [8], [2, 2, 2], [8, 8, 8])."""
    P = int(np.prod(dims))
    if algo == "B_CONT" and P & (P - 1):
        pytest.skip("B_CONT is power-of-two only")
    if ports > 1 and (len(dims) < 3 or P > 64):
        pytest.skip("ports > 1 on the small 3D tori only")
    count = 3 * ports * P + 5
    x = data((P, count), "float64", P + ports)
    ref = O.allreduce(x, dims, algo=algo, ports=ports)
    meta, items, pool, launches = plan_dump(dims, "allreduce", count, "float64", exec="STEP", in_place=True,
                                            num_ports=ports, allreduce_algo=algo, tile_elems=1 << 16)
    assert meta["extent"][0] == 0 and not meta["fused_ok"]
    for rev in (False, True):
        bufs = E.alloc_buffers(meta, np.float64, recv=x)
        E.run(meta, items, pool, launches, bufs, reverse=rev)
        assert same_bits(bufs[E.BUF_RECV], ref), (algo, rev)


def test_allreduce_in_place_needs_its_own_schedule():
    """The out-of-place L schedule on an odd step count writes RECV at its first step while
    other ranks still read SEND: run with SEND and RECV as one array, it goes wrong in at
    least one item order.  That is why in-place calls get a schedule of their own."""
    dims, P = [8], 8
    count = 3 * P + 5
    x = data((P, count), "float64", 3)
    ref = O.allreduce(x, dims, algo="L")
    meta, items, pool, launches = plan_dump(dims, "allreduce", count, "float64", exec="STEP", allreduce_algo="L",
                                            tile_elems=1 << 16)
    wrong = 0
    for rev in (False, True):
        bufs = E.alloc_buffers(meta, np.float64, send=x)
        bufs[E.BUF_RECV] = bufs[E.BUF_SEND]  # one array: MPI_IN_PLACE through the out-of-place schedule
        E.run(meta, items, pool, launches, bufs, reverse=rev)
        wrong += not same_bits(bufs[E.BUF_RECV], ref)
    assert wrong >= 1


def test_in_place_rejected_for_other_collectives():
    with pytest.raises(BineError):
        plan_dump([4, 4], "reduce_scatter", 4, "int32", in_place=True)


# ---- family DEFAULT (the reference's vendor-MPI fallback) ---------------------------------

DEFAULT_STANDIN = [("allreduce", [4, 4, 4], "RECDOUB"), ("allreduce", [6, 4], "BINE"), ("reduce_scatter", [4, 4], "RECDOUB"),
                   ("allgather", [2, 2, 2], "RECDOUB"), ("alltoall", [16], "BRUCK"), ("bcast", [8], "BINE"),
                   ("reduce", [16], "BINE"), ("scatter", [8], "BINE"), ("gather", [8], "BINE")]


@pytest.mark.parametrize("coll,dims,standin", DEFAULT_STANDIN)
def test_family_default_runs_its_stand_in(coll, dims, standin):
    """LIBBINE_<COLL>_ALGO_FAMILY=DEFAULT calls PMPI_* in the reference (lib/libbine.cc:1016): here it compiles
    the stand-in schedule, work item for work item."""
    P = int(np.prod(dims))
    count = 4 * P + (3 if coll in ("allreduce", "bcast", "reduce") else 0)
    a = plan_dump(dims, coll, count, "float32", exec="STEP", **{f"{coll}_family": "DEFAULT"})
    b = plan_dump(dims, coll, count, "float32", exec="STEP", **{f"{coll}_family": standin})
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert x.tobytes() == y.tobytes()


# ---- two-phase lowering ----------------------------------------------------------------

@pytest.mark.parametrize("dims,algo,ports", [([8], "B", 1), ([4, 4], "B", 1), ([4, 4, 4], "B_CONT", 1), ([2, 8, 2], "B", 1),
                                             ([4, 4], "B", 2), ([2, 2, 2], "B_COALESCE", 3), ([8, 8, 8], "B", 1)])
@pytest.mark.parametrize("mode", [1, 2, 5, 6])
@pytest.mark.parametrize("dtype", ["float32", "float64", "int32"])
def test_two_phase_allreduce(dims, algo, ports, mode, dtype):
    """Two-phase lowering of the fused allreduce (the schedules that fuse: powers of two; L and non-powers of two
    run per step): the first launch holds the reductions, each writing one output; the second holds copies of
    that output to the block's other P - 1 outputs (group by group in mode 1, tile by tile in mode 2; adding 4
    gives the reductions items of two tiles).  Every input is read once, every output written once, and each
    block's sum once more read back for the copies.  The emulated result equals the oracle bit for bit."""
    P = int(np.prod(dims))
    count = (40 if P <= 64 else 5) * P * ports + 3
    tile = 16
    x = data((P, count), dtype, P + mode)
    kw = dict(allreduce_algo=algo, num_ports=ports)
    meta, items, pool, launches = plan_dump(dims, "allreduce", count, dtype, exec="FUSED", tile_elems=tile,
                                            two_phase=mode, **kw)
    assert len(launches) == 2
    (f1, n1), (f2, n2) = launches
    first, second = items[f1:f1 + n1], items[f2:f2 + n2]
    red = first[first["lg"] > 0]
    assert len(red) > 0 and (red["nouts"] == 1).all()
    assert (red["n"] <= (2 if mode & 4 else 1) * tile).all()
    if mode & 4 and P <= 64:
        assert (red["n"] > tile).any()  # blocks of 2.5 tiles: some items span two
    assert (second["lg"] == 0).all() and (second["nouts"] == P - 1).all() and (second["n"] <= tile).all()
    reads = sum(int(it["n"]) << int(it["lg"]) for it in items)
    writes = sum(int(it["n"]) * int(it["nouts"]) for it in items)
    assert reads == (P + 1) * count and writes == P * count
    if mode & 3 == 1:  # group by group: one block's copies are consecutive
        assert (np.diff(second["leaf_off"].astype(np.int64)) >= 0).all()
    else:  # the same copies in another order
        m1 = plan_dump(dims, "allreduce", count, dtype, exec="FUSED", tile_elems=tile, two_phase=mode - 1, **kw)
        other = m1[1][m1[3][1][0]:]
        assert sorted(r.tobytes() for r in second) == sorted(r.tobytes() for r in other)
    _, bufs = run_plan(dims, "allreduce", count, dtype, "FUSED", send=x, tile=tile, two_phase=mode, **kw)
    assert same_bits(bufs[E.BUF_RECV], O.allreduce(x, dims, algo=algo, ports=ports))


@pytest.mark.parametrize("coll", ["reduce_scatter", "reduce", "allgather", "bcast"])
def test_two_phase_leaves_other_lists_alone(coll):
    """Lists without a reduction of several outputs lower the same with or without two phases."""
    dims, count = [4, 4], 64
    kw = {"root": 0} if coll in ("reduce", "bcast") else {}
    a = plan_dump(dims, coll, count, "float32", exec="FUSED", **kw)
    b = plan_dump(dims, coll, count, "float32", exec="FUSED", two_phase=2, **kw)
    assert len(a[3]) == len(b[3]) == 1
    assert a[1].tobytes() == b[1].tobytes() and a[2].tobytes() == b[2].tobytes()

"""Batched execution on the GPU: uneven buckets and mixed collectives in one
fused work list, eager and graph-replayed, against single calls and the oracle.

Every batched output must equal, bit for bit, what the single collective
call writes (itself checked against the oracle in test_gpu_parity.py), and a
sampled column of each output must equal the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import Batch, VComm  # noqa: E402
from bine_trees_fugaku_amd._lib import BineError  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def randn(shape, dtype, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randn(shape, dtype=dtype, device="cuda", generator=g)


def randint(shape, dtype, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    lo, hi = (-128, 128) if dtype == torch.int8 else (-1000, 1001)
    return torch.randint(lo, hi, shape, dtype=dtype, device="cuda", generator=g)


def same_bits(a, b):
    return torch.equal(a.view(torch.uint8), b.view(torch.uint8))


def bucket_sizes(n, seed, lo=4096, hi=1 << 20):
    rng = np.random.default_rng(seed)
    return (np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64) // 4).tolist()


@pytest.mark.parametrize("mode", ["launch", "graph"])
def test_uneven_buckets_equal_single_calls(gpu, mode):
    P, dims = 64, [4, 4, 4]
    sizes = bucket_sizes(24, 1) + [64, 65, 127, 4097]  # one element per block, ragged tails
    sends = [randn((P, n), torch.float32, 100 + i) for i, n in enumerate(sizes)]
    recvs = [torch.full_like(s, float("nan")) for s in sends]
    with VComm(dims, exec="FUSED") as comm, VComm(dims, exec="STEP") as step:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": n}
                         for s, r, n in zip(sends, recvs, sizes)])
        refs = [torch.empty_like(s) for s in sends]
        for s, r in zip(sends, refs):
            step.allreduce(s, r)
        replays = 3 if mode == "graph" else 1
        for k in range(replays):
            for r in recvs:
                r.fill_(float("nan"))  # eager work between replays
            torch.cuda.synchronize()
            b.launch_graph() if mode == "graph" else b.launch()
            torch.cuda.synchronize()
            for i, (r, ref) in enumerate(zip(recvs, refs)):
                assert same_bits(r, ref), (k, i, sizes[i])
        b.close()


def test_uneven_buckets_against_oracle(gpu):
    """Small buckets checked directly against the oracle (whole buckets, 0 ulp)."""
    P, dims = 8, [2, 2, 2]
    sizes = [8, 9, 15, 100, 1000, 4099]
    sends = [randn((P, n), torch.float64, 7 + i) for i, n in enumerate(sizes)]
    recvs = [torch.empty_like(s) for s in sends]
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": n}
                         for s, r, n in zip(sends, recvs, sizes)])
        b.launch_graph()
        torch.cuda.synchronize()
        for s, r in zip(sends, recvs):
            want = O.allreduce(s.cpu().numpy(), dims)
            assert np.array_equal(r.cpu().numpy().view(np.int64), want.view(np.int64))
        b.close()


def test_every_tuning_gives_the_same_bits(gpu):
    P, dims = 512, [8, 8, 8]
    count = 4096 + 3
    send = randn((P, count), torch.float64, 5)
    recv = torch.empty_like(send)
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": count}])
        b.set_tuning(1, 256, 0, 0)
        b.launch()
        ref = recv.clone()
        for vpt, thr, pers, nt in [(2, 128, 0, 0), (4, 64, 0, 1), (4, 64, 4096, 1), (1, 512, 512, 0), (2, 64, 1024, 1)]:
            recv.fill_(0)
            b.set_tuning(vpt, thr, pers, nt)
            b.launch_graph()
            torch.cuda.synchronize()
            assert same_bits(recv, ref), (vpt, thr, pers, nt)
        b.autotune(iters=1)
        recv.fill_(0)
        b.launch_graph()
        torch.cuda.synchronize()
        assert same_bits(recv, ref)
        b.close()


def mixed_items(P, seed, scale=1):
    """The 8x8x8 mixed workload: allreduce fp64 4 MiB per rank, alltoall int64
    4 KiB per pair, bcast int32 64 KiB, scatter fp32 256 KiB blocks, gather int32
    1 MiB blocks, reduce fp32 1 MiB (sizes divided by `scale`)."""
    c_ar, c_a2a = 524288 // scale, 512 // scale
    c_b, c_s, c_g, c_r = 16384 // scale, 65536 // scale, 262144 // scale, 262144 // scale
    t = {
        "allreduce": (randn((P, c_ar), torch.float64, seed), torch.empty((P, c_ar), dtype=torch.float64, device="cuda"), c_ar),
        "alltoall": (randint((P, P * c_a2a), torch.int64, seed + 1), torch.empty((P, P * c_a2a), dtype=torch.int64, device="cuda"), c_a2a),
        "bcast": (None, randint((P, c_b), torch.int32, seed + 2), c_b),
        "scatter": (randn((P * c_s,), torch.float32, seed + 3), torch.empty((P, c_s), dtype=torch.float32, device="cuda"), c_s),
        "gather": (randint((P, c_g), torch.int32, seed + 4), torch.empty((P * c_g,), dtype=torch.int32, device="cuda"), c_g),
        "reduce": (randn((P, c_r), torch.float32, seed + 5), torch.empty((c_r,), dtype=torch.float32, device="cuda"), c_r),
    }
    return t


def single_call(comm, coll, send, recv):
    if coll == "allreduce":
        comm.allreduce(send, recv)
    elif coll == "alltoall":
        comm.alltoall(send, recv)
    elif coll == "bcast":
        comm.bcast(recv)
    elif coll == "scatter":
        comm.scatter(send, recv)
    elif coll == "gather":
        comm.gather(send, recv)
    elif coll == "reduce":
        comm.reduce(send, recv)


@pytest.mark.parametrize("scale", [64, 1])
def test_mixed_collectives_8x8x8(gpu, scale):
    P, dims = 512, [8, 8, 8]
    t = mixed_items(P, 11, scale)
    bcast_in = t["bcast"][1].clone()
    with VComm(dims, exec="FUSED") as comm:
        items = [{"coll": c, "sendbuf": s, "recvbuf": r, "count": n} for c, (s, r, n) in t.items()]
        b = Batch(comm, items)
        print(f"mixed batch scale 1/{scale}: {b.stats()}", flush=True)
        b.launch_graph()
        torch.cuda.synchronize()
        got = {c: r.clone() for c, (s, r, n) in t.items()}
        # replay the captured graph and launch directly again, with eager work in between: every result must
        # match the first
        for k, how in enumerate(("graph", "graph", "launch", "graph")):
            t["bcast"][1].copy_(bcast_in)
            for c, (s, r, n) in t.items():
                if c != "bcast":
                    r.fill_(float("nan") if r.dtype.is_floating_point else -1)
            torch.cuda.synchronize()
            b.launch_graph() if how == "graph" else b.launch()
            torch.cuda.synchronize()
            for c, (s, r, n) in t.items():
                assert same_bits(got[c], r), (k, how, c)
        b.close()
        # single calls on fresh outputs
        t["bcast"][1].copy_(bcast_in)
        for c, (s, r, n) in t.items():
            if c != "bcast":
                r.fill_(0)
            single_call(comm, c, s, r)
        torch.cuda.synchronize()
        for c, (s, r, n) in t.items():
            assert same_bits(got[c], r), c
    # oracle on the small instance (whole buffers)
    if scale == 64:
        a = {c: (None if s is None else s.cpu().numpy(), r.cpu().numpy()) for c, (s, r, n) in t.items()}
        assert np.array_equal(got["allreduce"].cpu().numpy(), O.allreduce(a["allreduce"][0], dims))
        assert np.array_equal(got["alltoall"].cpu().numpy(), O.alltoall(a["alltoall"][0], P))
        assert np.array_equal(got["bcast"].cpu().numpy(), O.bcast(bcast_in.cpu().numpy(), 0))
        assert np.array_equal(got["scatter"].cpu().numpy(), O.scatter(a["scatter"][0], P, 0))
        assert np.array_equal(got["gather"].cpu().numpy().reshape(-1), O.gather(a["gather"][0], 0).reshape(-1))
        assert np.array_equal(got["reduce"].cpu().numpy().reshape(-1), O.reduce(a["reduce"][0], 0).reshape(-1))


def test_batch_alignment_per_bucket(gpu):
    """One batch mixing int8 buckets whose strides are not 4-byte aligned (scalar items) with int8, int32 and
    float32 buckets whose strides are 4-byte but not 16-byte aligned (vector items): every bucket against the
    oracle."""
    P, dims = 16, [4, 4]
    specs = [(torch.int8, 16 * 9 + 1), (torch.int8, 16 * 9 + 4), (torch.int32, 16 * 33 + 5),
             (torch.float32, 16 * 21 + 3), (torch.int8, 16 * 64 + 7), (torch.float32, 16 * 64)]
    sends = [(randint if dt != torch.float32 else randn)((P, n), dt, 40 + i) for i, (dt, n) in enumerate(specs)]
    recvs = [torch.zeros_like(s) for s in sends]
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": s.shape[1]}
                         for s, r in zip(sends, recvs)])
        for replay in range(2):
            for r in recvs:
                r.zero_()
            b.launch() if replay == 0 else b.launch_graph()
            torch.cuda.synchronize()
            for i, (s, r) in enumerate(zip(sends, recvs)):
                want = O.allreduce(s.cpu().numpy(), dims)
                assert r.cpu().numpy().tobytes() == want.tobytes(), (replay, i, specs[i])
        b.close()


def test_realigning_kernels_with_deep_trees(gpu):
    """A float32 list that writes far more than it reads (a bcast of many outputs per leaf) runs the realigning
    kernels; its odd-count allreduce then takes the realigned stores in both kernel paths: shallow trees, and deep
    trees split across the waves of 128- to 512-thread workgroups.  Every candidate shape writes the oracle's
    bytes."""
    from bine_trees_fugaku_amd.collectives import tuning_candidates
    P, dims = 64, [4, 4, 4]
    cb, ca = 64 * 61 + 5, 64 * 37 + 5
    root_row = randn((cb,), torch.float32, 8)
    bc = torch.zeros((P, cb), dtype=torch.float32, device="cuda")
    send = randn((P, ca), torch.float32, 9)
    recv = torch.zeros_like(send)
    want_ar = O.allreduce(send.cpu().numpy(), dims)
    want_bc = np.tile(root_row.cpu().numpy().reshape(1, -1), (P, 1))
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "bcast", "sendbuf": None, "recvbuf": bc, "count": cb},
                         {"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": ca}])
        st = b.stats()
        assert st["bytes_written"] >= 2 * st["bytes_read"]  # store-heavy: the realigning kernels run
        for vpt, thr, pers, nt in tuning_candidates():
            recv.zero_()
            bc.zero_()
            bc[0] = root_row
            b.set_tuning(int(vpt), int(thr), int(pers), int(nt))
            b.launch()
            torch.cuda.synchronize()
            assert recv.cpu().numpy().tobytes() == want_ar.tobytes(), (vpt, thr, pers, nt)
            assert bc.cpu().numpy().tobytes() == want_bc.tobytes(), (vpt, thr, pers, nt)
        b.close()


def test_batch_rejects_bad_items(gpu):
    P = 8
    s = randn((P, 16), torch.float32, 0)
    r = torch.empty((P, 15), dtype=torch.float32, device="cuda")
    with VComm([2, 2, 2]) as comm:
        with pytest.raises(ValueError):
            Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": 16}])
        with pytest.raises(ValueError):
            Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": s.double(), "count": 16}])
        small = randn((P, 7), torch.float32, 1)
        with pytest.raises(BineError):  # count < P: lib/libbine.cc:1036 asserts the same
            Batch(comm, [{"coll": "allreduce", "sendbuf": small, "recvbuf": torch.empty_like(small), "count": 7}])
        r2 = torch.empty_like(s)
        with pytest.raises(BineError):  # outputs of two items overlap
            Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r2, "count": 16},
                         {"coll": "allreduce", "sendbuf": s, "recvbuf": r2, "count": 16}])


@pytest.mark.parametrize("collective,type_name,count", [
    ("MPI_Allreduce", "INT32", 262144), ("MPI_Reduce_scatter", "FLOAT", 4096), ("MPI_Allgather", "INT", 1024),
    ("MPI_Bcast", "CHAR", 65536), ("MPI_Alltoall", "INT32", 64), ("MPI_Scatter", "FLOAT", 1024),
    ("MPI_Gather", "INT32", 1024), ("MPI_Reduce", "VOID", 65536), ("MPI_Allreduce", "DOUBLE", 4099),
    ("MPI_Reduce", "INT32", 1000), ("MPI_Allreduce", "CHAR", 4096), ("MPI_Allreduce", "FLOAT", 5000)])
def test_harness_runs_and_validates(gpu, collective, type_name, count, capsys):
    """bench/bench.cc's loop on the GPU: record printed, timed output equal to the MPI result computed directly
    with torch (exact for integer types, char sums wrapping) and to per-step execution."""
    from bine_trees_fugaku_amd import harness as H
    with VComm([2, 2, 2]) as comm:
        run = H.Run(collective, type_name, count, comm=comm)
        samples = run.time(3, warmup=1)
        assert len(samples) == 3 and all(s > 0 for s in samples)
        assert run.validate()
    rec = H.format_record(samples).splitlines()
    assert rec[0] == "highest" and len(rec) == 5


@pytest.mark.parametrize("coll", ["alltoall", "allgather", "scatter", "reduce_scatter", "allreduce_f32",
                                  "allreduce_f64", "allreduce_i8", "bcast"])
def test_every_candidate_shape_gives_the_same_bits(gpu, coll):
    """Copy runs and reductions at every autotuner candidate shape (vectors per
    thread, workgroup, persistent grid, non-temporal) write the oracle's bytes.
    The odd counts put partitions and blocks 4, 8 or 12 bytes past 16-byte
    boundaries: the store-heavy allgather and bcast lists run the realigning
    kernels (aligned chunks built with lane shuffles), the others store
    unaligned 16-byte vectors."""
    from bine_trees_fugaku_amd.collectives import tuning_candidates
    P, dims = 64, [4, 4, 4]
    c = {"alltoall": 96, "allgather": 1001, "bcast": 64 * 55 + 7, "scatter": 3001, "reduce_scatter": 777, "allreduce_f32": 64 * 37 + 5,
         "allreduce_f64": 64 * 29 + 3, "allreduce_i8": 64 * 80 + 12}[coll]
    if coll == "alltoall":
        send, recv = randint((P, P * c), torch.int64, 1), torch.zeros((P, P * c), dtype=torch.int64, device="cuda")
        want = O.alltoall(send.cpu().numpy(), P)
    elif coll == "allgather":
        send, recv = randn((P, c), torch.float32, 2), torch.zeros((P, P * c), dtype=torch.float32, device="cuda")
        want = O.allgather(send.cpu().numpy(), dims)
    elif coll == "scatter":
        send, recv = randn((P * c,), torch.float64, 3), torch.zeros((P, c), dtype=torch.float64, device="cuda")
        want = O.scatter(send.cpu().numpy(), P)
    elif coll == "bcast":
        send, recv = None, torch.zeros((P, c), dtype=torch.float32, device="cuda")
        root_row = randn((c,), torch.float32, 6)
        want = np.tile(root_row.cpu().numpy().reshape(1, -1), (P, 1))
    elif coll == "reduce_scatter":
        send, recv = randn((P, P * c), torch.float32, 4), torch.zeros((P, c), dtype=torch.float32, device="cuda")
        want = O.reduce_scatter(send.cpu().numpy(), dims)
    else:
        dt = {"allreduce_f32": torch.float32, "allreduce_f64": torch.float64, "allreduce_i8": torch.int8}[coll]
        send = (randint if dt == torch.int8 else randn)((P, c), dt, 5)
        recv = torch.zeros_like(send)
        want = O.allreduce(send.cpu().numpy(), dims)
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": coll.split("_")[0] if coll.startswith("allreduce") else coll,
                          "sendbuf": send, "recvbuf": recv, "count": c}])
        for vpt, thr, pers, nt in tuning_candidates():
            recv.zero_()
            if coll == "bcast":
                recv[0] = root_row
            b.set_tuning(int(vpt), int(thr), int(pers), int(nt))
            b.launch()
            torch.cuda.synchronize()
            got = recv.cpu().numpy()
            assert got.tobytes() == want.reshape(got.shape).tobytes(), (coll, vpt, thr, pers, nt)
        b.close()


def test_starting_shape_follows_block_length(gpu):
    """Before any tuning a list of long blocks in at least 1024 tiles starts at four vectors per thread, one-wave
    workgroups and streaming (4, 64, 0, 1); a list of one-tile blocks starts at the general shape (1, 256, 0, 0).
    Both write the bits the other shape writes."""
    P, dims = 64, [4, 4, 4]
    for count, want in [(16 * 1024 * 64, (4, 64, 0, 1)), (1024 * 64, (1, 256, 0, 0))]:
        send = randn((P, count), torch.float32, 31)
        recv = torch.empty_like(send)
        with VComm(dims, exec="FUSED") as comm:
            b = Batch(comm, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": count}])
            t = b.tuning()
            assert (t["vpt"], t["threads"], t["persistent_wgs"], t["nontemporal"]) == want, (count, t)
            assert t["two_phase"] == 0
            b.launch()
            torch.cuda.synchronize()
            first = recv.clone()
            other = (1, 256, 0, 0) if want != (1, 256, 0, 0) else (4, 64, 0, 1)
            b.set_tuning(*other)
            recv.zero_()
            b.launch()
            torch.cuda.synchronize()
            assert same_bits(recv, first), count
            b.close()

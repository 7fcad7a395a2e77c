"""Two-phase work lists on the GPU: the reductions write each block's sum once, a second launch copies it to the
block's other outputs.  Every launch shape in both copy orders, graph replays, uneven bucket lists, mixed batches
and an autotuned list must write the oracle's bytes, or the bytes of the one-launch list (itself checked against
the oracle)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import Batch, VComm  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.test_gpu_batch import bucket_sizes, mixed_items, randint, randn, same_bits, single_call  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("double", [0, 1])
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("coll", ["allreduce_f32", "allreduce_f64", "allreduce_i8"])
def test_every_candidate_shape_two_phase(gpu, coll, mode, double):
    """Odd counts on 4x4x4: blocks start 4, 8 or 12 bytes past 16-byte boundaries and end in ragged tails.
    double: reduction items of two tiles."""
    from bine_trees_fugaku_amd.collectives import tuning_candidates
    P, dims = 64, [4, 4, 4]
    dt, c = {"allreduce_f32": (torch.float32, 64 * 37 + 5), "allreduce_f64": (torch.float64, 64 * 29 + 3),
             "allreduce_i8": (torch.int8, 64 * 80 + 12)}[coll]
    send = (randint if dt == torch.int8 else randn)((P, c), dt, 5)
    recv = torch.zeros_like(send)
    want = O.allreduce(send.cpu().numpy(), dims)
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": c}])
        for vpt, thr, pers, nt in tuning_candidates():
            recv.zero_()
            b.set_tuning(int(vpt), int(thr), int(pers), int(nt), two_phase=mode, double_tiles=double)
            assert b.tuning()["two_phase"] == mode and b.tuning()["double_tiles"] == double
            b.launch()
            torch.cuda.synchronize()
            got = recv.cpu().numpy()
            assert got.tobytes() == want.reshape(got.shape).tobytes(), (coll, mode, double, vpt, thr, pers, nt)
        assert b.stats()["launches"] == 2
        b.close()


def test_two_phase_512_ranks_launch_and_graph(gpu):
    P, dims = 512, [8, 8, 8]
    count = 4096 + 3
    send = randn((P, count), torch.float64, 5)
    recv = torch.empty_like(send)
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": count}])
        b.set_tuning(1, 256, 0, 0)
        b.launch()
        torch.cuda.synchronize()
        ref = recv.clone()
        col = O.allreduce(send[:, :P].cpu().numpy().copy(), dims)  # a 512-element allreduce, one element per block
        for mode in (1, 2):
            for vpt, thr, pers, nt, dbl in [(4, 64, 0, 1, 0), (1, 512, 0, 0, 1), (2, 128, 1024, 1, 0), (4, 256, 0, 1, 1)]:
                b.set_tuning(vpt, thr, pers, nt, two_phase=mode, double_tiles=dbl)
                for k in range(3):  # graph replays with eager work between them
                    recv.fill_(float("nan"))
                    torch.cuda.synchronize()
                    b.launch_graph() if k else b.launch()
                    torch.cuda.synchronize()
                    assert same_bits(recv, ref), (mode, vpt, thr, pers, nt, dbl, k)
        b.close()
    # a two-phase list of one element per block against the oracle
    with VComm(dims, exec="FUSED") as comm:
        x = send[:, :P].contiguous()
        y = torch.empty_like(x)
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": x, "recvbuf": y, "count": P}])
        b.set_tuning(4, 64, 0, 1, two_phase=1)
        b.launch()
        torch.cuda.synchronize()
        assert y.cpu().numpy().tobytes() == col.tobytes()
        b.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_two_phase_uneven_buckets(gpu, mode):
    P, dims = 64, [4, 4, 4]
    sizes = bucket_sizes(24, 1) + [64, 65, 127, 4097]
    sends = [randn((P, n), torch.float32, 100 + i) for i, n in enumerate(sizes)]
    recvs = [torch.full_like(s, float("nan")) for s in sends]
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": n}
                         for s, r, n in zip(sends, recvs, sizes)])
        b.launch()
        torch.cuda.synchronize()
        refs = [r.clone() for r in recvs]
        for i in (0, 5, 24, 25, 27):
            assert recvs[i].cpu().numpy().tobytes() == O.allreduce(sends[i].cpu().numpy(), dims).tobytes(), i
        for shape, dbl in [((4, 64, 0, 1), 0), ((2, 256, 0, 1), 1), ((1, 512, 2048, 0), 0), ((4, 64, 0, 1), 1)]:
            b.set_tuning(*shape, two_phase=mode, double_tiles=dbl)
            for k in range(2):
                for r in recvs:
                    r.fill_(float("nan"))
                torch.cuda.synchronize()
                b.launch_graph() if k else b.launch()
                torch.cuda.synchronize()
                for i, (r, ref) in enumerate(zip(recvs, refs)):
                    assert same_bits(r, ref), (mode, shape, dbl, k, i, sizes[i])
        b.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_two_phase_mixed_batch(gpu, mode):
    P, dims = 512, [8, 8, 8]
    t = mixed_items(P, 11, scale=64)
    with VComm(dims, exec="FUSED") as comm:
        refs = {}
        for coll, (s, r, n) in t.items():
            ref = r.clone() if coll == "bcast" else torch.empty_like(r)
            single_call(comm, coll, s, ref)
            refs[coll] = ref
        b = Batch(comm, [{"coll": coll, "sendbuf": s, "recvbuf": r, "count": n} for coll, (s, r, n) in t.items()])
        b.set_tuning(2, 128, 0, 1, two_phase=mode)
        b.launch()
        torch.cuda.synchronize()
        for coll, (s, r, n) in t.items():
            assert same_bits(r, refs[coll]), (mode, coll)
        b.close()


def test_autotune_with_two_phase_trials(gpu):
    """A large 512-rank allreduce: the autotuner times the two-phase trials too; whatever it keeps writes the
    same bits as the one-launch list."""
    P, dims = 512, [8, 8, 8]
    count = 65536
    send = randn((P, count), torch.float64, 9)
    recv = torch.empty_like(send)
    with VComm(dims, exec="FUSED") as comm:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": count}])
        b.set_tuning(4, 64, 0, 1)
        b.launch()
        torch.cuda.synchronize()
        ref = recv.clone()
        b.autotune(iters=1)
        tun = b.tuning()
        recv.zero_()
        b.launch()
        torch.cuda.synchronize()
        assert same_bits(recv, ref), tun
        b.close()


@pytest.mark.parametrize("dims,count,dt", [([4, 4, 4], 64 * 4096 + 7, torch.float32), ([8, 8, 8], 512 * 256, torch.float64)])
def test_single_call_autotune_with_two_phase_trials(gpu, dims, count, dt):
    """bine_plan_autotune on a fused allreduce also times the two-phase trials; the call it leaves behind writes
    the bits it wrote before tuning, and the oracle's on a sampled block."""
    P = int(np.prod(dims))
    send = randn((P, count), dt, 21)
    recv = torch.empty_like(send)
    with VComm(dims, exec="FUSED") as comm:
        comm.allreduce(send, recv)
        torch.cuda.synchronize()
        ref = recv.clone()
        comm.autotune("allreduce", send, recv, count, exec="FUSED", iters=1)
        recv.zero_()
        comm.allreduce(send, recv)
        torch.cuda.synchronize()
        assert same_bits(recv, ref)
    cols = send[:, :P].contiguous()  # a P-element allreduce: one element per block, against the oracle
    out = torch.empty_like(cols)
    with VComm(dims, exec="FUSED") as comm:
        comm.autotune("allreduce", cols, out, P, exec="FUSED", iters=1)
        comm.allreduce(cols, out)
        torch.cuda.synchronize()
        assert out.cpu().numpy().tobytes() == O.allreduce(cols.cpu().numpy(), dims).tobytes()

"""The uneven-bucket workload at its full size, as a test.

256 fp32 buckets of log-uniform size 4 KiB .. 8 MiB per rank (seed 7, rounded
to 4 bytes; 278.7 MB per rank, 17.8 GB per buffer set), P = 64 on a 4x4x4
torus, N(0,1) data from torch's generator seeded 7: the same buckets as the
benchmark's bucketed leg.  The batch (one fused launch over all buckets),
its graph replay and the per-step schedule (one launch per bucket per step)
must write identical bits, and one element of every block of every bucket
must equal the oracle (an element's reduction tree depends only on its
block, lib/libbine.cc:971-983, so the oracle runs at one element per block).
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import Batch, VComm  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

DIMS, P, NB, SEED = [4, 4, 4], 64, 256, 7


def bucket_sizes():
    rng = np.random.default_rng(SEED)
    return np.exp(rng.uniform(np.log(4096), np.log(8 << 20), NB)).astype(np.int64) // 4


def block_starts(n):
    """First element of each of the P blocks of an n-element allreduce (one port)."""
    q, rem = divmod(n, P)
    return np.array([i * q + min(i, rem) for i in range(P)], dtype=np.int64)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_buckets_full_size(gpu):
    sizes = bucket_sizes()
    assert sizes.min() >= 1024 and sizes.max() <= (8 << 20) // 4
    say(f"buckets: {NB} fp32 buckets, {int(sizes.sum()) * 4} bytes per rank, P={P} on 4x4x4")
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED)
    sends = [torch.randn((P, int(n)), dtype=torch.float32, device="cuda", generator=g) for n in sizes]
    batched = [torch.full_like(s, float("nan")) for s in sends]
    stepped = [torch.full_like(s, float("nan")) for s in sends]
    with VComm(DIMS, exec="FUSED") as comm, VComm(DIMS, exec="STEP") as step:
        b = Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": int(n)}
                         for s, r, n in zip(sends, batched, sizes)])
        b.set_tuning(4, 64, 0, 1)  # the shape the autotuner picks for this list on an MI355X
        say("buckets: batch compiled")
        for s, r in zip(sends, stepped):
            step.allreduce(s, r)
        torch.cuda.synchronize()
        say("buckets: per-step done")
        b.launch()
        torch.cuda.synchronize()
        for i in range(NB):
            assert torch.equal(batched[i].view(torch.int32), stepped[i].view(torch.int32)), ("launch", i)
        say("buckets: batched launch equals per-step")
        for k in range(2):
            for r in batched:
                r.fill_(float("nan"))  # eager work between the replays
            torch.cuda.synchronize()
            b.launch_graph()
            torch.cuda.synchronize()
            for i in range(NB):
                assert torch.equal(batched[i].view(torch.int32), stepped[i].view(torch.int32)), ("graph", k, i)
        say("buckets: graph replays equal per-step")
        for i, n in enumerate(sizes):
            cols = torch.from_numpy(block_starts(int(n))).cuda()
            x = sends[i].index_select(1, cols).cpu().numpy()
            got = batched[i].index_select(1, cols).cpu().numpy()
            want = O.allreduce(np.ascontiguousarray(x), DIMS)
            assert np.array_equal(got.view(np.int32), want.view(np.int32)), ("oracle", i, int(n))
        say("buckets: one element per block of every bucket equals the oracle")
        b.close()

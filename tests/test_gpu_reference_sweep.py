"""The reference's own correctness sweep, run on the GPU against the oracle.

The reference validates with INT32, count 131072, every collective on N in
{4, 8, 64} ranks, the allreduce variants on the non-powers of two
{6, 10, 12, 14, 18} and on the virtual tori of validate/allreduce.sh:86-139
(the cases of the reference's validate/*.sh).  Here every case is compared bit
for bit with the oracle, in per-step and fused execution.  For N = 64 the three
collectives whose buffers grow with N^2 (reduce_scatter, allgather, alltoall)
use count 2048 so each case stays within a few hundred MiB.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import VComm  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
COUNT = 131072


def ints(shape, seed):
    return np.random.default_rng(seed).integers(-1000, 1001, size=shape).astype(np.int32)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def run_both(dims, fn, **cfg):
    outs = []
    for exec in ("STEP", "AUTO"):
        with VComm(dims, exec=exec, **cfg) as c:
            outs.append(fn(c))
    return outs


@pytest.mark.parametrize("N", [4, 8, 64])
@pytest.mark.parametrize("coll", ["allreduce", "reduce_scatter", "allgather", "alltoall", "bcast", "reduce",
                                  "scatter", "gather"])
def test_every_collective(gpu, N, coll):
    print(f"sweep: {coll} N={N}", flush=True)
    big = coll in ("reduce_scatter", "allgather", "alltoall") and N == 64
    c = 2048 if big else COUNT
    dims = [N]
    if coll == "allreduce":
        x = ints((N, c), N)
        ref = O.allreduce(x, dims)

        def f(cm):
            r = torch.zeros((N, c), dtype=torch.int32, device="cuda")
            cm.allreduce(dev(x), r)
            return host(r)
    elif coll == "reduce_scatter":
        x = ints((N, c * N), N + 1)
        ref = O.reduce_scatter(x, dims)

        def f(cm):
            r = torch.zeros((N, c), dtype=torch.int32, device="cuda")
            cm.reduce_scatter(dev(x), r)
            return host(r)
    elif coll == "allgather":
        x = ints((N, c), N + 2)
        ref = O.allgather(x, dims)

        def f(cm):
            r = torch.zeros((N, c * N), dtype=torch.int32, device="cuda")
            cm.allgather(dev(x), r)
            return host(r)
    elif coll == "alltoall":
        x = ints((N, c * N), N + 3)
        ref = O.alltoall(x, N)

        def f(cm):
            r = torch.zeros((N, c * N), dtype=torch.int32, device="cuda")
            cm.alltoall(dev(x), r)
            return host(r)
    elif coll == "bcast":
        x = np.zeros((N, c), np.int32)
        x[0] = ints((c,), N + 4)
        ref = O.bcast(x.copy(), 0)

        def f(cm):
            t = dev(x)
            cm.bcast(t)
            return host(t)
    elif coll == "reduce":
        x = ints((N, c), N + 5)
        ref = O.reduce(x, 0)

        def f(cm):
            r = torch.zeros((c,), dtype=torch.int32, device="cuda")
            cm.reduce(dev(x), r)
            return host(r)
    elif coll == "scatter":
        x = ints((N * c,), N + 6)
        ref = O.scatter(x, N, 0)

        def f(cm):
            r = torch.zeros((N, c), dtype=torch.int32, device="cuda")
            cm.scatter(dev(x), r)
            return host(r)
    else:
        x = ints((N, c), N + 7)
        ref = O.gather(x, 0)

        def f(cm):
            r = torch.zeros((N * c,), dtype=torch.int32, device="cuda")
            cm.gather(dev(x), r)
            return host(r)
    for out in run_both(dims, f):
        assert np.array_equal(out.reshape(-1), np.asarray(ref).reshape(-1))


@pytest.mark.parametrize("N", [6, 10, 12, 14, 18])
@pytest.mark.parametrize("algo", ["B", "B_COALESCE", "L"])
def test_allreduce_non_power_of_two(gpu, N, algo):
    x = ints((N, COUNT), 100 + N)
    ref = O.allreduce(x, [N], algo=algo)
    assert np.array_equal(ref[0], x.sum(0, dtype=np.int64).astype(np.int32))

    def f(cm):
        r = torch.zeros((N, COUNT), dtype=torch.int32, device="cuda")
        cm.allreduce(dev(x), r)
        return host(r)
    for out in run_both([N], f, allreduce_algo=algo):
        assert np.array_equal(out, ref)


TORI = [[2, 8, 2], [6, 6], [10, 10], [4, 6, 4], [6, 2, 6], [2, 6, 2], [2, 2, 6], [6, 2, 2], [4, 4, 4]]


@pytest.mark.parametrize("dims", TORI, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("algo", ["B", "L"])
def test_allreduce_virtual_tori(gpu, dims, algo):
    P = int(np.prod(dims))
    x = ints((P, COUNT), 200 + P)
    ref = O.allreduce(x, dims, algo=algo)

    def f(cm):
        r = torch.zeros((P, COUNT), dtype=torch.int32, device="cuda")
        cm.allreduce(dev(x), r)
        return host(r)
    for out in run_both(dims, f, allreduce_algo=algo):
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("dims", [[2, 8, 2], [4, 4, 4], [2, 2, 2]], ids=lambda d: "x".join(map(str, d)))
def test_allgather_virtual_tori(gpu, dims):
    P = int(np.prod(dims))
    c = COUNT // P
    x = ints((P, c), 300 + P)
    ref = O.allgather(x, dims)

    def f(cm):
        r = torch.zeros((P, c * P), dtype=torch.int32, device="cuda")
        cm.allgather(dev(x), r)
        return host(r)
    for out in run_both(dims, f):
        assert np.array_equal(out, ref)

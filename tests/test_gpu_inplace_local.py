"""GPU tests of MPI_IN_PLACE allreduce, the exported local reductions, and the
plan cache's lifetime rules (temporaries shared across streams, eviction,
work lists replaced under a captured graph).

In-place results are compared bit for bit with the oracle (every algorithm,
odd step counts, several ports, non-powers of two).  reduce_local is compared
with numpy for all five element types (integers wrap, floats are one IEEE
addition per element, so the results are exact)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import Batch, VComm  # noqa: E402
from bine_trees_fugaku_amd._lib import BineError  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"int8": np.int8, "int32": np.int32, "int64": np.int64, "float32": np.float32, "float64": np.float64}


def data(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == "int8":
        return rng.integers(-128, 128, size=shape).astype(np.int8)
    if dtype.startswith("int"):
        return rng.integers(-2**30, 2**30, size=shape).astype(NP[dtype])
    return (rng.standard_normal(shape) * 10.0 ** rng.uniform(-3, 3, shape)).astype(NP[dtype])


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def same_bits(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


# ---- MPI_IN_PLACE ----------------------------------------------------------------------

@pytest.mark.parametrize("dims", [[8], [2, 2, 2], [2, 8, 2], [8, 8, 8], [4, 4], [6], [10], [2, 6, 2]])
@pytest.mark.parametrize("algo", ["B", "B_COALESCE", "L", "B_CONT"])
def test_allreduce_in_place_every_algorithm(gpu, dims, algo):
    P = int(np.prod(dims))
    if algo == "B_CONT" and P & (P - 1):
        pytest.skip("B_CONT is power-of-two only")
    count = 5 * P + 3 if P <= 64 else 2 * P + 1
    x = data((P, count), "float64", P)
    ref = O.allreduce(x, dims, algo=algo)
    with VComm(dims, allreduce_algo=algo) as c:
        buf = dev(x)
        c.allreduce(buf, buf)
        assert same_bits(host(buf), ref)
        buf.copy_(dev(x))  # a second call on the cached in-place schedule
        c.allreduce(buf, buf)
        assert same_bits(host(buf), ref)
        with pytest.raises(BineError):  # fused execution cannot alias its inputs
            c.with_exec("FUSED").allreduce(buf, buf)


@pytest.mark.parametrize("algo", ["B", "L"])
def test_allreduce_in_place_multiport_int32(gpu, algo):
    dims, P, ports = [2, 2, 2], 8, 3
    count = 3 * ports * P + 7
    x = data((P, count), "int32", 11)
    with VComm(dims, num_ports=ports, allreduce_algo=algo) as c:
        buf = dev(x)
        c.allreduce(buf, buf)
        assert same_bits(host(buf), O.allreduce(x, dims, algo=algo, ports=ports))


def test_in_place_batch_item(gpu):
    """A batch mixing a fused item with an in-place allreduce item (run per step after the fused list)."""
    dims, P = [8], 8
    x = data((P, 77), "float64", 1)
    y = data((P, 90), "float64", 2)
    with VComm(dims, allreduce_algo="L") as c:
        a = dev(x)
        s, r = dev(y), torch.empty((P, 90), dtype=torch.float64, device="cuda")
        b = Batch(c, [{"coll": "allreduce", "sendbuf": a, "recvbuf": a, "count": 77},
                      {"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": 90}])
        for k in range(3):
            a.copy_(dev(x))
            r.fill_(0)
            torch.cuda.synchronize()
            b.launch_graph() if k else b.launch()
            assert same_bits(host(a), O.allreduce(x, dims, algo="L")), k
            assert same_bits(host(r), O.allreduce(y, dims, algo="L")), k
        b.close()


# ---- reduce_local ----------------------------------------------------------------------

@pytest.mark.parametrize("dtype", ["int8", "int32", "int64", "float32", "float64"])
@pytest.mark.parametrize("n", [0, 1, 17, 4099, (1 << 22) + 3])
def test_reduce_local_two_and_three_operands(gpu, dtype, n):
    a, b = data((n,), dtype, n), data((n,), dtype, n + 1)
    with np.errstate(over="ignore"):
        want = (a + b).astype(NP[dtype])
    with VComm([2]) as c:
        x, y = dev(a), dev(b)
        c.reduce_local(x, y)  # y = x + y
        assert same_bits(host(y), want)
        z = torch.zeros_like(x)
        c.reduce_local3(x, dev(b), z)  # z = x + b
        assert same_bits(host(z), want)
        y = dev(b)
        c.reduce_local(x, y, op="VOID")  # VOID: no effect
        assert same_bits(host(y), b)
        z.fill_(0)
        c.reduce_local3(x, y, z, op="VOID")  # VOID writes nothing
        assert not host(z).any()


@pytest.mark.parametrize("dtype", ["int8", "int32", "float32", "float64"])
def test_reduce_local_unaligned_and_aliased(gpu, dtype):
    """Inputs one element off any 16-byte boundary (scalar path for int8, vector path at dword
    alignment otherwise) and inbuf == inoutbuf."""
    n = 100003
    a, b = data((n + 1,), dtype, 3), data((n + 2,), dtype, 4)
    with np.errstate(over="ignore"):
        want = (a[1:] + b[2:]).astype(NP[dtype])
        twice = (a[1:] + a[1:]).astype(NP[dtype])
    with VComm([2]) as c:
        x, y = dev(a), dev(b)
        c.reduce_local(x[1:], y[2:])
        assert same_bits(host(y)[2:], want)
        assert same_bits(host(y)[:2], b[:2])
        c.reduce_local(x[1:], x[1:])
        assert same_bits(host(x)[1:], twice)


def test_reduce_local_rejects_bad_arguments(gpu):
    with VComm([2]) as c:
        x = torch.zeros(1000, dtype=torch.float32, device="cuda")
        with pytest.raises(BineError):  # partial overlap of output and input
            c.reduce_local(x[:900], x[100:])
        with pytest.raises(BineError):
            c.reduce_local3(x[:500], x[500:], x[10:510])
        lib = c._lib
        rc = lib.bine_reduce_local(x.data_ptr(), x.data_ptr(), 10, 2, 7, None)  # op 7 is neither SUM nor VOID
        assert rc == 1


# ---- plan cache lifetime ------------------------------------------------------------------

def test_per_step_temporaries_ordered_across_streams(gpu):
    """Allreduce L per step keeps its running value in a temporary shared by every call of the
    same size: calls on two streams, issued back to back, must not overlap on it."""
    dims, P, count = [4, 4, 4], 64, 1 << 16
    xs = [data((P, count), "float32", 40 + i) for i in range(4)]
    refs = [O.allreduce(x, dims, algo="L") for x in xs]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with VComm(dims, exec="STEP", allreduce_algo="L") as c:
        sends = [dev(x) for x in xs]
        recvs = [torch.empty_like(s) for s in sends]
        torch.cuda.synchronize()
        for i in range(4):
            c.allreduce(sends[i], recvs[i], stream=s1 if i % 2 == 0 else s2)
        torch.cuda.synchronize()
        for i in range(4):
            assert same_bits(host(recvs[i]), refs[i]), i
        assert c.cache_info()["tmp_bytes"] > 0


def test_plan_cache_limit_and_trim(gpu):
    dims, P = [2, 2, 2], 8
    with VComm(dims) as c:
        c.set_cache_limit(2)
        for n in (8, 16, 24, 32, 40):
            x = data((P, n), "int32", n)
            r = torch.empty((P, n), dtype=torch.int32, device="cuda")
            c.allreduce(dev(x), r)
            assert same_bits(host(r), O.allreduce(x, dims))
            assert c.cache_info()["plans"] <= 2
        c.trim()
        assert c.cache_info() == {"plans": 0, "tmp_bytes": 0}
        x = data((P, 8), "int32", 1)
        r = torch.empty((P, 8), dtype=torch.int32, device="cuda")
        c.allreduce(dev(x), r)  # recompiled after the trim
        assert same_bits(host(r), O.allreduce(x, dims))


def test_graph_survives_autotune_of_its_step_plan(gpu):
    """A batch graph names the per-step work list of its in-place item; autotuning that
    schedule afterwards must leave the graph's arrays alive (replays stay exact)."""
    dims, P, count = [8], 8, 4099
    x = data((P, count), "float64", 9)
    ref = O.allreduce(x, dims, algo="L")
    with VComm(dims, allreduce_algo="L") as c:
        a = dev(x)
        b = Batch(c, [{"coll": "allreduce", "sendbuf": a, "recvbuf": a, "count": count}])
        b.launch_graph()
        assert same_bits(host(a), ref)
        t = dev(x)
        c.autotune("allreduce", t, t, count, exec="STEP", iters=1)  # replaces the in-place work list
        for k in range(3):
            a.copy_(dev(x))
            torch.cuda.synchronize()
            b.launch_graph()
            assert same_bits(host(a), ref), k
        t.copy_(dev(x))
        c.allreduce(t, t)  # the retuned work list
        assert same_bits(host(t), ref)
        b.close()

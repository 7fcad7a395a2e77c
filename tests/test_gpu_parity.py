"""GPU parity: libbine_hip.so on an MI355X against the CPU oracle.

Every collective is run through the C ABI (via the Python mirror) and compared
bit for bit with the oracle on the same seeded inputs: exact for integers and,
because each output element replays the reference's reduction tree, exact for
float32/float64 as well (tolerance 0 ulp).  At full sizes the oracle is run
on one column per block: an element's association depends only on its block,
so the sampled problem (block size 1) fixes the same trees.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import VComm  # noqa: E402
from bine_trees_fugaku_amd._lib import BineError  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

NP = {"int8": np.int8, "int32": np.int32, "int64": np.int64, "float32": np.float32, "float64": np.float64}


def data(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype.startswith("int"):
        return rng.integers(-1000, 1001, size=shape).astype(NP[dtype])
    return rng.standard_normal(shape).astype(NP[dtype])


def per_rank(P, n, dtype, base_seed):
    """Rank r's buffer seeded with base_seed + r."""
    return np.stack([data((n,), dtype, base_seed + r) for r in range(P)])


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def same_bits(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


@pytest.mark.parametrize("exec", ["STEP", "FUSED"])
def test_allreduce_2x2x2_int32_full(gpu, exec):
    """Plain Bine allreduce, P=8 on 2x2x2, int32 SUM, 262144 elements per rank, uniform [-1000, 1000]."""
    x = per_rank(8, 262144, "int32", 1000)
    ref = O.allreduce(x, [2, 2, 2])
    with VComm([2, 2, 2], exec=exec) as c:
        s, r = dev(x), torch.zeros_like(dev(x))
        c.allreduce(s, r)
        assert same_bits(host(r), ref)
        assert same_bits(host(r)[0], x.sum(0, dtype=np.int64).astype(np.int32))


@pytest.mark.parametrize("dims", [[2, 2, 2], [4, 4], [4, 4, 4], [16], [6, 6], [2, 6, 2]])
@pytest.mark.parametrize("algo", ["B", "L", "B_CONT"])
@pytest.mark.parametrize("dtype", ["float32", "float64", "int32"])
@pytest.mark.parametrize("exec", ["STEP", "AUTO"])
def test_allreduce_small(gpu, dims, algo, dtype, exec):
    P = int(np.prod(dims))
    if algo == "B_CONT" and P & (P - 1):
        pytest.skip("B_CONT is power-of-two only")
    count = 37 * P + 3  # uneven blocks and unaligned vectors
    x = per_rank(P, count, dtype, 7)
    ref = O.allreduce(x, dims, algo=algo)
    with VComm(dims, exec=exec, allreduce_algo=algo) as c:
        r = torch.zeros((P, count), dtype=getattr(torch, dtype), device="cuda")
        c.allreduce(dev(x), r)
        assert same_bits(host(r), ref)


@pytest.mark.parametrize("count_per_rank", [37, 40])  # int8 strides 37P+3 (not 4-aligned: scalar) and 40P+4
@pytest.mark.parametrize("exec", ["STEP", "FUSED"])
def test_allreduce_int8_alignment(gpu, count_per_rank, exec):
    """int8 partitions whose stride is not a multiple of 4 bytes take the scalar path; a 4-byte stride keeps the
    16-byte vector path.  Both wrap like the reference's char sum."""
    dims = [4, 4]
    P = 16
    count = count_per_rank * P + (3 if count_per_rank % 2 else 4)
    x = per_rank(P, count, "int8", 13)
    ref = O.allreduce(x, dims)
    with VComm(dims, exec=exec) as c:
        r = torch.zeros((P, count), dtype=torch.int8, device="cuda")
        c.allreduce(dev(x), r)
        assert same_bits(host(r), ref)


@pytest.mark.parametrize("ports", [2, 3, 6])  # 6 ports on 3 dimensions: three mirrored (lib/libbine_coll.cc:15-31)
def test_allreduce_multiport(gpu, ports):
    x = per_rank(64, 64 * ports * 5 + 1, "float32", 3)
    ref = O.allreduce(x, [4, 4, 4], ports=ports)
    for exec in ("STEP", "FUSED"):
        with VComm([4, 4, 4], num_ports=ports, exec=exec) as c:
            r = torch.zeros_like(dev(x))
            c.allreduce(dev(x), r)
            assert same_bits(host(r), ref)


def test_allreduce_in_place_and_void(gpu):
    x = per_rank(16, 160, "int32", 5)
    with VComm([4, 4]) as c:
        buf = dev(x)
        c.allreduce(buf, buf)  # sendbuf == recvbuf runs the step schedule in place
        assert same_bits(host(buf), O.allreduce(x, [4, 4]))
        r = torch.zeros_like(dev(x))
        c.allreduce(dev(x), r, op="VOID")
        assert same_bits(host(r), O.allreduce(x, [4, 4], op="VOID"))


def sampled_oracle_check(out, x, dims, P, count):
    """Compare one column per block with the oracle run at block size 1."""
    q = count // P
    cols = np.arange(P) * q
    ref = O.allreduce(np.ascontiguousarray(x[:, cols]), dims)
    assert same_bits(out[:, cols], ref)


def test_allreduce_4x4x4_fp32_full(gpu):
    """P=64 on 4x4x4, fp32 N(0,1), 4M elements (16 MiB) per rank: STEP == FUSED bitwise, sampled oracle check."""
    P, count = 64, 4 * 1024 * 1024
    x = per_rank(P, count, "float32", 2000)
    s = dev(x)
    outs = {}
    for exec in ("STEP", "FUSED"):
        with VComm([4, 4, 4], exec=exec) as c:
            r = torch.empty_like(s)
            c.allreduce(s, r)
            torch.cuda.synchronize()
            outs[exec] = r
    assert torch.equal(outs["STEP"].view(torch.int32), outs["FUSED"].view(torch.int32))
    out = host(outs["FUSED"])
    assert (out == out[0]).all()
    sampled_oracle_check(out, x, [4, 4, 4], P, count)
    assert np.allclose(out[0], x.sum(0, dtype=np.float64), atol=1e-3)


def test_allreduce_8x8x8_fp64_full(gpu):
    """P=512 on 8x8x8, fp64 N(0,1), 524288 elements (4 MiB) per rank."""
    P, count = 512, 524288
    g = torch.Generator(device="cuda")
    g.manual_seed(42)
    s = torch.randn((P, count), dtype=torch.float64, device="cuda", generator=g)
    outs = {}
    for exec in ("STEP", "FUSED"):
        with VComm([8, 8, 8], exec=exec) as c:
            r = torch.empty_like(s)
            c.allreduce(s, r)
            torch.cuda.synchronize()
            outs[exec] = r
    assert torch.equal(outs["STEP"].view(torch.int64), outs["FUSED"].view(torch.int64))
    assert torch.equal(outs["FUSED"][0], outs["FUSED"][P - 1])
    q = count // P
    cols = torch.arange(P, device="cuda") * q
    x_s = s[:, cols].cpu().numpy()
    out_s = outs["FUSED"][:, cols].cpu().numpy()
    assert same_bits(out_s, O.allreduce(np.ascontiguousarray(x_s), [8, 8, 8]))
    del outs


@pytest.mark.parametrize("dims", [[8], [4, 4], [4, 4, 4]])
@pytest.mark.parametrize("dtype", ["int32", "float32", "float64"])
def test_reduce_scatter(gpu, dims, dtype):
    P = int(np.prod(dims))
    c = 33
    x = per_rank(P, c * P, dtype, 11)
    ref = O.reduce_scatter(x, dims)
    for exec in ("STEP", "FUSED"):
        with VComm(dims, exec=exec) as cm:
            r = torch.zeros((P, c), dtype=getattr(torch, dtype), device="cuda")
            cm.reduce_scatter(dev(x), r)
            assert same_bits(host(r), ref)


@pytest.mark.parametrize("dims", [[8], [16], [4, 4], [4, 4, 4]])
@pytest.mark.parametrize("algo", ["VEC_DOUBLING_CONT_PERMUTE", "VEC_DOUBLING_CONT_SEND", "VEC_DOUBLING_BLOCKS"])
def test_allgather(gpu, dims, algo):
    if algo == "VEC_DOUBLING_BLOCKS" and len(dims) > 1:
        pytest.skip("1D only")
    P = int(np.prod(dims))
    a = per_rank(P, 21, "float32", 13)
    for exec in ("STEP", "FUSED"):
        with VComm(dims, exec=exec, allgather_algo=algo) as c:
            r = torch.zeros((P, 21 * P), dtype=torch.float32, device="cuda")
            c.allgather(dev(a), r)
            assert same_bits(host(r), np.tile(a.reshape(1, -1), (P, 1)))


@pytest.mark.parametrize("P", [2, 8, 64, 512])
@pytest.mark.parametrize("family", ["BINE", "BRUCK"])
def test_alltoall(gpu, P, family):
    c = 512 if P == 512 else 5  # 512 int64 = 4 KiB per rank pair at P=512
    g = torch.Generator(device="cuda")
    g.manual_seed(P)
    s = torch.randint(-2**62, 2**62, (P, P * c), dtype=torch.int64, device="cuda", generator=g)
    exp = s.view(P, P, c).transpose(0, 1).contiguous().view(P, P * c)
    for exec in ("STEP", "FUSED"):
        with VComm([P], exec=exec, alltoall_family=family) as cm:
            r = torch.zeros_like(s)
            cm.alltoall(s, r)
            torch.cuda.synchronize()
            assert torch.equal(r, exp)
    if P <= 64:
        ref = O.alltoall(s.cpu().numpy(), P, family=family)
        assert same_bits(host(r), ref)


@pytest.mark.parametrize("P", [2, 8, 64, 512])
def test_bcast_reduce_scatter_gather(gpu, P):
    rng_root = [0] if P == 2 else [0, 2, P - 1]
    for root in rng_root:
        b = np.zeros((P, 1000), np.float32)
        b[root] = data((1000,), "float32", root)
        ref = O.bcast(b, root)
        for exec in ("STEP", "FUSED"):
            with VComm([P], exec=exec) as c:
                t = dev(b)
                c.bcast(t, root)
                assert same_bits(host(t), ref)
        s = per_rank(P, 777, "float64", root)
        ref = O.reduce(s, root)
        for exec in ("STEP", "FUSED"):
            with VComm([P], exec=exec) as c:
                out = torch.zeros(777, dtype=torch.float64, device="cuda")
                c.reduce(dev(s), out, root=root)
                assert same_bits(host(out), ref)
    sr = data((P * 100,), "int32", 9)
    g = per_rank(P, 100, "int32", 19)
    for exec in ("STEP", "FUSED"):
        with VComm([P], exec=exec) as c:
            r = torch.zeros((P, 100), dtype=torch.int32, device="cuda")
            c.scatter(dev(sr), r, root=0)
            assert same_bits(host(r), O.scatter(sr, P, 0))
            out = torch.zeros(P * 100, dtype=torch.int32, device="cuda")
            c.gather(dev(g), out, root=0)
            assert same_bits(host(out), O.gather(g, 0))


def test_rooted_variants(gpu):
    P = 64
    b = np.zeros((P, 1001), np.int32)
    b[0] = data((1001,), "int32", 1)
    with VComm([P], bcast_algo="SCATTER_ALLGATHER", bcast_distance="INCREASING") as c:
        t = dev(b)
        c.bcast(t, 0)
        assert same_bits(host(t), O.bcast(b, 0, "SCATTER_ALLGATHER"))
    s = per_rank(P, 1001, "float32", 2)
    with VComm([P], reduce_algo="REDUCE_SCATTER_GATHER") as c:
        out = torch.zeros(1001, dtype=torch.float32, device="cuda")
        c.reduce(dev(s), out)
        assert same_bits(host(out), O.reduce(s, 0, "REDUCE_SCATTER_GATHER"))


def test_errors_do_not_launch(gpu):
    with VComm([4, 4]) as c:
        s = torch.zeros((16, 64), dtype=torch.float32, device="cuda")
        small = torch.zeros((8, 64), dtype=torch.float32, device="cuda")
        with pytest.raises(ValueError):
            c.allreduce(s, small)  # the mirror refuses a buffer that does not split into P ranks
        from bine_trees_fugaku_amd import _lib
        L = _lib.load()
        # a host pointer is not HIP device memory: the C ABI refuses before any launch
        host_buf = np.zeros((16, 64), np.float32)
        rc = L.bine_allreduce(s.data_ptr(), host_buf.ctypes.data, 64, 2, 0, c.handle, None)
        assert rc == _lib.BINE_ERR_ARG
        with pytest.raises(BineError):
            c.gather(s[:, :4].contiguous(), torch.zeros(64, device="cuda"), root=1)  # odd root unsupported
    torch.cuda.synchronize()


def test_family_default_stand_in(gpu):
    """Family DEFAULT (the reference's PMPI_* fallback, lib/libbine.cc:1016) runs recursive doubling for the
    allreduce on a power-of-two torus and Bruck for the alltoall: compared with the oracle of those schedules."""
    x = per_rank(64, 64 * 5 + 3, "float64", 77)
    with VComm([4, 4, 4], allreduce_family="DEFAULT") as c:
        r = torch.zeros_like(dev(x))
        c.allreduce(dev(x), r)
        assert same_bits(host(r), O.allreduce(x, [4, 4, 4], family="RECDOUB"))
    a = per_rank(16, 16 * 7, "int64", 78)
    with VComm([16], alltoall_family="DEFAULT") as c:
        r = torch.zeros_like(dev(a))
        c.alltoall(dev(a), r)
        assert same_bits(host(r), O.alltoall(a, 16, family="BRUCK"))

"""bine_allreduce_noalloc, called through the C ABI, against bine_allreduce.

The allocation-free entry point runs the cached schedule with its per-step
temporaries in the library's per-device scratch array.  Each case makes two
calls in a row, the second with a smaller input, and each output must equal
the regular entry point's output bit for bit (and the oracle's)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import VComm  # noqa: E402
from bine_trees_fugaku_amd._lib import BINE_ERR_ARG, BINE_ERR_UNSUPPORTED, BINE_OK  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

DT = {np.float32: (2, torch.float32), np.float64: (3, torch.float64)}
OP_SUM = 0


def data(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) * 10.0 ** rng.uniform(-3, 3, shape)).astype(dtype)


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def regular(c, x, dtype, in_place):
    code, _ = DT[dtype]
    P, n = x.shape
    s = torch.from_numpy(x).cuda()
    r = s if in_place else torch.empty_like(s)
    rc = c._lib.bine_allreduce(s.data_ptr(), r.data_ptr(), n, code, OP_SUM, c.handle, stream_ptr())
    assert rc == BINE_OK, c._lib.bine_last_error()
    torch.cuda.synchronize()
    return r.cpu().numpy()


def noalloc(c, x, dtype, in_place):
    code, _ = DT[dtype]
    P, n = x.shape
    s = torch.from_numpy(x).cuda()
    r = s if in_place else torch.full_like(s, float("nan"))
    torch.cuda.synchronize()
    rc = c._lib.bine_allreduce_noalloc(s.data_ptr(), r.data_ptr(), n, code, OP_SUM, c.handle, stream_ptr())
    assert rc == BINE_OK, c._lib.bine_last_error()
    torch.cuda.synchronize()
    return r.cpu().numpy()


@pytest.mark.parametrize("dims,algo,exec,in_place,dtype,counts", [
    ([8], "L", "AUTO", True, np.float64, (4099, 1000)),         # in place, 3 steps: the value passes through TMP
    ([4, 4, 4], "L", "STEP", False, np.float32, (65536 + 3, 4099)),
    ([2, 2, 2], "B", "AUTO", True, np.float64, (8 * 1000 + 5, 8 * 10)),
    ([6], "B", "AUTO", False, np.float32, (6000, 600)),          # non power of two: per step
    ([4, 4], "B", "AUTO", False, np.float64, (16 * 300, 16 * 3 + 1)),  # fused: no temporaries
])
def test_noalloc_equals_allreduce_twice(gpu, dims, algo, exec, in_place, dtype, counts):
    P = int(np.prod(dims))
    with VComm(dims, exec=exec, allreduce_algo=algo) as c:
        for k, n in enumerate(counts):
            print(f"noalloc case dims={dims} algo={algo} exec={exec} in_place={in_place} n={n} call={k}", flush=True)
            x = data((P, n), dtype, 100 * k + P)
            want = regular(c, x, dtype, in_place)  # compiles and caches the schedule: the set-up
            got = noalloc(c, x, dtype, in_place)
            assert got.tobytes() == want.tobytes(), (k, n)
            assert got.tobytes() == O.allreduce(x, dims, algo=algo).tobytes(), (k, n)


def test_noalloc_refuses_what_it_cannot_do_without_allocating(gpu):
    with VComm([4, 4, 4], exec="STEP", allreduce_algo="L") as c:
        s = torch.zeros((64, 1024), dtype=torch.float32, device="cuda")
        r = torch.empty_like(s)
        # nothing cached for this count: no compile, no allocation, an error
        assert c._lib.bine_allreduce_noalloc(s.data_ptr(), r.data_ptr(), 1024, 2, OP_SUM, c.handle,
                                             stream_ptr()) == BINE_ERR_UNSUPPORTED
        # 64 ranks x 200000 doubles of per-step temporaries exceed the 64 MiB workspace
        big = torch.zeros((64, 200000), dtype=torch.float64, device="cuda")
        out = torch.empty_like(big)
        c.allreduce(big, out)
        assert c._lib.bine_allreduce_noalloc(big.data_ptr(), out.data_ptr(), 200000, 3, OP_SUM, c.handle,
                                             stream_ptr()) == BINE_ERR_ARG
        assert b"64 MiB" in c._lib.bine_last_error() or b"67108864" in c._lib.bine_last_error()

"""Several host threads calling collectives on one virtual communicator at once, each on a stream of its own.

The communicator's plan cache is guarded by a mutex; per-step runs of one schedule share its temporaries and are
ordered on the GPU by the event the previous run recorded; an evicted schedule that a call still holds stays
alive until that call drops it.  Integer sums are exact, so every result is compared with numpy.  The ctypes calls
release the interpreter lock, so the threads' calls into the library overlap."""
import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import VComm  # noqa: E402

pytestmark = pytest.mark.gpu

P_DIMS = [2, 2, 2]
P = 8


def run_threads(comm, counts_of, iters, nthreads=4):
    errors = []

    def work(tid):
        try:
            st = torch.cuda.Stream()
            rng = np.random.default_rng(100 + tid)
            for it in range(iters):
                n = counts_of(tid, it)
                x = rng.integers(-1000, 1001, size=(P, n)).astype(np.int32)
                with torch.cuda.stream(st):
                    xs = torch.from_numpy(x).cuda()
                    ys = torch.full_like(xs, 7)
                    comm.allreduce(xs, ys, stream=st)
                    got = ys.cpu()  # ordered after the collective on this stream
                st.synchronize()
                want = np.broadcast_to(x.sum(axis=0, dtype=np.int64).astype(np.int32), (P, n))
                if not np.array_equal(got.numpy(), want):
                    errors.append((tid, it, n, "wrong sums"))
                    return
        except Exception as e:  # reported by the main thread
            errors.append((tid, -1, -1, repr(e)))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(nthreads)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors


@pytest.mark.parametrize("exec_mode", ["STEP", "FUSED"])
def test_threads_share_plans(gpu, exec_mode):
    # threads 0 and 1 use the same count: one cached schedule, its per-step temporaries shared across streams
    counts = [4096, 4096, 12288, 777]
    with VComm(P_DIMS, exec=exec_mode) as c:
        run_threads(c, lambda tid, it: counts[tid], iters=12)


def test_threads_with_constant_eviction(gpu):
    # a cache of two schedules and a new count on almost every call: schedules are evicted while other threads'
    # calls still hold them
    with VComm(P_DIMS, exec="STEP") as c:
        c.set_cache_limit(2)
        run_threads(c, lambda tid, it: 1024 + 8 * ((tid * 7 + it) % 9), iters=16)
        assert c.cache_info()["plans"] <= 2

"""A seeded random sweep of the collectives on the GPU against the oracle, bit for bit.

Each case draws a collective, its algorithm or variant, a torus (powers of two and not), a port count, an element
type, an uneven count, a root and an execution mode (per step, fused, automatic) from a fixed seed, runs it through
the C ABI and compares every output byte with the oracle.  Configurations the reference itself refuses (the
library's BINE_ERR_UNSUPPORTED cases: non-powers of two for the rooted collectives and the alltoall, odd gather
roots, more than one port for the reduce-scatter, ...) are drawn too and must be refused, not run.  The same cases
run on the CPU through the host-compiled work lists (tests/test_random_sweep_cpu.py)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bine_trees_fugaku_amd import VComm  # noqa: E402
from bine_trees_fugaku_amd._lib import BineError  # noqa: E402
from tests.random_cases import N_CASES, draw, oracle_output  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def same_bits(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def gpu_output(case):
    k = case
    tdt = getattr(torch, k["dtype"])
    with VComm(k["dims"], num_ports=k["ports"], exec=k["exec"], **k["overrides"]) as c:
        coll, P = k["coll"], k["P"]
        if coll == "allreduce":
            r = torch.zeros_like(dev(k["send"]))
            c.allreduce(dev(k["send"]), r)
        elif coll == "reduce_scatter":
            r = torch.zeros((P, k["count"]), dtype=tdt, device="cuda")
            c.reduce_scatter(dev(k["send"]), r)
        elif coll == "allgather":
            r = torch.zeros((P, P * k["count"]), dtype=tdt, device="cuda")
            c.allgather(dev(k["send"]), r)
        elif coll == "alltoall":
            r = torch.zeros_like(dev(k["send"]))
            c.alltoall(dev(k["send"]), r)
        elif coll == "bcast":
            r = dev(k["send"])
            c.bcast(r, k["root"])
        elif coll == "reduce":
            r = torch.zeros(k["count"], dtype=tdt, device="cuda")
            c.reduce(dev(k["send"]), r, root=k["root"])
        elif coll == "scatter":
            r = torch.zeros((P, k["count"]), dtype=tdt, device="cuda")
            c.scatter(dev(k["send"]), r, root=k["root"])
        else:
            r = torch.zeros(P * k["count"], dtype=tdt, device="cuda")
            c.gather(dev(k["send"]), r, root=k["root"])
        return host(r)


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_case(gpu, seed):
    case = draw(seed)
    print(case["name"], flush=True)
    try:
        got = gpu_output(case)
    except BineError as e:
        assert e.code == 2, f"{case['name']}: {e}"  # BINE_ERR_UNSUPPORTED: a configuration the reference refuses
        pytest.skip(f"refused as in the reference: {case['name']}")
    want = oracle_output(case)
    assert got.shape == want.shape, case["name"]
    assert same_bits(got, want), case["name"]

"""The seeded random collective cases of the GPU sweep, run on the CPU: each schedule is compiled on the host
(plan_dump), executed by the numpy model of the gather-reduce kernel and compared byte for byte with the oracle.
A configuration the library refuses must be refused with BINE_ERR_UNSUPPORTED (the reference refuses it too, or
the schedule cannot be fused when FUSED execution is asked for)."""
import numpy as np
import pytest

from bine_trees_fugaku_amd import plan_dump
from bine_trees_fugaku_amd._lib import BineError
from tests import emulator as E
from tests.random_cases import NP, N_CASES, draw, oracle_output


def cpu_output(k, two_phase=0):
    meta, items, pool, launches = plan_dump(k["dims"], k["coll"], k["count"], k["dtype"], root=k["root"],
                                            exec=k["exec"], num_ports=k["ports"], tile_elems=1 << 12,
                                            two_phase=two_phase, **k["overrides"])
    dt = NP[k["dtype"]]
    if k["coll"] == "bcast":
        bufs = E.alloc_buffers(meta, dt, recv=k["send"])
    else:
        bufs = E.alloc_buffers(meta, dt, send=k["send"])
    E.run(meta, items, pool, launches, bufs)
    out = bufs[E.BUF_RECV]
    return out.reshape(-1) if k["coll"] in ("reduce", "gather") else out


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_case_cpu(seed):
    """Every third case is lowered in two phases (copies block by block or tile by tile); lists without a
    reduction of several outputs lower the same either way."""
    case = draw(seed)
    try:
        got = cpu_output(case, two_phase=seed % 3)
    except BineError as e:
        assert e.code == 2, f"{case['name']}: {e}"
        pytest.skip(f"refused: {case['name']}")
    want = oracle_output(case)
    assert got.shape == want.shape, case["name"]
    assert np.ascontiguousarray(got).tobytes() == np.ascontiguousarray(want).tobytes(), case["name"]

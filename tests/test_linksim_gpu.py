"""Link-traffic simulation on the GPU, bit-exact against the oracle.

All cases of the five tori (2x2x2, 4x4x4, 8x8x8, 64x64, 32x256) and every
algorithm defined for each collective run in one GPU pass.  Every class's
per-link message counts must equal the oracle's routing of the host-side
destinations (themselves pinned against the oracle's traces and closed forms
by the CPU tests), and on the small tori the whole collective's link bytes
must equal the routed trace of the oracle's restated reference collective.
"""
import numpy as np
import pytest

from bine_trees_fugaku_amd import linksim as ls
from oracle import oracle as O
from tests.linksim_ref import TRACED, per_step_bytes, simpy_to_lib, trace_case

pytestmark = pytest.mark.gpu

TORI = [[2, 2, 2], [4, 4, 4], [8, 8, 8], [64, 64], [32, 256]]


def all_cases():
    cases = []
    for dims in TORI:
        for coll, algos in ls.DEFINED.items():
            for algo in algos:
                if algo == "simpy" and len(set(dims)) != 1:
                    continue
                cases.append((dims, coll, algo))
    return cases


@pytest.fixture(scope="module")
def sim(gpu):
    cases = all_cases()
    s = ls.LinkSim(cases)
    print(f"linksim: {len(cases)} cases", flush=True)
    s.run()
    yield s, cases
    s.close()


def expected_counts(dims, coll, algo, cls):
    P = int(np.prod(dims))
    dest = ls.peers(dims, coll, algo, cls).astype(np.int64)
    src = np.arange(P)
    keep = dest != ls.NONE
    src, dest = src[keep], dest[keep]
    if algo == "simpy":
        src, dest = simpy_to_lib(dims, src), simpy_to_lib(dims, dest)
    lb, longest = O.route_bytes(dims, src, dest, 1)
    return lb.astype(np.uint32), longest


def test_link_counts_bit_exact(sim):
    s, cases = sim
    checked = 0
    for i, (dims, coll, algo) in enumerate(cases):
        st = s.stats(i)
        for c in range(len(st.num)):
            want, longest = expected_counts(dims, coll, algo, c)
            got = s.link_counts(i, c)
            assert np.array_equal(got, want), (dims, coll, algo, c)
            assert st.max_mult[c] == want.max() and st.sum_mult[c] == want.astype(np.uint64).sum()
            assert st.max_hops[c] == longest
            checked += 1
    print(f"linksim: {checked} classes bit-exact", flush=True)


@pytest.mark.parametrize("dims", [[2, 2, 2], [4, 4, 4], [4, 8], [16]], ids=lambda d: "x".join(map(str, d)))
def test_collective_link_bytes_equal_oracle_trace(gpu, dims):
    """Whole-collective bytes per link: GPU counts x shares == the oracle
    collective's own messages routed over the torus."""
    cases = [(dims, c, a) for c, a in TRACED]
    s = ls.LinkSim(cases)
    s.run()
    P = int(np.prod(dims))
    for i, (d, coll, algo) in enumerate(cases):
        S, msgs = trace_case(d, coll, algo)
        steps = per_step_bytes(msgs, P)
        src = np.concatenate([v[0] for v in steps.values()])
        dst = np.concatenate([v[1] for v in steps.values()])
        nb = np.concatenate([v[2] for v in steps.values()])
        want, _ = O.route_bytes(d, src, dst, nb)
        assert np.array_equal(s.link_bytes(i, S), want), (coll, algo)
    s.close()


@pytest.mark.parametrize("dims", [[4, 4, 4], [8, 8, 8], [16, 16], [64, 64]], ids=lambda d: "x".join(map(str, d)))
def test_simpy_stats_bit_exact(gpu, dims):
    """simulator.py run_collectives stats (sum_of_distances, bw_new, bw_ideal)
    from the GPU counts equal the oracle's restatement exactly."""
    P = int(np.prod(dims))
    s = ls.LinkSim([(dims, "allreduce", "simpy")])
    s.run()
    got = ls.simpy_stats(s.stats(0), P)
    data = np.random.default_rng(0).standard_normal((P, P))
    _, want = O.sim_run(dims, data, "ALLREDUCE")
    assert got == (want[0], want[2], want[4])
    s.close()


def test_rerun_is_idempotent(gpu):
    """Counters are cleared by every run: two runs give the same counts."""
    cases = [([8, 8, 8], "alltoall", "bine"), ([32, 256], "allreduce", "ring")]
    s = ls.LinkSim(cases)
    s.run()
    a = [s.link_counts(i, 0).copy() for i in range(2)]
    s.run()
    s.run()
    b = [s.link_counts(i, 0) for i in range(2)]
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    s.close()

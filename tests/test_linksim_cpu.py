"""Link-traffic simulation, host side: the step-class model against the oracle.

No GPU: these tests pin everything the GPU kernel is handed (the class list,
every rank's destination, the message shares) and the oracle's routing, so
the GPU test only has to show that the kernel counts what the host describes.
"""
from collections import deque

import numpy as np
import pytest

from bine_trees_fugaku_amd import linksim as ls
from oracle import oracle as O
from tests.linksim_ref import TRACED, closed_form_dest, per_step_bytes, simpy_to_lib, trace_case

SMALL_TORI = [[2, 2, 2], [4, 4, 4], [4, 8], [2, 4, 8], [16], [8, 8]]


def model_messages(dims, coll, algo, S):
    cl = ls.classes(dims, coll, algo)
    out = {}
    for c in range(len(cl["num"])):
        dest = ls.peers(dims, coll, algo, c).astype(np.int64)
        src = np.arange(len(dest))
        keep = (dest != ls.NONE) & (dest != src)
        b = S * int(cl["num"][c]) // int(cl["den"][c]) * int(cl["rep"][c])
        out[c] = (src[keep], dest[keep], b)
    return out


@pytest.mark.parametrize("dims", SMALL_TORI, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("coll,algo", TRACED)
def test_class_model_equals_oracle_trace(dims, coll, algo):
    """Every class's (src, dst, bytes) set equals the messages the oracle's
    restated reference collective sends in the same step."""
    P = int(np.prod(dims))
    S, msgs = trace_case(dims, coll, algo)
    got = per_step_bytes(msgs, P)
    want = model_messages(dims, coll, algo, S)
    assert sorted(got) == sorted(k for k, v in want.items() if len(v[0]))
    for step, (src, dst, nb) in got.items():
        ws, wd, wb = want[step]
        order = np.lexsort((wd, ws))
        assert np.array_equal(src, ws[order]) and np.array_equal(dst, wd[order]), (step,)
        assert np.all(nb == wb), (step, np.unique(nb), wb)


def test_allreduce_8x8x8_matches_trace():
    """The headline torus at full P = 512 (one int32 per block)."""
    dims = [8, 8, 8]
    S, msgs = trace_case(dims, "allreduce", "bine", c=1)
    got = per_step_bytes(msgs, 512)
    want = model_messages(dims, "allreduce", "bine", S)
    for step, (src, dst, nb) in got.items():
        ws, wd, wb = want[step]
        order = np.lexsort((wd, ws))
        assert np.array_equal(dst, wd[order]) and np.all(nb == wb)


CLOSED = [("allreduce", "ring"), ("reduce_scatter", "ring"), ("allgather", "ring"), ("alltoall", "recdoub"),
          ("bcast", "recdoub"), ("scatter", "recdoub"), ("reduce", "recdoub"), ("gather", "recdoub")]


@pytest.mark.parametrize("dims", [[2, 2, 2], [4, 4, 4], [8, 8, 8], [64, 64], [32, 256]],
                         ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("coll,algo", CLOSED)
def test_closed_form_destinations(dims, coll, algo):
    L = int(np.log2(np.prod(dims)))
    cl = ls.classes(dims, coll, algo)
    for c in range(len(cl["num"])):
        want = closed_form_dest(dims, coll, algo, c, L)
        got = ls.peers(dims, coll, algo, c).astype(np.int64)
        got[got == ls.NONE] = -1
        assert np.array_equal(got, want), c


@pytest.mark.parametrize("dims", [[4, 4, 4], [8, 8, 8], [16, 16], [64, 64], [32]], ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("coll", ["allreduce", "reduce_scatter"])
def test_simpy_destinations(dims, coll):
    """simulator.py's get_peer (:70-76), restated in the oracle, in its own id order."""
    cl = ls.classes(dims, coll, "simpy")
    P = int(np.prod(dims))
    assert len(cl["num"]) == int(np.log2(P))
    for c in range(len(cl["num"])):
        want = np.array([O.sim_peer(r, c, dims) for r in range(P)])
        assert np.array_equal(ls.peers(dims, coll, "simpy", c).astype(np.int64), want)


def test_undefined_cases_rejected():
    from bine_trees_fugaku_amd._lib import BineError
    for dims, coll, algo in (([4, 4, 4], "bcast", "ring"), ([32, 256], "allreduce", "simpy"),
                             ([6, 4], "allreduce", "bine"), ([4, 4], "allgather", "simpy")):
        with pytest.raises(BineError):
            ls.classes(dims, coll, algo)


def test_simpy_beyond_its_distance_table_rejected():
    """simulator.py takes step s's distance from a 10-entry table (simulator.py:195): 10 steps per dimension
    (a ring of 1024, or 32x32) work, 11 (a ring of 2048) raise IndexError there and are refused here."""
    from bine_trees_fugaku_amd._lib import BineError
    assert len(ls.classes([1024], "allreduce", "simpy")["kind"]) == 10
    assert len(ls.classes([32, 32], "reduce_scatter", "simpy")["kind"]) == 10
    for dims in ([2048], [2048, 2048]):
        with pytest.raises(BineError):
            ls.classes(dims, "allreduce", "simpy")


def test_class_shares():
    """Closed forms: reduce-scatter halves (sum (P-1)/P), allgather doubles, ring repeats P-1 per phase."""
    from fractions import Fraction as F
    for dims in ([8, 8, 8], [64, 64], [32, 256]):
        P = int(np.prod(dims))
        cl = ls.classes(dims, "allreduce", "bine")
        fr = [F(int(n), int(d)) for n, d in zip(cl["num"], cl["den"])]
        L = len(fr) // 2
        assert sum(fr[:L]) == F(P - 1, P) and sum(fr[L:]) == F(P - 1, P)
        r = ls.classes(dims, "allreduce", "ring")
        assert r["rep"].tolist() == [2 * (P - 1)] and F(int(r["num"][0]), int(r["den"][0])) == F(1, P)


# ---- the oracle's routing, pinned by brute force -----------------------------

def bfs_dist(dims, src):
    P = int(np.prod(dims))
    dist = [-1] * P
    dist[src] = 0
    q = deque([src])
    while q:
        u = q.popleft()
        c = list(np.unravel_index(u, dims))
        for d in range(len(dims)):
            for s in (1, -1):
                cc = list(c)
                cc[d] = (cc[d] + s) % dims[d]
                v = int(np.ravel_multi_index(cc, dims))
                if dist[v] < 0:
                    dist[v] = dist[u] + 1
                    q.append(v)
    return dist


@pytest.mark.parametrize("dims", [[4, 4], [2, 3, 4], [5], [6, 2], [3, 3, 3]], ids=lambda d: "x".join(map(str, d)))
def test_route_lengths_are_shortest_paths(dims):
    P = int(np.prod(dims))
    for s in range(P):
        dist = bfs_dist(dims, s)
        for t in range(P):
            lb, longest = O.route_bytes(dims, [s], [t], [1])
            assert int(lb.sum()) == dist[t] == longest


def test_route_links_form_a_path():
    """Each route is a walk of neighbouring nodes: link (node, d, dir) leaves node
    along dimension d; dimension 0 first; ties go the positive way."""
    dims = [4, 6, 2]
    P = int(np.prod(dims))
    rng = np.random.default_rng(3)
    for s, t in rng.integers(0, P, (200, 2)):
        lb, _ = O.route_bytes(dims, [s], [t], [1])
        links = np.flatnonzero(lb)
        cur = np.array(np.unravel_index(int(s), dims))
        seen = 0
        for d in range(3):
            fwd = (np.unravel_index(int(t), dims)[d] - cur[d]) % dims[d]
            neg = 2 * fwd > dims[d]
            for _ in range(dims[d] - fwd if neg else fwd):
                node = int(np.ravel_multi_index(cur, dims))
                assert lb[(node * 3 + d) * 2 + int(neg)] == 1
                cur[d] = (cur[d] + (-1 if neg else 1)) % dims[d]
                seen += 1
        assert seen == len(links)
        assert int(np.ravel_multi_index(cur, dims)) == t


def test_route_byte_hops_identity():
    dims = [8, 8, 8]
    rng = np.random.default_rng(5)
    src = rng.integers(0, 512, 1000)
    dst = rng.integers(0, 512, 1000)
    nb = rng.integers(1, 1 << 20, 1000).astype(np.uint64)
    lb, _ = O.route_bytes(dims, src, dst, nb)
    hops = np.array([O.route_bytes(dims, [a], [b], [1])[0].sum() for a, b in zip(src, dst)], np.uint64)
    assert int(lb.sum()) == int((nb * hops).sum())


@pytest.mark.parametrize("dims", [[4, 4, 4], [8, 8, 8], [16, 16]], ids=lambda d: "x".join(map(str, d)))
def test_simpy_stats_from_route_lengths(dims):
    """simulator.py's stats follow from route lengths: the minimal route of a
    step's message is |distance| hops, so sum_mult / P = |distance| and
    bw_new = sum 1/2^(s+1) * |d_s| reproduces run_collectives (:160-175)."""
    P = int(np.prod(dims))
    data = np.random.default_rng(0).standard_normal((P, P))
    _, want = O.sim_run(dims, data, "ALLREDUCE")
    cl = ls.classes(dims, "allreduce", "simpy")
    sums, hops = [], []
    for c in range(len(cl["num"])):
        dest = simpy_to_lib(dims, ls.peers(dims, "allreduce", "simpy", c))
        src = simpy_to_lib(dims, np.arange(P))
        lb, longest = O.route_bytes(dims, src, dest, 1)
        sums.append(int(lb.sum()))
        hops.append(longest)
    st = ls.ClassStats(cl["num"], cl["den"], cl["rep"], np.zeros_like(cl["num"]), np.array(sums, np.uint64),
                       np.array(hops, np.uint32))
    sum_d, bw_new, bw_ideal = ls.simpy_stats(st, P)
    assert (sum_d, bw_new, bw_ideal) == (want[0], want[2], want[4])

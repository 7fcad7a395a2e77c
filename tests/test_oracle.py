"""Pins the CPU oracle (the checker) before anything is compared against it.

The reference ships no golden vectors (its checks compare against the vendor
MPI at run time, bench/bench.cc:270,328-344), so the oracle is pinned by:
  * the reference's own simulation/remap_rank.c, compiled unchanged under
    oracle/_ref (remap_rank);
  * textbook base -2 arithmetic (btonb/nbtob, simulation/negbinary.py:10-26);
  * closed forms: rho_k = (1-(-2)^(k+1))/3, round-robin step schedule,
    block partition, MPI result semantics (sums and permutations);
  * invariants that fix the schedule uniquely on power-of-two tori
    (|S| = P/2^(s+1), S and R disjoint, S_r = R_peer, each step a matching);
  * an independent replay of the reduction tree built from the peer function
    alone, which fixes the floating-point association of allreduce B (the
    other reducing collectives: test_oracle_association.py).
"""
import math
import os
import shutil

import numpy as np
import pytest

from oracle import oracle as O

VALIDATE_TORI = [[4], [8], [64], [6], [10], [12], [14], [18], [2, 8, 2], [6, 6], [10, 10], [4, 6, 4],
                 [6, 2, 6], [2, 6, 2], [2, 2, 6], [6, 2, 2], [4, 4, 4]]  # validate/allreduce.sh:111-139
P2_TORI = [[2], [4], [8], [16], [64], [2, 2, 2], [4, 4], [4, 4, 4], [2, 8, 2], [8, 8, 8]]


def textbook_negabinary(v: int) -> int:
    """Digits of v in base -2 read as a binary number (simulation/binomial.py:66-81)."""
    bits, pos = 0, 0
    while v != 0:
        v, rem = divmod(v, -2)
        if rem < 0:
            v, rem = v + 1, rem + 2
        bits |= rem << pos
        pos += 1
    return bits


@pytest.mark.parametrize("v", list(range(-700, 700)))
def test_btonb_textbook(v):
    assert O.btonb(v) == textbook_negabinary(v)
    assert O.nbtob(O.btonb(v)) == v


def test_rho_closed_form():
    peers = O.compute_peers([1 << 12], 0)
    for k, p in enumerate(peers):
        rho = (1 - (-2) ** (k + 1)) // 3
        assert p == rho % (1 << 12)


@pytest.mark.skipif(not os.path.exists(O.REF_REMAP) and not shutil.which("gcc"), reason="no reference build")
def test_remap_rank_against_reference_binary():
    if not os.path.exists(O.REF_REMAP):
        pytest.skip("oracle/_ref/remap_rank not built (reference tree absent)")
    cases = [(P, r) for P in (2, 4, 8, 16, 32) for r in range(P)]
    cases += [(P, r) for P in (6, 10, 12, 24) for r in range(P)]  # non power of two: the 80000000 quirk
    rng = np.random.default_rng(0)
    cases += [(P, int(r)) for P in (512, 4096, 8192) for r in rng.integers(0, P, 12)]
    for P, r in cases:
        assert O.remap_rank(r, P) == O.ref_remap_rank(P, r), (P, r)


def test_remap_rank_is_permutation_p2():
    for P in (2, 4, 8, 16, 64, 512):
        assert sorted(O.remap_rank(r, P) for r in range(P)) == list(range(P))


def test_ceil_log2():
    for x in range(1, 5000):
        assert O.ceil_log2(x) == math.ceil(math.log2(x))


def test_step_info_round_robin():
    assert O.step_info([2, 2, 2]) == [(0, 0), (1, 0), (2, 0)]
    assert O.step_info([8, 8, 8]) == [(d, s) for s in range(3) for d in range(3)]
    si = O.step_info([32, 256])
    assert si[:10] == [(d, s) for s in range(5) for d in range(2)]
    assert si[10:] == [(1, 5), (1, 6), (1, 7)]
    assert O.step_info([4, 4, 4], port=1)[0] == (1, 0)


def test_peers_8x8x8_rank0():
    # rank 0 is even in every coordinate, so the peers follow +rho on (x, y, z) in turn
    assert O.compute_peers([8, 8, 8], 0) == [64, 8, 1, 448, 56, 7, 192, 24, 3]


@pytest.mark.parametrize("dims", [[2, 2, 2], [4, 4, 4], [8, 8, 8], [64, 64], [32, 256], [16], [6, 6], [10]])
def test_each_step_is_a_matching(dims):
    P = int(np.prod(dims))
    ranks = range(P) if P <= 512 else np.random.default_rng(1).integers(0, P, 256)
    table = {}
    for r in ranks:
        table[int(r)] = O.compute_peers(dims, int(r))
    for r, pr in table.items():
        for s, p in enumerate(pr):
            assert O.compute_peers(dims, p)[s] == r
            # the peer differs from r in exactly one coordinate
            cr = np.unravel_index(r, dims)
            cp = np.unravel_index(p, dims)
            assert sum(a != b for a, b in zip(cr, cp)) == 1


@pytest.mark.parametrize("dims", [[2, 2, 2], [4, 4, 4], [16], [4, 8], [2, 8, 2]])
def test_bitmap_invariants_p2(dims):
    P = int(np.prod(dims))
    L = sum(int(math.log2(d)) for d in dims)
    for r in range(P):
        peers = O.compute_peers(dims, r)
        for s in range(L):
            S, R = O.step_bitmaps(dims, r, s)
            assert S.sum() == P >> (s + 1) and R.sum() == P >> (s + 1)
            assert not (S & R).any()
            S_peer, _ = O.step_bitmaps(dims, peers[s], s)
            assert (S_peer == R).all()
        bs = O.block_step(dims, r)
        assert bs[r] == L  # own block never sent


def replay_allreduce_tree(x, dims, family="BINE"):
    """Independent replay: for block b, holders before step s are T_s with
    T_L = {b} and T_s = T_{s+1} + {peer_s(h)}; at step s each h in T_{s+1}
    adds the partial of peer_s(h).  Built from the peer function only."""
    P, n = x.shape
    L = sum(int(math.log2(d)) for d in dims)
    peers = [O.compute_peers(dims, r, family=family) for r in range(P)]
    q, rem = divmod(n, P)
    out = np.empty(n, dtype=x.dtype)
    off = 0
    for b in range(P):
        cnt = q + (1 if b < rem else 0)
        T = [None] * (L + 1)
        T[L] = [b]
        for s in range(L - 1, -1, -1):
            T[s] = T[s + 1] + [peers[h][s] for h in T[s + 1]]
        val = {r: x[r, off:off + cnt].copy() for r in range(P)}
        for s in range(L):
            for h in T[s + 1]:
                val[h] = val[peers[h][s]] + val[h]
        out[off:off + cnt] = val[b]
        off += cnt
    return out


@pytest.mark.parametrize("dims", [[2, 2, 2], [4, 4, 4], [16], [2, 8, 2], [4, 8]])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_allreduce_b_fp_association_pinned(dims, dtype):
    P = int(np.prod(dims))
    rng = np.random.default_rng(P)
    x = (rng.standard_normal((P, 3 * P + 5)) * 10 ** rng.uniform(-3, 3, (P, 1))).astype(dtype)
    y = O.allreduce(x, dims, algo="B")
    assert (y == y[0]).all()  # every rank holds identical bits
    ref = replay_allreduce_tree(x, dims)
    assert np.array_equal(y[0].view(np.uint8), ref.view(np.uint8))
    # a sequential sum is a different association: make sure the data can tell them apart
    assert not np.array_equal(y[0], np.cumsum(x, axis=0, dtype=dtype)[-1])


def test_allreduce_recdoub_fp_association_pinned():
    dims = [4, 4]
    x = np.random.default_rng(3).standard_normal((16, 40))
    y = O.allreduce(x, dims, algo="B", family="RECDOUB")
    assert np.array_equal(y[0], replay_allreduce_tree(x, dims, family="RECDOUB"))


@pytest.mark.parametrize("dims", VALIDATE_TORI)
@pytest.mark.parametrize("algo", ["B", "B_COALESCE", "L", "B_CONT"])
def test_allreduce_int32_sum(dims, algo):
    P = int(np.prod(dims))
    if algo == "B_CONT" and P & (P - 1):
        pytest.skip("B_CONT is power-of-two only (lib/libbine_common.cc:1247-1251)")
    rng = np.random.default_rng(P)
    x = rng.integers(-1000, 1001, size=(P, 2 * P + 7), dtype=np.int32)
    y = O.allreduce(x, dims, algo=algo)
    assert (y == x.sum(0, dtype=np.int64).astype(np.int32)).all()


@pytest.mark.parametrize("ports", [2, 3, 6])
def test_allreduce_multiport(ports):
    dims = [4, 4, 4]
    x = np.random.default_rng(ports).integers(-1000, 1001, size=(64, 64 * ports + 11), dtype=np.int32)
    for algo in ("B", "L"):
        y = O.allreduce(x, dims, algo=algo, ports=ports)
        assert (y == x.sum(0).astype(np.int32)).all()


def test_allreduce_void_is_diagonal_allgather():
    dims = [4, 4]
    x = np.random.default_rng(5).integers(-9, 9, size=(16, 16 * 3), dtype=np.int32)
    y = O.allreduce(x, dims, op="VOID")
    diag = np.concatenate([x[b, 3 * b:3 * b + 3] for b in range(16)])
    assert (y == diag).all()


def test_allreduce_l_identical_on_all_ranks_fp():
    x = np.random.default_rng(9).standard_normal((64, 100))
    y = O.allreduce(x, [4, 4, 4], algo="L")
    assert (y == y[0]).all()
    assert np.allclose(y[0], x.sum(0))


@pytest.mark.parametrize("dims", [[8], [4, 4], [2, 2, 2], [4, 4, 4], [16]])
def test_reduce_scatter_semantics(dims):
    P = int(np.prod(dims))
    x = np.random.default_rng(P).integers(-1000, 1001, size=(P, 5 * P), dtype=np.int32)
    assert (O.reduce_scatter(x, dims) == x.sum(0).astype(np.int32).reshape(P, 5)).all()


@pytest.mark.parametrize("dims", [[8], [4, 4], [2, 2, 2], [4, 4, 4], [16]])
@pytest.mark.parametrize("variant", ["CONT_PERMUTE", "CONT_SEND", "BLOCKS"])
def test_allgather_semantics(dims, variant):
    if variant == "BLOCKS" and len(dims) > 1:
        pytest.skip("bine_allgather_blocks_mpi is 1D")
    P = int(np.prod(dims))
    a = np.random.default_rng(P).integers(-1000, 1001, size=(P, 5), dtype=np.int32)
    assert (O.allgather(a, dims, variant=variant) == a.reshape(1, -1)).all()


def test_allgather_multiport():
    a = np.random.default_rng(1).integers(-1000, 1001, size=(64, 12), dtype=np.int32)
    assert (O.allgather(a, [4, 4, 4], ports=3) == a.reshape(1, -1)).all()


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64, 512])
@pytest.mark.parametrize("family", ["BINE", "BRUCK"])
def test_alltoall_semantics(P, family):
    x = np.random.default_rng(P).integers(-1000, 1001, size=(P, 3 * P), dtype=np.int32)
    exp = x.reshape(P, P, 3).transpose(1, 0, 2).reshape(P, 3 * P)
    assert (O.alltoall(x, P, family=family) == exp).all()


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64])
def test_rooted_semantics(P):
    rng = np.random.default_rng(P)
    for root in sorted({0, 1, P // 2, P - 1}):
        b = np.zeros((P, 7), np.int32)
        b[root] = rng.integers(-5, 5, 7)
        assert (O.bcast(b, root) == b[root]).all()
        s = rng.integers(-1000, 1001, size=(P, 7), dtype=np.int32)
        assert (O.reduce(s, root) == s.sum(0)).all()
    b = np.zeros((P, 13), np.int32)
    b[0] = rng.integers(-5, 5, 13)
    assert (O.bcast(b, 0, "SCATTER_ALLGATHER") == b[0]).all()
    s = rng.integers(-1000, 1001, size=(P, 13), dtype=np.int32)
    assert (O.reduce(s, 0, "REDUCE_SCATTER_GATHER") == s.sum(0)).all()
    sr = rng.integers(-1000, 1001, size=(P * 4,), dtype=np.int32)
    assert (O.scatter(sr, P, 0) == sr.reshape(P, 4)).all()
    s = rng.integers(-1000, 1001, size=(P, 4), dtype=np.int32)
    for root in range(0, P, 2):
        assert (O.gather(s, root) == s.reshape(-1)).all()


def test_gather_odd_root_reference_quirk():
    """bine_gather_mpi grows its window by rank parity (lib/libbine_coll_gather.cc:221-224)
    while the tree is rooted at vrank = rank - root; for odd roots the blocks land
    in the wrong place.  The oracle restates that faithfully."""
    s = np.arange(8, dtype=np.int32).reshape(4, 2)
    assert not (O.gather(s, 1) == s.reshape(-1)).all()


def test_gather_non_p2_deadlocks_in_reference():
    """P=6: rank 0 waits at mask 2 for rank 5, which already left the loop
    (lib/libbine_coll_gather.cc:219-262): the schedule cannot complete."""
    with pytest.raises(O.OracleError):
        O.gather(np.zeros((6, 2), np.int32), 0)


@pytest.mark.parametrize("dims", [[8], [16], [4, 4], [8, 8], [4, 4, 4]])
def test_tree_remap_is_permutation(dims):
    P = int(np.prod(dims))
    for root in (0, 1, P - 1):
        tr = O.get_tree(dims, root)
        assert sorted(tr["remapped_ranks"]) == list(range(P))
        assert tr["remapped_ranks"][root] == 0
        assert tr["parent"][root] == 0xFFFFFFFF
        # every subtree occupies a contiguous range [remapped, remapped_max]
        for r in range(P):
            lo, hi = tr["remapped_ranks"][r], tr["remapped_ranks_max"][r]
            members = [q for q in range(P) if lo <= tr["remapped_ranks"][q] <= hi]
            for q in members:
                a = q
                while a != r and a != root:
                    a = tr["parent"][a]
                assert a == r or r == root


# ---- simulator.py restatement ---------------------------------------------------

def sim_peer_closed_form(sender, step, dims):
    dist = O.JACOBSTHAL
    D = len(dims)
    dim, rel = step % D, step // D
    if D == 1:
        coord = [sender]
    elif D == 2:
        coord = [sender // dims[1], sender % dims[1]]
    else:
        coord = [(sender // dims[1]) % dims[0], sender % dims[1], sender // (dims[0] * dims[1])]
    sign = (-1) ** rel * (1 if coord[dim] % 2 == 0 else -1)
    coord[dim] = (coord[dim] + sign * dist[rel]) % dims[dim]
    if D == 1:
        return coord[0]
    if D == 2:
        return coord[0] * dims[1] + coord[1]
    return coord[2] * dims[0] * dims[1] + coord[0] * dims[1] + coord[1]


@pytest.mark.parametrize("dims", [[8], [16], [8, 8], [4, 4, 4], [8, 8, 8]])
def test_sim_peers_closed_form(dims):
    P = int(np.prod(dims))
    for step in range(int(math.log2(P))):
        for r in range(P):
            assert O.sim_peer(r, step, dims) == sim_peer_closed_form(r, step, dims)


@pytest.mark.parametrize("dims", [[8], [16], [8, 8], [4, 4, 4]])
def test_sim_run_results_and_stats(dims):
    P = int(np.prod(dims))
    L = int(math.log2(P))
    D = len(dims)
    data = np.random.default_rng(P).integers(0, 101, size=(P, P)).astype(np.float64)
    out, stats = O.sim_run(dims, data, "ALLREDUCE")
    assert (out == data.sum(0)).all()  # compute_expected_data, simulator.py:35-38
    out, stats = O.sim_run(dims, data, "REDUCESCATTER")
    assert all(out[i, i] == data[:, i].sum() for i in range(P))  # simulator.py:39-44
    dist = [abs(O.JACOBSTHAL[s // D]) for s in range(L)]
    assert stats[0] == sum(dist)
    assert stats[1] == sum(2 ** (s / D) for s in range(L))
    assert stats[2] == sum(d / 2 ** (s + 1) for s, d in enumerate(dist))
    assert stats[4] == 1 - 2.0 ** -L
    # bw_ref (simulator.py:168,172) = sum_s 2^-(s+1) 2^(s/D): a geometric series with ratio q = 2^(1/D - 1),
    # (1 - q^L) / (2 (1 - q)); for D = 1 every term is 1/2
    q = 2.0 ** (1.0 / D - 1.0)
    closed = L / 2 if D == 1 else (1 - q ** L) / (2 * (1 - q))
    assert abs(stats[3] - closed) <= 1e-14 * closed
    if D == 1:
        assert stats[3] == L / 2



def test_remap_rank_against_golden_fixture():
    """Fixture written by tests/golden/make_remap_rank.py from the reference's
    own simulation/remap_rank.c (runs where the reference tree is absent)."""
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "remap_rank.json")
    with open(path) as f:
        rows = json.load(f)["P_rank_remap"]
    assert len(rows) > 150
    for P, r, want in rows:
        assert O.remap_rank(r, P) == want, (P, r)

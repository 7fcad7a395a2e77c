"""Pins the floating-point association of every reducing collective of the CPU
oracle on power-of-two sizes.

IEEE addition is commutative, so the bits a reduction produces depend only on
its association: which partial sums are added together, in which tree.  Each
replay below builds that tree from the peer function alone (checked against
rho_k = (1 - (-2)^(k+1)) / 3 in test_oracle.py) and, where the algorithm
places blocks by a remapped rank, from remap_rank (checked against the
reference's own simulation/remap_rank.c) or from a post-order numbering
written here.  None of them runs the oracle's message engine, its bitmaps,
its chunk ranges or its negabinary masks.

The input data spans six decades per rank, so different trees give different
bits; every test also shows that a plausible wrong tree (a reversed step
order, a left-to-right sum, the neighbour's owner) gives different bits.

Replayed schedules (reference file:line):
  * allreduce B, multi-port: lib/libbine.cc:986-996 (ports split first),
    lib/libbine_common.cc:1631-1735;
  * allreduce B_CONT: lib/libbine_common.cc:1162-1237, 1524-1550, dfs :1981-1993;
  * allreduce L (butterfly), multi-port: lib/libbine_common.cc:733-846;
  * reduce_scatter (contiguous): lib/libbine_coll_reduce_scatter.cc:660-758;
  * reduce, binomial tree: lib/libbine_coll_reduce.cc:575-615;
  * reduce, reduce-scatter then gather: lib/libbine_coll_reduce.cc:1042-1152.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

P2_TORI = [[2], [8], [16], [64], [2, 2, 2], [4, 4], [4, 8], [2, 8, 2], [4, 4, 4]]


def spread_data(P, n, dtype, seed):
    """N(0,1) values scaled per rank over six decades: associations differ in their bits."""
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((P, n)) * 10.0 ** rng.uniform(-3, 3, (P, 1))).astype(dtype)


def nsteps(dims):
    return sum(int(math.log2(d)) for d in dims)


def holder_tree(owner, peer, L):
    """Holder sets of one block: T[L] = {owner}; T[s] = T[s+1] + peers of T[s+1] at step s.

    At step s every h in T[s+1] adds the partial of peer(h, s) to its own.  The
    sets must partition the ranks: each step doubles the holders."""
    T = [None] * (L + 1)
    T[L] = [owner]
    for s in range(L - 1, -1, -1):
        add = [peer(h, s) for h in T[s + 1]]
        assert not set(add) & set(T[s + 1]), "peer of a holder is a holder"
        T[s] = T[s + 1] + add
    return T


def replay(vals, owner, peer, L):
    """Value of one block at its owner: vals[r] is rank r's input for the block."""
    T = holder_tree(owner, peer, L)
    assert sorted(T[0]) == list(range(len(vals)))
    acc = {r: vals[r].copy() for r in range(len(vals))}
    for s in range(L):
        for h in T[s + 1]:
            acc[h] = acc[peer(h, s)] + acc[h]
    return acc[owner]


def port_blocks(count, P, ports):
    """Block partition of the allreduce: ports first, then P blocks per port, the
    first remainders one element longer (lib/libbine.cc:971-996).  [(port, block, off, cnt)]"""
    out, off = [], 0
    for q in range(ports):
        cq = count // ports + (1 if q < count % ports else 0)
        for i in range(P):
            cb = cq // P + (1 if i < cq % P else 0)
            out.append((q, i, off, cb))
            off += cb
    return out


def peer_table(dims, ports, family="BINE"):
    P = int(np.prod(dims))
    return [[O.compute_peers(dims, r, port=q, family=family) for r in range(P)] for q in range(ports)]


# ---- allreduce B, several ports ------------------------------------------------

def replay_allreduce_b(x, dims, ports, family="BINE"):
    P, n = x.shape
    L = nsteps(dims)
    pt = peer_table(dims, ports, family)
    out = np.empty(n, x.dtype)
    for q, b, off, cnt in port_blocks(n, P, ports):
        vals = [x[r, off:off + cnt] for r in range(P)]
        out[off:off + cnt] = replay(vals, b, lambda h, s: pt[q][h][s], L)
    return out


@pytest.mark.parametrize("dims,ports", [([4, 4, 4], 2), ([4, 4, 4], 3), ([4, 4, 4], 6), ([2, 8, 2], 3),
                                        ([4, 8], 2), ([4, 8], 4), ([16], 2), ([8, 8], 4)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_allreduce_b_multiport_association(dims, ports, dtype):
    P = int(np.prod(dims))
    x = spread_data(P, 2 * ports * P + 7, dtype, P + ports)
    y = O.allreduce(x, dims, algo="B", ports=ports)
    assert (y == y[0]).all()
    want = replay_allreduce_b(x, dims, ports)
    assert np.array_equal(y[0].view(np.uint8), want.view(np.uint8))
    # every port on port 0's peers is a different tree
    assert not np.array_equal(y[0], replay_allreduce_b(x, dims, 1))


@pytest.mark.parametrize("dims", [[4, 4], [2, 2, 2], [16]])
def test_allreduce_b_coalesce_same_association_as_b(dims):
    P = int(np.prod(dims))
    x = spread_data(P, 3 * P + 1, np.float64, 11)
    assert np.array_equal(O.allreduce(x, dims, algo="B_COALESCE")[0], replay_allreduce_b(x, dims, 1))


# ---- allreduce B_CONT ----------------------------------------------------------

def post_order_numbering(dims, port, family="BINE"):
    """Rank -> its index in the post-order walk of the binomial tree rooted at rank 0,
    a node reached at step i having children at steps i+1..L-1 in step order (the walk
    of lib/libbine_common.cc:1981-1993, written iteratively here)."""
    P = int(np.prod(dims))
    L = nsteps(dims)
    peers = [O.compute_peers(dims, r, port=port, family=family) for r in range(P)]
    number = {}
    stack = [(0, 0, False)]  # (rank, first child step, children done)
    while stack:
        r, first, done = stack.pop()
        if done:
            number.setdefault(r, len(number))
            continue
        stack.append((r, first, True))
        for i in range(L - 1, first - 1, -1):  # pushed last-first so step `first` is walked first
            stack.append((peers[r][i], i + 1, False))
    assert sorted(number) == list(range(P))
    return [number[r] for r in range(P)]


def replay_allreduce_b_cont(x, dims, ports, family="BINE", rev=False):
    P, n = x.shape
    L = nsteps(dims)
    pt = peer_table(dims, ports, family)
    if rev:  # the steps in reverse order: a wrong reading, for contrast
        pt = [[row[::-1] for row in tab] for tab in pt]
    owners = []
    for q in range(ports):
        m = post_order_numbering(dims, q, family)
        owners.append({m[r]: r for r in range(P)})  # block m ends at the rank numbered m
    out = np.empty(n, x.dtype)
    for q, b, off, cnt in port_blocks(n, P, ports):
        vals = [x[r, off:off + cnt] for r in range(P)]
        out[off:off + cnt] = replay(vals, owners[q][b], lambda h, s: pt[q][h][s], L)
    return out


@pytest.mark.parametrize("dims", P2_TORI)
def test_cont_numbering_halves_split_at_every_step(dims):
    """The kept half of a rank and of its peer are complementary at every step: the
    numbering is what makes the contiguous halving exchange a reduce-scatter."""
    P = int(np.prod(dims))
    L = nsteps(dims)
    m = post_order_numbering(dims, 0)
    assert [O.cont_remapped_rank(dims, r) for r in range(P)] == m
    peers = [O.compute_peers(dims, r) for r in range(P)]
    for r in range(P):
        lo, hi = 0, P
        for s in range(L):
            mid = (lo + hi) // 2
            mine, theirs = m[r] >= mid, m[peers[r][s]] >= mid
            assert mine != theirs
            assert lo <= m[peers[r][s]] < hi
            lo, hi = (mid, hi) if mine else (lo, mid)
        assert (lo, hi) == (m[r], m[r] + 1)


@pytest.mark.parametrize("dims", P2_TORI)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_allreduce_b_cont_association(dims, dtype):
    P = int(np.prod(dims))
    x = spread_data(P, 3 * P + 5, dtype, P + 1)
    y = O.allreduce(x, dims, algo="B_CONT")
    assert (y == y[0]).all()
    want = replay_allreduce_b_cont(x, dims, 1)
    assert np.array_equal(y[0].view(np.uint8), want.view(np.uint8))
    if nsteps(dims) >= 2:  # the steps taken in reverse order build another tree
        assert not np.array_equal(y[0], replay_allreduce_b_cont(x, dims, 1, rev=True))


@pytest.mark.parametrize("dims,ports", [([4, 4, 4], 3), ([4, 4, 4], 6), ([4, 8], 2)])
def test_allreduce_b_cont_multiport_association(dims, ports):
    P = int(np.prod(dims))
    x = spread_data(P, 2 * ports * P + 3, np.float64, 5)
    y = O.allreduce(x, dims, algo="B_CONT", ports=ports)
    assert np.array_equal(y[0], replay_allreduce_b_cont(x, dims, ports))


# ---- allreduce L ---------------------------------------------------------------

def replay_allreduce_l(x, dims, ports, family="BINE", rev=False):
    """Butterfly: at step s every rank adds its step-s peer's running value (per port,
    on that port's element range and peers)."""
    P, n = x.shape
    L = nsteps(dims)
    pt = peer_table(dims, ports, family)
    if rev:  # the steps in reverse order: a wrong reading, for contrast
        pt = [[row[::-1] for row in tab] for tab in pt]
    out = np.empty((P, n), x.dtype)
    off = 0
    for q in range(ports):
        cq = n // ports + (1 if q < n % ports else 0)
        v = x[:, off:off + cq].copy()
        for s in range(L):
            v = np.stack([v[pt[q][r][s]] + v[r] for r in range(P)])
        out[:, off:off + cq] = v
        off += cq
    return out


@pytest.mark.parametrize("dims,ports", [(d, 1) for d in P2_TORI] + [([4, 4, 4], 2), ([4, 4, 4], 3),
                                                                    ([4, 4, 4], 6), ([4, 8], 4)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_allreduce_l_association(dims, ports, dtype):
    P = int(np.prod(dims))
    x = spread_data(P, 5 * ports + 3, dtype, P * 7 + ports)
    y = O.allreduce(x, dims, algo="L", ports=ports)
    want = replay_allreduce_l(x, dims, ports)
    assert np.array_equal(y.view(np.uint8), want.view(np.uint8))
    if nsteps(dims) >= 2:  # the steps taken in reverse order build another tree
        assert not np.array_equal(y, replay_allreduce_l(x, dims, ports, rev=True))
    if ports > 1:  # every port on port 0's peers is another tree
        assert not np.array_equal(y, replay_allreduce_l(x, dims, 1))


def test_allreduce_l_recdoub_association():
    dims = [4, 4]
    x = spread_data(16, 23, np.float64, 4)
    assert np.array_equal(O.allreduce(x, dims, algo="L", family="RECDOUB"),
                          replay_allreduce_l(x, dims, 1, family="RECDOUB"))


# ---- reduce_scatter --------------------------------------------------------------

def replay_reduce_scatter(x, dims, dist, family="BINE"):
    """Block b ends at rank b; at step s the exchange partner is peers[s]
    (INCREASING) or peers[L-1-s] (DECREASING)."""
    P, n = x.shape
    c = n // P
    L = nsteps(dims)
    pt = peer_table(dims, 1, family)[0]
    order = list(range(L)) if dist == "INCREASING" else list(range(L - 1, -1, -1))
    out = np.empty((P, c), x.dtype)
    for b in range(P):
        vals = [x[r, b * c:(b + 1) * c] for r in range(P)]
        out[b] = replay(vals, b, lambda h, s: pt[h][order[s]], L)
    return out


@pytest.mark.parametrize("dims", P2_TORI)
@pytest.mark.parametrize("dist", ["INCREASING", "DECREASING"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reduce_scatter_association(dims, dist, dtype):
    P = int(np.prod(dims))
    x = spread_data(P, 3 * P, dtype, P + 17)
    y = O.reduce_scatter(x, dims, dist=dist)
    want = replay_reduce_scatter(x, dims, dist)
    assert np.array_equal(y.view(np.uint8), want.view(np.uint8))
    if nsteps(dims) >= 2:  # the other step order is another tree
        other = "DECREASING" if dist == "INCREASING" else "INCREASING"
        assert not np.array_equal(y, replay_reduce_scatter(x, dims, other))


def test_reduce_scatter_recdoub_association():
    x = spread_data(16, 48, np.float64, 2)
    assert np.array_equal(O.reduce_scatter(x, [4, 4], family="RECDOUB"),
                          replay_reduce_scatter(x, [4, 4], "INCREASING", family="RECDOUB"))


def test_reduce_scatter_increasing_matches_allreduce_b_blocks():
    """With INCREASING distance the reduce-scatter builds block b at rank b over the
    same peers as allreduce B's first half: the same trees, the same bits."""
    dims = [4, 4, 4]
    P = 64
    x = spread_data(P, 2 * P, np.float32, 3)
    rs = O.reduce_scatter(x, dims)
    ar = O.allreduce(x, dims, algo="B")[0]
    assert np.array_equal(rs.reshape(-1), ar)


# ---- reduce: binomial tree ---------------------------------------------------------

def reduce_tree_sets(P, root):
    """Holder sets of the 1D binomial reduce in rank space: virtual rank v = rank - root
    exchanges at step k with its 1D torus peer (v + rho_k for even v, v - rho_k for odd)."""
    L = nsteps([P])
    vp = [O.compute_peers([P], v) for v in range(P)]
    return holder_tree(0, lambda h, s: vp[h][s], L), vp


def replay_reduce_tree(x, root):
    P = x.shape[0]
    L = nsteps([P])
    vp = [O.compute_peers([P], v) for v in range(P)]
    vals = [x[(v + root) % P] for v in range(P)]
    return replay(vals, 0, lambda h, s: vp[h][s], L)


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64, 256])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reduce_tree_association(P, dtype):
    x = spread_data(P, 29, dtype, P)
    for root in sorted({0, 1, P // 2 + 1, P - 1} & set(range(P))):
        got = O.reduce(x, root)
        want = replay_reduce_tree(x, root)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), root
    if P >= 4:  # a left-to-right sum, and the untranslated tree of root 0, are other trees
        assert not np.array_equal(O.reduce(x, 0), np.cumsum(x, axis=0, dtype=dtype)[-1])
        assert not np.array_equal(O.reduce(x, 1), replay_reduce_tree(x, 0))


@pytest.mark.parametrize("P", [4, 8, 16, 64])
def test_reduce_tree_parent_is_get_tree_parent(P):
    """The rank each rank sends to in the reduce is its parent in get_tree's tree of the
    translated root 0 with DECREASING distance (a child attached at step a has its own
    children at steps below a); for even roots that is get_tree of the root itself."""
    L = nsteps([P])
    T, vp = reduce_tree_sets(P, 0)
    vparent = {}
    for s in range(L):
        for h in T[s + 1]:
            vparent[vp[h][s]] = h
    tree0 = O.get_tree([P], 0, dist="DECREASING")["parent"]
    assert all(tree0[v] == vparent[v] for v in range(1, P))
    for root in range(0, P, 2):
        tr = O.get_tree([P], root, dist="DECREASING")["parent"]
        for r in range(P):
            if r != root:
                assert tr[r] == (vparent[(r - root) % P] + root) % P, (root, r)


# ---- reduce: reduce-scatter then gather ---------------------------------------------

def replay_reduce_rsg(x, rev=False):
    """Data block j (displacements of lib/libbine_coll_reduce.cc:1056-1060) is reduced at
    the rank whose remap_rank is j, over the 1D torus peers, then copied to root 0.
    rev=True takes the steps in reverse order (a wrong reading, for contrast)."""
    P, n = x.shape
    L = nsteps([P])
    pt = [O.compute_peers([P], r)[::-1 if rev else 1] for r in range(P)]
    owner = {O.remap_rank(r, P): r for r in range(P)}
    q, rem = divmod(n, P)
    out = np.empty(n, x.dtype)
    for j in range(P):
        off, cnt = q * j + min(j, rem), q + (1 if j < rem else 0)
        vals = [x[r, off:off + cnt] for r in range(P)]
        out[off:off + cnt] = replay(vals, owner[j], lambda h, s: pt[h][s], L)
    return out


@pytest.mark.parametrize("P", [2, 4, 8, 16, 64, 256])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reduce_rsg_association(P, dtype):
    x = spread_data(P, 3 * P + 2, dtype, P + 3)
    got = O.reduce(x, 0, "REDUCE_SCATTER_GATHER")
    want = replay_reduce_rsg(x)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    if P >= 4:  # the steps taken in reverse order build another tree
        assert not np.array_equal(got, replay_reduce_rsg(x, rev=True))


# ---- one tree per peer function ------------------------------------------------------

@pytest.mark.parametrize("dims", [[2], [4], [4, 4], [2, 2, 2], [4, 4, 4], [2, 4, 2]])
def test_small_dimensions_give_one_tree(dims):
    """Where no dimension exceeds 4, the peers of steps 0..s-1 split the ranks into groups
    of 2^s that every step pairs whole: a partial sum over a group is the same expression
    tree whichever rank holds it.  So B, B_COALESCE, B_CONT, L and the INCREASING
    reduce_scatter produce identical bits there, block owners aside."""
    P = int(np.prod(dims))
    x = spread_data(P, 2 * P, np.float64, 21)
    ref = O.allreduce(x, dims, algo="B")[0]
    for algo in ("B_COALESCE", "B_CONT", "L"):
        assert np.array_equal(O.allreduce(x, dims, algo=algo)[0], ref), algo
    assert np.array_equal(O.reduce_scatter(x, dims).reshape(-1), ref)
    # the groups: ranks reachable within steps < s form a partition refined step by step
    L = nsteps(dims)
    pt = peer_table(dims, 1)[0]
    group = {r: frozenset([r]) for r in range(P)}
    for s in range(L):
        nxt = {r: group[r] | group[pt[r][s]] for r in range(P)}
        for r in range(P):
            assert not group[r] & group[pt[r][s]]
            assert all(nxt[h] == nxt[r] for h in nxt[r])
        group = nxt
    assert group[0] == frozenset(range(P))


@pytest.mark.parametrize("dims", [[8], [16], [4, 8], [2, 8, 2], [8, 8]])
def test_wide_dimensions_give_one_tree_per_owner(dims):
    """Along a dimension of 8 or more the Bine windows slide: after two steps rank 0 of a
    ring of 16 has summed ranks 14, 15, 0, 1 and rank 1 ranks 0..3.  The groups overlap
    without nesting, so the tree depends on the rank that ends with the sum: allreduce
    L leaves rank-dependent bits, and B_CONT (block b at the rank numbered b) differs
    from B (block b at rank b).  Each matches its own replay above."""
    P = int(np.prod(dims))
    x = spread_data(P, 2 * P, np.float64, 21)
    y = O.allreduce(x, dims, algo="L")
    assert not (y == y[0]).all()
    assert not np.array_equal(O.allreduce(x, dims, algo="B")[0], O.allreduce(x, dims, algo="B_CONT")[0])
    pt = peer_table(dims, 1)[0]
    group = {r: frozenset([r]) for r in range(P)}
    nested = True
    for s in range(nsteps(dims)):
        group = {r: group[r] | group[pt[r][s]] for r in range(P)}
        nested = nested and all(group[h] == group[r] for r in range(P) for h in group[r])
    assert not nested and group[0] == frozenset(range(P))

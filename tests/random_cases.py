"""Seeded random collective cases shared by the GPU sweep and its CPU counterpart.

draw(seed) returns one case: the collective, its torus or rank count, algorithm or variant overrides, port count,
element type, per-rank count, root, execution mode, the input array and a printable name.  oracle_output(case) is
the oracle's result for it."""
import numpy as np

from oracle import oracle as O

NP = {"int8": np.int8, "int32": np.int32, "int64": np.int64, "float32": np.float32, "float64": np.float64}
TORI = [[2], [4], [8], [16], [32], [6], [10], [12], [2, 2], [4, 4], [2, 8], [6, 6], [2, 2, 2], [4, 4, 4],
        [2, 8, 2], [2, 6, 2], [4, 2, 4]]
POW2_1D = [2, 4, 8, 16, 32, 64]
N_CASES = 400
COLLS = ["allreduce", "allreduce", "reduce_scatter", "allgather", "alltoall", "bcast", "reduce", "scatter", "gather"]


def data(shape, dtype, rng):
    if dtype == "int8":
        return rng.integers(-128, 128, size=shape).astype(np.int8)
    if dtype.startswith("int"):
        return rng.integers(-1000, 1001, size=shape).astype(NP[dtype])
    return (rng.standard_normal(shape) * 10.0 ** rng.uniform(-2, 2, shape)).astype(NP[dtype])


def pick(rng, seq):
    return seq[int(rng.integers(len(seq)))]


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    coll = COLLS[seed % len(COLLS)]
    k = {"coll": coll, "exec": pick(rng, ["STEP", "FUSED", "AUTO"]), "ports": 1, "root": 0, "overrides": {}}
    if coll in ("allreduce", "reduce_scatter"):
        dims = pick(rng, TORI)
    elif coll == "allgather":
        dims = pick(rng, TORI) if rng.random() < 0.7 else pick(rng, [t for t in TORI if len(t) == 1])
    elif coll == "alltoall":
        dims = [pick(rng, [2, 4, 8, 16, 32, 64, 6, 12])]
    else:
        dims = [pick(rng, POW2_1D) if rng.random() < 0.85 else pick(rng, [6, 12])]
    P = int(np.prod(dims))
    k.update(dims=dims, P=P)
    if coll == "allreduce":
        k["algo"] = pick(rng, ["B", "B_COALESCE", "B_CONT", "L"])
        k["overrides"] = {"allreduce_algo": k["algo"]}
        k["ports"] = int(rng.integers(1, 2 * len(dims) + 1)) if rng.random() < 0.4 else 1
        k["dtype"] = pick(rng, ["int8", "int32", "int64", "float32", "float64"])
        k["count"] = P * k["ports"] * int(rng.integers(1, 9)) + int(rng.integers(0, P * k["ports"]))
        k["send"] = data((P, k["count"]), k["dtype"], rng)
        detail = f"{k['algo']} ports={k['ports']}"
    elif coll == "reduce_scatter":
        k["dist"] = pick(rng, ["INCREASING", "DECREASING"])
        k["overrides"] = {"reduce_scatter_distance": k["dist"]}
        k["dtype"] = pick(rng, ["int32", "float32", "float64"])
        k["count"] = int(rng.integers(1, 40))
        k["send"] = data((P, P * k["count"]), k["dtype"], rng)
        detail = k["dist"]
    elif coll == "allgather":
        k["variant"] = pick(rng, ["CONT_PERMUTE", "CONT_SEND"] + (["BLOCKS"] if len(dims) == 1 else []))
        k["overrides"] = {"allgather_algo": "VEC_DOUBLING_" + k["variant"]}
        k["dtype"] = pick(rng, ["int8", "int32", "float64"])
        k["count"] = int(rng.integers(1, 40))
        k["send"] = data((P, k["count"]), k["dtype"], rng)
        detail = k["variant"]
    elif coll == "alltoall":
        k["family"] = pick(rng, ["BINE", "BRUCK"])
        k["overrides"] = {"alltoall_family": k["family"]}
        k["dtype"] = pick(rng, ["int8", "int32", "int64", "float32"])
        k["count"] = int(rng.integers(1, 20))
        k["send"] = data((P, P * k["count"]), k["dtype"], rng)
        detail = k["family"]
    else:
        # mostly roots the reference serves (root 0 for scatter and the two-phase variants, even gather roots),
        # sometimes any root, which must then be refused where the reference refuses it
        k["root"] = 0 if rng.random() < 0.4 else int(rng.integers(P))
        if coll == "gather" and rng.random() < 0.7:
            k["root"] -= k["root"] % 2
        k["count"] = int(rng.integers(1, 300))
        if coll == "bcast":
            k["variant"] = pick(rng, ["BINOMIAL_TREE", "SCATTER_ALLGATHER"])
            k["overrides"] = {"bcast_algo": k["variant"]}
            k["dtype"] = pick(rng, ["int32", "float64", "int8"])
            b = np.zeros((P, k["count"]), NP[k["dtype"]])
            b[k["root"]] = data((k["count"],), k["dtype"], rng)
            k["send"] = b
            detail = k["variant"]
        elif coll == "reduce":
            k["variant"] = pick(rng, ["BINOMIAL_TREE", "REDUCE_SCATTER_GATHER"])
            k["overrides"] = {"reduce_algo": k["variant"]}
            k["dtype"] = pick(rng, ["int32", "float32", "float64"])
            if k["variant"] == "REDUCE_SCATTER_GATHER":
                k["count"] = P * int(rng.integers(1, 20))
            k["send"] = data((P, k["count"]), k["dtype"], rng)
            detail = k["variant"]
        elif coll == "scatter":
            k["dtype"] = pick(rng, ["int32", "float32", "int64"])
            k["send"] = data((P * k["count"],), k["dtype"], rng)
            detail = ""
        else:
            k["dtype"] = pick(rng, ["int32", "float64", "int8"])
            k["send"] = data((P, k["count"]), k["dtype"], rng)
            detail = ""
        detail = f"root={k['root']} {detail}"
    k["name"] = f"{seed}: {coll} {dims} {detail} {k['dtype']} n={k['count']} {k['exec']}"
    return k


def oracle_output(k):
    coll, dims, P = k["coll"], k["dims"], k["P"]
    if coll == "allreduce":
        return O.allreduce(k["send"], dims, algo=k["algo"], ports=k["ports"])
    if coll == "reduce_scatter":
        return O.reduce_scatter(k["send"], dims, dist=k["dist"])
    if coll == "allgather":
        return O.allgather(k["send"], dims, variant=k["variant"])
    if coll == "alltoall":
        return O.alltoall(k["send"], P, family=k["family"])
    if coll == "bcast":
        return O.bcast(k["send"], k["root"], k["variant"])
    if coll == "reduce":
        return O.reduce(k["send"], k["root"], k["variant"])
    if coll == "scatter":
        return O.scatter(k["send"], P, k["root"])
    return O.gather(k["send"], k["root"])

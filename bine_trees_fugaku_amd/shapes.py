"""Learned launch shapes: predict the fastest gather-reduce launch shape of a
batch from its description, trained on measured kernel timings.

The autotuner (bine_batch_autotune) times every candidate shape (vectors per
thread, workgroup size, persistent grid, non-temporal access) on the real
buffers: about 28 x (iters + 1) launches before the first useful one.  This
model predicts the time of every shape from a few features of the batch, so a
new batch starts at a good shape without any timing run; the autotuner remains
the fallback and the yardstick of the model's regret.

The network is the selector's fp32 MLP (HIP forward/backward/Adam), inputs
below, one output per candidate shape: log2 of the launch time, standardised.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import selector as sel
from .collectives import Batch, VComm, tuning_candidates

COLLS = list(sel.COLLS)
N_FEATURES = 16


def buffer_elems(coll: str, P: int, count: int):
    """(sendbuf, recvbuf) elements of one batch item (layout of the single call)."""
    return {
        "allreduce": (P * count, P * count), "reduce_scatter": (P * P * count, P * count),
        "allgather": (P * count, P * P * count), "bcast": (0, P * count), "alltoall": (P * P * count, P * P * count),
        "scatter": (P * count, P * count), "gather": (P * count, P * count), "reduce": (P * count, count),
    }[coll]


def features(coll: str, P: int, ds: int, count: int, stats: dict, buckets: int = 1) -> np.ndarray:
    ns, nr = buffer_elems(coll, P, count)
    f = np.zeros(N_FEATURES, np.float32)
    f[COLLS.index(coll)] = 1.0
    f[8] = 1.0 if ds == 8 else 0.0
    f[9] = math.log2(P) / 14.0
    f[10] = math.log2(max(1, count * ds)) / 30.0                      # bytes per rank and block/pair unit
    f[11] = math.log2(max(1, (ns + nr) * ds)) / 34.0                  # bytes touched
    f[12] = math.log2(max(1, stats["max_tree_leaves"])) / 10.0
    f[13] = math.log2(max(1, stats["items"])) / 24.0
    f[14] = math.log2(max(1, stats["bytes_read"] + stats["bytes_written"])) / 34.0
    f[15] = math.log2(buckets) / 10.0
    return f


@dataclass
class Timings:
    X: np.ndarray                 # (n, N_FEATURES)
    T: np.ndarray                 # (n, shapes) ms per launch
    shapes: np.ndarray            # (shapes, 4)
    keys: list = field(default_factory=list)


SIZES = (16 << 10, 64 << 10, 256 << 10, 1 << 20, 2 << 20, 4 << 20, 8 << 20)
TORI = ([8], [2, 2, 2], [4, 4, 4], [64], [16, 16], [4, 8, 8], [8, 8, 8])


def default_workloads(max_bytes: int = 1 << 30, tori=TORI, sizes=SIZES):
    """(coll, dims, dtype, count) over the tori, the 8 collectives, fp32 and fp64
    and per-rank buffer sizes `sizes`, keeping every buffer within max_bytes."""
    out = []
    for dims in tori:
        P = int(np.prod(dims))
        for coll in COLLS:
            for dt, ds in (("float32", 4), ("float64", 8)):
                for S in sizes:
                    count = S // ds
                    if coll in ("reduce_scatter", "allgather", "alltoall", "scatter", "gather"):
                        count = max(1, count // P)  # per block / pair: the rank's buffer is S
                    if coll in ("allreduce", "reduce_scatter") and count * (P if coll == "reduce_scatter" else 1) < P:
                        continue
                    ns, nr = buffer_elems(coll, P, count)
                    if max(ns, nr) * ds > max_bytes:
                        continue
                    out.append((coll, dims, dt, count))
    return out


def measure(workloads, iters: int = 3, device: int = 0, seed: int = 0, stream=None, log=None) -> Timings:
    """Time every candidate shape on every workload (fused execution, one batch item);
    log(i, key) is called before each workload."""
    import torch
    shapes = tuning_candidates(device)
    Xs, Ts, keys = [], [], []
    g = torch.Generator(device=f"cuda:{device}")
    g.manual_seed(seed)
    for i, (coll, dims, dt, count) in enumerate(workloads):
        if log:
            log(i, (coll, dims, dt, count))
        P = int(np.prod(dims))
        tdt = getattr(torch, dt)
        ns, nr = buffer_elems(coll, P, count)
        send = torch.randn(max(ns, 1), dtype=tdt, device=f"cuda:{device}", generator=g) if ns else None
        recv = torch.randn(nr, dtype=tdt, device=f"cuda:{device}", generator=g)
        with VComm(dims, exec="FUSED", device=device) as comm:
            b = Batch(comm, [{"coll": coll, "sendbuf": send, "recvbuf": recv, "count": count}])
            st = b.stats()
            ms = b.time_shapes(shapes, iters=iters, stream=stream)
            b.close()
        Xs.append(features(coll, P, torch.tensor([], dtype=tdt).element_size(), count, st))
        Ts.append(ms.astype(np.float64))
        keys.append((coll, tuple(dims), dt, count))
        del send, recv
    return Timings(np.stack(Xs), np.stack(Ts), shapes, keys)


class ShapeModel:
    """Predicts, per candidate shape, log2 of its launch time relative to the
    mean over shapes; picks the fastest."""

    # 256 x 3 trained for 800 epochs: held-out regret 2.5-2.6 % over three seeds, against 4.4-4.8 % at
    # 128 x 3 and 400 epochs, which underfit (3.5 % regret on its own training rows)
    def __init__(self, n_shapes: int, hidden: int = 256, depth: int = 3, seed: int = 0, device: int = 0):
        self.net = sel.MLP([N_FEATURES] + [hidden] * depth + [n_shapes], max_batch=4096, seed=seed, device=device)
        self.device = device
        self.mu, self.sigma = 0.0, 1.0

    def fit(self, data: Timings, epochs: int = 800, lr: float = 2e-3, seed: int = 0):
        torch = self.net.torch
        # targets relative to each row's mean: the choice depends only on differences between shapes
        y = np.log2(data.T)
        y = y - y.mean(axis=1, keepdims=True)
        self.mu, self.sigma = float(y.mean()), float(y.std())
        dev = f"cuda:{self.device}"
        X = torch.as_tensor(data.X).to(dev)
        T = torch.as_tensor(((y - self.mu) / self.sigma).astype(np.float32)).to(dev)
        M = torch.ones_like(T)
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
        n = X.shape[0]
        batch = min(n, 256)
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(dev)
            cur = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * ep / epochs)))
            acc, nb = 0.0, 0
            for i in range(0, n, batch):
                idx = perm[i:i + batch]
                self.net.train_step(X[idx].contiguous(), T[idx].contiguous(), M[idx].contiguous(), cur, loss)
                nb += 1
            acc = float(loss.item())
            losses.append(acc)
        return losses

    def predict(self, X: np.ndarray) -> np.ndarray:
        torch = self.net.torch
        y = self.net.forward(torch.as_tensor(np.ascontiguousarray(X, np.float32)).to(f"cuda:{self.device}"))
        return y.cpu().numpy().astype(np.float64) * self.sigma + self.mu

    def choose(self, X: np.ndarray) -> np.ndarray:
        return np.argmin(self.predict(np.atleast_2d(X)), axis=1)

    def tune(self, batch: Batch, coll: str, P: int, ds: int, count: int, shapes=None) -> tuple:
        """Set a single-collective batch to the predicted fastest shape, without timing runs."""
        shapes = tuning_candidates(self.device) if shapes is None else shapes
        x = features(coll, P, ds, count, batch.stats())
        vpt, threads, pers, nt = (int(v) for v in shapes[int(self.choose(x)[0])])
        batch.set_tuning(vpt, threads, pers, nt)
        return vpt, threads, pers, nt

    def close(self):
        self.net.close()


def regret(data: Timings, choice: np.ndarray) -> np.ndarray:
    rows = np.arange(len(choice))
    return data.T[rows, choice] / data.T.min(axis=1) - 1.0


def best_static(data: Timings) -> int:
    """The one shape with the smallest mean time relative to each row's best."""
    return int(np.argmin((data.T / data.T.min(axis=1, keepdims=True)).mean(axis=0)))


def split(data: Timings, frac: float = 0.3, seed: int = 0):
    """Held-out rows (seeded); returns (train, test)."""
    rng = np.random.default_rng(seed)
    test = rng.random(len(data.X)) < frac
    pick = lambda m: Timings(data.X[m], data.T[m], data.shapes, [k for k, t in zip(data.keys, m) if t])  # noqa: E731
    return pick(~test), pick(test)

"""Learned algorithm selector: which algorithm and how many pipeline segments
to use for a collective of a given size on a given torus.

The reference fixes one algorithm per collective (lib/libbine.cc:28-73, all
Bine, LIBBINE_SEGMENT_SIZE = 0) and leaves selection to the user; Fujitsu MPI
ships a hand-written decision tree (fjmpi-dectree.conf).  Here a small fp32
network (3 hidden layers of 256, ReLU), trained on the GPU by the library's
own HIP kernels, predicts the time of every (algorithm, segments) option from
the link-traffic simulation; exhaustive simulation is the fallback when the
network is unsure, and the yardstick of its regret.

Cost model of one option (times in microseconds):

    seg      = ceil(S / k)                      (k pipeline segments)
    stage_c  = alpha + beta * hops_c + mult_c * floor(seg * num_c / den_c) / bw
    T(S, k)  = sum_c rep_c * stage_c + (k - 1) * max_c stage_c

where for step class c of the schedule hops_c is the longest route,
mult_c the message count on the most loaded link (both from the GPU link
counters) and num_c/den_c the share of the vector each message carries.
alpha = 1 us per message step, beta = 0.1 us per hop and bw = 6800 bytes/us
(6.8 GB/s per link direction, the Tofu-D figure) are model constants.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import linksim as ls

COLLS = list(ls.COLLS)
ALGOS = ("bine", "recdoub", "ring", "bruck")  # algorithms the selector may return
SEGMENTS = (1, 2, 4, 8, 16, 32, 64)
OPTIONS = [(a, k) for a in ALGOS for k in SEGMENTS]
ALPHA_US, BETA_US, BW_BYTES_PER_US = 1.0, 0.1, 6800.0
HELD_OUT = ([2, 2, 2], [4, 4, 4], [8, 8, 8], [64, 64], [32, 256])
N_FEATURES = 18


def option_valid(coll: str) -> np.ndarray:
    ok = np.zeros(len(OPTIONS), bool)
    for i, (a, _) in enumerate(OPTIONS):
        ok[i] = a in ls.DEFINED[coll]
    return ok


def case_times(st: ls.ClassStats, sizes) -> np.ndarray:
    """T(S, k) in microseconds for every size (rows) and segment count (columns)."""
    S = np.asarray(sizes, np.uint64)[:, None]
    k = np.asarray(SEGMENTS, np.uint64)[None, :]
    seg = (S + k - np.uint64(1)) // k  # (n_s, n_k)
    num = st.num.astype(np.uint64)
    den = st.den.astype(np.uint64)
    share = (seg[..., None] * num) // den  # (n_s, n_k, C), exact integers
    stage = ALPHA_US + BETA_US * st.max_hops.astype(np.float64) + \
        st.max_mult.astype(np.float64) * share.astype(np.float64) / BW_BYTES_PER_US
    total = (stage * st.rep.astype(np.float64)).sum(-1)
    return total + (k.astype(np.float64) - 1.0) * stage.max(-1)


def features(dims, coll: str, nbytes) -> np.ndarray:
    """Network inputs for every size of one (torus, collective)."""
    nbytes = np.atleast_1d(np.asarray(nbytes, np.float64))
    P = float(np.prod(dims))
    lg = [math.log2(d) for d in dims] + [0.0] * (3 - len(dims))
    f = np.zeros((len(nbytes), N_FEATURES), np.float32)
    f[:, COLLS.index(coll)] = 1.0
    f[:, 8] = np.log2(nbytes) / 28.0
    f[:, 9] = math.log2(P) / 14.0
    f[:, 10] = np.log2(nbytes / P) / 28.0
    f[:, 11 + len(dims) - 1] = 1.0
    f[:, 14:17] = np.array(lg) / 9.0
    f[:, 17] = max(lg) / 9.0
    return f


@dataclass
class Dataset:
    X: np.ndarray          # (n, N_FEATURES) float32
    T: np.ndarray          # (n, len(OPTIONS)) float64 microseconds, inf where invalid
    keys: list = field(default_factory=list)  # (dims, coll, nbytes) per row

    @property
    def mask(self) -> np.ndarray:
        return np.isfinite(self.T)

    @property
    def samples(self) -> int:
        """(collective, algorithm, segments, torus, bytes) points evaluated."""
        return int(self.mask.sum())


def simulate(tori, sizes, device: int = 0, chunk: int = 12) -> Dataset:
    """Exhaustive simulation: GPU link counts for every (torus, collective,
    algorithm), then the cost model at every size and segment count."""
    sizes = np.asarray(sizes, np.uint64)
    Xs, Ts, keys = [], [], []
    tori = [list(t) for t in tori]
    for c0 in range(0, len(tori), chunk):
        group = tori[c0:c0 + chunk]
        cases = [(d, coll, a) for d in group for coll in COLLS for a in ALGOS if a in ls.DEFINED[coll]]
        sim = ls.LinkSim(cases, device)
        sim.run()
        times = {}
        for i, (d, coll, a) in enumerate(cases):
            times[(tuple(d), coll, a)] = case_times(sim.stats(i), sizes)
        sim.close()
        for d in group:
            for coll in COLLS:
                T = np.full((len(sizes), len(OPTIONS)), np.inf)
                for j, a in enumerate(ALGOS):
                    if a in ls.DEFINED[coll]:
                        T[:, j * len(SEGMENTS):(j + 1) * len(SEGMENTS)] = times[(tuple(d), coll, a)]
                Xs.append(features(d, coll, sizes))
                Ts.append(T)
                keys.extend((tuple(d), coll, int(s)) for s in sizes)
    return Dataset(np.concatenate(Xs), np.concatenate(Ts), keys)


def training_tori(exclude=HELD_OUT, max_p: int = 16384):
    """Power-of-two tori of 1 to 3 dimensions, 8 <= P <= max_p, without the
    held-out shapes or any permutation of them."""
    ex = {tuple(sorted(t)) for t in exclude}
    pw = [2 ** i for i in range(1, 10)]
    out = []
    for n in (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192):
        if n <= max_p:
            out.append([n])
    for a in pw:
        for b in pw:
            if 8 <= a * b <= max_p and tuple(sorted((a, b))) not in ex:
                out.append([a, b])
    for a in pw[:5]:
        for b in pw[:5]:
            for c in pw[:5]:
                if 8 <= a * b * c <= max_p and tuple(sorted((a, b, c))) not in ex:
                    out.append([a, b, c])
    return out


def default_sizes(per_octave: int = 4, lo: int = 3, hi: int = 28) -> np.ndarray:
    """Message sizes 2^lo .. 2^hi bytes, per_octave points per doubling (rounded to bytes)."""
    e = np.arange(lo * per_octave, hi * per_octave + 1) / per_octave
    return np.unique(np.round(2.0 ** e).astype(np.uint64))


# ---- the network ---------------------------------------------------------------

class MLP:
    """fp32 MLP on the GPU (bine_mlp_*): ReLU hidden layers, linear output."""

    def __init__(self, dims, max_batch: int = 8192, seed: int = 0, device: int = 0):
        import torch
        self.torch = torch
        self.dims = [int(d) for d in dims]
        self.device = device
        self.max_batch = int(max_batch)
        arr = (ctypes.c_int32 * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().bine_mlp_create(arr, len(self.dims), self.max_batch, seed, device, ctypes.byref(h)),
                   "bine_mlp_create")
        self._h = h
        n = ctypes.c_uint64()
        _lib.check(_lib.load().bine_mlp_num_params(self._h, ctypes.byref(n)), "bine_mlp_num_params")
        self.num_params = n.value
        self._loss = torch.zeros(1, dtype=torch.float32, device=f"cuda:{device}")

    def _dev(self, a):
        t = self.torch.as_tensor(np.ascontiguousarray(a, np.float32))
        return t.to(f"cuda:{self.device}")

    @staticmethod
    def _s(stream):
        return None if stream is None else ctypes.c_void_p(stream.cuda_stream)

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        _lib.check(_lib.load().bine_mlp_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "bine_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        _lib.check(_lib.load().bine_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "bine_mlp_set_params")

    def forward(self, x, stream=None):
        """x: device tensor (n, inputs) float32 -> device tensor (n, outputs)."""
        n = x.shape[0]
        y = self.torch.empty((n, self.dims[-1]), dtype=self.torch.float32, device=x.device)
        for i in range(0, n, self.max_batch):
            m = min(self.max_batch, n - i)
            _lib.check(_lib.load().bine_mlp_forward(self._h, x[i:].data_ptr(), m, y[i:].data_ptr(), self._s(stream)),
                       "bine_mlp_forward")
        return y

    def gradients(self, x, t, mask, stream=None):
        """(loss, flat gradient) of the masked MSE on one batch (device tensors)."""
        g = self.torch.empty(self.num_params, dtype=self.torch.float32, device=x.device)
        _lib.check(_lib.load().bine_mlp_gradients(self._h, x.data_ptr(), t.data_ptr(), mask.data_ptr(), x.shape[0],
                                                  g.data_ptr(), self._loss.data_ptr(), self._s(stream)),
                   "bine_mlp_gradients")
        return self._loss, g

    def train_step(self, x, t, mask, lr: float, loss_out=None, stream=None):
        _lib.check(_lib.load().bine_mlp_train_step(self._h, x.data_ptr(), t.data_ptr(), mask.data_ptr(), x.shape[0],
                                                   float(lr), None if loss_out is None else loss_out.data_ptr(),
                                                   self._s(stream)), "bine_mlp_train_step")

    def close(self):
        if self._h:
            _lib.check(_lib.load().bine_mlp_destroy(self._h), "bine_mlp_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Selector:
    """Predicts log2 time of every option; picks the fastest valid one."""

    # 3 x 256 trained for 150 epochs: held-out mean regret 0.42 %, p95 3.1-3.4 % over two seeds; 3 x 128 for 60
    # epochs gave 0.6-1.0 % and p95 4.2-5.9 %
    def __init__(self, hidden: int = 256, depth: int = 3, seed: int = 0, device: int = 0, max_batch: int = 8192):
        self.net = MLP([N_FEATURES] + [hidden] * depth + [len(OPTIONS)], max_batch=max_batch, seed=seed,
                       device=device)
        self.device = device
        self.mu, self.sigma = 0.0, 1.0
        self._sim_cache = {}

    def _targets(self, data: Dataset):
        y = np.where(data.mask, np.log2(np.where(data.mask, data.T, 1.0)), 0.0)
        return ((y - self.mu) / self.sigma).astype(np.float32), data.mask.astype(np.float32)

    def fit(self, data: Dataset, epochs: int = 150, batch: int = 4096, lr: float = 2e-3, seed: int = 0):
        """Adam on the GPU; returns the per-epoch mean training loss."""
        torch = self.net.torch
        y = np.log2(data.T[data.mask])
        self.mu, self.sigma = float(y.mean()), float(y.std())
        t, m = self._targets(data)
        dev = f"cuda:{self.device}"
        X = torch.as_tensor(data.X).to(dev)
        T = torch.as_tensor(t).to(dev)
        M = torch.as_tensor(m).to(dev)
        n = X.shape[0]
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        losses = []
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
        acc = torch.zeros(1, dtype=torch.float32, device=dev)
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(dev)
            Xp, Tp, Mp = X[perm].contiguous(), T[perm].contiguous(), M[perm].contiguous()
            acc.zero_()
            nb = 0
            for i in range(0, n - batch + 1, batch):
                # cosine decay to lr / 20
                frac = (ep + i / n) / epochs
                cur = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * frac)))
                self.net.train_step(Xp[i:i + batch], Tp[i:i + batch], Mp[i:i + batch], cur, loss)
                acc += loss
                nb += 1
            losses.append(float(acc.item()) / max(nb, 1))
        return losses

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Predicted log2 microseconds for every option (rows of X)."""
        torch = self.net.torch
        y = self.net.forward(torch.as_tensor(np.ascontiguousarray(X, np.float32)).to(f"cuda:{self.device}"))
        return y.cpu().numpy().astype(np.float64) * self.sigma + self.mu

    def choose_rows(self, data: Dataset, margin: float = 0.0):
        """Chosen option per row of a dataset; rows whose two best predictions
        are within `margin` (log2 units) fall back to the simulated times."""
        pred = np.where(data.mask, self.predict(data.X), np.inf)
        order = np.argsort(pred, axis=1)
        best, second = order[:, 0], order[:, 1]
        rows = np.arange(len(pred))
        unsure = (pred[rows, second] - pred[rows, best]) < margin
        exact = np.argmin(data.T, axis=1)
        return np.where(unsure, exact, best), unsure

    def choose(self, dims, coll: str, nbytes: int, margin: float = 0.0):
        """(algorithm, segments, used_simulation) for one call."""
        x = features(dims, coll, [nbytes])
        pred = self.predict(x)[0]
        pred[~option_valid(coll)] = np.inf
        order = np.argsort(pred)
        if pred[order[1]] - pred[order[0]] >= margin:
            return OPTIONS[order[0]] + (False,)
        T = exhaustive(dims, coll, nbytes, self.device, self._sim_cache)
        return OPTIONS[int(np.argmin(T))] + (True,)

    def close(self):
        self.net.close()


def exhaustive(dims, coll: str, nbytes: int, device: int = 0, cache=None) -> np.ndarray:
    """Simulated time of every option (inf where invalid) for one call."""
    key = (tuple(dims), coll)
    if cache is not None and key in cache:
        stats = cache[key]
    else:
        algos = [a for a in ALGOS if a in ls.DEFINED[coll]]
        sim = ls.LinkSim([(list(dims), coll, a) for a in algos], device)
        sim.run()
        stats = {a: sim.stats(i) for i, a in enumerate(algos)}
        sim.close()
        if cache is not None:
            cache[key] = stats
    T = np.full(len(OPTIONS), np.inf)
    for j, a in enumerate(ALGOS):
        if a in stats:
            T[j * len(SEGMENTS):(j + 1) * len(SEGMENTS)] = case_times(stats[a], [nbytes])[0]
    return T


def regret(data: Dataset, choice: np.ndarray) -> np.ndarray:
    """T(chosen) / T(best) - 1 per row."""
    rows = np.arange(len(choice))
    return data.T[rows, choice] / data.T.min(axis=1) - 1.0


# ---- hand-written selectors the reference relies on ---------------------------

def reference_default_choice(coll: str, P: int, nbytes: int) -> tuple:
    """lib/libbine.cc:28-73: Bine for every collective, LIBBINE_SEGMENT_SIZE = 0."""
    return ("bine", 1)


def _seg(nbytes: int, segsize: int) -> int:
    if segsize <= 0:
        return 1
    k = max(1, -(-int(nbytes) // segsize))
    return min(SEGMENTS, key=lambda s: abs(math.log2(s) - math.log2(k)))


def fjmpi_choice(coll: str, P: int, nbytes: int) -> tuple:
    """Fujitsu MPI's decision tree (fjmpi-dectree.conf) for the process counts
    in use, reduced to the first algorithm of each rule that has a
    counterpart here: recursive_doubling / recursive_halving / binomial and the
    segmented binary/trinary trees -> recdoub (with the rule's segment size),
    ring / neighbor -> ring, modified_bruck -> bruck, pairwise-type alltoalls
    -> recdoub.  msg_size is the per-rank block for allgather, alltoall,
    scatter and gather, the whole vector otherwise."""
    S = int(nbytes)
    if coll == "allreduce":  # :93-229
        if P <= 32:
            return ("recdoub", 1) if S / P < 8192 else ("ring", 1)
        if S <= 98304:
            return ("recdoub", 1)
        return ("ring", _seg(S, 4096 if S <= 393216 else 16384 if S <= 6291456 else 65536))
    if coll == "reduce_scatter":  # :523-544
        if P <= 32 and S > 262144:
            return ("ring", 1)
        return ("recdoub", 1)
    if coll == "allgather":  # :30-62
        m = S // P
        if m < 8192 and m * P <= 65536:
            return ("recdoub", 1)
        return ("ring", 1)
    if coll == "alltoall":  # :230-258
        return ("bruck", 1) if S // P <= 192 else ("recdoub", 1)
    if coll == "bcast":  # :267-385
        if S <= 768:
            return ("recdoub", 1)
        if S <= 24576:
            return ("recdoub", _seg(S, 1024))
        if S <= 49152:
            return ("recdoub", _seg(S, 8192))
        return ("recdoub", _seg(S, 4096 if S <= 393216 else 16384 if S <= 50331648 else 65536))
    if coll == "reduce":  # :416-522
        if S <= 6144:
            return ("recdoub", 1)
        return ("recdoub", _seg(S, 8192 if S <= 12288 else 65536))
    if coll in ("gather", "scatter"):  # :386-412, :545-565
        return ("recdoub", 1)
    raise ValueError(coll)


def rule_choice_rows(data: Dataset, rule) -> np.ndarray:
    idx = {o: i for i, o in enumerate(OPTIONS)}
    out = np.zeros(len(data.keys), np.int64)
    for r, (dims, coll, nbytes) in enumerate(data.keys):
        out[r] = idx[rule(coll, int(np.prod(dims)), nbytes)]
    return out

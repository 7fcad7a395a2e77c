#!/usr/bin/env python3
"""Benchmark: emulated Bine allreduce on one MI355X.

Workload: P = 512 virtual ranks on an 8x8x8 torus, fp64 SUM, 524288 elements
(4 MiB) per rank, inputs N(0,1) resident in HBM.  One step is one complete
allreduce of all 512 ranks.  The timed region is bracketed by device
synchronisations; `value`, `ms_per_step` and `roofline.achieved` all come from
HIP events recorded on the launch stream around it (the wall clock of the same
region is reported beside them).  The measured path is the batched execution: the
whole Bine schedule compiled into one fused gather-reduce work list (replaying
the reference's reduction tree for every element), launch shape picked by the
autotuner, issued by bine_batch_launch.  The autotuner may keep the two-phase
lowering: one launch reduces every block into one rank's output, a second
copies it to the other ranks' outputs, so HBM serves a read phase and then a
write phase (same bits).  The same batch replayed as a captured
HIP graph and the plain per-step path (one kernel per schedule step, 1 + 9 + 9
launches) are timed beside it.

Algorithmic bytes per allreduce: 2 * P * S = 4 GiB (read every rank's input
once, write every rank's output once).  GB/s = algorithmic bytes / time.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "emulated Bine allreduce GB/s & %HBM peak, P=512 x 4 MiB; batched vs per-step"
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, 8 TB/s
DIMS = [8, 8, 8]
P = 512
COUNT = 524288  # fp64 elements per rank = 4 MiB
ELEM = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-buckets", action="store_true", help="skip the uneven-bucket batched vs per-step run")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-collective batch run")
    ap.add_argument("--no-linksim", action="store_true", help="skip the link-traffic simulation run")
    ap.add_argument("--no-selector", action="store_true", help="skip training and scoring the selector")
    ap.add_argument("--no-small", action="store_true", help="skip the 2x2x2 and 4x4x4 allreduce configurations")
    ap.add_argument("--no-share", action="store_true", help="skip the one-GPU-share (8x8, 64 ranks) allreduce")
    ap.add_argument("--no-collectives", action="store_true", help="skip the per-collective 8x8x8 runs")
    ap.add_argument("--cpu-warmup", type=int, default=10, help="untimed oracle runs before the CPU baseline")
    ap.add_argument("--cpu-runs", type=int, default=10, help="timed oracle runs of the CPU baseline")
    ap.add_argument("--only", choices=["batched", "step", "copy", "alltoall", "buckets", "c2"], default=None,
                    help="profiling aid: run only that part, --steps times, with a fixed launch shape")
    ap.add_argument("--tuning", default="2,128,0,0",
                    help="vpt,threads,persistent,nontemporal[,two_phase] for --only batched/alltoall/buckets/c2")
    return ap.parse_args()


def time_launches(fn, stream, steps, warmup):
    """Wall time per step (sync-bracketed) and HIP-event time on the launch stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall * 1e3, e0.elapsed_time(e1) / steps


def time_gpu(fn, stream, steps, warmup, one_by_one=0):
    """GPU milliseconds per launch with the host's issue time and the GPU's idle spells taken out: a short
    device-side spin (torch's sleep kernel, on the current stream, which must be `stream`) runs first, so every
    launch of the timed region is already queued, and the GPU busy, when it reaches the start event.  Returns the mean over `steps` back-to-back
    launches and, when one_by_one > 0, that many launches timed singly the same way."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()

    def timed(n):
        torch.cuda._sleep(4_000_000)  # about 2 ms at 2.4 GHz: longer than enqueueing the launches below
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    mean = timed(steps)
    singles = [timed(1) for _ in range(one_by_one)]
    return mean, singles


def max_ulp(a, b):
    """Largest distance in units in the last place between two float64 tensors (0 when bit-identical): the
    sign-magnitude bit patterns are mapped onto one monotone integer line and subtracted."""
    import torch
    lo = torch.iinfo(torch.int64).min
    ia, ib = a.view(torch.int64), b.view(torch.int64)
    ia = torch.where(ia < 0, lo - ia, ia)
    ib = torch.where(ib < 0, lo - ib, ib)
    return int((ia - ib).abs().max().item())


def cpu_baseline(warmup, runs):
    """The oracle (C restatement of the reference's MPI-layer allreduce B over P
    virtual ranks; message copies and reductions spread over OpenMP threads) on the
    headline workload itself: 512 ranks on 8x8x8, fp64, 4 MiB per rank.  Timed as
    bench/bench.cc:289-318 times the reference: untimed warm-up runs, then timed
    runs, wall time per run (the mean is the value, the median is reported)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    x = np.random.default_rng(0).standard_normal((P, COUNT))
    for _ in range(warmup):
        O.allreduce(x, DIMS)
    per = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.allreduce(x, DIMS)
        per.append(time.perf_counter() - t0)
    mean = float(np.mean(per))
    threads = O.threads()
    return {"value": round(2.0 * P * COUNT * ELEM / mean / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"oracle allreduce B, P=512 on 8x8x8, fp64, {COUNT} elements (4 MiB) per rank (the headline "
                      f"workload), {warmup} untimed warm-up runs, then {runs} timed runs: mean {mean:.3f} s, median "
                      f"{float(np.median(per)):.3f} s, {threads} OpenMP threads; GB/s = 2*P*S / mean time"}


def tuning_words(text):
    """vpt,threads,persistent,nontemporal[,two_phase] -> the arguments of Batch.set_tuning."""
    w = [int(v) for v in text.split(",")]
    return w[:4] + [w[4] if len(w) > 4 else 0]


def profile_only(args, stream, send, recv, recv_step):
    """Run one part of the benchmark for a profiler: the batched fused launch
    with a fixed shape, the per-step launches, a plain device copy of the
    same 2 GiB (the reference point for HBM counters), the fused alltoall
    on the same torus (int64, 4 KiB per rank pair: copy runs only), or the
    256 uneven fp32 buckets at P = 64 of the bucketed leg, with a fixed shape."""
    import torch
    from bine_trees_fugaku_amd import Batch, VComm
    with torch.cuda.stream(stream):
        if args.only == "batched":
            comm = VComm(DIMS, exec="FUSED")
            batch = Batch(comm, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": COUNT}])
            batch.set_tuning(*tuning_words(args.tuning))
            for _ in range(args.steps):
                batch.launch(stream)
        elif args.only == "alltoall":
            g = torch.Generator(device="cuda")
            g.manual_seed(99)
            a = torch.randint(-2**40, 2**40, (P, P * 512), dtype=torch.int64, device="cuda", generator=g)
            b = torch.empty_like(a)
            comm = VComm(DIMS, exec="FUSED")
            batch = Batch(comm, [{"coll": "alltoall", "sendbuf": a, "recvbuf": b, "count": 512}])
            batch.set_tuning(*tuning_words(args.tuning))
            for _ in range(args.steps):
                batch.launch(stream)
        elif args.only == "buckets":
            import numpy as np
            rng = np.random.default_rng(7)
            sizes = np.exp(rng.uniform(np.log(4096), np.log(8 << 20), 256)).astype(np.int64) // 4
            g = torch.Generator(device="cuda")
            g.manual_seed(7)
            sends = [torch.randn((64, int(n)), dtype=torch.float32, device="cuda", generator=g) for n in sizes]
            recvs = [torch.empty_like(x) for x in sends]
            comm = VComm([4, 4, 4], exec="FUSED")
            batch = Batch(comm, [{"coll": "allreduce", "sendbuf": x, "recvbuf": y, "count": int(n)}
                                 for x, y, n in zip(sends, recvs, sizes)])
            batch.set_tuning(*tuning_words(args.tuning))
            for _ in range(args.steps):
                batch.launch(stream)
        elif args.only == "c2":
            g = torch.Generator(device="cuda")
            g.manual_seed(64)
            x = torch.randn((64, 4194304), dtype=torch.float32, device="cuda", generator=g)
            y = torch.empty_like(x)
            comm = VComm([4, 4, 4], exec="FUSED")
            batch = Batch(comm, [{"coll": "allreduce", "sendbuf": x, "recvbuf": y, "count": 4194304}])
            batch.set_tuning(*tuning_words(args.tuning))
            for _ in range(args.steps):
                batch.launch(stream)
        elif args.only == "step":
            comm = VComm(DIMS, exec="STEP")
            for _ in range(args.steps):
                comm.allreduce(send, recv_step, stream=stream)
        else:
            for _ in range(args.steps):
                recv.copy_(send)
    torch.cuda.synchronize()
    print(json.dumps({"only": args.only, "steps": args.steps}), flush=True)
    return 0


def bucketed_p64(stream, seed=7, nb=256):
    """Batched vs per-step on uneven buckets: P=64 on 4x4x4, 256 fp32 buckets of
    log-uniform size 4 KiB..8 MiB per rank (rounded to 4 bytes), N(0,1)."""
    import numpy as np
    import torch
    from bine_trees_fugaku_amd import Batch, VComm
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(4096), np.log(8 << 20), nb)).astype(np.int64) // 4  # fp32 elements
    P64 = 64
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    sends = [torch.randn((P64, int(n)), dtype=torch.float32, device="cuda", generator=g) for n in sizes]
    recvs = [torch.empty_like(s) for s in sends]
    recvs2 = [torch.empty_like(s) for s in sends]
    b_alg = int(2 * P64 * sizes.sum() * 4)
    with torch.cuda.stream(stream):
        comm = VComm([4, 4, 4], exec="FUSED")
        batch = Batch(comm, [{"coll": "allreduce", "sendbuf": s, "recvbuf": r, "count": int(n)}
                             for s, r, n in zip(sends, recvs, sizes)])
        batch.autotune(iters=3, stream=stream)
        b_wall, _ = time_launches(lambda: batch.launch(stream), stream, 5, 2)
        g_wall, _ = time_launches(lambda: batch.launch_graph(stream), stream, 5, 2)
        step = VComm([4, 4, 4], exec="STEP")

        def per_step():
            for s, r in zip(sends, recvs2):
                step.allreduce(s, r, stream=stream)
        s_wall, _ = time_launches(per_step, stream, 3, 1)
    torch.cuda.synchronize()
    same = all(torch.equal(a.view(torch.int32), c.view(torch.int32)) for a, c in zip(recvs, recvs2))
    out = {"workload": "bucketed_allreduce_p64_4x4x4_fp32_256_buckets", "buckets": nb,
           "bytes_per_rank": int(sizes.sum() * 4), "algorithmic_bytes": b_alg,
           "batched_ms": round(b_wall, 3), "graph_replay_ms": round(g_wall, 3), "per_step_ms": round(s_wall, 3),
           "batched_GBps": round(b_alg / (b_wall * 1e-3) / 1e9, 1), "per_step_GBps": round(b_alg / (s_wall * 1e-3) / 1e9, 1),
           "batched_vs_per_step": round(s_wall / b_wall, 2), "per_step_launches": 13 * nb,
           "tuning": batch.tuning(), "bitwise_equal": same}
    batch.close()
    comm.close()
    step.close()
    return out


def one_gpu_share(stream, seed=64):
    """The in-HBM part one GPU would hold if the 8x8x8 torus were sharded over 8
    GPUs by its first coordinate: 64 ranks (an 8x8 slab) of fp64, 4 MiB per rank,
    run here as the fused Bine allreduce over that 8x8 torus.  A proxy for the
    per-GPU fused segments between the cross-GPU steps (same ranks, same bytes
    per rank); checked bitwise against the per-step execution."""
    import torch
    from bine_trees_fugaku_amd import Batch, VComm
    dims, count = [8, 8], COUNT
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn((64, count), dtype=torch.float64, device="cuda", generator=g)
    y, z = torch.empty_like(x), torch.empty_like(x)
    b_alg = 2 * 64 * count * 8
    with torch.cuda.stream(stream):
        comm = VComm(dims, exec="FUSED")
        bt = Batch(comm, [{"coll": "allreduce", "sendbuf": x, "recvbuf": y, "count": count}])
        bt.autotune(iters=5, stream=stream)
        w, _ = time_launches(lambda: bt.launch(stream), stream, 20, 3)
        step = VComm(dims, exec="STEP")
        step.allreduce(x, z, stream=stream)
    torch.cuda.synchronize()
    out = {"workload": "allreduce_8x8_fp64_4MiB_per_rank (64 ranks: one GPU's share of 8x8x8 over 8 GPUs)",
           "algorithmic_bytes": b_alg, "batched_ms": round(w, 4), "batched_GBps": round(b_alg / (w * 1e-3) / 1e9, 1),
           "tuning": bt.tuning(), "bitwise_equal_step": bool(torch.equal(y.view(torch.int64), z.view(torch.int64)))}
    bt.close()
    comm.close()
    step.close()
    return out


def mixed_8x8x8(stream, seed=21):
    """Mixed collective batch on 8x8x8 (P=512): allreduce fp64 4 MiB per rank,
    alltoall int64 4 KiB per rank pair, bcast int32 64 KiB, scatter fp32 256 KiB
    blocks, gather int32 1 MiB blocks, reduce fp32 1 MiB; seeded uniform
    integers in [-1000, 1000] and N(0,1).  Batched: one fused work list per
    element type, issued by bine_batch_launch (and replayed as a captured HIP
    graph beside it).  Baseline: the six collectives as
    six single calls (each one fused launch; per-step execution of the rooted
    collectives at these sizes needs the reference's full-size temporary on
    every rank, P x 512 MiB for the gather)."""
    import torch
    from bine_trees_fugaku_amd import Batch, VComm
    Pm = 512
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)

    def ri(shape, dt):
        return torch.randint(-1000, 1001, shape, dtype=dt, device="cuda", generator=g)

    def rn(shape, dt):
        return torch.randn(shape, dtype=dt, device="cuda", generator=g)
    c_ar, c_a2a, c_b, c_s, c_g, c_r = 524288, 512, 16384, 65536, 262144, 262144
    t = {
        "allreduce": (rn((Pm, c_ar), torch.float64), torch.empty((Pm, c_ar), dtype=torch.float64, device="cuda"), c_ar),
        "alltoall": (ri((Pm, Pm * c_a2a), torch.int64), torch.empty((Pm, Pm * c_a2a), dtype=torch.int64, device="cuda"), c_a2a),
        "bcast": (None, ri((Pm, c_b), torch.int32), c_b),
        "scatter": (rn((Pm * c_s,), torch.float32), torch.empty((Pm, c_s), dtype=torch.float32, device="cuda"), c_s),
        "gather": (ri((Pm, c_g), torch.int32), torch.empty((Pm * c_g,), dtype=torch.int32, device="cuda"), c_g),
        "reduce": (rn((Pm, c_r), torch.float32), torch.empty((c_r,), dtype=torch.float32, device="cuda"), c_r),
    }
    # compulsory bytes per collective: read every input once, write every output once
    b_alg = (2 * Pm * c_ar * 8 + 2 * Pm * Pm * c_a2a * 8 + (Pm + 1) * c_b * 4 + 2 * Pm * c_s * 4 + 2 * Pm * c_g * 4
             + (Pm + 1) * c_r * 4)
    with torch.cuda.stream(stream):
        comm = VComm(DIMS, exec="FUSED")
        batch = Batch(comm, [{"coll": c, "sendbuf": s, "recvbuf": r, "count": n} for c, (s, r, n) in t.items()])
        batch.autotune(iters=5, stream=stream)
        b_wall, _ = time_launches(lambda: batch.launch(stream), stream, 10, 3)
        g_wall, _ = time_launches(lambda: batch.launch_graph(stream), stream, 10, 3)
        outs = {c: r.clone() for c, (s, r, n) in t.items() if c != "bcast"}
        step = VComm(DIMS, exec="AUTO")

        def singles():
            step.allreduce(t["allreduce"][0], t["allreduce"][1], stream=stream)
            step.alltoall(t["alltoall"][0], t["alltoall"][1], stream=stream)
            step.bcast(t["bcast"][1], stream=stream)
            step.scatter(t["scatter"][0], t["scatter"][1], stream=stream)
            step.gather(t["gather"][0], t["gather"][1], stream=stream)
            step.reduce(t["reduce"][0], t["reduce"][1], stream=stream)
        s_wall, _ = time_launches(singles, stream, 3, 1)
    torch.cuda.synchronize()
    same = all(torch.equal(outs[c].view(torch.uint8), t[c][1].view(torch.uint8)) for c in outs)
    out = {"workload": "mixed_8x8x8_allreduce_fp64_4MiB_alltoall_int64_4KiB_bcast_scatter_gather_reduce",
           "algorithmic_bytes": int(b_alg), "batched_ms": round(b_wall, 3), "graph_replay_ms": round(g_wall, 3),
           "per_call_ms": round(s_wall, 3),
           "batched_GBps": round(b_alg / (b_wall * 1e-3) / 1e9, 1), "per_call_GBps": round(b_alg / (s_wall * 1e-3) / 1e9, 1),
           "batched_vs_per_call": round(s_wall / b_wall, 2), "stats": batch.stats(), "tuning": batch.tuning(),
           "bitwise_equal": same}
    batch.close()
    comm.close()
    step.close()
    return out


def small_configs(stream, cpu=True):
    """Allreduce on 2x2x2 int32 (262144 elements per rank, uniform integers in
    [-1000, 1000]) and on 4x4x4 fp32 (4M elements per rank, N(0,1)): GPU per-step,
    GPU fused single call, GPU batched graph, and the oracle on the host (3 untimed
    warm-up runs, then the mean of 5 timed runs of the whole configuration).
    GB/s = 2*P*S / time."""
    import numpy as np
    import torch
    from bine_trees_fugaku_amd import Batch, VComm
    out = {}
    for name, dims, dt, count in (("allreduce_2x2x2_int32_1MiB", [2, 2, 2], torch.int32, 262144),
                                  ("allreduce_4x4x4_fp32_16MiB", [4, 4, 4], torch.float32, 4194304)):
        Pc = int(np.prod(dims))
        g = torch.Generator(device="cuda")
        g.manual_seed(Pc)
        if dt == torch.int32:
            x = torch.randint(-1000, 1001, (Pc, count), dtype=dt, device="cuda", generator=g)
        else:
            x = torch.randn((Pc, count), dtype=dt, device="cuda", generator=g)
        y = torch.empty_like(x)
        b_alg = 2 * Pc * count * x.element_size()
        res = {"algorithmic_bytes": b_alg}
        with torch.cuda.stream(stream):
            for exec in ("STEP", "FUSED"):
                c = VComm(dims, exec=exec)
                c.autotune("allreduce", x, y, count, exec=exec, iters=2, stream=stream)
                w, _ = time_launches(lambda: c.allreduce(x, y, stream=stream), stream, 20, 3)
                res[f"{exec.lower()}_ms"] = round(w, 4)
                res[f"{exec.lower()}_GBps"] = round(b_alg / (w * 1e-3) / 1e9, 1)
                c.close()
            c = VComm(dims, exec="FUSED")
            bt = Batch(c, [{"coll": "allreduce", "sendbuf": x, "recvbuf": y, "count": count}])
            bt.autotune(iters=2, stream=stream)
            w, _ = time_launches(lambda: bt.launch(stream), stream, 20, 3)
            res["batched_ms"] = round(w, 4)
            res["batched_GBps"] = round(b_alg / (w * 1e-3) / 1e9, 1)
            w, _ = time_launches(lambda: bt.launch_graph(stream), stream, 20, 3)
            res["graph_replay_ms"] = round(w, 4)
            # the same launches with the host's issue time taken out (all queued behind a device-side spin)
            gpu_ms, _ = time_gpu(lambda: bt.launch(stream), stream, 20, 3)
            res["batched_gpu_only_ms"] = round(gpu_ms, 4)
            res["batched_gpu_only_GBps"] = round(b_alg / (gpu_ms * 1e-3) / 1e9, 1)
            gpu_ms, _ = time_gpu(lambda: bt.launch_graph(stream), stream, 20, 3)
            res["graph_replay_gpu_only_ms"] = round(gpu_ms, 4)
            bt.close()
            c.close()
            # the per-step schedule (one launch per step) issued directly and as one captured graph
            c = VComm(dims, exec="STEP")
            bt = Batch(c, [{"coll": "allreduce", "sendbuf": x, "recvbuf": y, "count": count}])
            w, _ = time_launches(lambda: bt.launch(stream), stream, 20, 3)
            res["step_direct_ms"] = round(w, 4)
            w, _ = time_launches(lambda: bt.launch_graph(stream), stream, 20, 3)
            res["step_graph_ms"] = round(w, 4)
            gpu_ms, _ = time_gpu(lambda: bt.launch(stream), stream, 20, 3)
            res["step_direct_gpu_only_ms"] = round(gpu_ms, 4)
            gpu_ms, _ = time_gpu(lambda: bt.launch_graph(stream), stream, 20, 3)
            res["step_graph_gpu_only_ms"] = round(gpu_ms, 4)
            bt.close()
            c.close()
        torch.cuda.synchronize()
        if cpu:
            from oracle import oracle as O
            xh = x.cpu().numpy()
            for _ in range(3):
                ref = O.allreduce(xh, dims)
            per = []
            for _ in range(5):
                t0 = time.perf_counter()
                ref = O.allreduce(xh, dims)
                per.append(time.perf_counter() - t0)
            el = float(np.mean(per))
            res["cpu_oracle_s"] = round(el, 4)
            res["cpu_oracle_GBps"] = round(b_alg / el / 1e9, 3)
            res["cpu_threads"] = O.threads()
            res["bitwise_equal_oracle"] = bool(np.array_equal(y.cpu().numpy().view(np.uint8), ref.view(np.uint8)))
        out[name] = res
        del x, y
    return out


def per_collective_8x8x8(stream):
    """Each collective alone on 8x8x8 (P=512), fused single call, autotuned:
    GB/s of its compulsory bytes (read every input once, write every output once).
    allreduce/reduce_scatter/allgather move 4 MiB per rank, alltoall 4 KiB per
    rank pair (int64), bcast/scatter/gather/reduce 1 MiB per rank."""
    import torch
    from bine_trees_fugaku_amd import VComm
    Pc = 512
    g = torch.Generator(device="cuda")
    g.manual_seed(99)
    out = {}
    M = 1 << 20
    # (name, dtype, count, send elems, recv elems, compulsory bytes, call)
    cases = [
        ("allreduce", torch.float32, M, Pc * M, Pc * M, 2 * Pc * M * 4),
        ("reduce_scatter", torch.float32, M // Pc, Pc * M, M, (Pc + 1) * M * 4),
        ("allgather", torch.float32, M // Pc, M, Pc * M, (M + Pc * M) * 4),
        ("alltoall", torch.int64, 512, Pc * Pc * 512, Pc * Pc * 512, 2 * Pc * Pc * 512 * 8),
        ("bcast", torch.int32, M // 4, 0, Pc * (M // 4), (Pc + 1) * M),
        ("scatter", torch.float32, M // 4, Pc * (M // 4), Pc * (M // 4), 2 * Pc * M),
        ("gather", torch.int32, M // 4, Pc * (M // 4), Pc * (M // 4), 2 * Pc * M),
        ("reduce", torch.float32, M // 4, Pc * (M // 4), M // 4, (Pc + 1) * M),
    ]
    with torch.cuda.stream(stream):
        comm = VComm(DIMS, exec="FUSED")
        for name, dt, count, ns, nr, b_alg in cases:
            send = (torch.randn(ns, device="cuda", generator=g).to(dt) if ns else None)
            recv = torch.randn(nr, device="cuda", generator=g).to(dt)
            if name == "allreduce":
                fn = lambda: comm.allreduce(send, recv, stream=stream)  # noqa: E731
            elif name == "reduce_scatter":
                fn = lambda: comm.reduce_scatter(send, recv, stream=stream)  # noqa: E731
            elif name == "allgather":
                fn = lambda: comm.allgather(send, recv, stream=stream)  # noqa: E731
            elif name == "alltoall":
                fn = lambda: comm.alltoall(send, recv, stream=stream)  # noqa: E731
            elif name == "bcast":
                fn = lambda: comm.bcast(recv, stream=stream)  # noqa: E731
            elif name == "scatter":
                fn = lambda: comm.scatter(send, recv, stream=stream)  # noqa: E731
            elif name == "gather":
                fn = lambda: comm.gather(send, recv, stream=stream)  # noqa: E731
            else:
                fn = lambda: comm.reduce(send, recv, stream=stream)  # noqa: E731
            comm.autotune(name, send, recv, count, exec="FUSED", iters=5, stream=stream)
            # GPU time only: timed from events after a device-side spin, so that neither the host's issue time nor
            # the GPU's resumption after an idle spell falls inside the region.  Five regions of 10 back-to-back
            # launches; `ms` is the median region's mean, so that one rare slow launch does not stand for the
            # collective (every region's mean is reported beside it)
            regions = []
            for k in range(5):
                ev, one = time_gpu(fn, stream, 10, 2 if k == 0 else 0, one_by_one=9 if k == 4 else 0)
                regions.append(ev)
            ev = sorted(regions)[len(regions) // 2]
            one = sorted(one)
            out[name] = {"algorithmic_bytes": int(b_alg), "ms": round(ev, 4),
                         "GBps": round(b_alg / (ev * 1e-3) / 1e9, 1),
                         "pct_hbm_peak": round(100 * b_alg / (ev * 1e-3) / 1e9 / HBM_PEAK_GBPS, 1),
                         "region_means_ms": [round(r, 4) for r in regions],
                         "single_launch_ms_min_median_max": [round(one[0], 4), round(one[len(one) // 2], 4),
                                                             round(one[-1], 4)]}
            del send, recv
        comm.close()
    torch.cuda.synchronize()
    return out


def linksim_tori(stream):
    """Link-traffic simulation over the five tori 2x2x2, 4x4x4, 8x8x8, 64x64 and
    32x256: every algorithm defined for each of the 8 collectives, one GPU pass of
    integer atomics.  A sample is one (collective, algorithm, torus, message size)
    point at the 26 sizes 2^3..2^28 bytes."""
    import torch
    from bine_trees_fugaku_amd import linksim as ls
    tori = ([2, 2, 2], [4, 4, 4], [8, 8, 8], [64, 64], [32, 256])
    cases = [(d, c, a) for d in tori for c, algos in ls.DEFINED.items() for a in algos
             if not (a == "simpy" and len(set(d)) != 1)]
    sim = ls.LinkSim(cases)
    reps = 20
    sim.run(stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        sim.run(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    classes = crossings = 0
    for i in range(len(cases)):
        st = sim.stats(i)
        classes += len(st.num)
        crossings += int(st.sum_mult.sum())
    sim.close()
    samples = len(cases) * 26
    # Link counts do not depend on the message size (bytes are count * floor(S * num / den), computed on the
    # host), so the GPU work per pass is the cases and their step classes; samples/s multiplies by the 26
    # sizes each pass answers.
    return {"workload": "linksim_5_tori_all_algorithms", "cases": len(cases), "classes": classes,
            "link_crossings_per_pass": crossings, "ms_per_pass": round(ms, 4),
            "gpu_cases_per_s": round(len(cases) / (ms * 1e-3), 1), "gpu_classes_per_s": round(classes / (ms * 1e-3), 1),
            "samples_per_s": round(samples / (ms * 1e-3), 1), "sizes_per_case": 26,
            "link_crossings_per_s": round(crossings / (ms * 1e-3), 1)}


def selector_regret():
    """Learned selector: train the 3x256 network on the GPU on 199 training tori,
    measure mean regret on the five held-out tori against exhaustive simulation."""
    import numpy as np
    from bine_trees_fugaku_amd import selector as sel
    t0 = time.perf_counter()
    train = sel.simulate(sel.training_tori(), sel.default_sizes(2))
    test = sel.simulate(sel.HELD_OUT, sel.default_sizes(4))
    t1 = time.perf_counter()
    s = sel.Selector(seed=0)
    s.fit(train)
    t2 = time.perf_counter()
    r = sel.regret(test, s.choose_rows(test)[0])
    ch, unsure = s.choose_rows(test, margin=0.05)
    rf = sel.regret(test, ch)
    base = {name: round(float(sel.regret(test, sel.rule_choice_rows(test, rule)).mean()), 4)
            for name, rule in (("reference_defaults", sel.reference_default_choice), ("fjmpi", sel.fjmpi_choice))}
    s.close()
    return {"train_rows": int(len(train.X)), "train_samples": train.samples, "simulate_s": round(t1 - t0, 2),
            "train_s": round(t2 - t1, 2), "held_out_mean_regret": round(float(r.mean()), 4),
            "held_out_p95_regret": round(float(np.percentile(r, 95)), 4),
            "with_fallback_mean_regret": round(float(rf.mean()), 4), "fallback_rate": round(float(unsure.mean()), 4),
            "baseline_mean_regret": base}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the P virtual ranks live in one MI355X's HBM; "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    import torch
    from bine_trees_fugaku_amd import Batch, VComm

    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    send = torch.randn((P, COUNT), dtype=torch.float64, device="cuda", generator=g)
    recv = torch.empty_like(send)
    recv_step = torch.empty_like(send)
    b_alg = 2 * P * COUNT * ELEM

    if args.only:
        return profile_only(args, stream, send, recv, recv_step)

    with torch.cuda.stream(stream):
        fused = VComm(DIMS, exec="FUSED")
        batch = Batch(fused, [{"coll": "allreduce", "sendbuf": send, "recvbuf": recv, "count": COUNT}])
        batch.autotune(iters=3, stream=stream)
        tuning = batch.tuning()
        bstats = batch.stats()
        wall_ms, ev_ms = time_launches(lambda: batch.launch(stream), stream, args.steps, args.warmup)
        # the same batch replayed as a captured HIP graph (each replay adds a fence of its own)
        g_wall, g_ev = time_launches(lambda: batch.launch_graph(stream), stream, args.steps, args.warmup)

        step = VComm(DIMS, exec="STEP")
        step.autotune("allreduce", send, recv_step, COUNT, exec="STEP", iters=2, stream=stream)
        sstats = step.plan_stats("allreduce", COUNT, "float64", exec="STEP")
        s_wall, s_ev = time_launches(lambda: step.allreduce(send, recv_step, stream=stream), stream,
                                     max(3, args.steps // 4), 2)
        # reference point: a plain device copy of the same 2 GiB (reads S*P, writes S*P)
        c_wall, c_ev = time_launches(lambda: recv_step.copy_(send), stream, 5, 2)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        step.allreduce(send, recv_step, stream=stream)
    torch.cuda.synchronize()
    same = bool(torch.equal(recv.view(torch.int64), recv_step.view(torch.int64)))
    ulp = max_ulp(recv, recv_step)

    # one timer for value, ms_per_step and roofline.achieved: the HIP events around the timed launches
    gbps = b_alg / (ev_ms * 1e-3) / 1e9
    ach = gbps
    step_gbps = b_alg / (s_ev * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_allreduce_8x8x8_fp64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_step")

    line = {
        "metric": METRIC, "value": round(gbps, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ev_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic N(0,1), seeded",
        "config": {
            "workload": "allreduce_8x8x8_fp64_4MiB_per_rank", "P": P, "torus": "8x8x8", "elements_per_rank": COUNT,
            "bytes_per_rank": COUNT * ELEM, "algorithmic_bytes": b_alg,
            "exec": "batched: the autotuned fused work list (bine_batch_launch): one launch, or two when the "
                    "two-phase lowering wins (reductions writing one output per block, then copies to the others)",
            "timer": "HIP events on the launch stream around the timed launches, synchronised on both sides",
            "wall_ms_per_step": round(wall_ms, 4),
            "graph_replay_ms": round(g_ev, 4), "graph_replay_GBps": round(b_alg / (g_ev * 1e-3) / 1e9, 2),
            "tuning": tuning, "batched_launches": bstats["launches"], "batched_items": bstats["items"],
            "per_step_GBps": round(step_gbps, 2), "per_step_ms": round(s_ev, 4),
            "per_step_launches": sstats["launches"], "per_step_schedule_bytes": sstats["bytes_read"] + sstats["bytes_written"],
            "batched_vs_per_step": round(s_ev / ev_ms, 3), "pct_hbm_peak": round(100 * gbps / HBM_PEAK_GBPS, 2),
            "batched_equals_per_step_bitwise": same, "max_ulp_batched_vs_per_step": ulp,
            "device_copy_GBps": round(b_alg / (c_ev * 1e-3) / 1e9, 2),
        },
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": traffic},
        "cpu_baseline": None,
    }
    del recv_step
    torch.cuda.empty_cache()
    if not args.no_buckets:
        line["config"]["bucketed"] = bucketed_p64(stream)
        torch.cuda.empty_cache()
    if not args.no_mixed:
        line["config"]["mixed_8x8x8"] = mixed_8x8x8(stream)
        torch.cuda.empty_cache()
    if not args.no_small:
        line["config"]["small_configs"] = small_configs(stream, cpu=not args.no_cpu_baseline)
        torch.cuda.empty_cache()
    if not args.no_share:
        line["config"]["one_gpu_share"] = one_gpu_share(stream)
        torch.cuda.empty_cache()
    if not args.no_collectives:
        line["config"]["collectives_8x8x8"] = per_collective_8x8x8(stream)
        torch.cuda.empty_cache()
    if not args.no_linksim:
        line["config"]["linksim"] = linksim_tori(stream)
    if not args.no_selector:
        line["config"]["selector"] = selector_regret()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_warmup, args.cpu_runs)
    print(json.dumps(line), flush=True)
    batch.close()
    fused.close()
    step.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

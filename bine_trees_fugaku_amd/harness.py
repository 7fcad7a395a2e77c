"""Benchmark harness in the reference's record formats.

GPU counterpart of bench/bench.cc (:107-395): the same command line
(collective, type, count in elements, iterations), 10 untimed warm-up
iterations, every timed iteration bracketed by a device synchronisation (the
reference brackets with two barriers), a validation of the result, and the
same stdout record (bench/bench.cc:376-387):

    highest
    <ns of iteration 0>
    ...
    Average runtime: <seconds>

"highest" is the per-iteration maximum over ranks in the reference; here all
P virtual ranks run in one call, so an iteration's time is the collective's.

The metadata CSV follows bench/generate_metadata.py: the same header
(:18-29) and the same environment variables (:65-75).  The reference writes
four keys that its header lacks (:41-47), which csv.DictWriter rejects; only
the header's fields are written here.

    LIBBINE_DIMENSIONS=8x8x8 python -m bine_trees_fugaku_amd.harness MPI_Allreduce INT32 262144 20
"""
from __future__ import annotations

import argparse
import csv
import os
import sys
import time

TYPES = {  # bench/bench.cc:244-260, plus the 8-byte types the library adds
    "CHAR": ("int8", "SUM"), "BYTE": ("int8", "SUM"), "FLOAT": ("float32", "SUM"), "INT": ("int32", "SUM"),
    "INT32": ("int32", "SUM"), "VOID": ("int32", "VOID"), "DOUBLE": ("float64", "SUM"), "INT64": ("int64", "SUM"),
}
COLLECTIVES = ("MPI_Allreduce", "MPI_Reduce_scatter", "MPI_Allgather", "MPI_Bcast", "MPI_Alltoall", "MPI_Scatter",
               "MPI_Gather", "MPI_Reduce")
METADATA_FIELDS = ["timestamp", "test_id", "nnodes", "collective_type", "mpi_lib", "mpi_lib_version",
                   "libbine_version", "CUDA", "output_level", "MPI_Op", "notes"]
METADATA_ENV = {"timestamp": "TIMESTAMP", "nnodes": "N_NODES", "collective_type": "COLLECTIVE_TYPE",
                "mpi_lib": "MPI_LIB", "mpi_lib_version": "MPI_LIB_VERSION", "libbine_version": "LIBBINE_VERSION",
                "CUDA": "CUDA", "output_level": "OUTPUT_LEVEL", "MPI_Op": "MPI_OP", "notes": "NOTES"}


def buffer_counts(collective: str, count: int, P: int):
    """Elements per rank of (sendbuf, recvbuf), as allocate_buffers sizes them (bench/bench.cc:178-210)."""
    table = {
        "MPI_Allreduce": (count, count), "MPI_Bcast": (count, count), "MPI_Reduce": (count, count),
        "MPI_Reduce_scatter": (count * P, count), "MPI_Allgather": (count, count * P),
        "MPI_Scatter": (count * P, count), "MPI_Gather": (count, count * P), "MPI_Alltoall": (count * P, count * P),
    }
    if collective not in table:
        raise ValueError(f"Unknown collective {collective}")
    return table[collective]


def format_record(samples_s) -> str:
    """The stdout record of bench/bench.cc:376-387 for per-iteration times in seconds."""
    lines = ["highest"]
    for s in samples_s:
        lines.append(str(int(s * 1e9)))
    avg = sum(samples_s) / len(samples_s)
    lines.append(f"Average runtime: {avg:f}")
    return "\n".join(lines) + "\n"


def write_metadata(results_dir: str, system_name: str, test_id: int, values: dict) -> str:
    """Append one row to <results_dir>/<system_name>_metadata.csv (header written once)."""
    path = os.path.join(results_dir, f"{system_name}_metadata.csv")
    new = not os.path.isfile(path)
    row = {k: values.get(k, "") for k in METADATA_FIELDS}
    row["test_id"] = int(test_id)
    if row.get("notes") in ("", None):
        row["notes"] = "null"
    with open(path, "a", newline="") as f:
        w = csv.DictWriter(f, fieldnames=METADATA_FIELDS)
        if new:
            w.writeheader()
        w.writerow(row)
    return path


def metadata_from_env() -> dict:
    return {k: os.environ.get(v, "") for k, v in METADATA_ENV.items()}


class Run:
    """One benchmark configuration on the GPU: buffers, the timed call, validation."""

    def __init__(self, collective: str, type_name: str, count: int, comm=None, seed: int = 0):
        import torch
        from .collectives import VComm
        if type_name not in TYPES:
            raise ValueError(f"Unknown type {type_name}")
        self.torch = torch
        self.collective, self.count = collective, int(count)
        self.dtype_name, self.op = TYPES[type_name]
        self.comm = comm if comm is not None else VComm.from_env()
        P = self.comm.size
        ns, nr = buffer_counts(collective, self.count, P)
        dt = getattr(torch, self.dtype_name)
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        # bench/bench.cc:213-219 fills bytes with rand() % 1024; here every element is a seeded integer in [0, 1024)
        send = torch.randint(0, 1024 if self.dtype_name != "int8" else 128, (P, ns), generator=g, device="cuda")
        self.send = send.to(dt)
        if collective == "MPI_Scatter":
            self.send = self.send[0].contiguous()  # the root's buffer
        self.recv = torch.zeros((P, nr) if collective not in ("MPI_Gather", "MPI_Reduce") else (nr,), dtype=dt,
                                device="cuda")
        if collective == "MPI_Bcast":
            self.recv = self.send.clone()

    def call(self, comm=None):
        c = comm or self.comm
        k = self.collective
        if k == "MPI_Allreduce":
            c.allreduce(self.send, self.recv, op=self.op)
        elif k == "MPI_Reduce_scatter":
            c.reduce_scatter(self.send, self.recv, op=self.op)
        elif k == "MPI_Allgather":
            c.allgather(self.send, self.recv)
        elif k == "MPI_Bcast":
            c.bcast(self.recv)
        elif k == "MPI_Alltoall":
            c.alltoall(self.send, self.recv)
        elif k == "MPI_Scatter":
            c.scatter(self.send, self.recv)
        elif k == "MPI_Gather":
            c.gather(self.send, self.recv)
        elif k == "MPI_Reduce":
            c.reduce(self.send, self.recv, op=self.op)

    def time(self, iterations: int, warmup: int = 10):
        torch = self.torch
        samples = []
        for i in range(-warmup, iterations):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.call()
            torch.cuda.synchronize()
            if i >= 0:
                samples.append(time.perf_counter() - t0)
        return samples

    def expected(self):
        """The collective's MPI result computed directly with torch on the same device buffers: an
        implementation independent of the library, as the reference validates against the system MPI
        (bench/bench.cc:270, 328-344).  Integer sums are formed in int64 and wrapped to the element type;
        float sums in float64."""
        torch = self.torch
        P = self.comm.size
        k = self.collective
        x = self.send
        wide = torch.float64 if x.dtype.is_floating_point else torch.int64

        def narrow(v):
            if x.dtype.is_floating_point:
                return v
            bits = torch.iinfo(x.dtype).bits
            v = torch.remainder(v + (1 << (bits - 1)), 1 << bits) - (1 << (bits - 1))
            return v.to(x.dtype)

        if k == "MPI_Allreduce":
            return narrow(x.to(wide).sum(0)).expand(P, -1)
        if k == "MPI_Reduce":
            return narrow(x.to(wide).sum(0))
        if k == "MPI_Reduce_scatter":
            return narrow(x.to(wide).sum(0)).reshape(P, -1)
        if k == "MPI_Allgather":
            return x.reshape(1, -1).expand(P, -1)
        if k == "MPI_Gather":
            return x.reshape(-1)
        if k == "MPI_Bcast":
            return x[0].expand(P, -1)
        if k == "MPI_Scatter":
            return x.reshape(P, -1)
        c = x.shape[1] // P  # MPI_Alltoall: block j of rank i lands as block i of rank j
        return x.reshape(P, P, c).transpose(0, 1).reshape(P, P * c)

    def validate(self) -> bool:
        """Compare the timed output with the directly computed MPI result (exact for integer types; for
        floats within the rounding of P - 1 additions, |err| <= (P - 1) * eps * sum |x|), and bit for bit
        with the same collective run one launch per reference step on fresh buffers.  Skipped for the VOID
        op, as bench/bench.cc:334 does."""
        if self.op == "VOID":
            return True
        torch = self.torch
        got = self.recv.clone()
        want = self.expected()
        if got.dtype.is_floating_point and self.collective in ("MPI_Allreduce", "MPI_Reduce", "MPI_Reduce_scatter"):
            mag = self.send.to(torch.float64).abs().sum(0)
            if self.collective == "MPI_Reduce_scatter":
                mag = mag.reshape(self.comm.size, -1)
            eps = torch.finfo(got.dtype).eps
            ok = bool(((got.to(torch.float64) - want).abs() <= (self.comm.size - 1) * eps * mag).all())
        else:
            ok = torch.equal(got.reshape(want.shape).view(torch.uint8), want.contiguous().view(torch.uint8))
        if not ok:
            return False
        saved = self.recv
        self.recv = torch.zeros_like(saved) if self.collective != "MPI_Bcast" else self.send.clone()
        with self.comm.with_exec("STEP") as step:
            self.call(step)
        torch.cuda.synchronize()
        ok = torch.equal(got.view(torch.uint8), self.recv.view(torch.uint8))
        self.recv = saved
        return ok


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="bench/bench.cc on one GPU: collective type count iterations")
    ap.add_argument("collective", choices=COLLECTIVES)
    ap.add_argument("type", choices=sorted(TYPES))
    ap.add_argument("count", type=int)
    ap.add_argument("iterations", type=int)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--metadata", nargs=2, metavar=("RESULTS_DIR", "TEST_ID"),
                    help="also append a row to <RESULTS_DIR>/<LOCATION>_metadata.csv")
    a = ap.parse_args(argv)
    run = Run(a.collective, a.type, a.count)
    samples = run.time(a.iterations, a.warmup)
    skip = os.environ.get("LIBBINE_SKIP_VALIDATION")
    if (skip is None or skip == "0") and not run.validate():
        print(f"Rank 0: Validation failed for {a.collective}", file=sys.stderr)
        return 1
    sys.stdout.write(format_record(samples))
    if a.metadata:
        system = os.environ.get("LOCATION")
        if not system:
            print("LOCATION is not set", file=sys.stderr)
            return 1
        write_metadata(a.metadata[0], system, int(a.metadata[1]), metadata_from_env())
    return 0


if __name__ == "__main__":
    sys.exit(main())

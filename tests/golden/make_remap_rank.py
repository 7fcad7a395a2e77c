"""Write remap_rank.json: the outputs of the reference's simulation/remap_rank.c,
compiled unchanged under oracle/_ref by oracle/Makefile, for every rank of
P = 2..32 and of the non-powers of two 6, 10, 12, 24 (where the decimal
80000000 test acts), and seeded samples of P = 512, 4096, 8192.

    make -C oracle && python tests/golden/make_remap_rank.py
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "remap_rank")


def ref(P, r):
    out = subprocess.run([REF, str(P), str(r)], capture_output=True, text=True, check=True).stdout
    return int(out.strip().split()[-1])  # "rank: r, remap_rank: v"


def main():
    cases = [(P, r) for P in (2, 4, 8, 16, 32) for r in range(P)]
    cases += [(P, r) for P in (6, 10, 12, 24) for r in range(P)]
    rng = np.random.default_rng(0)
    cases += [(P, int(r)) for P in (512, 4096, 8192) for r in rng.integers(0, P, 24)]
    data = [[P, r, ref(P, r)] for P, r in cases]
    with open(os.path.join(HERE, "remap_rank.json"), "w") as f:
        json.dump({"source": "simulation/remap_rank.c via oracle/_ref/remap_rank", "P_rank_remap": data}, f)


if __name__ == "__main__":
    main()

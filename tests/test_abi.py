"""The C-ABI library loads on the CPU and exports every symbol include/bine_hip.h declares."""
import ctypes

import numpy as np
import os
import re

import pytest

from bine_trees_fugaku_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bine_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|uint32_t|void|const char\*|char\*)\s*\**\s*(bine_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_the_mpi_entry_points():
    names = declared_functions()
    for n in ("bine_allreduce", "bine_reduce_scatter", "bine_allgather", "bine_bcast", "bine_alltoall",
              "bine_scatter", "bine_gather", "bine_reduce"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTED) == declared_functions()


def test_config_defaults_follow_init_env():
    cfg = _lib.VCommConfig()
    _lib.check(_lib.load().bine_vcomm_config_init(ctypes.byref(cfg), 64), "init")
    assert cfg.dimensions_num == 1 and cfg.dimensions[0] == 64 and cfg.num_ports == 1
    assert cfg.allreduce.family == 1 and cfg.allreduce.algo == 1 and cfg.allreduce.distance == 0
    assert cfg.allgather.algo == 0 and cfg.allgather.distance == 1  # CONT_PERMUTE, DECREASING
    assert cfg.bcast.distance == 1 and cfg.gather.distance == 1


def test_config_from_env(monkeypatch):
    monkeypatch.setenv("LIBBINE_DIMENSIONS", "8x8x8")
    monkeypatch.setenv("LIBBINE_NUM_PORTS", "3")
    monkeypatch.setenv("LIBBINE_ALLREDUCE_ALGO", "B_CONT")
    monkeypatch.setenv("LIBBINE_ALLTOALL_ALGO_FAMILY", "BRUCK")
    monkeypatch.setenv("LIBBINE_EXEC", "STEP")
    cfg = _lib.VCommConfig()
    L = _lib.load()
    _lib.check(L.bine_vcomm_config_init(ctypes.byref(cfg), 2), "init")
    _lib.check(L.bine_vcomm_config_from_env(ctypes.byref(cfg)), "env")
    assert list(cfg.dimensions) == [8, 8, 8] and cfg.dimensions_num == 3 and cfg.num_ports == 3
    assert cfg.allreduce.algo == 3 and cfg.alltoall.family == 3 and cfg.exec == 1
    monkeypatch.setenv("LIBBINE_ALLREDUCE_ALGO", "NOPE")
    assert L.bine_vcomm_config_from_env(ctypes.byref(cfg)) == _lib.BINE_ERR_ARG


def test_schedule_helpers_match_oracle():
    from oracle import oracle as O
    L = _lib.load()
    for dims in ([8], [4, 4], [8, 8, 8], [2, 8, 2], [6, 6], [32, 256]):
        d = (ctypes.c_uint32 * 3)(*(list(dims) + [1] * (3 - len(dims))))
        P = 1
        for x in dims:
            P *= x
        steps = len(O.step_info(dims))
        ranks = range(P) if P <= 64 else [0, 1, 7, P // 2 + 3, P - 1]
        for port in range(min(3, 2 * len(dims))):
            sd = (ctypes.c_uint32 * steps)()
            sk = (ctypes.c_uint32 * steps)()
            _lib.check(L.bine_compute_step_info(d, len(dims), port, sd, sk), "step_info")
            assert [(sd[i], sk[i]) for i in range(steps)] == O.step_info(dims, port)
            for r in ranks:
                out = (ctypes.c_uint32 * steps)()
                _lib.check(L.bine_compute_peers(d, len(dims), r, port, 1, out), "peers")
                assert list(out) == O.compute_peers(dims, r, port)
                if P <= 512:
                    bs = (ctypes.c_uint32 * P)()
                    _lib.check(L.bine_block_step(d, len(dims), r, port, 1, bs), "block_step")
                    assert list(bs) == list(O.block_step(dims, r, port))


def test_get_peer_and_mirroring_port():
    """get_peer_c (lib/libbine_coll.cc:89-107): one step from a coordinate.  Expected values from the definition:
    step i moves dimension d of compute_step_info's (d, k) by rho_k = (1 - (-2)^(k+1)) / 3 for BINE (negated on an
    odd coordinate) or to c ^ 2^k for RECDOUB, negated again on a mirrored port; on tori of equal dimensions the
    result is compute_peers' peer of the rank.  get_mirroring_port (lib/libbine_coll.cc:44-52): the first mirrored
    port, ports >= 3 on 3D tori, 2-3 on 2D, odd ones on 1D."""
    from oracle import oracle as O
    L = _lib.load()
    i32 = ctypes.c_int32

    def mirrored(port, nd):
        return port >= 3 if nd == 3 else (port in (2, 3) if nd == 2 else port % 2 == 1)

    for dims in ([8], [16], [4, 4], [6, 6], [8, 8, 8], [2, 8, 2], [4, 16]):
        nd = len(dims)
        d = (ctypes.c_uint32 * 3)(*(list(dims) + [1] * (3 - nd)))
        P = int(np.prod(dims))
        for family, fam_code in (("BINE", 1), ("RECDOUB", 2)):
            if family == "RECDOUB" and any(x & (x - 1) for x in dims):
                continue
            for port in range(2 * nd):
                steps = O.step_info(dims, port)
                for r in (range(P) if P <= 64 else [0, 1, 5, 77, P - 1]):
                    coord = []
                    x = r
                    for n in reversed(dims):
                        coord.append(x % n)
                        x //= n
                    coord = coord[::-1]
                    peers = O.compute_peers(dims, r, port, family) if len(set(dims)) == 1 else None
                    for i, (dd, k) in enumerate(steps):
                        c = coord[dd]
                        if family == "BINE":
                            dist = (1 - (-2) ** (k + 1)) // 3 * (-1 if c % 2 else 1)
                        else:
                            dist = (c ^ (1 << k)) - c
                        if mirrored(port, nd):
                            dist = -dist
                        want = list(coord)
                        want[dd] = (c + dist) % dims[dd]
                        out = (i32 * nd)()
                        _lib.check(L.bine_get_peer(d, nd, (i32 * nd)(*coord), i, port, fam_code, out), "get_peer")
                        assert list(out) == want, (dims, family, port, r, i)
                        if peers is not None:
                            rid = 0
                            for n, x in zip(dims, want):
                                rid = rid * n + x
                            assert rid == peers[i], (dims, family, port, r, i)
    assert L.bine_get_peer(d, 3, (i32 * 3)(8, 0, 0), 0, 0, 1, (i32 * 3)()) == _lib.BINE_ERR_ARG
    for nd in (1, 2, 3):
        for n in range(0, 7):
            first = next((p for p in range(n) if mirrored(p, nd)), -1)
            assert L.bine_get_mirroring_port(n, nd) == first, (n, nd)


def test_trees_and_remap_match_oracle():
    from oracle import oracle as O
    L = _lib.load()
    for dims in ([8], [16], [4, 4], [4, 4, 4]):
        P = 1
        for x in dims:
            P *= x
        d = (ctypes.c_uint32 * 3)(*(list(dims) + [1] * (3 - len(dims))))
        for root in (0, 1, P - 1):
            for dist in (0, 1):
                arrs = [(ctypes.c_uint32 * P)() for _ in range(5)]
                _lib.check(L.bine_get_tree(d, len(dims), root, 0, 1, dist, *arrs), "get_tree")
                ref = O.get_tree(dims, root, dist=["INCREASING", "DECREASING"][dist])
                for a, key in zip(arrs, ("parent", "reached_at_step", "remapped_ranks", "remapped_ranks_max",
                                         "subtree_roots")):
                    assert list(a) == list(ref[key]), (dims, root, dist, key)
    for P in (2, 4, 8, 16, 64, 512, 6, 10, 12):
        for r in range(min(P, 64)):
            assert L.bine_remap_rank(r, P) == O.remap_rank(r, P)


def test_no_device_is_reported_not_crashed():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    L = _lib.load()
    cfg = _lib.VCommConfig()
    _lib.check(L.bine_vcomm_config_init(ctypes.byref(cfg), 8), "init")
    h = ctypes.c_void_p()
    assert L.bine_vcomm_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.BINE_ERR_NO_DEVICE


def test_remap_rank_export_against_golden_fixture():
    import json
    import os
    from bine_trees_fugaku_amd import _lib
    L = _lib.load()
    with open(os.path.join(os.path.dirname(__file__), "golden", "remap_rank.json")) as f:
        rows = json.load(f)["P_rank_remap"]
    for P, r, want in rows:  # including non-powers of two, where the decimal 80000000 test acts
        assert L.bine_remap_rank(r, P) == want, (P, r)


def test_torus_dimensions_are_one_to_three():
    """LIBBINE_MAX_SUPPORTED_DIMENSIONS is 3 (lib/libbine_common.h:19): a fourth dimension is refused on both
    sides of the boundary, in Python before a config is built and in C by plan_dump."""
    from bine_trees_fugaku_amd import make_config
    with pytest.raises(ValueError):
        make_config([4, 4, 4, 4])
    with pytest.raises(ValueError):
        make_config([])
    cfg = make_config([4, 4, 4])
    cfg.dimensions_num = 4
    info = _lib.PlanDumpInfo()
    L = _lib.load()
    assert L.bine_plan_dump(ctypes.byref(cfg), 0, 256, 2, 0, 0, 0, 1, 0, 0, ctypes.byref(info), None, None,
                            None) == _lib.BINE_ERR_ARG


def test_host_only_entry_points_check_their_arguments():
    """Argument errors of the simulator, tuning and network entry points are
    reported before any device call (so they run without a GPU)."""
    L = _lib.load()
    n = ctypes.c_uint32()
    assert L.bine_sim_classes(None, ctypes.byref(n), None, None, None, None, None, None) == _lib.BINE_ERR_ARG
    c = _lib.SimCase()
    c.dimensions[0], c.dimensions_num, c.coll, c.algo = 8, 1, 0, 0
    assert L.bine_sim_classes(ctypes.byref(c), ctypes.byref(n), None, None, None, None, None, None) == _lib.BINE_OK
    assert n.value == 6  # 3 reduce-scatter + 3 allgather steps
    dest = (ctypes.c_uint32 * 8)()
    assert L.bine_sim_peers(ctypes.byref(c), 6, dest) == _lib.BINE_ERR_ARG  # no class 6
    c.coll, c.algo = 3, 2  # bcast has no ring schedule
    assert L.bine_sim_classes(ctypes.byref(c), ctypes.byref(n), None, None, None, None, None, None) == _lib.BINE_ERR_UNSUPPORTED
    cnt = ctypes.c_int()
    assert L.bine_tuning_candidates(0, None, 0, ctypes.byref(cnt)) == _lib.BINE_OK and cnt.value >= 20
    assert L.bine_tuning_candidates(0, None, 4, ctypes.byref(cnt)) == _lib.BINE_ERR_ARG  # cap without a buffer
    h = ctypes.c_void_p()
    dims = (ctypes.c_int32 * 1)(18)
    assert L.bine_mlp_create(dims, 1, 16, 0, 0, ctypes.byref(h)) == _lib.BINE_ERR_ARG  # a network needs 2+ widths
    dims2 = (ctypes.c_int32 * 2)(18, 0)
    assert L.bine_mlp_create(dims2, 2, 16, 0, 0, ctypes.byref(h)) == _lib.BINE_ERR_ARG
    assert L.bine_mlp_forward(None, None, 1, None, None) == _lib.BINE_ERR_ARG
    assert L.bine_sim_run(None, None) == _lib.BINE_ERR_ARG
    assert L.bine_batch_time_shapes(None, None, 1, None, 1, None) == _lib.BINE_ERR_ARG


C_CALLER = r"""
#include <stdio.h>
#include "bine_hip.h"

int main(void) {
    bine_vcomm_config_t cfg;
    if (bine_vcomm_config_init(&cfg, 512) != BINE_OK) { fprintf(stderr, "%s\n", bine_last_error()); return 1; }
    printf("dims %u %u\n", cfg.dimensions[0], cfg.dimensions_num);
    printf("remap %u\n", bine_remap_rank(5, 8));
    const uint32_t dims[3] = {8, 8, 8};
    uint32_t peers[9];
    if (bine_compute_peers(dims, 3, 0, 0, BINE_FAMILY_BINE, peers) != BINE_OK) return 2;
    printf("peers");
    for (int i = 0; i < 9; ++i) printf(" %u", peers[i]);
    printf("\n");
    return 0;
}
"""


def test_c_caller_compiles_links_and_runs(tmp_path):
    """A plain C11 program (the binding a C caller of the reference's MPI layer would write) compiles against
    include/bine_hip.h without warnings, links libbine_hip.so and gets the reference's values from the host-only
    entry points: remap_rank(5, 8) = 7 (simulation/remap_rank.c) and rank 0's negabinary peers on 8x8x8."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    _lib.load()  # built
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(_lib.LIB_PATH)
    inc = os.path.dirname(HEADER)
    r = subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{inc}", str(src), "-o", str(exe),
                        f"-L{libdir}", "-lbine_hip", f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "dims 512 1"
    assert lines[1] == "remap 7"
    assert lines[2] == "peers 64 8 1 448 56 7 192 24 3"

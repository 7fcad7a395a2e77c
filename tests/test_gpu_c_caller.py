"""The C-ABI end to end from a plain C program on the GPU, with no Python or
PyTorch in the process: hipMalloc'd partitions, bine_allreduce on a 2x2x2
torus, results checked against the exact integer sums (the layout and calls
of the caller binding described for bench/bench.cc:107-395)."""
import os
import shutil
import subprocess

import pytest

from bine_trees_fugaku_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <stdio.h>
#include <stdlib.h>
#include <hip/hip_runtime_api.h>
#include "bine_hip.h"

#define P 8
#define COUNT 10003  /* uneven blocks, partitions off 16-byte boundaries */

int main(void) {
    bine_vcomm_config_t cfg;
    bine_vcomm_t comm;
    if (bine_vcomm_config_init(&cfg, P) != BINE_OK) return 1;
    cfg.dimensions[0] = cfg.dimensions[1] = cfg.dimensions[2] = 2;
    cfg.dimensions_num = 3;
    if (bine_vcomm_create(&cfg, &comm) != BINE_OK) { fprintf(stderr, "%s\n", bine_last_error()); return 2; }
    size_t n = (size_t)P * COUNT;
    int32_t* h = (int32_t*)malloc(n * sizeof(int32_t));
    int32_t* out = (int32_t*)malloc(n * sizeof(int32_t));
    for (size_t i = 0; i < n; ++i) h[i] = (int32_t)((i * 2654435761u) % 2001u) - 1000;
    void *send = NULL, *recv = NULL;
    if (hipMalloc(&send, n * sizeof(int32_t)) != hipSuccess || hipMalloc(&recv, n * sizeof(int32_t)) != hipSuccess)
        return 3;
    if (hipMemcpy(send, h, n * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) return 4;
    int rc = bine_allreduce(send, recv, COUNT, BINE_INT32, BINE_OP_SUM, comm, NULL);
    if (rc != BINE_OK) { fprintf(stderr, "%s\n", bine_last_error()); return 5; }
    if (hipDeviceSynchronize() != hipSuccess) return 6;
    if (hipMemcpy(out, recv, n * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) return 7;
    for (size_t j = 0; j < COUNT; ++j) {
        int32_t want = 0;
        for (int r = 0; r < P; ++r) want += h[(size_t)r * COUNT + j];
        for (int r = 0; r < P; ++r)
            if (out[(size_t)r * COUNT + j] != want) { fprintf(stderr, "mismatch rank %d elem %zu\n", r, j); return 8; }
    }
    if (bine_allreduce(send, recv, COUNT, BINE_INT32, BINE_OP_SUM, comm, (void*)0) != BINE_OK) return 9;
    /* a host pointer is refused before anything is launched */
    if (bine_allreduce(h, recv, COUNT, BINE_INT32, BINE_OP_SUM, comm, NULL) != BINE_ERR_ARG) return 10;
    hipFree(send);
    hipFree(recv);
    bine_vcomm_destroy(comm);
    free(h);
    free(out);
    printf("c caller: allreduce 2x2x2 int32 ok\n");
    return 0;
}
"""


def test_c_program_allreduce(gpu, tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    libdir = os.path.dirname(_lib.LIB_PATH)
    src = tmp_path / "caller.c"
    src.write_text(SRC)
    exe = tmp_path / "caller"
    r = subprocess.run([cc, "-std=c11", "-O1", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.join(ROOT, 'include')}",
                        f"-I{os.path.join(rocm, 'include')}", str(src), "-o", str(exe), f"-L{libdir}", "-lbine_hip",
                        f"-L{os.path.join(rocm, 'lib')}", "-lamdhip64", f"-Wl,-rpath,{libdir}",
                        f"-Wl,-rpath,{os.path.join(rocm, 'lib')}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "c caller: allreduce 2x2x2 int32 ok" in r.stdout

"""Record formats of the benchmark harness (bench/bench.cc:376-387,
bench/generate_metadata.py:18-29) -- host only."""
import csv

import pytest

from bine_trees_fugaku_amd import harness as H


def test_stdout_record_matches_reference_layout():
    out = H.format_record([1.5e-3, 2.25e-3, 1e-9])
    lines = out.splitlines()
    assert lines[0] == "highest"
    assert lines[1:4] == ["1500000", "2250000", "1"]  # (int64_t)(seconds * 1e9)
    assert lines[4] == "Average runtime: %f" % ((1.5e-3 + 2.25e-3 + 1e-9) / 3)
    assert out.endswith("\n")


def test_buffer_counts_follow_allocate_buffers():
    P, c = 64, 10
    assert H.buffer_counts("MPI_Allreduce", c, P) == (c, c)
    assert H.buffer_counts("MPI_Reduce_scatter", c, P) == (c * P, c)
    assert H.buffer_counts("MPI_Allgather", c, P) == (c, c * P)
    assert H.buffer_counts("MPI_Alltoall", c, P) == (c * P, c * P)
    assert H.buffer_counts("MPI_Scatter", c, P) == (c * P, c)
    assert H.buffer_counts("MPI_Gather", c, P) == (c, c * P)
    with pytest.raises(ValueError):
        H.buffer_counts("MPI_Scan", c, P)


def test_metadata_csv_header_and_rows(tmp_path, monkeypatch):
    for k, v in {"TIMESTAMP": "2026_12_02___10_00_00", "N_NODES": "1", "COLLECTIVE_TYPE": "ALLREDUCE",
                 "MPI_LIB": "libbine_hip", "MPI_LIB_VERSION": "7.2", "LIBBINE_VERSION": "1.0", "CUDA": "no",
                 "OUTPUT_LEVEL": "summarized", "MPI_OP": "MPI_SUM"}.items():
        monkeypatch.setenv(k, v)
    monkeypatch.delenv("NOTES", raising=False)
    path = H.write_metadata(str(tmp_path), "mi355x", 3, H.metadata_from_env())
    H.write_metadata(str(tmp_path), "mi355x", 4, H.metadata_from_env())
    with open(path) as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["timestamp", "test_id", "nnodes", "collective_type", "mpi_lib", "mpi_lib_version",
                       "libbine_version", "CUDA", "output_level", "MPI_Op", "notes"]
    assert rows[1] == ["2026_12_02___10_00_00", "3", "1", "ALLREDUCE", "libbine_hip", "7.2", "1.0", "no",
                       "summarized", "MPI_SUM", "null"]
    assert rows[2][1] == "4" and len(rows) == 3


def test_types_map_to_library_types():
    assert H.TYPES["VOID"] == ("int32", "VOID") and H.TYPES["FLOAT"] == ("float32", "SUM")
    assert set(H.COLLECTIVES) == {"MPI_Allreduce", "MPI_Reduce_scatter", "MPI_Allgather", "MPI_Bcast",
                                  "MPI_Alltoall", "MPI_Scatter", "MPI_Gather", "MPI_Reduce"}

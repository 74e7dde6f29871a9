"""The C-ABI library loads and exports every symbol the header declares
(no compute calls: runs without a GPU)."""
import ctypes
import subprocess

import pytest

from parafem_amd import api


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = api.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_status_strings_and_version():
    L = api.lib()
    assert L.pf_version() >= 100
    for s in range(0, 9):
        assert L.pf_status_string(s)


def test_reduce_grid_contract():
    # G = clamp(ceil(n/256), 1, 16384): fixes the dot-product summation order
    for n, g in ((0, 1), (1, 1), (256, 1), (257, 2), (524288, 2048), (4194304, 16384),
                 (10 ** 8, 16384)):
        assert api.reduce_grid(n) == g


def test_bad_arguments_are_rejected_without_a_device():
    L = api.lib()
    assert L.pf_create(0, None) == api.PF_ERR_ARG
    assert L.pf_destroy(None) == api.PF_OK
    assert L.pf_sync(None) == api.PF_ERR_ARG
    assert L.pf_mesh_set(None, 20, 1, None) == api.PF_ERR_ARG


def test_no_cpu_fallback_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(api, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(api, "_lib", None)
    with pytest.raises(api.LibraryNotBuilt):
        api.Context(0)


def test_bench_byte_counts_match_the_storage_layouts():
    """bench.py's algorithmic bytes per element: FP64 K_e in full, the packed
    symmetric FP32 layout's ntot(ntot+1)/2 entries padded to a multiple of 4
    (1832 for hex20, 300 for hex8 elastic, 36 for hex8 Laplace), twice for
    the split precision, plus the int32 steering."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    assert bench.ebe_bytes(60, "fp64") == 60 * 60 * 8 + 240
    assert bench.ebe_bytes(60, "fp32", "mixed") == 1832 * 4 + 240
    assert bench.ebe_bytes(60, "fp32", "split") == 2 * 1832 * 4 + 240
    assert bench.ebe_bytes(24, "fp32", "mixed") == 300 * 4 + 96
    assert bench.ebe_bytes(8, "fp32", "mixed") == 36 * 4 + 32
    assert bench.iteration_bytes(10, 8, 7, "fp32") == 10 * (36 * 4 + 32) + 15 * 7 * 8

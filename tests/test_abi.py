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

"""Post-solve outputs on the CPU: the oracle's SCATTER_NODES
(gather_scatter.f90:1786-1929) pinned by counting, exactness and
conservation, and the EnSight Gold displacement file p121 writes
(p121.f90:125-137, DISMSH_ENSI_P output.f90:2983-3111) with Fortran's
E12.d edit descriptor checked against the Fortran compiler's own output."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from parafem_amd import mesh as M

FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def _mesh(n=3):
    m = M.box_hex20(n, n, n)
    return m, m.g_num.astype(np.int32)


def test_scatter_nodes_counts_elements_per_node():
    m, g = _mesh(3)
    cnt = O.scatter_nodes(g, np.ones(g.shape + (1,)), m.nn, flag=0)[:, 0]
    brute = np.bincount(g.ravel() - 1, minlength=m.nn)
    assert np.array_equal(cnt, brute.astype(np.float64))
    # hex20 bricks: corner nodes shared by up to 8 elements, mid-side by 4
    assert cnt.max() == 8 and cnt.min() == 1


def test_scatter_nodes_average_of_equal_values():
    """A field gathered from nodal values averages back to those values:
    exactly where a node has one or two elements (v + v = 2v and 2v / 2 are
    exact), within a few units in the last place elsewhere."""
    m, g = _mesh(3)
    rng = np.random.default_rng(1)
    nodal = rng.normal(size=(m.nn, 3))
    contrib = nodal[g - 1]                                  # (nels, nod, 3)
    avg = O.scatter_nodes(g, contrib, m.nn, flag=1)
    cnt = np.bincount(g.ravel() - 1, minlength=m.nn)
    few = cnt <= 2
    assert few.any() and np.array_equal(avg[few], nodal[few])
    assert np.all(np.abs(avg - nodal) <= 4 * np.spacing(np.abs(nodal)))


def test_scatter_nodes_sums_are_conserved_and_ordered():
    m, g = _mesh(2)
    rng = np.random.default_rng(2)
    contrib = rng.normal(size=g.shape + (6,))
    tot = O.scatter_nodes(g, contrib, m.nn, flag=0)
    assert np.allclose(tot.sum(axis=0), contrib.reshape(-1, 6).sum(axis=0), rtol=0, atol=1e-12)
    # the sum for one node in ascending element, then element-node, order
    node = g[0, 2]
    terms = [contrib[e, j] for e in range(g.shape[0]) for j in range(g.shape[1]) if g[e, j] == node]
    acc = np.zeros(6)
    for t in terms:
        acc = acc + t
    assert np.array_equal(tot[node - 1], acc)


def test_scatter_nodes_skips_foreign_nodes_and_marks_unused():
    g = np.array([[1, 2, 5], [2, 3, 0]], dtype=np.int32)     # node 5 > nn, 0 < 1: skipped
    c = np.arange(6, dtype=np.float64).reshape(2, 3, 1) + 1.0
    s = O.scatter_nodes(g, c, 4, flag=0)[:, 0]
    assert np.array_equal(s, [1.0, 2.0 + 4.0, 5.0, 0.0])
    a = O.scatter_nodes(g, c, 4, flag=1)[:, 0]
    assert np.array_equal(a[:3], [1.0, 3.0, 5.0]) and np.isnan(a[3])   # 0/0 as the reference


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler")
def test_fortran_e_matches_the_fortran_runtime(tmp_path):
    src = tmp_path / "fe.f90"
    src.write_text(
        "program fe\n  implicit none\n  integer, parameter :: dp = selected_real_kind(15, 300)\n"
        "  real(dp) :: v\n  integer :: ios\n  do\n    read(*, *, iostat=ios) v\n"
        "    if (ios /= 0) exit\n    write(*, '(e12.5)') v\n    write(*, '(e12.4)') v\n  end do\n"
        "end program\n")
    exe = tmp_path / "fe"
    r = subprocess.run([FLANG, "-O0", str(src), "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(3)
    vals = list(rng.normal(size=400) * 10.0 ** rng.integers(-30, 30, 400))
    vals += [0.0, -0.0, 1.0, -1.0, 0.9999996, 0.99999951, 9.99995e-5, 1e-99, 1e-100, -1e150, 123455.0,
             0.123455, 0.123465, 2.5e-5, 0.015180]
    vals += list(np.round(rng.uniform(-1, 1, 200), 6))            # decimal ties after rounding
    inp = "".join(f"{float(v)!r}\n" for v in vals)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, timeout=60).stdout.splitlines()
    assert len(out) == 2 * len(vals)
    for i, v in enumerate(vals):
        assert M.fortran_e(v, 12, 5) == out[2 * i], (v, out[2 * i])
        assert M.fortran_e(v, 12, 4) == out[2 * i + 1], (v, out[2 * i + 1])


def test_ensi_vector_file_roundtrip(tmp_path):
    rng = np.random.default_rng(4)
    a = rng.normal(size=(50, 3)) * 1e-2
    a[3] = 0.0
    path = tmp_path / "x.ensi.DISPL-000001"
    M.write_ensi_vector(str(path), a)
    lines = path.read_text().split("\n")
    assert lines[:4] == list(M.ENSI_VECTOR_HEADER)
    assert len(lines) == 4 + 150 + 1 and all(len(t) == 12 for t in lines[4:-1])
    b = M.read_ensi_vector(str(path))
    assert b.shape == a.shape
    assert np.all(np.abs(b - a) <= 0.5e-5 * 10.0 ** np.ceil(np.log10(np.maximum(np.abs(a), 1e-300))) + 1e-300)

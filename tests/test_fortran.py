"""The Fortran binding (parafem_amd/fortran/pf_gpu.f90) and the p121 driver
built on it (parafem_amd/bin/pf_p121).

The driver reads a p121 deck and solves it either with p121's own loop
statement for statement (GATHER / element products / SCATTER /
DOT_PRODUCT_P / checon_par and the whole-array statements on device vectors,
p121.f90:86-104), or with the fused library PCG.  Bars: the two give
bit-identical displacements and iteration counts, equal to the oracle in the
device's reduction order; on xx3-tiny the iteration count and the printed
job data equal the reference's own results file (xx3-tiny.res).
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

from parafem_amd import mesh as M, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parafem_amd", "bin", "pf_p121")
EXE124 = os.path.join(ROOT, "parafem_amd", "bin", "pf_p124")
EXE122 = os.path.join(ROOT, "parafem_amd", "bin", "pf_p122")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="pf_p121 not built (no amdflang)")


def say(msg):
    print(f"[fortran-test] {msg}", flush=True)
    sys.stdout.flush()


def run(deck, mode, *extra):
    r = subprocess.run([EXE, os.path.basename(deck), mode, *extra], cwd=os.path.dirname(deck),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, f"pf_p121 {mode} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    res = open(deck + ".pf.res").read()
    x = np.fromfile(deck + ".pf.x", dtype=np.float64)
    iters = int(re.search(r"convergence was\s+(\d+)", res).group(1))
    return x, iters, res


@needs_exe
def test_driver_usage_without_arguments():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout


@needs_exe
def test_driver_rejects_unknown_mode(tmp_path):
    r = subprocess.run([EXE, "deck", "bogus"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown mode" in r.stdout


@needs_exe
@pytest.mark.gpu
def test_fortran_dropin_equals_fused_xx3_tiny(tmp_path, golden):
    import oracle as O
    src = os.path.join(golden, "xx3_tiny", "xx3-tiny")
    for ext in ("dat", "d", "bnd", "lds"):
        shutil.copy(f"{src}.{ext}", tmp_path)
    deck = str(tmp_path / "xx3-tiny")
    say("pf_p121 dropin xx3-tiny")
    xa, ia, res = run(deck, "dropin")
    say("pf_p121 fused xx3-tiny")
    xb, ib, _ = run(deck, "fused")
    assert ia == ib == 79
    assert np.array_equal(xa, xb)
    ref = open(src + ".res").read()
    neq = int(re.search(r"equations solved\s+(\d+)", ref).group(1))
    assert re.search(rf"restrained and\s+{neq} equations", res)
    assert "The total load is: -0.1000E+03" in res
    d = M.read_deck(src)
    nf, neq2 = M.steering(d["nn"], 3, d["rest"])
    g_g = M.g_g_from(nf, d["g_num"])
    gc = M.element_coords(d["g_num"], d["coords"])
    r = M.nodal_loads_to_rhs(nf, d["load_node"], d["load_val"], neq2)
    Ko = O.assemble_elastic(gc, d["e"], d["v"])
    xo, it, conv = O.pcg(g_g, Ko, r, O.diag(g_g, Ko, neq2), d["tol"], d["limit"],
                         dot_order=O.DOT_DEVICE)
    assert neq2 == neq == xa.size and it == ia
    assert np.array_equal(xa, xo)
    # post-solve outputs (p121.f90:112-139): element 1's centroid stresses
    # as printed, and the EnSight displacement file line for line
    sig = O.centroid_stress(gc, g_g, xo, d["e"], d["v"])[0]
    assert "The Centroid point stresses for element 1 are\nPoint     1\n" in res
    assert "".join(M.fortran_e(s_, 12, 4) for s_ in sig) in res
    assert re.search(r"This analysis took  :\s+[0-9.]+", res)
    disp = O.scatter_nodes(d["g_num"], O.gather(g_g, xo).reshape(d["nels"], d["nod"], 3), d["nn"], 1)
    lines = open(deck + ".pf.ensi.DISPL-000001").read().split("\n")
    assert lines[:4] == list(M.ENSI_VECTOR_HEADER)
    assert lines[4:-1] == [M.fortran_e(v, 12, 5) for v in disp.T.ravel()]


@needs_exe
@pytest.mark.gpu
def test_fortran_host_arrays_equal_fused_xx3_tiny(tmp_path, golden):
    """p121.f90:89-104 as the reference writes it, on host arrays, against
    the module's GATHER / SCATTER / DOT_PRODUCT_P / checon_par generics on
    REAL(iwp) arrays (host mode): the fused device solve's bits and
    iteration count.  With the reference's sequential dot order (hostseq,
    fusedseq): bit-identical to each other and to the oracle in that
    order."""
    import oracle as O
    src = os.path.join(golden, "xx3_tiny", "xx3-tiny")
    for ext in ("dat", "d", "bnd", "lds"):
        shutil.copy(f"{src}.{ext}", tmp_path)
    deck = str(tmp_path / "xx3-tiny")
    say("pf_p121 host xx3-tiny")
    xh, ih, _ = run(deck, "host")
    say("pf_p121 fused xx3-tiny")
    xf, if_, _ = run(deck, "fused")
    assert ih == if_ == 79
    assert np.array_equal(xh, xf)
    say("pf_p121 hostseq / fusedseq xx3-tiny")
    xhs, ihs, _ = run(deck, "hostseq")
    xfs, ifs, _ = run(deck, "fusedseq")
    assert ihs == ifs and np.array_equal(xhs, xfs)
    d = M.read_deck(src)
    nf, neq = M.steering(d["nn"], 3, d["rest"])
    g_g = M.g_g_from(nf, d["g_num"])
    gc = M.element_coords(d["g_num"], d["coords"])
    r = M.nodal_loads_to_rhs(nf, d["load_node"], d["load_val"], neq)
    Ko = O.assemble_elastic(gc, d["e"], d["v"])
    xo, it, conv = O.pcg(g_g, Ko, r, O.diag(g_g, Ko, neq), d["tol"], d["limit"], dot_order=O.DOT_SEQ)
    assert conv and ihs == it == 79
    assert np.array_equal(xhs, xo)


@needs_exe
@pytest.mark.gpu
def test_fortran_generated_abaqus_deck_and_mixed(tmp_path):
    """A 6^3 hex20 cube deck in Abaqus node order (ABAQUS2SG applied on
    read): dropin and fused bit-identical; the mixed-precision solve within
    1e-8 of the FP64 solution; PCG_KM bit-identical to the oracle."""
    pr = P.elastic_cube(6)
    rest = M.base_fixed_rest(pr.mesh)
    nodes, vals = M.top_pressure_loads(pr.mesh, -1.0)
    deck = str(tmp_path / "cube6")
    M.write_p121_deck(deck, pr.mesh, rest, nodes, vals, 100.0, 0.3, 1e-9, 2000)
    say("pf_p121 dropin cube6")
    xa, ia, _ = run(deck, "dropin")
    say("pf_p121 fused cube6")
    xb, ib, _ = run(deck, "fused")
    assert ia == ib and np.array_equal(xa, xb)
    say("pf_p121 mixed cube6")
    xm, im, _ = run(deck, "mixed", "1e-11")
    assert np.abs(xm - xa).max() <= 1e-8 * np.abs(xa).max()
    # PCG_KM with element 1's matrix for all elements and the driver's
    # inverted assembled diagonal: bit-identical to the oracle doing the same
    import oracle as O
    say("pf_p121 km cube6")
    xk, ik, _ = run(deck, "km")
    d = M.read_deck(deck)
    nf, neq = M.steering(d["nn"], 3, d["rest"])
    g_g = M.g_g_from(nf, d["g_num"])
    gc = M.element_coords(d["g_num"], d["coords"])
    r = M.nodal_loads_to_rhs(nf, d["load_node"], d["load_val"], neq)
    Ko = O.assemble_elastic(gc, d["e"], d["v"])
    full = np.repeat(Ko[:1], Ko.shape[0], axis=0)
    xo, it, _ = O.pcg(g_g, full, r, O.diag(g_g, Ko, neq), d["tol"], d["limit"], dot_order=O.DOT_DEVICE)
    assert ik == it and np.array_equal(xk, xo)
    assert np.abs(xk - xa).max() <= 1e-6 * np.abs(xa).max()


@needs_exe
@pytest.mark.gpu
def test_fortran_p124_demo(tmp_path, golden):
    """pf_p124 on the p124 demo deck (regenerated exactly from its
    generator): the temperature history prints as p124_demo.res prints it
    (same digits, same iteration counts); the last temperatures are the
    oracle's bit for bit (device dot order) and every NDTTR file equals the
    oracle's nodal averages line for line; the fields agree with the
    demo's printed ones."""
    sys.path.insert(0, os.path.dirname(__file__))
    import oracle as O
    from test_p124 import _demo_deck, oracle_p124
    meta, base = _demo_deck(golden, tmp_path)
    say("pf_p124 demo deck")
    r = subprocess.run([EXE124, os.path.basename(base)], cwd=os.path.dirname(base), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, f"pf_p124 failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    res = open(base + ".pf.res").read().splitlines()
    ref = open(os.path.join(golden, "p124_demo.json")).read()
    hist_lines = [ln for ln in res if re.match(r"^\s+\d\.\d{4}E[+-]\d\d\s+\S+", ln)]
    pub = __import__("json").loads(ref)["history"]
    assert len(hist_lines) == len(pub)
    for ln, (t, v, it) in zip(hist_lines, pub):
        expect = M.fortran_e(t, 12, 4) + M.fortran_e(v, 12, 4) + (f"{it:10d}" if it is not None else "")
        assert ln == expect, (ln, expect)
    assert re.search(r"There are\s+17576 nodes\s+1951 restrained and\s+15625 equations", res[1])
    d = M.read_deck(base, "p124")
    keep = tuple(__import__("json").loads(ref)["ttr_steps"])
    hist, fields, xo, (nf, neq, g_g) = oracle_p124(d, O.DOT_DEVICE, keep)
    x = np.fromfile(base + ".pf.x", dtype=np.float64)
    assert np.array_equal(x, xo)
    ttr = np.load(os.path.join(golden, "p124_demo_ttr.npz"))
    for step in keep:
        eld = O.gather(g_g, fields[step]).reshape(d["nels"], 8, 1)
        nodal = O.scatter_nodes(d["g_num"], eld, d["nn"], 1)[:, 0]
        lines = open(f"{base}.pf.ensi.NDTTR-{step:06d}").read().split("\n")
        assert lines[:4] == ["Alya Ensight Gold --- Scalar per-node variable file", "part", "    1", "coordinates"]
        assert lines[4:-1] == [M.fortran_e(v, 12, 5) for v in nodal], step
        refv = ttr[f"step{step}"].astype(np.float64)
        unit = 10.0 ** (np.floor(np.log10(np.maximum(np.abs(refv), 1e-30))) - 4)
        assert np.all(np.abs(nodal - refv) <= unit + 1e-4 * np.abs(refv).max()), step


@needs_exe
@pytest.mark.gpu
def test_fortran_p122_demo(tmp_path, golden):
    """pf_p122 on the p122 demo deck (1152 bricks, 19 prescribed
    displacements, 10 increments): the critical time step, per increment the
    displacement and the three stresses print as p122_demo.res prints them,
    with the same CG and plastic iteration counts; the total displacements
    agree with the oracle's run in the device's dot order to 1e-9."""
    sys.path.insert(0, os.path.dirname(__file__))
    import json
    import oracle as O
    from test_p122 import demo, oracle_p122
    meta, base = demo(golden, tmp_path)
    say("pf_p122 demo deck")
    r = subprocess.run([EXE122, os.path.basename(base)], cwd=os.path.dirname(base), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, f"pf_p122 failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    res = open(base + ".pf.res").read().splitlines()
    assert f"The critical timestep is   {M.fortran_e(meta['res_dt'], 12, 4)}" in res
    incs = []
    for k, line in enumerate(res):
        if line.startswith("Load Increment"):
            incs.append((res[k + 1], res[k + 3], int(res[k + 4].split()[-1]), int(res[k + 5].split()[-1])))
    pub = meta["increments"]
    assert len(incs) == len(pub)
    for (ld, ls, cj, pl), inc in zip(incs, pub):
        assert ld == "The displacement is  " + M.fortran_e(inc["displacement"], 12, 4), ld
        assert ls == "".join(M.fortran_e(v, 12, 4) for v in inc["sigma_zxy"]), ls
        assert (cj, pl) == (inc["cj_iterations"], inc["plastic_iterations"]), (inc["increment"], cj, pl)
    d = M.read_deck(base, "p122")
    _, totd, _ = oracle_p122(d, dot=lambda a, b: O.dot_device(a, b))
    x = np.fromfile(base + ".pf.x", dtype=np.float64)
    assert np.abs(x - totd).max() <= 1e-9 * np.abs(totd).max()
    lines = open(base + ".pf.ensi.DISPL-000010").read().split("\n")
    assert lines[:4] == ["Alya Ensight Gold --- Vector per-node variable file", "part", "    1", "coordinates"]


EXE123 = os.path.join(ROOT, "parafem_amd", "bin", "pf_p123")


def _p123_deck(tmp_path):
    """A p123 deck as p12meshgen's p123 case lays it out (GEOMETRY_8BXZ
    mesh, BOX_BC8 restraints, the load 10 on equation nres) on a 24 x 16 x
    20 box, plus one fixed freedom to exercise the penalty rows."""
    pr = P.p123_book(24, 16, 20, 0.005, 2.0)
    nres = pr.info["nres"]
    node = int(np.nonzero(pr.nf[:, 0] == pr.neq // 3)[0][0]) + 1
    base = str(tmp_path / "p123_box")
    M.write_p123_deck(base, pr.mesh, pr.info["rest"], 2.0, 2.0, 2.0, 1e-5, 2000, nres,
                      loads=[(nres, 10.0)], fix=[(node, 1, 5.0)])
    return pr, base, node


@needs_exe
@pytest.mark.gpu
def test_fortran_p123_deck(tmp_path):
    """pf_p123 on a p123 deck: job data, iteration count, total load and the
    four potentials print as p123 prints them, xnew equals the oracle's in
    the device's dot order bit for bit, the NDPTL file holds the oracle's
    nodal averages; the mixed mode reaches relres 1e-10 within 1e-8 of the
    FP64 solution."""
    sys.path.insert(0, os.path.dirname(__file__))
    import oracle as O
    pr, base, node = _p123_deck(tmp_path)
    say(f"pf_p123 deck neq={pr.neq}")
    r = subprocess.run([EXE123, os.path.basename(base)], cwd=os.path.dirname(base), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, f"pf_p123 failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    res = open(base + ".pf.res").read().splitlines()
    d = M.read_deck(base, "p123")
    nf, neq = M.steering(d["nn"], 1, d["rest"])
    g_g = M.g_g_from(nf, d["g_num"])
    assert neq == pr.neq and np.array_equal(g_g, pr.g_g)
    assert res[1] == f"There are {d['nn']:12d} nodes{d['nr']:12d} restrained and   {neq:12d} equations"
    K = O.assemble_laplace(M.element_coords(d["g_num"], d["coords"]), d["kx"], d["ky"], d["kz"])
    dg = O.diag(g_g, K, neq)
    no_f = nf[d["fix_node"] - 1, 0]
    store = dg[no_f - 1] + 1.0e20
    dg[no_f - 1] = store
    rhs = np.zeros(neq)
    rhs[d["load_node"] - 1] = d["load_val"][:, 0]
    q = rhs.sum()
    rhs[no_f - 1] = store * d["fix_val"]
    xo, it, conv = O.pcg(g_g, K, rhs, dg, d["tol"], d["limit"], no_f=no_f, store=store,
                         dot_order=O.DOT_DEVICE)
    x = np.fromfile(base + ".pf.x", dtype=np.float64)
    assert conv and np.array_equal(x, xo)
    assert res[3] == f"The number of iterations to convergence was {it:5d}"
    assert res[4] == "The total load is " + M.fortran_e(q, 12, 4)
    nres = d["nres"]
    assert res[7:11] == [f"{nres + i:8d}     " + M.fortran_e(xo[nres + i - 1], 12, 4) for i in range(4)]
    lines = open(base + ".pf.ensi.NDPTL-000001").read().split("\n")
    assert lines[:4] == ["Alya Ensight Gold --- Scalar per-node variable file", "part", "    1", "coordinates"]
    nodal = O.scatter_nodes(d["g_num"], O.gather(g_g, xo).reshape(d["nels"], 8, 1), d["nn"], 1)[:, 0]
    assert lines[4:-1] == [M.fortran_e(v, 12, 5) for v in nodal]
    r = subprocess.run([EXE123, os.path.basename(base), "mixed", "1e-10"], cwd=os.path.dirname(base),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    xm = np.fromfile(base + ".pf.x", dtype=np.float64)
    x64, _, conv64 = O.pcg(g_g, K, rhs, dg, 1e-12, 20000, stop=O.STOP_RELRES, no_f=no_f, store=store)
    assert conv64 and np.abs(xm - x64).max() <= 1e-8 * np.abs(x64).max()


@needs_exe
@pytest.mark.gpu
def test_fortran_p123_deck_without_restraints(tmp_path):
    """pf_p123 with nr = 0 (g_g_pp = g_num_pp, p123.f90:53-55) and the
    temperature held by penalty fixed freedoms on x = 0, unit flux on x = max
    (the conduction cube of the benchmark, uniform k): xnew equals the
    oracle's in the device's dot order bit for bit."""
    sys.path.insert(0, os.path.dirname(__file__))
    import oracle as O
    pr = P.laplace_cube(6, lognormal_k=False, nels_xyz=(6, 5, 4))
    m = pr.mesh
    fix_nodes = np.nonzero(np.abs(m.coords[:, 0]) < 1e-12)[0] + 1
    base = str(tmp_path / "p123_free")
    loads = [(int(q) + 1, float(v)) for q, v in enumerate(pr.r0) if v != 0.0]
    M.write_p123_deck(base, m, np.zeros((0, 2), dtype=np.int32), 1.0, 1.0, 1.0, 1e-6, 2000, 1,
                      loads=loads, fix=[(int(n), 1, 0.0) for n in fix_nodes])
    say(f"pf_p123 nr=0 neq={pr.neq}")
    r = subprocess.run([EXE123, os.path.basename(base)], cwd=os.path.dirname(base), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, f"pf_p123 failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    d = M.read_deck(base, "p123")
    nf, neq = M.steering(d["nn"], 1, None)
    g_g = M.g_g_from(nf, d["g_num"])
    assert d["nr"] == 0 and neq == d["nn"] and np.array_equal(g_g, d["g_num"])
    K = O.assemble_laplace(M.element_coords(d["g_num"], d["coords"]), 1.0, 1.0, 1.0)
    dg = O.diag(g_g, K, neq)
    no_f = d["fix_node"]
    store = dg[no_f - 1] + 1.0e20
    dg[no_f - 1] = store
    rhs = np.zeros(neq)
    rhs[d["load_node"] - 1] = d["load_val"][:, 0]
    rhs[no_f - 1] = store * d["fix_val"]
    xo, it, conv = O.pcg(g_g, K, rhs, dg, d["tol"], d["limit"], no_f=no_f, store=store,
                         dot_order=O.DOT_DEVICE)
    x = np.fromfile(base + ".pf.x", dtype=np.float64)
    assert conv and np.array_equal(x, xo)
    res = open(base + ".pf.res").read().splitlines()
    assert res[3] == f"The number of iterations to convergence was {it:5d}"

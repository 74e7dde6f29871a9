"""p124 (transient heat conduction, 8-node bricks, theta-method) on the CPU:
the oracle's element matrices pinned by closed forms and against the p123
element, and the p124 deck reader on the reference's demo deck."""
import json
import os

import numpy as np
import pytest

import oracle as O
from parafem_amd import mesh as M


def _box_coord(a=0.3, b=0.5, c=0.7):
    m = M.box_hex8(1, 1, 1, a, b, c)
    return m.g_coord()[0].T.copy()          # (8, 3)


def test_shape_fun_partition_of_unity_and_delta():
    pts, _ = O.sample_hex8()
    for i in range(1, 9):
        assert abs(O.shape_fun_hex8(pts, i).sum() - 1.0) < 1e-15
    # at the corners (+-1) each node's function is 1 at its own corner
    corners = np.array([[-1, -1, -1], [-1, -1, 1], [1, -1, 1], [1, -1, -1],
                        [-1, 1, -1], [-1, 1, 1], [1, 1, 1], [1, 1, -1]], dtype=float)
    for i in range(1, 9):
        assert np.array_equal(O.shape_fun_hex8(corners, i), np.eye(8)[i - 1])


def test_transient_conductivity_is_the_p123_element():
    """rho = 0, theta = dtim = 1: storka is kc, the same matrix p123 builds
    (another summation order: equal to rounding)."""
    rng = np.random.default_rng(1)
    coord = _box_coord() + rng.uniform(-0.05, 0.05, (8, 3))
    kay = np.array([1.3, 0.7, 2.1])
    ka, kb = O.transient_km(coord, kay, 0.0, 1.0, 1.0, 1.0)
    kc = O.laplace_kc(coord, *kay)
    assert np.allclose(ka, kc, rtol=0, atol=1e-14 * np.abs(kc).max())
    assert np.array_equal(kb, np.zeros((8, 8)) - ka * 0.0)          # pm - kc*0*1 = 0


def test_transient_capacity_closed_form():
    """kay = 0: storka = pm, the consistent capacity matrix of a brick,
    rho*cp*V/216 * 2^(3 - differing coordinates); its entries sum to
    rho*cp*V (partition of unity)."""
    a, b, c, rho, cp = 0.3, 0.5, 0.7, 2.0, 3.0
    coord = _box_coord(a, b, c)
    ka, kb = O.transient_km(coord, np.zeros(3), rho, cp, 0.5, 0.01)
    assert np.array_equal(ka, kb)
    V = a * b * c
    assert abs(ka.sum() - rho * cp * V) < 1e-14
    s = np.array([[-1, -1, -1], [-1, -1, 1], [1, -1, 1], [1, -1, -1],
                  [-1, 1, -1], [-1, 1, 1], [1, 1, 1], [1, 1, -1]])
    diff = (s[:, None, :] != s[None, :, :]).sum(axis=2)
    exact = rho * cp * V / 216.0 * 2.0 ** (3 - diff)
    assert np.allclose(ka, exact, rtol=1e-14, atol=0)


def test_transient_theta_combination():
    rng = np.random.default_rng(2)
    coord = _box_coord() + rng.uniform(-0.05, 0.05, (8, 3))
    kay, rho, cp, theta, dtim = np.array([1.0, 2.0, 0.5]), 1.5, 0.8, 0.5, 0.01
    ka, kb = O.transient_km(coord, kay, rho, cp, theta, dtim)
    pm, _ = O.transient_km(coord, np.zeros(3), rho, cp, theta, dtim)
    kc, _ = O.transient_km(coord, kay, 0.0, 1.0, 1.0, 1.0)
    assert np.allclose(ka, pm + kc * theta * dtim, rtol=1e-14, atol=1e-16)
    assert np.allclose(kb, pm - kc * (1 - theta) * dtim, rtol=1e-14, atol=1e-16)
    # symmetric, and storka + storkb = 2 pm + (2 theta - 1) kc dtim
    assert np.array_equal(ka, ka.T) or np.allclose(ka, ka.T, rtol=0, atol=1e-17)
    g = M.box_hex8(2, 1, 1).g_coord()
    KA, KB = O.assemble_transient(g, [[*kay, rho, cp]], [1, 1], theta, dtim)
    k0a, k0b = O.transient_km(g[0].T, kay, rho, cp, theta, dtim)
    assert np.array_equal(KA[0].T, k0a) and np.array_equal(KB[0].T, k0b)


def _demo_deck(golden, tmp_path):
    meta = json.load(open(os.path.join(golden, "p124_demo.json")))
    g, c = meta["mg"], meta["control"]
    m = M.box_hex8(g["nxe"], g["nye"], g["nze"], g["aa"], g["bb"], g["cc"])
    rest = M.p124_box_rest(m)
    base = str(tmp_path / "p124_demo")
    M.write_p124_deck(base, m, rest, meta["prop"], np.ones(m.nels, dtype=np.int32), c["val0"], c["dtim"],
                      c["nstep"], c["npri"], c["theta"], c["tol"], c["limit"], c["nres"])
    return meta, base


def test_p124_demo_deck_regenerated_exactly(golden, tmp_path):
    """The demo deck written from the box generator and the three-face
    restraint reads back to the same arrays (digest) and control data as
    the reference's p124_demo deck."""
    import hashlib
    meta, base = _demo_deck(golden, tmp_path)
    d = M.read_deck(base, "p124")
    h = hashlib.sha256()
    for k in ("g_num", "coords", "rest", "etype"):
        a = np.ascontiguousarray(d[k])
        if a.dtype.kind == "f":
            a = a + 0.0
        h.update(k.encode())
        h.update(str(a.dtype).encode())
        h.update(a.tobytes())
    assert h.hexdigest() == meta["deck_sha256"]
    for k, v in meta["control"].items():
        assert d[k] == v, k
    assert np.array_equal(d["prop"], np.array(meta["prop"]))
    nf, neq = M.steering(d["nn"], 1, d["rest"])
    assert neq == meta["res_neq"]


PENALTY = 1e20


def oracle_p124(d, dot_order=O.DOT_DEVICE, keep=()):
    """p124's time loop (p124.f90:126-228) on the oracle: returns the
    printed history [(time, x(nres), iters)], the fields x at the steps in
    `keep` and the last step's x."""
    nf, neq = M.steering(d["nn"], 1, d["rest"])
    g_g = M.g_g_from(nf, d["g_num"])
    gc = M.element_coords(d["g_num"], d["coords"])
    KA, KB = O.assemble_transient(gc, d["prop"], d["etype"], d["theta"], d["dtim"], nthreads=8)
    dg = O.diag(g_g, KA, neq)
    nfix = d["fixed_freedoms"]
    no_f = nf[d["fix_node"] - 1, 0] if nfix else np.zeros(0, dtype=np.int32)
    val_f = d["fix_val"] if nfix else np.zeros(0)
    store = dg[no_f - 1] + PENALTY
    dg[no_f - 1] = store
    x = np.full(neq, d["val0"])
    x[no_f - 1] = val_f
    nres = d["nres"]
    hist = [(0.0, x[nres - 1], None)]
    fields = {0: x.copy()} if 0 in keep else {}
    xnew = np.zeros(neq)
    for j in range(1, d["nstep"] + 1):
        loads = np.zeros(neq)
        if d["loaded_nodes"]:
            loads[d["load_node"] - 1] = d["load_val"][:, 0] * d["dtim"]
        if j != 1:
            u = O.apply(g_g, KB, xnew, nthreads=8)
            u[no_f - 1] = store * val_f
        else:
            u = O.apply(g_g, KA, x, nthreads=8)
        loads = loads + u
        r = np.zeros(neq)
        r[no_f - 1] = store * val_f
        r = loads - r
        xnew, it, conv = O.pcg(g_g, KA, r, dg, d["tol"], d["limit"], stop=O.STOP_CHECON,
                               no_f=no_f if nfix else None, store=store if nfix else None,
                               nthreads=8, dot_order=dot_order)
        if j % d["npri"] == 0:
            hist.append((j * d["dtim"], xnew[nres - 1], it))
            if j in keep:
                fields[j] = xnew.copy()
    return hist, fields, xnew, (nf, neq, g_g)


def test_p124_demo_history_matches_the_reference(golden, tmp_path):
    """The oracle's p124 run of the demo deck against p124_demo.res and the
    NDTTR fields (printed E12.4 / e12.5; the published run used two
    processes, whose dot products sum in another order): every printed
    temperature within half a printed unit (so it prints the same), the
    same PCG iteration count at every printed step; fields within a printed
    unit."""
    meta, base = _demo_deck(golden, tmp_path)
    d = M.read_deck(base, "p124")
    keep = tuple(meta["ttr_steps"])
    hist, fields, _, (nf, neq, g_g) = oracle_p124(d, O.DOT_SEQ, keep)
    pub = meta["history"]
    assert len(hist) == len(pub)
    for (t, v, it), (tp, vp, itp) in zip(hist, pub):
        assert abs(t - tp) < 1e-12
        unit = 10.0 ** (np.floor(np.log10(abs(vp))) - 3)
        assert abs(v - vp) <= 0.5 * unit * 1.001, (t, v, vp)
        if itp is not None:
            assert it == itp, (t, it, itp)
    ttr = np.load(os.path.join(golden, "p124_demo_ttr.npz"))
    for step in keep:
        eld = O.gather(g_g, fields[step]).reshape(d["nels"], 8, 1)
        nodal = O.scatter_nodes(d["g_num"], eld, d["nn"], 1)[:, 0]
        ref = ttr[f"step{step}"].astype(np.float64)
        unit = 10.0 ** (np.floor(np.log10(np.maximum(np.abs(ref), 1e-30))) - 4)
        assert np.all(np.abs(nodal - ref) <= 1.0 * unit + 1e-4 * np.abs(ref).max()), step

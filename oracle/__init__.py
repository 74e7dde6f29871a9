"""CPU oracle for the EBE-PCG path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``liboracle.so`` (built from ``pf_oracle.c`` by
``oracle/Makefile``), a plain-C FP64 restatement of ParaFEM's p121/p123
element assembly and PCG.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg use it, and only as the checker; the product
(``parafem_amd``) never imports it.

Every function below names the reference routine it restates; the C file
cites the reference line ranges.

Pinning status (how each function is checked in ``tests/test_oracle.py``):
  sample_hex8, deemat6, shape_der, beemat6 -- bit for bit against the
      reference's own new_library.f90 compiled in oracle/_ref
      (tests/test_ref.py) and against closed forms (Gauss-Legendre, Hooke's
      law, partition of unity, finite differences of the shape functions);
  determinant3, invert3 -- closed forms / numpy on random matrices;
  elastic_km, laplace_kc -- symmetry, rigid-body null space, patch test,
      the hex8 unit-cube closed form, numpy restatement of B^T D B;
  gather/scatter -- adjointness and brute force;
  elem_matvec, diag, apply -- numpy and the scipy-assembled global matrix;
  checon -- exact binary cases;
  dot_device, sum_p -- math.fsum within rounding; total load of the fixtures;
  centroid_stress -- element 1 of the p121 demo against its results file;
  plastic_update, mocouf, mocouq, formm6 -- closed forms (Tresca limit
      of MOCOUF at phi = 0, FORMM's sum m1+m2+m3 acting on the stress,
      MOCOUQ against finite differences of the potential), an elastic
      state that stays elastic, and the p122 demo's history (.res);
  shape_fun_hex8 -- bit for bit against the reference's SHAPE_FUN in
      oracle/_ref; partition of unity and the nodal delta property;
  transient_km, assemble_transient -- kc equals the p123 element (another
      summation order) to rounding, pm's entries sum to rho*cp*volume and
      its closed form on a box, storka/storkb combine them as stated, and
      the p124 demo's temperature history (.res) and fields (NDTTR);
  scatter_nodes -- element counts per node by brute force, exact averages
      of equal contributions where the count is a power of two, the sums
      conserved, and the p121 demo's printed displacement field;
  pcg -- the reference's committed fixtures (xx3-tiny: 1640 equations, 79
      iterations, displacements; p121 demo: 98360 equations, total load
      -100, 295 iterations, displacement field; p121 book: 777,520
      equations, 569 iterations, x(1); p123 book: 8,000,000
      equations, 180 iterations, the four printed potentials, with
      laplace_kc on the full 200^3 model) and scipy direct solves.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STOP_CHECON = 0
STOP_RELRES = 1
DOT_SEQ = 0        # DOT_PRODUCT order at one rank (the reference)
DOT_THREADS = 1    # per-thread contiguous chunks (like per-rank partial sums)
DOT_DEVICE = 2     # the device library's fixed reduction order


def build() -> None:
    import subprocess
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        d = C.POINTER(C.c_double)
        i = C.POINTER(C.c_int)
        L.or_centroid_stress.argtypes = [C.c_int, C.c_int, d, i, d, C.c_double, C.c_double, d, d]
        L.or_sample_hex8.argtypes = [d, d]
        L.or_scatter_nodes.argtypes = [C.c_int, C.c_int, C.c_int, i, C.c_int, d, d, C.c_int]
        L.or_shape_fun_hex8.argtypes = [d, d, C.c_int, C.c_int]
        L.or_plastic_update.argtypes = [C.c_int, d, i, d, d, d, d] + [C.c_double] * 6 + [C.c_int]
        L.or_mocouf.restype = C.c_double
        L.or_mocouf.argtypes = [C.c_double] * 5
        L.or_mocouq.argtypes = [C.c_double, C.c_double, C.c_double, d]
        L.or_formm6.argtypes = [d, d]
        L.or_transient_km.argtypes = [d, d, C.c_double, C.c_double, C.c_double, C.c_double, d, d]
        L.or_assemble_transient.argtypes = [C.c_int, d, d, i, C.c_double, C.c_double, d, d, C.c_int]
        L.or_deemat6.argtypes = [d, C.c_double, C.c_double]
        L.or_shape_der.argtypes = [d, d, C.c_int, C.c_int, C.c_int]
        L.or_shape_der.restype = C.c_int
        L.or_determinant3.argtypes = [d]
        L.or_determinant3.restype = C.c_double
        L.or_invert3.argtypes = [d]
        L.or_beemat6.argtypes = [d, d, C.c_int]
        L.or_elastic_km.argtypes = [d, C.c_int, d, d]
        L.or_laplace_kc.argtypes = [d, C.c_int, C.c_double, C.c_double, C.c_double, d]
        L.or_assemble_elastic.argtypes = [C.c_int, C.c_int, d, d, C.c_double, C.c_double, d, C.c_int]
        L.or_assemble_laplace.argtypes = [C.c_int, C.c_int, d, d, C.c_double, C.c_double, C.c_double, d, C.c_int]
        L.or_gather.argtypes = [C.c_int, C.c_int, i, d, d]
        L.or_scatter.argtypes = [C.c_int, C.c_int, i, d, d]
        L.or_elem_matvec.argtypes = [C.c_int, C.c_int, d, d, d, C.c_int]
        L.or_dot.argtypes = [C.c_long, d, d, C.c_int]
        L.or_dot.restype = C.c_double
        L.or_sum.argtypes = [C.c_long, d]
        L.or_sum.restype = C.c_double
        L.or_checon.argtypes = [C.c_long, d, C.c_double, d]
        L.or_checon.restype = C.c_int
        L.or_pcg.argtypes = [C.c_int, C.c_int, C.c_int, i, d, d, d, C.c_int, i, d,
                             C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, d, i, d]
        L.or_dot_dev.argtypes = [C.c_long, d, d]
        L.or_dot_dev.restype = C.c_double
        L.or_pcg.restype = C.c_int
        L.or_diag.argtypes = [C.c_int, C.c_int, C.c_int, i, d, d]
        L.or_apply.argtypes = [C.c_int, C.c_int, C.c_int, i, d, d, d, C.c_int]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def sample_hex8():
    """SAMPLE('hexahedron', points, weights), nip = 8 -> (points (8,3), w (8,))."""
    pts = np.zeros(24)
    w = np.zeros(8)
    lib().or_sample_hex8(_d(pts), _d(w))
    return pts.reshape(3, 8).T.copy(), w


def deemat6(e: float, v: float):
    """DEEMAT for nst = 6 -> (6,6)."""
    out = np.zeros(36)
    lib().or_deemat6(_d(out), e, v)
    return out.reshape(6, 6).T.copy()


def shape_der(points: np.ndarray, i: int, nod: int):
    """SHAPE_DER(der, points, i) -> der (3, nod); i is 1-based."""
    pts = _f64(np.asarray(points).T)           # column-major (nip,3)
    der = np.zeros(3 * nod)
    rc = lib().or_shape_der(_d(der), _d(pts), points.shape[0], i, nod)
    if rc:
        raise ValueError(nod)
    return der.reshape(nod, 3).T.copy()


def determinant3(jac):
    a = _f64(np.asarray(jac).T)
    return lib().or_determinant3(_d(a))


def invert3(jac):
    a = _f64(np.asarray(jac).T)
    lib().or_invert3(_d(a))
    return a.reshape(3, 3).T.copy()


def beemat6(deriv):
    nod = deriv.shape[1]
    dv = _f64(np.asarray(deriv).T)
    bee = np.zeros(6 * 3 * nod)
    lib().or_beemat6(_d(bee), _d(dv), nod)
    return bee.reshape(3 * nod, 6).T.copy()


def elastic_km(coord, dee):
    """One element of p121.f90:56-64; coord (nod,3) -> km (ntot,ntot)."""
    nod = coord.shape[0]
    c = _f64(np.asarray(coord).T)
    dd = _f64(np.asarray(dee).T)
    km = np.zeros(9 * nod * nod)
    lib().or_elastic_km(_d(c), nod, _d(dd), _d(km))
    return km.reshape(3 * nod, 3 * nod).T.copy()


def centroid_stress(g_coord, g_g, x, e, v, e_elem=None):
    """Centroid stresses of every element (p121.f90:112-123) -> (nels, 6)."""
    g = _f64(g_coord)
    gg = _i32(g_g)
    nels, ntot = gg.shape
    xx = _f64(x)
    ee = _f64(e_elem) if e_elem is not None else None
    out = np.zeros(6 * nels)
    lib().or_centroid_stress(nels, ntot // 3, _d(g), _i(gg), _d(xx), float(e), float(v),
                             _d(ee) if ee is not None else None, _d(out))
    return out.reshape(nels, 6)


def scatter_nodes(g_num, contrib, nn, flag=1):
    """SCATTER_NODES at one rank (gather_scatter.f90:1786-1929): g_num
    (nels, nod) 1-based, contrib (nels, nod, numvar) -> (nn, numvar)."""
    g = _i32(g_num)
    nels, nod = g.shape
    c = _f64(contrib)
    numvar = c.shape[2]
    out = np.zeros(nn * numvar)
    lib().or_scatter_nodes(int(nn), nod, nels, _i(g), numvar, _d(c), _d(out), int(flag))
    return out.reshape(nn, numvar)


def shape_fun_hex8(points: np.ndarray, i: int):
    """SHAPE_FUN for the 8-node brick at Gauss point i (1-based)."""
    p = np.asfortranarray(points, dtype=np.float64)
    out = np.zeros(8)
    lib().or_shape_fun_hex8(_d(out), _d(p), p.shape[0], int(i))
    return out


def transient_km(coord, kay, rho, cp, theta, dtim):
    """p124's storka, storkb of one 8-node brick (coord (8,3))."""
    c = _f64(np.asarray(coord).T)
    k3 = _f64(kay)
    ka, kb = np.zeros(64), np.zeros(64)
    lib().or_transient_km(_d(c), _d(k3), float(rho), float(cp), float(theta), float(dtim), _d(ka), _d(kb))
    return ka.reshape(8, 8).T.copy(), kb.reshape(8, 8).T.copy()


def assemble_transient(g_coord, prop, etype, theta, dtim, nthreads=0):
    """p124's storka_pp and storkb_pp as (nels, 8, 8) in storkm order."""
    g = _f64(g_coord)
    nels = g.shape[0]
    pr = np.ascontiguousarray(prop, dtype=np.float64)          # (np_types, 5)
    et = _i32(etype)
    ka, kb = np.zeros((nels, 8, 8)), np.zeros((nels, 8, 8))
    lib().or_assemble_transient(nels, _d(g), _d(pr), _i(et), float(theta), float(dtim), _d(ka), _d(kb),
                                int(nthreads))
    return ka, kb


def plastic_update(g_coord, g_g, x, evpt, tensor, e, v, phi, c, psi, dt, final):
    """p122's stress update (p122.f90:164-199): updates evpt and tensor
    ((nels, 8, 6) arrays, in place) and returns bload (nels, 24)."""
    g = _f64(g_coord)
    gg = _i32(g_g)
    nels = gg.shape[0]
    assert evpt.dtype == np.float64 and evpt.flags.c_contiguous and tensor.flags.c_contiguous
    out = np.zeros((nels, 24))
    lib().or_plastic_update(nels, _d(g), _i(gg), _d(_f64(x)), _d(evpt), _d(tensor), _d(out), float(e), float(v),
                            float(phi), float(c), float(psi), float(dt), int(bool(final)))
    return out


def mocouf(phi, c, sigm, dsbar, theta):
    return lib().or_mocouf(float(phi), float(c), float(sigm), float(dsbar), float(theta))


def mocouq(psi, dsbar, theta):
    dq = np.zeros(3)
    lib().or_mocouq(float(psi), float(dsbar), float(theta), _d(dq))
    return dq


def formm6(stress):
    m = np.zeros(108)
    lib().or_formm6(_d(_f64(stress)), _d(m))
    return [m[36 * k: 36 * (k + 1)].reshape(6, 6).T.copy() for k in range(3)]


def laplace_kc(coord, kx=1.0, ky=1.0, kz=1.0):
    nod = coord.shape[0]
    c = _f64(np.asarray(coord).T)
    kc = np.zeros(nod * nod)
    lib().or_laplace_kc(_d(c), nod, kx, ky, kz, _d(kc))
    return kc.reshape(nod, nod).T.copy()


def assemble_elastic(g_coord, e, v, e_elem=None, nthreads=0):
    """storkm_pp for every element.  g_coord is (nels, 3, nod) -- the
    reference's (nod,3,nels) column-major memory.  Returns (nels, ntot, ntot)
    whose [e] slice in memory order is storkm_pp(:,:,e) (column-major)."""
    g = _f64(g_coord)
    nels, _, nod = g.shape
    ntot = 3 * nod
    out = np.empty((nels, ntot, ntot))
    ee = None if e_elem is None else _f64(e_elem)
    lib().or_assemble_elastic(nels, nod, _d(g), None if ee is None else _d(ee), e, v, _d(out), nthreads)
    return out


def assemble_laplace(g_coord, kx=1.0, ky=1.0, kz=1.0, k_elem=None, nthreads=0):
    g = _f64(g_coord)
    nels, _, nod = g.shape
    out = np.empty((nels, nod, nod))
    kk = None if k_elem is None else _f64(k_elem)
    lib().or_assemble_laplace(nels, nod, _d(g), None if kk is None else _d(kk), kx, ky, kz, _d(out), nthreads)
    return out


def gather(g_g, p):
    g = _i32(g_g)
    nels, ntot = g.shape
    out = np.zeros((nels, ntot))
    pp = _f64(p)
    lib().or_gather(nels, ntot, _i(g), _d(pp), _d(out))
    return out


def scatter(g_g, utemp, neq, u=None):
    g = _i32(g_g)
    nels, ntot = g.shape
    out = np.zeros(neq) if u is None else _f64(u).copy()
    ut = _f64(utemp)
    lib().or_scatter(nels, ntot, _i(g), _d(ut), _d(out))
    return out


def elem_matvec(storkm, pmul, nthreads=0):
    k = _f64(storkm)
    nels, ntot, _ = k.shape
    pm = _f64(pmul)
    out = np.empty((nels, ntot))
    lib().or_elem_matvec(nels, ntot, _d(k), _d(pm), _d(out), nthreads)
    return out


def dot(a, b, nthreads=1):
    a = _f64(a)
    b = _f64(b)
    return lib().or_dot(a.size, _d(a), _d(b), nthreads)


def dot_device(a, b=None):
    """Sum of a*b (or of a) in the device library's reduction order."""
    a = _f64(a)
    bb = None if b is None else _f64(b)
    return lib().or_dot_dev(a.size, _d(a), None if bb is None else _d(bb))


def sum_p(a):
    a = _f64(a)
    return lib().or_sum(a.size, _d(a))


def checon(xnew, tol, x):
    """checon_par: returns (converged, x updated to xnew)."""
    xn = _f64(xnew)
    xo = _f64(x).copy()
    c = lib().or_checon(xn.size, _d(xn), tol, _d(xo))
    return bool(c), xo


def diag(g_g, storkm, neq):
    g = _i32(g_g)
    k = _f64(storkm)
    out = np.zeros(neq)
    lib().or_diag(g.shape[0], g.shape[1], neq, _i(g), _d(k), _d(out))
    return out


def apply(g_g, storkm, x, nthreads=0):
    g = _i32(g_g)
    k = _f64(storkm)
    xx = _f64(x)
    y = np.zeros(xx.size)
    lib().or_apply(g.shape[0], g.shape[1], xx.size, _i(g), _d(k), _d(xx), _d(y), nthreads)
    return y


def pcg(g_g, storkm, r0, diag_assembled, tol, limit, stop=STOP_CHECON,
        no_f=None, store=None, nthreads=1, history=False, dot_order=DOT_SEQ):
    """The p121/p123 PCG loop.  Returns (x, iters, converged[, hist])."""
    g = _i32(g_g)
    k = _f64(storkm)
    nels, ntot = g.shape
    r = _f64(r0)
    dg = _f64(diag_assembled)
    neq = r.size
    x = np.zeros(neq)
    it = np.zeros(1, dtype=np.int32)
    nfix = 0 if no_f is None else len(no_f)
    nf = _i32(no_f if nfix else np.zeros(1))
    st = _f64(store if nfix else np.zeros(1))
    hist = np.zeros(max(limit, 1))
    conv = lib().or_pcg(nels, ntot, neq, _i(g), _d(k), _d(r), _d(dg), nfix, _i(nf), _d(st),
                        tol, limit, stop, nthreads, dot_order, _d(x), _i(it), _d(hist))
    if history:
        return x, int(it[0]), bool(conv), hist[: int(it[0])]
    return x, int(it[0]), bool(conv)


# --------------------------------------------------------------------------
# oracle/_ref: the reference's own element routines (new_library.f90,
# geometry.f90) compiled by `make -C oracle ref` where the reference tree is
# present.  Used by tests only, to pin the restatements above.
# --------------------------------------------------------------------------
_REF = None
REF_PATH = os.path.join(_HERE, "_ref", "libparafem_ref.so")


def ref():
    """The compiled reference routines, or None when oracle/_ref is not built."""
    global _REF
    if _REF is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        d = C.POINTER(C.c_double)
        i = C.POINTER(C.c_int)
        ci, cd = C.c_int, C.c_double
        L.ref_sample_hex.argtypes = [ci, d, d]
        L.ref_shape_der.argtypes = [ci, ci, d, ci, d]
        L.ref_shape_fun.argtypes = [ci, ci, d, ci, d]
        L.ref_deemat.argtypes = [cd, cd, d]
        L.ref_beemat.argtypes = [ci, d, d]
        L.ref_geometry_20bxz.argtypes = [ci, ci, ci, cd, cd, cd, d, i]
        L.ref_geometry_8bxz.argtypes = [ci, ci, ci, cd, cd, cd, d, i]
        L.ref_cube_bc20.argtypes = [ci, ci, ci, ci, i]
        L.ref_steering.argtypes = [ci, ci, ci, ci, i, i, i]
        L.ref_abaqus2sg.argtypes = [ci, ci, i]
        for f in ("ref_sample_hex", "ref_shape_der", "ref_shape_fun", "ref_deemat", "ref_beemat",
                  "ref_geometry_20bxz", "ref_geometry_8bxz", "ref_cube_bc20", "ref_steering",
                  "ref_abaqus2sg"):
            getattr(L, f).restype = None
        _REF = L
    return _REF

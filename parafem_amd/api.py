"""Python binding of libparafem_amd (ctypes over the C ABI in include/).

The names mirror the reference's Fortran call surface for the EBE-PCG path
(``GATHER``, ``SCATTER``, ``SCATTER_NOADD``, ``DOT_PRODUCT_P``, ``SUM_P``,
``checon_par``; ``src/modules/mpi/gather_scatter.f90``,
``src/modules/mpi/maths.f90``) with the same argument meaning.  Arrays use
the reference's memory order (Fortran column-major): a numpy array of shape
(nels, ntot) holds ``g_g_pp(ntot, nels)``, (nels, 3, nod) holds
``g_coord_pp(nod, 3, nels)`` and (nels, ntot, ntot) holds
``storkm_pp(ntot, ntot, nels)`` where element e's block is the column-major
ntot x ntot matrix, i.e. ``K[e].T`` is the mathematical matrix (K is
symmetric, so the two agree).

The shared library is required: importing works without it, but creating a
:class:`Context` raises :class:`LibraryNotBuilt` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparafem_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "parafem_amd.h")

PF_OK = 0
PF_ERR_ARG = 1
PF_ERR_HIP = 2
PF_ERR_OOM = 3
PF_ERR_STATE = 4
PF_NOT_CONVERGED = 5
PF_ERR_BREAKDOWN = 6
PF_ERR_NODEVICE = 7
PF_ERR_TOO_LARGE = 8

STORE = {"fp64": 0, "fp32": 1}
STOP = {"checon": 0, "relres": 1}
PREC = {"fp64": 0, "mixed": 1, "mixed64": 2, "split": 3}
DOT_ORDER = {"tree": 0, "seq": 1}
SCATTER_MODE = {"slots": 0, "ranges": 1}


class LibraryNotBuilt(RuntimeError):
    pass


class PFError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: status {status} ({status_string(status)})")


class PcgOpts(C.Structure):
    _fields_ = [("limit", C.c_int), ("tol", C.c_double), ("stop_mode", C.c_int),
                ("precision", C.c_int), ("check_every", C.c_int), ("refine_every", C.c_int),
                ("refine_delta", C.c_double)]


class PcgResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("crit", C.c_double),
                ("rr0", C.c_double), ("rr", C.c_double), ("refinements", C.c_int),
                ("status", C.c_int)]


class BlockOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("inner_tol", C.c_double), ("inner_limit", C.c_int),
                ("outer_limit", C.c_int), ("check_every", C.c_int)]


class Tune(C.Structure):
    _fields_ = [("ebe_epl", C.c_int), ("blk_waves", C.c_int), ("blk_rows", C.c_int)]


class BlockResult(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner_iters", C.c_int), ("converged", C.c_int),
                ("max_relres", C.c_double), ("status", C.c_int)]


_lib = None


def lib():
    """Load libparafem_amd.so (raises LibraryNotBuilt when absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt(f"{LIB_PATH} is missing: run `python -m parafem_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl, sz = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
    pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "pf_version": (i32, []),
        "pf_status_string": (C.c_char_p, [i32]),
        "pf_create": (i32, [i32, C.POINTER(vp)]),
        "pf_destroy": (i32, [vp]),
        "pf_set_stream": (i32, [vp, vp]),
        "pf_get_stream": (vp, [vp]),
        "pf_sync": (i32, [vp]),
        "pf_dev_alloc": (i32, [vp, sz, C.POINTER(vp)]),
        "pf_dev_free": (i32, [vp, vp]),
        "pf_memcpy_h2d": (i32, [vp, vp, vp, sz]),
        "pf_memcpy_d2h": (i32, [vp, vp, vp, sz]),
        "pf_mem_used": (i32, [vp, C.POINTER(sz)]),
        "pf_mesh_set": (i32, [vp, i32, i32, pd]),
        "pf_steer_set": (i32, [vp, i32, i32, i32, pi]),
        "pf_assemble_elastic": (i32, [vp, dbl, dbl, pd, i32]),
        "pf_assemble_laplace": (i32, [vp, dbl, dbl, dbl, pd, i32]),
        "pf_storkm_set": (i32, [vp, pd, i32]),
        "pf_storkm_get": (i32, [vp, pd]),
        "pf_diag_get": (i32, [vp, pd]),
        "pf_fixed_set": (i32, [vp, i32, pi, dbl, pd]),
        "pf_gather": (i32, [vp, vp, vp]),
        "pf_elem_matvec": (i32, [vp, vp, vp]),
        "pf_scatter": (i32, [vp, vp, vp]),
        "pf_scatter_noadd": (i32, [vp, vp, vp]),
        "pf_dot_product_p": (i32, [vp, vp, vp, i64, pd]),
        "pf_sum_p": (i32, [vp, vp, i64, pd]),
        "pf_checon_par": (i32, [vp, vp, dbl, pi, vp, i64]),
        "pf_storkm_shared": (i32, [vp, pd]),
        "pf_precon_set": (i32, [vp, pd]),
        "pf_pcg_r": (i32, [vp, pd, i32]),
        "pf_pcg_run_graph": (i32, [vp, i32]),
        "pf_pcg_query": (i32, [vp, vp, pi]),
        "pf_reset": (i32, [vp]),
        "pf_block_solve": (i32, [vp, i32, pd, pd, vp, vp]),
        "pf_block_apply": (i32, [vp, i32, pd, pd, i32]),
        "pf_block_time_ebe": (i32, [vp, i32, pd]),
        "pf_stress_centroid": (i32, [vp, pd, dbl, dbl, pd, pd]),
        "pf_scatter_nodes": (i32, [vp, i32, i32, i32, pi, i32, pd, pd, i32]),
        "pf_trim": (i32, [vp]),
        "pf_steer_build": (i32, [vp, i32, i32, i32, pi, i32, i32, pi, pi, pi]),
        "pf_assemble_transient": (i32, [vp, dbl, dbl, i32, pd, pi, i32]),
        "pf_apply_raw": (i32, [vp, i32, pd, pd]),
        "pf_plastic_update": (i32, [vp, vp, vp, vp, vp, dbl, dbl, dbl, dbl, dbl, dbl, i32]),
        "pf_fixed_rows": (i32, [vp, vp, vp, i32]),
        "pf_pcg_solve_warm": (i32, [vp, vp, vp, vp, i32, C.POINTER(PcgOpts), C.POINTER(PcgResult)]),
        "pf_steer_get": (i32, [vp, pi]),
        "pf_tune_set": (i32, [vp, C.POINTER(Tune)]),
        "pf_tune_get": (i32, [vp, C.POINTER(Tune)]),
        "pf_device_info": (i32, [vp, C.c_char_p, i32, pi]),
        "pf_vec_fill": (i32, [vp, vp, dbl, i64]),
        "pf_vec_waxpy": (i32, [vp, vp, vp, vp, dbl, i64]),
        "pf_vec_mul": (i32, [vp, vp, vp, vp, i64]),
        "pf_vec_recip": (i32, [vp, vp, vp, i64]),
        "pf_pcg_solve": (i32, [vp, pd, pd, C.POINTER(PcgOpts), C.POINTER(PcgResult)]),
        "pf_pcg_begin": (i32, [vp, vp, i32, C.POINTER(PcgOpts)]),
        "pf_pcg_run": (i32, [vp, i32]),
        "pf_pcg_poll": (i32, [vp, C.POINTER(PcgResult)]),
        "pf_pcg_x": (i32, [vp, vp, i32]),
        "pf_apply": (i32, [vp, pd, pd]),
        "pf_reduce_grid": (i32, [i64]),
        "pf_pcg_timing_enable": (i32, [vp, i32]),
        "pf_pcg_timing_get": (i32, [vp, pd, pi]),
        "pf_dev_copy": (i32, [vp, vp, vp, sz]),
        "pf_dot_order_set": (i32, [vp, i32]),
        "pf_apply_noalloc": (i32, [vp, pd, pd]),
        "pf_batch_create": (i32, [i32, C.POINTER(vp)]),
        "pf_batch_destroy": (i32, [vp]),
        "pf_batch_run": (i32, [vp, i32, C.POINTER(vp), i32]),
        "pf_batch_run_graph": (i32, [vp, i32, C.POINTER(vp), i32]),
        "pf_batch_run_persistent": (i32, [vp, i32, C.POINTER(vp), i32]),
        "pf_batch_query": (i32, [vp, vp, pi]),
        "pf_batch_solo_set": (i32, [vp, C.c_long]),
        "pf_dot_order_get": (i32, [vp, pi]),
        "pf_scatter_mode_set": (i32, [vp, i32]),
        "pf_scatter_mode_get": (i32, [vp, pi]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def header_symbols() -> list[str]:
    """Function names declared in include/parafem_amd.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|const char\s*\*)\s*(pf_\w+)\s*\(", txt, re.M)))


def status_string(s: int) -> str:
    try:
        return lib().pf_status_string(int(s)).decode()
    except LibraryNotBuilt:
        return "library not built"


def _check(rc: int, what: str, ok=(PF_OK,)):
    if rc not in ok:
        raise PFError(rc, what)
    return rc


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class DeviceArray:
    """A device buffer of float64 owned by a context (pf_dev_alloc)."""

    def __init__(self, ctx: "Context", n: int):
        self.ctx = ctx
        self.n = int(n)
        p = C.c_void_p()
        _check(lib().pf_dev_alloc(ctx.h, max(self.n, 1) * 8, C.byref(p)), "pf_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, a):
        a = _f64(a).reshape(-1)
        d = cls(ctx, a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = _f64(a).reshape(-1)
        assert a.size == self.n
        _check(lib().pf_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "pf_memcpy_h2d")

    def download(self):
        out = np.empty(self.n)
        _check(lib().pf_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "pf_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().pf_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


@dataclass
class SolveResult:
    iters: int
    converged: bool
    crit: float
    rr0: float
    rr: float
    refinements: int
    status: int


class Context:
    """One GPU's EBE-PCG state (pf_create .. pf_destroy)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        _check(L.pf_create(int(device), C.byref(h)), "pf_create")
        self.h = h.value
        self.nels = self.ntot = self.neq = self.nod = 0

    def close(self):
        if getattr(self, "h", None):
            lib().pf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ plumbing
    @property
    def stream(self) -> int:
        return lib().pf_get_stream(self.h) or 0

    def set_stream(self, stream_handle: int | None):
        _check(lib().pf_set_stream(self.h, stream_handle or None), "pf_set_stream")

    def sync(self):
        _check(lib().pf_sync(self.h), "pf_sync")

    def dev_copy(self, dst: "DeviceArray", src: "DeviceArray"):
        """dst = src on the device with the streaming copy kernel."""
        if dst.n != src.n:
            raise ValueError("size mismatch")
        _check(lib().pf_dev_copy(self.h, dst.ptr, src.ptr, src.n * 8), "pf_dev_copy")

    @property
    def dot_order(self) -> str:
        """Summation order of dots and sums: "tree" (the device's fixed
        two-level order, default) or "seq" (the reference's sequential
        DOT_PRODUCT order at one rank, for parity runs)."""
        o = C.c_int()
        _check(lib().pf_dot_order_get(self.h, C.byref(o)), "pf_dot_order_get")
        return {v: k for k, v in DOT_ORDER.items()}[o.value]

    @dot_order.setter
    def dot_order(self, order: str):
        _check(lib().pf_dot_order_set(self.h, DOT_ORDER[order]), "pf_dot_order_set")

    @property
    def scatter_mode(self) -> str:
        """How the mixed solver sums element products into u: "slots"
        (utemp, then each equation's slots in ascending element order;
        default) or "ranges" (range-fused products and sums, no utemp)."""
        o = C.c_int()
        _check(lib().pf_scatter_mode_get(self.h, C.byref(o)), "pf_scatter_mode_get")
        return {v: k for k, v in SCATTER_MODE.items()}[o.value]

    @scatter_mode.setter
    def scatter_mode(self, mode: str):
        _check(lib().pf_scatter_mode_set(self.h, SCATTER_MODE[mode]), "pf_scatter_mode_set")

    def mem_used(self) -> int:
        n = C.c_size_t()
        _check(lib().pf_mem_used(self.h, C.byref(n)), "pf_mem_used")
        return n.value

    # -------------------------------------------------------------- set-up
    def mesh_set(self, g_coord: np.ndarray):
        """g_coord_pp as (nels, 3, nod)."""
        g = _f64(g_coord)
        nels, three, nod = g.shape
        assert three == 3
        _check(lib().pf_mesh_set(self.h, nod, nels, _pd(g)), "pf_mesh_set")
        self.nels, self.nod = nels, nod

    def steer_set(self, g_g: np.ndarray, neq: int):
        """g_g_pp as (nels, ntot), 1-based equations, 0 = restrained."""
        g = _i32(g_g)
        nels, ntot = g.shape
        _check(lib().pf_steer_set(self.h, ntot, nels, int(neq), _pi(g)), "pf_steer_set")
        self.nels, self.ntot, self.neq = nels, ntot, int(neq)

    def assemble_transient(self, theta: float, dtim: float, prop, etype, store: str = "fp64"):
        """p124's element matrices: prop (np_types, 5) kx, ky, kz, rho, cp;
        etype (nels,) 1-based types."""
        pr = np.ascontiguousarray(prop, dtype=np.float64)
        et = _i32(etype)
        assert pr.ndim == 2 and pr.shape[1] == 5 and et.size == self.nels
        _check(lib().pf_assemble_transient(self.h, float(theta), float(dtim), pr.shape[0], _pd(pr), _pi(et),
                                           STORE[store]), "pf_assemble_transient")

    def plastic_update(self, x: "DeviceArray", evpt: "DeviceArray", tensor: "DeviceArray", utemp: "DeviceArray",
                       e: float, v: float, phi: float, c: float, psi: float, dt: float, final: bool):
        """p122's stress update on device arrays (x: neq; evpt, tensor:
        6*8*nels; utemp: 24*nels out)."""
        _check(lib().pf_plastic_update(self.h, x.ptr, evpt.ptr, tensor.ptr, utemp.ptr, float(e), float(v),
                                       float(phi), float(c), float(psi), float(dt), int(bool(final))),
               "pf_plastic_update")

    def fixed_rows(self, u: "DeviceArray", p: "DeviceArray | None", mode: int):
        _check(lib().pf_fixed_rows(self.h, u.ptr, None if p is None else p.ptr, int(mode)), "pf_fixed_rows")

    def apply_raw(self, which: int, x) -> np.ndarray:
        """storka (which 0) or storkb (which 1) times x, no fixed rows."""
        xx = _f64(x)
        y = np.empty(self.neq)
        _check(lib().pf_apply_raw(self.h, int(which), _pd(xx), _pd(y)), "pf_apply_raw")
        return y

    def steer_build(self, g_num: np.ndarray, rest, nn: int, nodof: int = 3):
        """Steering built on the device from g_num (nels, nod) 1-based node
        numbers and rest rows (node, nodof freedom flags; 0 = restrained):
        returns (neq, nf (nn, nodof))."""
        g = _i32(g_num)
        nels, nod = g.shape
        rr = _i32(np.zeros((0, nodof + 1)) if rest is None else rest)
        assert rr.ndim == 2 and rr.shape[1] == nodof + 1
        neq = C.c_int()
        nf = np.empty((int(nn), nodof), dtype=np.int32)
        _check(lib().pf_steer_build(self.h, nod, nels, nodof, _pi(g), int(nn), rr.shape[0],
                                    _pi(rr) if rr.size else None, C.byref(neq), _pi(nf)), "pf_steer_build")
        self.nels, self.ntot, self.neq = nels, nod * nodof, neq.value
        return neq.value, nf

    def steer_get(self) -> np.ndarray:
        """The device's g_g_pp as (nels, ntot)."""
        g = np.empty((self.nels, self.ntot), dtype=np.int32)
        _check(lib().pf_steer_get(self.h, _pi(g)), "pf_steer_get")
        return g

    def assemble_elastic(self, e: float, v: float, e_elem=None, store: str = "fp64"):
        ee = None if e_elem is None else _f64(e_elem)
        _check(lib().pf_assemble_elastic(self.h, float(e), float(v), None if ee is None else _pd(ee),
                                         STORE[store]), "pf_assemble_elastic")

    def assemble_laplace(self, kx=1.0, ky=1.0, kz=1.0, k_elem=None, store: str = "fp64"):
        kk = None if k_elem is None else _f64(k_elem)
        _check(lib().pf_assemble_laplace(self.h, float(kx), float(ky), float(kz),
                                         None if kk is None else _pd(kk), STORE[store]),
               "pf_assemble_laplace")

    def storkm_set(self, storkm: np.ndarray, store: str = "fp64"):
        k = _f64(storkm)
        assert k.shape == (self.nels, self.ntot, self.ntot)
        _check(lib().pf_storkm_set(self.h, _pd(k), STORE[store]), "pf_storkm_set")

    def storkm_shared(self, km: np.ndarray):
        """One element matrix for every element (PCG_KM's km), numpy order
        (column, row) like one storkm_pp block."""
        k = _f64(km)
        assert k.shape == (self.ntot, self.ntot)
        _check(lib().pf_storkm_shared(self.h, _pd(k)), "pf_storkm_shared")

    def precon_set(self, diag_precon: np.ndarray):
        """Caller's inverted Jacobi preconditioner (neq)."""
        d = _f64(diag_precon)
        assert d.shape == (self.neq,)
        _check(lib().pf_precon_set(self.h, _pd(d)), "pf_precon_set")

    def storkm_get(self) -> np.ndarray:
        out = np.empty((self.nels, self.ntot, self.ntot))
        _check(lib().pf_storkm_get(self.h, _pd(out)), "pf_storkm_get")
        return out

    def diag_get(self) -> np.ndarray:
        out = np.empty(self.neq)
        _check(lib().pf_diag_get(self.h, _pd(out)), "pf_diag_get")
        return out

    def fixed_set(self, no_f, penalty: float = 1.0e20) -> np.ndarray:
        """p123 penalty fixed freedoms; returns store_pp."""
        nf = _i32(no_f)
        store = np.zeros(max(nf.size, 1))
        _check(lib().pf_fixed_set(self.h, nf.size, _pi(nf) if nf.size else None, float(penalty),
                                  _pd(store)), "pf_fixed_set")
        return store[: nf.size]

    # ------------------------------------------------------ drop-in layer
    @staticmethod
    def _need(buf: DeviceArray, n: int):
        """Host-side size check of a device buffer before a kernel uses it."""
        if not isinstance(buf, DeviceArray) or buf.n != n or not buf.ptr:
            raise ValueError(f"device buffer of {n} float64 expected, got {getattr(buf, 'n', None)}")

    def gather(self, p_pp: DeviceArray, pmul_pp: DeviceArray):
        """GATHER(p_pp, pmul_pp)."""
        self._need(p_pp, self.neq); self._need(pmul_pp, self.ntot * self.nels)
        _check(lib().pf_gather(self.h, p_pp.ptr, pmul_pp.ptr), "pf_gather")

    def elem_matvec(self, pmul_pp: DeviceArray, utemp_pp: DeviceArray):
        self._need(pmul_pp, self.ntot * self.nels); self._need(utemp_pp, self.ntot * self.nels)
        _check(lib().pf_elem_matvec(self.h, pmul_pp.ptr, utemp_pp.ptr), "pf_elem_matvec")

    def scatter(self, u_pp: DeviceArray, utemp_pp: DeviceArray):
        """SCATTER(u_pp, utemp_pp): u_pp accumulates."""
        self._need(u_pp, self.neq); self._need(utemp_pp, self.ntot * self.nels)
        _check(lib().pf_scatter(self.h, u_pp.ptr, utemp_pp.ptr), "pf_scatter")

    def scatter_noadd(self, utemp_pp: DeviceArray, u_pp: DeviceArray):
        """SCATTER_NOADD(utemp_pp, u_pp): assignment (note the argument order)."""
        self._need(u_pp, self.neq); self._need(utemp_pp, self.ntot * self.nels)
        _check(lib().pf_scatter_noadd(self.h, utemp_pp.ptr, u_pp.ptr), "pf_scatter_noadd")

    def dot_product_p(self, a: DeviceArray, b: DeviceArray) -> float:
        self._need(b, a.n)
        out = C.c_double()
        _check(lib().pf_dot_product_p(self.h, a.ptr, b.ptr, a.n, C.byref(out)), "pf_dot_product_p")
        return out.value

    def sum_p(self, a: DeviceArray) -> float:
        out = C.c_double()
        _check(lib().pf_sum_p(self.h, a.ptr, a.n, C.byref(out)), "pf_sum_p")
        return out.value

    def checon_par(self, xnew: DeviceArray, tol: float, x: DeviceArray) -> bool:
        """checon_par(loads, tol, converged, oldlds): returns converged, x = xnew."""
        self._need(x, xnew.n)
        conv = C.c_int()
        _check(lib().pf_checon_par(self.h, xnew.ptr, float(tol), C.byref(conv), x.ptr, xnew.n),
               "pf_checon_par")
        return bool(conv.value)

    def vec_fill(self, x: DeviceArray, value: float):
        """x = value"""
        _check(lib().pf_vec_fill(self.h, x.ptr, float(value), x.n), "pf_vec_fill")

    def vec_waxpy(self, w: DeviceArray, x: DeviceArray, y: DeviceArray, a: float):
        """w = x + y*a (xnew = x + p*alpha; r = r - u*alpha as a = -alpha)."""
        self._need(x, w.n); self._need(y, w.n)
        _check(lib().pf_vec_waxpy(self.h, w.ptr, x.ptr, y.ptr, float(a), w.n), "pf_vec_waxpy")

    def vec_mul(self, z: DeviceArray, x: DeviceArray, y: DeviceArray):
        """z = x*y (d = diag_precon*r)."""
        self._need(x, z.n); self._need(y, z.n)
        _check(lib().pf_vec_mul(self.h, z.ptr, x.ptr, y.ptr, z.n), "pf_vec_mul")

    def vec_recip(self, y: DeviceArray, x: DeviceArray):
        """y = 1/x (diag_precon = 1/diag_precon)."""
        self._need(x, y.n)
        _check(lib().pf_vec_recip(self.h, y.ptr, x.ptr, y.n), "pf_vec_recip")

    # ----------------------------------------------------------- the solve
    @staticmethod
    def _opts(limit, tol, stop, precision, check_every=0, refine_every=0, refine_delta=0.0):
        return PcgOpts(int(limit), float(tol), STOP[stop], PREC[precision], int(check_every),
                       int(refine_every), float(refine_delta))

    def pcg_solve(self, r0, limit: int, tol: float, stop: str = "checon", precision: str = "fp64",
                  check_every: int = 0, refine_every: int = 0, refine_delta: float = 0.0):
        """The p121/p123 PCG loop; returns (xnew, SolveResult).  precision
        "mixed" needs FP32 storage and stop "relres"."""
        r = _f64(r0)
        assert r.size == self.neq
        x = np.zeros(self.neq)
        o = self._opts(limit, tol, stop, precision, check_every, refine_every, refine_delta)
        res = PcgResult()
        rc = lib().pf_pcg_solve(self.h, _pd(r), _pd(x), C.byref(o), C.byref(res))
        _check(rc, "pf_pcg_solve", ok=(PF_OK, PF_NOT_CONVERGED, PF_ERR_BREAKDOWN))
        return x, self._res(res)

    @staticmethod
    def _res(res):
        return SolveResult(res.iters, bool(res.converged), res.crit, res.rr0, res.rr,
                           res.refinements, res.status)

    def pcg_solve_warm(self, r0: "DeviceArray", x0: "DeviceArray", xnew: "DeviceArray", fixed_mode: int,
                       limit: int, tol: float, stop: str = "checon", precision: str = "fp64",
                       check_every: int = 0) -> "SolveResult":
        """p122's warm-started CG on device vectors: x = x0, r = r0 (the
        caller's loads - K x0), fixed rows u = p*store (fixed_mode 1) or 0
        (fixed_mode 0); xnew receives the last iterate."""
        o = self._opts(limit, tol, stop, precision, check_every)
        res = PcgResult()
        rc = lib().pf_pcg_solve_warm(self.h, r0.ptr, x0.ptr, xnew.ptr, int(fixed_mode), C.byref(o), C.byref(res))
        _check(rc, "pf_pcg_solve_warm", ok=(PF_OK, PF_NOT_CONVERGED, PF_ERR_BREAKDOWN))
        return self._res(res)

    def pcg_begin(self, r0, limit: int, tol: float, stop: str = "checon", precision: str = "fp64",
                  refine_every: int = 0, refine_delta: float = 0.0, on_device: bool = False):
        o = self._opts(limit, tol, stop, precision, 0, refine_every, refine_delta)
        if on_device:
            _check(lib().pf_pcg_begin(self.h, C.c_void_p(int(r0)), 1, C.byref(o)), "pf_pcg_begin")
        else:
            r = _f64(r0)
            _check(lib().pf_pcg_begin(self.h, r.ctypes.data, 0, C.byref(o)), "pf_pcg_begin")

    def pcg_run(self, n: int):
        _check(lib().pf_pcg_run(self.h, int(n)), "pf_pcg_run")

    def pcg_poll(self) -> SolveResult:
        res = PcgResult()
        _check(lib().pf_pcg_poll(self.h, C.byref(res)), "pf_pcg_poll")
        return self._res(res)

    def pcg_x(self) -> np.ndarray:
        x = np.empty(self.neq)
        _check(lib().pf_pcg_x(self.h, x.ctypes.data, 0), "pf_pcg_x")
        return x

    def pcg_run_graph(self, n: int):
        """Enqueue n iterations as one replayed HIP graph (no host sync)."""
        _check(lib().pf_pcg_run_graph(self.h, int(n)), "pf_pcg_run_graph")

    def pcg_query(self):
        """Non-blocking poll: a PcgResult once the state snapshot taken at
        the first call after the last enqueued work has arrived, else None."""
        res = PcgResult()
        ready = C.c_int(0)
        _check(lib().pf_pcg_query(self.h, C.byref(res), C.byref(ready)), "pf_pcg_query")
        return res if ready.value else None

    def trim(self):
        """Give the context's cached device blocks back to the device."""
        _check(lib().pf_trim(self.h), "pf_trim")

    def reset(self):
        """Free the current problem's device buffers; keep the context."""
        _check(lib().pf_reset(self.h), "pf_reset")
        self.nels = self.ntot = self.neq = self.nod = 0

    def scatter_nodes(self, g_num, contrib, nn: int, flag: int = 1) -> np.ndarray:
        """SCATTER_NODES at one rank (gather_scatter.f90:1786-1929): g_num
        (nels, nod) 1-based node numbers, contrib (nels, nod, numvar) ->
        (nn, numvar) sums, or averages with flag 1."""
        g = np.ascontiguousarray(g_num, dtype=np.int32)
        c = np.ascontiguousarray(contrib, dtype=np.float64)
        nels, nod = g.shape
        assert c.shape[:2] == (nels, nod)
        numvar = c.shape[2]
        out = np.empty(int(nn) * numvar)
        _check(lib().pf_scatter_nodes(self.h, int(nn), nod, nels, _pi(g), numvar, _pd(c), _pd(out), int(flag)),
               "pf_scatter_nodes")
        return out.reshape(int(nn), numvar)

    def stress_centroid(self, x, e: float, v: float, e_elem=None) -> np.ndarray:
        """Centroid stresses (nels, 6) of the displacement x (p121.f90:112-123)."""
        xx = _f64(x)
        assert xx.size == self.neq
        out = np.zeros(6 * self.nels)
        ee = _f64(e_elem) if e_elem is not None else None
        _check(lib().pf_stress_centroid(self.h, _pd(xx), float(e), float(v),
                                        _pd(ee) if ee is not None else None, _pd(out)), "pf_stress_centroid")
        return out.reshape(self.nels, 6)

    def block_solve(self, f, tol: float = 1e-10, inner_tol: float = 0.0, inner_limit: int = 0,
                    outer_limit: int = 0, check_every: int = 0):
        """Multi-RHS solve: f (neq, k), k <= 64 -> (x (neq, k), BlockResult)."""
        F = np.asfortranarray(f, dtype=np.float64)
        assert F.ndim == 2 and F.shape[0] == self.neq and 1 <= F.shape[1] <= 64
        X = np.zeros(F.shape, order="F")
        o = BlockOpts(float(tol), float(inner_tol), int(inner_limit), int(outer_limit), int(check_every))
        res = BlockResult()
        rc = lib().pf_block_solve(self.h, F.shape[1], _pd(F), _pd(X), C.byref(o), C.byref(res))
        _check(rc, "pf_block_solve", ok=(PF_OK, PF_NOT_CONVERGED))
        return X, res

    def block_apply(self, x, precision: str = "fp64") -> np.ndarray:
        """Y = K X through the matrix-core block product (x (neq, k))."""
        Xf = np.asfortranarray(x, dtype=np.float64)
        assert Xf.ndim == 2 and Xf.shape[0] == self.neq and 1 <= Xf.shape[1] <= 64
        Y = np.zeros(Xf.shape, order="F")
        _check(lib().pf_block_apply(self.h, Xf.shape[1], _pd(Xf), _pd(Y), PREC[precision]), "pf_block_apply")
        return Y

    def block_time_ebe(self, reps: int = 10) -> float:
        """Average ms of the FP32 matrix-core element product over reps
        launches (HIP events on the context's stream) on the block of the
        last block_apply(..., "mixed") or block_solve."""
        ms = C.c_double()
        _check(lib().pf_block_time_ebe(self.h, int(reps), C.byref(ms)), "pf_block_time_ebe")
        return ms.value

    def tune_set(self, ebe_epl: int = 0, blk_waves: int = 0, blk_rows: int = 0):
        """Launch shapes of the element products (0 = built-in choice)."""
        t = Tune(int(ebe_epl), int(blk_waves), int(blk_rows))
        _check(lib().pf_tune_set(self.h, C.byref(t)), "pf_tune_set")

    def tune_get(self) -> dict:
        t = Tune()
        _check(lib().pf_tune_get(self.h, C.byref(t)), "pf_tune_get")
        return {"ebe_epl": t.ebe_epl, "blk_waves": t.blk_waves, "blk_rows": t.blk_rows}

    def device_info(self):
        """(architecture name, compute units) of the context's device."""
        buf = C.create_string_buffer(256)
        cus = C.c_int()
        _check(lib().pf_device_info(self.h, buf, 256, C.byref(cus)), "pf_device_info")
        return buf.value.decode(), cus.value

    def pcg_r(self) -> np.ndarray:
        """Recurrence residual at the end of the solve."""
        r = np.empty(self.neq)
        _check(lib().pf_pcg_r(self.h, _pd(r), 0), "pf_pcg_r")
        return r

    def timing_enable(self, on: bool = True):
        """Per-kernel HIP-event timing of pcg_run (resets the accumulators)."""
        _check(lib().pf_pcg_timing_enable(self.h, int(bool(on))), "pf_pcg_timing_enable")

    def timing_get(self):
        """{kernel group: (total ms, launches)} for ebe, scatter, update,
        pupdate and refine (the mixed solver's true-residual replacement)."""
        ms = np.zeros(5)
        cnt = np.zeros(5, dtype=np.int32)
        _check(lib().pf_pcg_timing_get(self.h, _pd(ms), _pi(cnt)), "pf_pcg_timing_get")
        return {k: (float(ms[i]), int(cnt[i]))
                for i, k in enumerate(("ebe", "scatter", "update", "pupdate", "refine"))}

    def apply(self, x) -> np.ndarray:
        """y = K x (FP64, with p123 fixed rows)."""
        xx = _f64(x)
        y = np.empty(self.neq)
        _check(lib().pf_apply(self.h, _pd(xx), _pd(y)), "pf_apply")
        return y

    def apply_noalloc(self, x) -> np.ndarray:
        """apply() through the allocation-free entry point (the library's
        fixed per-GPU scratch; PFError with PF_ERR_TOO_LARGE beyond it)."""
        xx = _f64(x)
        assert xx.size == self.neq
        y = np.empty(self.neq)
        _check(lib().pf_apply_noalloc(self.h, _pd(xx), _pd(y)), "pf_apply_noalloc")
        return y


BATCH_MAX = 64


class Batch:
    """Many begun solves advanced by one launch per kernel (pf_batch_*):
    each solve's iterates equal what pf_pcg_run on its context alone gives,
    bit for bit."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().pf_batch_create(int(device), C.byref(h)), "pf_batch_create")
        self.h = h.value
        self._last = []
        self._snap = None

    def close(self):
        if getattr(self, "h", None):
            lib().pf_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _handles(ctxs):
        arr = (C.c_void_p * len(ctxs))(*[c.h for c in ctxs])
        return arr

    def solo_set(self, min_elems: int):
        """Solves of at least min_elems elements (0: none) run their own
        iteration launches inside the batch's runs (pf_batch_solo_set)."""
        _check(lib().pf_batch_solo_set(self.h, int(min_elems)), "pf_batch_solo_set")

    def run(self, ctxs, n: int, graph: bool = True, persistent: bool = False):
        """Enqueue n iterations of every context in ctxs (no host sync):
        a replayed HIP graph (graph), direct launches, or one persistent
        cooperative kernel that stops early once every solve has stopped."""
        if not 1 <= len(ctxs) <= BATCH_MAX:
            raise ValueError(f"1..{BATCH_MAX} contexts per batch")
        arr = self._handles(ctxs)
        f = (lib().pf_batch_run_persistent if persistent else
             lib().pf_batch_run_graph if graph else lib().pf_batch_run)
        _check(f(self.h, len(ctxs), arr, int(n)), "pf_batch_run")
        self._last = list(ctxs)

    def query(self):
        """Non-blocking poll: (contexts, [PcgResult]) of the snapshot taken
        at the first call after the previous one was delivered -- the
        contexts of the last run at that point, in its order -- once it has
        arrived, else None.  Runs enqueued after the snapshot do not change
        it."""
        if self._snap is None:
            self._snap = self._last
        res = (PcgResult * len(self._snap))()
        ready = C.c_int(0)
        _check(lib().pf_batch_query(self.h, C.cast(res, C.c_void_p), C.byref(ready)), "pf_batch_query")
        if not ready.value:
            return None
        snap, self._snap = self._snap, None
        return snap, [res[k] for k in range(len(snap))]


def reduce_grid(n: int) -> int:
    return lib().pf_reduce_grid(int(n))

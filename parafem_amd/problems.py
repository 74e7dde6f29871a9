"""Problem builders and the p121 / p123 driver flows on the device.

A *problem* holds what the reference drivers have in memory just before
their element loop: per-element coordinates ``g_coord_pp``, the steering
``g_g_pp``, ``neq``, the loaded right-hand side ``r_pp`` and (p123) the fixed
freedoms.  The flows then follow ``src/programs/5th_ed/p121/p121.f90`` and
``src/programs/5th_ed/p123/p123.f90`` step for step on a
:class:`parafem_amd.api.Context`.

Workloads (all synthetic, generated here from documented seeds):
  * ``elastic_cube`` -- hex20 cube, unit edges, base z = -n fixed, uniform unit
    pressure on the top face z = 0 as consistent nodal loads; optional
    per-element log-normal E (mu = ln 100, sigma = 0.5, seed 1);
  * ``laplace_cube`` -- hex8 cube, T = 0 on x = 0 by penalty fixed freedoms,
    unit flux on x = n, per-element log-normal k (mu = 0, sigma = 1, seed 2);
  * ``uniaxial_cube`` -- hex20 cube with symmetry rollers and a uniform top
    pressure, whose exact solution is linear (patch test).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh as M

PENALTY = 1.0e20   # p123.f90:17


@dataclass
class Problem:
    kind: str                      # "elastic" or "laplace"
    mesh: M.Mesh
    g_g: np.ndarray                # (nels, ntot) int32
    neq: int
    r0: np.ndarray                 # (neq,) loaded right-hand side
    nf: np.ndarray                 # (nn, nodof) node freedom table
    e: float = 100.0
    v: float = 0.3
    e_elem: np.ndarray | None = None
    kx: float = 1.0
    ky: float = 1.0
    kz: float = 1.0
    k_elem: np.ndarray | None = None
    fix_eq: np.ndarray | None = None    # p123 fixed equations (1-based)
    fix_val: np.ndarray | None = None
    info: dict = field(default_factory=dict)

    @property
    def nels(self):
        return self.mesh.nels

    @property
    def ntot(self):
        return self.g_g.shape[1]

    def g_coord(self):
        return self.mesh.g_coord()


def elastic_cube(n: int, e: float = 100.0, v: float = 0.3, lognormal_e: bool = False,
                 seed: int = 1, pressure: float = 1.0, nels_xyz=None) -> Problem:
    """p121-style elastic cube (unit edge hex20), base fixed, top pressure."""
    nx, ny, nz = nels_xyz if nels_xyz else (n, n, n)
    m = M.box_hex20(nx, ny, nz, 1.0, 1.0, 1.0)
    rest = M.base_fixed_rest(m)
    nf, neq = M.steering(m.nn, 3, rest)
    g_g = M.g_g_from(nf, m.g_num)
    nodes, vals = M.top_pressure_loads(m, -pressure)
    r0 = M.nodal_loads_to_rhs(nf, nodes, vals, neq)
    e_elem = M.lognormal(seed, m.nels, np.log(100.0), 0.5) if lognormal_e else None
    return Problem("elastic", m, g_g, neq, r0, nf, e=e, v=v, e_elem=e_elem,
                   info={"rest_nodes": rest[:, 0], "load_nodes": nodes})


def uniaxial_cube(n: int, e: float = 100.0, v: float = 0.3, q: float = 1.0) -> Problem:
    """Hex20 cube with rollers on x = 0 (ux), y = 0 (uy) and the base (uz),
    uniform pressure q on the top.  Exact solution: uz = -q (z + n)/E,
    ux = v q x / E, uy = v q y / E (uniform uniaxial stress)."""
    m = M.box_hex20(n, n, n, 1.0, 1.0, 1.0)
    c = m.coords
    free = np.ones((m.nn, 3), dtype=bool)
    free[np.abs(c[:, 0]) < 1e-12, 0] = False
    free[np.abs(c[:, 1]) < 1e-12, 1] = False
    free[np.abs(c[:, 2] + n) < 1e-12, 2] = False
    sel = np.nonzero(~free.all(axis=1))[0]
    rest = np.zeros((sel.size, 4), dtype=np.int32)
    rest[:, 0] = sel + 1
    rest[:, 1:] = free[sel]
    nf, neq = M.steering(m.nn, 3, rest)
    g_g = M.g_g_from(nf, m.g_num)
    nodes, vals = M.top_pressure_loads(m, -q)
    r0 = M.nodal_loads_to_rhs(nf, nodes, vals, neq)
    exact = np.zeros((m.nn, 3))
    exact[:, 0] = v * q * c[:, 0] / e
    exact[:, 1] = v * q * c[:, 1] / e
    exact[:, 2] = -q * (c[:, 2] + n) / e
    xe = np.zeros(neq)
    msk = nf > 0
    xe[nf[msk] - 1] = exact[msk]
    return Problem("elastic", m, g_g, neq, r0, nf, e=e, v=v, info={"exact": xe})


def laplace_cube(n: int, lognormal_k: bool = True, seed: int = 2, flux: float = 1.0,
                 nels_xyz=None) -> Problem:
    """p123-style hex8 conduction cube: T = 0 on x = 0 by penalty fixed
    freedoms (nr = 0 so g_g_pp = g_num_pp, p123.f90:54), unit flux on x = n."""
    nx, ny, nz = nels_xyz if nels_xyz else (n, n, n)
    m = M.box_hex8(nx, ny, nz, 1.0, 1.0, 1.0)
    nf, neq = M.steering(m.nn, 1, None)
    g_g = M.g_g_from(nf, m.g_num)
    nodes, vals = M.face_flux_loads(m, float(nx), flux)
    r0 = np.zeros(neq)
    r0[nf[nodes - 1, 0] - 1] = vals            # p123.f90:115-117: r(no) = val
    fixed_nodes = np.nonzero(np.abs(m.coords[:, 0]) < 1e-12)[0] + 1
    fix_eq = nf[fixed_nodes - 1, 0].astype(np.int32)
    k_elem = M.lognormal(seed, m.nels, 0.0, 1.0) if lognormal_k else None
    return Problem("laplace", m, g_g, neq, r0, nf, k_elem=k_elem, fix_eq=fix_eq,
                   fix_val=np.zeros(fix_eq.size))


def p121_mg(nels: int, nxe: int, nze: int, aa: float, bb: float, cc: float, e: float = 100.0,
            v: float = 0.3) -> Problem:
    """A p121 model as p12meshgen's p121 case builds it for 20-node bricks
    (``src/tools/preprocessing/p12meshgen/p12meshgen.f90:118-174``):
    nye = nels / nxe / nze, GEOMETRY_20BXZ mesh, CUBE_BC20 restraints and
    LOAD_P121's pressure patch (nle = nxe / 5) scaled by
    -aa bb 25/12.  The reference's demo (``examples/5th_ed/p121/demo``) is
    p121_mg(8000, 20, 20, .5, .5, .5) and its book run
    (``examples/5th_ed/p121/book``) p121_mg(64000, 40, 40, .25, .25, .25)."""
    nye = nels // nxe // nze
    m = M.box_hex20(nxe, nye, nze, aa, bb, cc)
    rest = M.cube_bc20(nxe, nye, nze)
    node, val = M.load_p121(nxe // 5, 20, nxe, nze)
    val = -np.asarray(val) * aa * bb * (25.0 / 12.0)
    nf, neq = M.steering(m.nn, 3, rest)
    g_g = M.g_g_from(nf, m.g_num)
    vals = np.zeros((len(node), 3))
    vals[:, 2] = val
    r0 = M.nodal_loads_to_rhs(nf, np.asarray(node), vals, neq)
    return Problem("elastic", m, g_g, neq, r0, nf, e=e, v=v, info={"rest": rest, "load_node": node})


def p123_book(nxe: int = 200, nye: int = 200, nze: int = 200, h: float = 0.005, k: float = 2.0,
              load: float = 10.0) -> Problem:
    """The p123 deck of the reference's book example
    (``examples/5th_ed/p123/book/p123.mg``: 200^3 bricks of edge 0.005,
    kx = ky = kz = 2) as p12meshgen's p123 case builds it
    (``src/tools/preprocessing/p12meshgen/p12meshgen.f90:658-760``):
    GEOMETRY_8BXZ mesh, BOX_BC8 restraints, one loaded freedom, equation
    nres = nxe (nze - 1) + 1, with value 10 (used as an equation number,
    p123.f90:108-118).  info["nres"] is that equation."""
    m = M.box_hex8(nxe, nye, nze, h, h, h)
    rest = M.box_bc8_rest(nxe, nye, nze)
    nf, neq = M.steering(m.nn, 1, rest)
    g_g = M.g_g_from(nf, m.g_num)
    nres = nxe * (nze - 1) + 1
    r0 = np.zeros(neq)
    r0[nres - 1] = load
    return Problem("laplace", m, g_g, neq, r0, nf, kx=k, ky=k, kz=k,
                   info={"nres": nres, "rest": rest})


def block_cube(n: int, k: int = 64, seed: int = 3, e: float = 100.0, v: float = 0.3):
    """Multi-RHS workload: hex20 cube (unit edges, E, nu uniform, base
    fixed) with k load cases; case c is a top-face traction whose three
    components are drawn U[-1,1] per top-face node (numpy default_rng(seed),
    array (ntop, 3, k)), turned into consistent nodal loads.
    Returns (Problem with r0 = case 0, F (neq, k))."""
    m = M.box_hex20(n, n, n, 1.0, 1.0, 1.0)
    rest = M.base_fixed_rest(m)
    nf, neq = M.steering(m.nn, 3, rest)
    g_g = M.g_g_from(nf, m.g_num)
    top = np.nonzero(np.abs(m.coords[:, 2]) < 1e-12)[0]
    rng = np.random.default_rng(seed)
    t = np.zeros((m.nn, 3, k))
    t[top] = rng.uniform(-1.0, 1.0, size=(top.size, 3, k))
    f = M.face_traction_loads(m, t)
    F = np.zeros((neq, k))
    msk = nf > 0
    F[nf[msk] - 1] = f[msk]
    pr = Problem("elastic", m, g_g, neq, F[:, 0].copy(), nf, e=e, v=v)
    return pr, F


def setup(ctx, prob: Problem, store: str = "fp64", assemble: bool = True):
    """Upload mesh and steering and build the element matrices on the device
    (p121.f90:40-69 / p123.f90:43-92).  For p123 also the penalty rows and the
    starting r (p123.f90:120-132).  Returns the starting r."""
    ctx.mesh_set(prob.g_coord())
    ctx.steer_set(prob.g_g, prob.neq)
    if assemble:
        if prob.kind == "elastic":
            ctx.assemble_elastic(prob.e, prob.v, prob.e_elem, store=store)
        else:
            ctx.assemble_laplace(prob.kx, prob.ky, prob.kz, prob.k_elem, store=store)
    return start_r(ctx, prob)


def start_r(ctx, prob: Problem):
    r = prob.r0.copy()
    if prob.fix_eq is not None and prob.fix_eq.size:
        store = ctx.fixed_set(prob.fix_eq, PENALTY)
        r[prob.fix_eq - 1] = store * prob.fix_val
    return r


def solve(ctx, prob: Problem, tol: float, limit: int, stop: str = "checon",
          precision: str = "fp64", store: str | None = None, **kw):
    """Full flow: set-up, assembly, PCG.  Returns (x, SolveResult).  The
    mixed solver stores the element matrices as FP32 hi/lo."""
    if store is None:
        store = "fp64" if precision == "fp64" else "fp32"
    r = setup(ctx, prob, store=store)
    return ctx.pcg_solve(r, limit, tol, stop=stop, precision=precision, **kw)


def nodal_field(prob: Problem, x: np.ndarray) -> np.ndarray:
    """Scatter an equation vector back to nodes (restrained = 0)."""
    nodof = prob.nf.shape[1]
    out = np.zeros((prob.mesh.nn, nodof))
    m = prob.nf > 0
    out[m] = x[prob.nf[m] - 1]
    return out

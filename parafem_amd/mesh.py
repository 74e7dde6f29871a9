"""Structured hexahedral meshes, boundary conditions, loads and ParaFEM decks.

Everything here is host-side set-up for the EBE-PCG path: it produces the
arrays the reference drivers read from their decks (``g_num_pp``,
``g_coord_pp``, ``rest``, loaded nodes/values) so that the device path can be
fed exactly what p121/p123 would see.

Node numbering and element node order follow ParaFEM's box generators
(``src/modules/shared/geometry.f90:70-166`` GEOMETRY_8BXZ and ``:172-286``
GEOMETRY_20BXZ): elements are counted along x, then down z, then along y
("x-z planes counted in the y-direction"); the top face is z = 0 and z grows
negative downwards.

Arrays use C order with the *element* index first, i.e. ``g_num[e, l]`` is
the reference's ``g_num_pp(l, e)`` (Fortran column-major), so the memory
layouts are identical and a buffer can be handed to the C ABI unchanged.
"""
from __future__ import annotations

import os
import shlex
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Mesh", "box_hex20", "box_hex8", "cube_bc20", "load_p121", "abaqus_to_sg",
    "sg_to_abaqus", "base_fixed_rest", "top_pressure_loads", "face_flux_loads",
    "steering", "nodal_loads_to_rhs", "read_deck", "write_p121_deck",
    "splitmix64_uniform", "lognormal", "element_coords", "fortran_e", "write_ensi_vector",
    "read_ensi_vector", "write_p124_deck", "p124_box_rest",
]


@dataclass
class Mesh:
    """A mesh in ParaFEM's arrays (0-based numpy views of 1-based ids)."""
    nod: int
    g_num: np.ndarray          # (nels, nod) int32, 1-based node ids
    coords: np.ndarray         # (nn, 3) float64 nodal coordinates
    nxe: int = 0
    nye: int = 0
    nze: int = 0
    aa: float = 1.0
    bb: float = 1.0
    cc: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def nels(self) -> int:
        return int(self.g_num.shape[0])

    @property
    def nn(self) -> int:
        return int(self.coords.shape[0])

    def g_coord(self) -> np.ndarray:
        """Per-element coordinates in the reference ``g_coord_pp(nod,3,nels)``
        memory order: returned as (nels, 3, nod) C-order float64.  A
        generator that formed them element by element keeps them (read-only)
        for as long as ``coords`` is the array it made."""
        gc = self.extra.get("g_coord")
        if gc is not None and self.extra.get("g_coord_of") is self.coords:
            return gc
        return element_coords(self.g_num, self.coords)


def element_coords(g_num: np.ndarray, coords: np.ndarray) -> np.ndarray:
    """``g_coord_pp(:,:,iel) = g_coord(g_num_pp(:,iel), :)`` as (nels,3,nod)."""
    c = coords[g_num - 1]                    # (nels, nod, 3)
    return np.ascontiguousarray(np.transpose(c, (0, 2, 1)))


def _elem_ijk(nels: int, nxe: int, nze: int):
    iel = np.arange(1, nels + 1, dtype=np.int64)
    iq = (iel - 1) // (nxe * nze) + 1
    iplane = iel - (iq - 1) * nxe * nze
    is_ = (iplane - 1) // nxe + 1
    ip = iplane - (is_ - 1) * nxe
    return ip, iq, is_


def box_hex20(nxe: int, nye: int, nze: int, aa: float = 1.0, bb: float = 1.0,
              cc: float = 1.0) -> Mesh:
    """20-node brick box in ParaFEM's node numbering (S&G element order).

    Follows GEOMETRY_20BXZ (``geometry.f90:172-286``) for every element at
    once: five base node numbers per element, each local node one of them
    plus a fixed offset; the element's coordinates pick, per local node and
    axis, the lower face, the upper face or their midpoint (the same
    expressions for every element sharing a node, so nodal coordinates are
    the element-local ones scattered to nodes and the per-element array is
    kept as ``g_coord``).
    """
    nels = nxe * nye * nze
    ip, iq, is_ = _elem_ijk(nels, nxe, nze)
    per_plane = (2 * nxe + 1) * (nze + 1) + (2 * nze + 1) * (nxe + 1)
    fac1 = per_plane * (iq - 1)
    fac2 = per_plane * iq
    nn = ((2 * nxe + 1) * (nze + 1) + (nxe + 1) * nze) * (nye + 1) \
        + (nxe + 1) * (nze + 1) * nye
    it = np.int32 if nn < 2 ** 31 - 1 else np.int64
    base = np.empty((nels, 5), dtype=it)
    base[:, 0] = fac1 + (3 * nxe + 2) * is_ + 2 * ip - 1       # local nodes 1-3, 4-8
    base[:, 1] = fac1 + (3 * nxe + 2) * is_ - nxe + ip - 1     # 2, 6
    base[:, 2] = fac2 - (nxe + 1) * (nze + 1) + (nxe + 1) * is_ + ip   # 9-12
    base[:, 3] = fac2 + (3 * nxe + 2) * is_ + 2 * ip - 1       # 13, 15-17, 19, 20
    base[:, 4] = fac2 + (3 * nxe + 2) * is_ - nxe + ip - 1     # 14, 18
    which = np.array([0, 1, 0, 0, 0, 1, 0, 0, 2, 2, 2, 2, 3, 4, 3, 3, 3, 4, 3, 3])
    step = np.array([0, 0, -3 * nxe - 2, -3 * nxe - 1, -3 * nxe, 1, 2, 1,
                     0, -nxe - 1, -nxe, 1,
                     0, 0, -3 * nxe - 2, -3 * nxe - 1, -3 * nxe, 1, 2, 1], dtype=it)
    n = base[:, which]
    n += step

    # per element and axis: lower face, upper face, midpoint
    def faces(lo, hi):
        f = np.empty((nels, 3), dtype=np.float64)
        f[:, 0] = lo
        f[:, 1] = hi
        f[:, 2] = 0.5 * (f[:, 0] + f[:, 1])
        return f

    fx = faces((ip - 1) * aa, ip * aa)
    fy = faces((iq - 1) * bb, iq * bb)
    fz = faces(-is_ * cc, -(is_ - 1) * cc)        # z decreases downwards
    sx = np.array([0, 0, 0, 2, 1, 1, 1, 2, 0, 0, 1, 1, 0, 0, 0, 2, 1, 1, 1, 2])
    sy = np.array([0] * 8 + [2] * 4 + [1] * 8)
    sz = np.array([0, 2, 1, 1, 1, 2, 0, 0, 0, 1, 1, 0, 0, 2, 1, 1, 1, 2, 0, 0])
    gc = np.empty((nels, 3, 20), dtype=np.float64)
    gc[:, 0, :] = fx[:, sx]
    gc[:, 1, :] = fy[:, sy]
    gc[:, 2, :] = fz[:, sz]
    coords = np.zeros((nn, 3), dtype=np.float64)
    idx = n.reshape(-1) - 1
    for k in range(3):
        coords[idx, k] = gc[:, k, :].reshape(-1)
    gc.flags.writeable = False
    m = Mesh(20, n.astype(np.int32, copy=False), coords, nxe, nye, nze, aa, bb, cc)
    m.extra["g_coord"] = gc
    m.extra["g_coord_of"] = coords
    return m


def box_hex8(nxe: int, nye: int, nze: int, aa: float = 1.0, bb: float = 1.0,
             cc: float = 1.0) -> Mesh:
    """8-node brick box, GEOMETRY_8BXZ (``geometry.f90:70-166``)."""
    nels = nxe * nye * nze
    ip, iq, is_ = _elem_ijk(nels, nxe, nze)
    n = np.empty((nels, 8), dtype=np.int64)
    n[:, 0] = (iq - 1) * (nxe + 1) * (nze + 1) + is_ * (nxe + 1) + ip
    n[:, 1] = n[:, 0] - nxe - 1
    n[:, 2] = n[:, 1] + 1
    n[:, 3] = n[:, 0] + 1
    n[:, 4] = n[:, 0] + (nxe + 1) * (nze + 1)
    n[:, 5] = n[:, 4] - nxe - 1
    n[:, 6] = n[:, 5] + 1
    n[:, 7] = n[:, 4] + 1
    c = np.empty((nels, 8, 3), dtype=np.float64)
    for l in (0, 1, 4, 5):
        c[:, l, 0] = (ip - 1) * aa
    for l in (2, 3, 6, 7):
        c[:, l, 0] = ip * aa
    c[:, 0:4, 1] = ((iq - 1) * bb)[:, None]
    c[:, 4:8, 1] = (iq * bb)[:, None]
    for l in (0, 3, 4, 7):
        c[:, l, 2] = -is_ * cc
    for l in (1, 2, 5, 6):
        c[:, l, 2] = -(is_ - 1) * cc
    nn = (nxe + 1) * (nye + 1) * (nze + 1)
    coords = np.zeros((nn, 3), dtype=np.float64)
    coords[n.reshape(-1) - 1] = c.reshape(-1, 3)
    return Mesh(8, n.astype(np.int32), coords, nxe, nye, nze, aa, bb, cc)


# Abaqus <-> Smith & Griffiths node order.  ABAQUS2SG
# (``new_library.f90:3578-3619``) sets sg(:, k) = abq(:, perm[k]).
_ABQ2SG = {
    20: np.array([4, 12, 1, 9, 2, 10, 3, 11, 20, 17, 18, 19, 8, 16, 5, 13,
                  6, 14, 7, 15]) - 1,
    8: np.array([1, 5, 6, 2, 4, 8, 7, 3]) - 1,
}


def abaqus_to_sg(g_num: np.ndarray) -> np.ndarray:
    """Element node lists read in Abaqus order -> S&G order (meshgen == 2)."""
    return np.ascontiguousarray(g_num[:, _ABQ2SG[g_num.shape[1]]])


def sg_to_abaqus(g_num: np.ndarray) -> np.ndarray:
    inv = np.argsort(_ABQ2SG[g_num.shape[1]])
    return np.ascontiguousarray(g_num[:, inv])


def cube_bc20(nxe: int, nye: int, nze: int) -> np.ndarray:
    """Symmetry/roller restraints of the p121 demonstrator, CUBE_BC20
    (``geometry.f90:425-500``).  Returns rest (nr, 4): node, fx, fy, fz with
    1 = free, 0 = fixed, ascending node ids."""
    face1 = 3 * nxe * nze + 2 * (nxe + nze) + 1
    face2 = (nxe + 1) * (nze + 1)
    face = face1 + face2
    l = nze * (nxe + 1)
    m = 3 * nxe + 2
    n = 3 * nxe * nze + 2 * nze
    out = []

    def on_side(k):
        return ((k + m - 1) % m == 0 or (k + nxe + 1) % m == 0
                or (k + nxe) % m == 0 or k % m == 0)

    for i in range(0, nye + 1):
        for j in range(i * face + 1, i * face + face1 + 1):
            k = j - i * face
            if i == 0 or i == nye:
                if k <= n and on_side(k):
                    out.append((j, 0, 0, 1))
                elif k <= n:
                    out.append((j, 1, 0, 1))
                else:
                    out.append((j, 0, 0, 0))
            else:
                if k <= n and on_side(k):
                    out.append((j, 0, 1, 1))
                elif k > n:
                    out.append((j, 0, 0, 0))
        if i < nye:
            for j in range(i * face + face1 + 1, (i + 1) * face + 1):
                k = j - (face1 + i * face)
                if k <= l and ((k + nxe) % (nxe + 1) == 0 or k % (nxe + 1) == 0):
                    out.append((j, 0, 1, 1))
                elif k > l:
                    out.append((j, 0, 0, 0))
    return np.array(out, dtype=np.int32).reshape(-1, 4)


def load_p121(nle: int, nod: int, nxe: int, nze: int):
    """Loaded nodes and relative values of LOAD_P121
    (``src/modules/mpi/loading.f90:386-546``)."""
    node, val = [], []
    if nod == 20:
        f1 = (2 * nxe + 1) * (nze + 1) + (nxe + 1) * nze
        f2 = (nxe + 1) * (nze + 1)
        for i in range(1, 2 * nle + 2):
            node.append(i)
            val.append(-1.0 if i in (1, 2 * nle + 1) else (4.0 if i % 2 == 0 else -2.0))
        for j in range(nle):
            for i in range(1, nle + 2):
                node.append(i + f1 + j * (f1 + f2))
                val.append(4.0 if i in (1, nle + 1) else 8.0)
            for i in range(1, 2 * nle + 2):
                node.append(i + f1 + (j + 1) * f2 + j * f1)
                end = i in (1, 2 * nle + 1)
                if j != nle - 1:
                    val.append(-2.0 if end else (8.0 if i % 2 == 0 else -4.0))
                else:
                    val.append(-1.0 if end else (4.0 if i % 2 == 0 else -2.0))
    elif nod == 8:
        f1 = (nxe + 1) * (nze + 1)
        for i in range(1, nle + 2):
            node.append(i)
            val.append(-6.25 if i in (1, nle + 1) else -12.5)
        for j in range(nle):
            for i in range(1, nle + 2):
                node.append(i + (j + 1) * f1)
                end = i in (1, nle + 1)
                if j != nle - 1:
                    val.append(-12.5 if end else -25.0)
                else:
                    val.append(-6.25 if end else -12.5)
    else:
        raise ValueError("load_p121 handles nod = 8 or 20")
    return np.array(node, dtype=np.int32), np.array(val, dtype=np.float64)


def base_fixed_rest(mesh: Mesh, nodof: int = 3) -> np.ndarray:
    """rest rows fixing every dof of the nodes on the base z = -nze*cc."""
    zbase = -mesh.nze * mesh.cc
    nodes = np.nonzero(np.abs(mesh.coords[:, 2] - zbase) <= 1e-9 * max(1.0, abs(zbase)))[0] + 1
    rest = np.zeros((nodes.size, nodof + 1), dtype=np.int32)
    rest[:, 0] = nodes
    return rest


def top_pressure_loads(mesh: Mesh, q) -> tuple[np.ndarray, np.ndarray]:
    """Consistent nodal loads of a pressure on the top face z = 0.

    ``q`` is the z-traction (negative = downwards), either a scalar or an
    array (ncases,) / a callable is not supported; for a per-node traction
    field use :func:`face_traction_loads`.  For hex20 faces (8-node
    serendipity) the consistent loads are -A/12 at corners and A/3 at
    mid-sides (the reference's ``load_p121`` ratios -1:4); for hex8 faces
    A/4 per corner.  Returns (nodes (nl,), values (nl, 3)).
    """
    top_local, corner_w, mid_w = _top_face(mesh.nod)
    A = mesh.aa * mesh.bb
    is_ = _elem_ijk(mesh.nels, mesh.nxe, mesh.nze)[2]
    top = np.nonzero(is_ == 1)[0]
    fz = np.zeros(mesh.nn, dtype=np.float64)
    for l, w in zip(top_local, corner_w + mid_w):
        np.add.at(fz, mesh.g_num[top, l] - 1, w * A)
    nodes = np.nonzero(fz != 0.0)[0]
    val = np.zeros((nodes.size, 3), dtype=np.float64)
    val[:, 2] = fz[nodes] * q
    return (nodes + 1).astype(np.int32), val


def _top_face(nod):
    if nod == 20:
        # z = 0 face of a GEOMETRY_20BXZ element: corners 3,5,15,17,
        # mid-sides 4,16,10,11 (1-based S&G local numbers)
        loc = [2, 4, 14, 16, 3, 15, 9, 10]
        return loc, [-1.0 / 12] * 4, [1.0 / 3] * 4
    if nod == 8:
        return [1, 2, 5, 6], [0.25] * 4, []
    raise ValueError(nod)


def face_traction_loads(mesh: Mesh, node_traction: np.ndarray):
    """Consistent loads of a traction field given at top-face nodes.

    ``node_traction`` is (nn, 3, k): the traction vector per node and load
    case; it is interpolated with the face shape functions, so each element
    face contributes sum_m W_lm t_m.  Used for the multi-RHS workload where
    traction components are drawn per node.
    Returns f (nn, 3, k) nodal loads.
    """
    if mesh.nod != 20:
        raise ValueError("hex20 only")
    W = _serendipity_face_mass(mesh.aa, mesh.bb)
    loc = [2, 3, 4, 10, 16, 15, 14, 9]   # counter-clockwise face ring
    is_ = _elem_ijk(mesh.nels, mesh.nxe, mesh.nze)[2]
    top = np.nonzero(is_ == 1)[0]
    k = node_traction.shape[2]
    f = np.zeros((mesh.nn, 3, k), dtype=np.float64)
    ids = mesh.g_num[top][:, loc] - 1            # (ntop, 8)
    t = node_traction[ids]                        # (ntop, 8, 3, k)
    contrib = np.einsum("lm,emdk->eldk", W, t)
    np.add.at(f, ids.reshape(-1), contrib.reshape(-1, 3, k))
    return f


def _serendipity_face_mass(a, b):
    """W_lm = int N_l N_m dA over an a x b 8-node serendipity face, ring order
    corner, mid, corner, mid, ... (3x3 Gauss, exact for degree-4 products)."""
    ring = [(-1, -1), (0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0)]
    g = np.sqrt(0.6)
    pts = [-g, 0.0, g]
    wts = [5 / 9, 8 / 9, 5 / 9]
    W = np.zeros((8, 8))
    for xi, wx in zip(pts, wts):
        for et, wy in zip(pts, wts):
            N = np.empty(8)
            for l, (x0, y0) in enumerate(ring):
                if x0 != 0 and y0 != 0:
                    N[l] = 0.25 * (1 + xi * x0) * (1 + et * y0) * (xi * x0 + et * y0 - 1)
                elif x0 == 0:
                    N[l] = 0.5 * (1 - xi * xi) * (1 + et * y0)
                else:
                    N[l] = 0.5 * (1 + xi * x0) * (1 - et * et)
            W += wx * wy * np.outer(N, N)
    return W * (a * b / 4.0)


def face_flux_loads(mesh: Mesh, x_face: float, flux: float = 1.0):
    """Consistent nodal flux on the plane x = x_face of a hex8 mesh
    (bilinear faces: A/4 per face corner).  Returns (nodes, values)."""
    if mesh.nod != 8:
        raise ValueError("hex8 only")
    on = np.abs(mesh.coords[:, 0] - x_face) <= 1e-9 * max(1.0, abs(x_face))
    ip = _elem_ijk(mesh.nels, mesh.nxe, mesh.nze)[0]
    sel = np.nonzero(ip == mesh.nxe if x_face > 0 else ip == 1)[0]
    A = mesh.bb * mesh.cc
    f = np.zeros(mesh.nn)
    for l in range(8):
        ids = mesh.g_num[sel, l] - 1
        mask = on[ids]
        np.add.at(f, ids[mask], 0.25 * A * flux)
    nodes = np.nonzero(f != 0.0)[0]
    return (nodes + 1).astype(np.int32), f[nodes]


def steering(nn: int, nodof: int, rest: np.ndarray | None):
    """Node freedom table: equation numbers by ascending node then freedom,
    skipping restrained freedoms (0).  This is what REARRANGE + FIND_G3
    (``new_library.f90:3059-3212``) and REARRANGE_2 + FIND_G4
    (``:3118-3124,3249-3271``) produce when they are in their correct regime;
    nodes absent from ``rest`` are free in every direction.
    Returns (nf (nn, nodof) int32, neq)."""
    free = np.ones((nn, nodof), dtype=bool)
    if rest is not None and len(rest):
        rest = np.asarray(rest)
        free[rest[:, 0] - 1] = rest[:, 1:1 + nodof] != 0
    nf = np.zeros((nn, nodof), dtype=np.int64)
    flat = free.reshape(-1)
    nf.reshape(-1)[flat] = np.arange(1, int(flat.sum()) + 1)
    return nf.astype(np.int32), int(flat.sum())


def g_g_from(nf: np.ndarray, g_num: np.ndarray) -> np.ndarray:
    """Element steering ``g_g_pp(ntot, nels)`` as (nels, nod*nodof)."""
    nodof = nf.shape[1]
    return np.ascontiguousarray(nf[g_num - 1].reshape(g_num.shape[0], g_num.shape[1] * nodof))


def nodal_loads_to_rhs(nf: np.ndarray, nodes: np.ndarray, vals: np.ndarray, neq: int):
    """``LOAD`` (``src/modules/mpi/loading.f90:36-142``): r(eq) = val for the
    free freedoms of loaded nodes (assignment, not accumulation)."""
    r = np.zeros(neq, dtype=np.float64)
    vals = np.asarray(vals, dtype=np.float64).reshape(len(nodes), -1)
    eqs = nf[np.asarray(nodes) - 1]
    m = eqs > 0
    r[eqs[m] - 1] = vals[m]
    return r


# --------------------------------------------------------------------------
# Deterministic random numbers (documented so that any language can redo
# them): SplitMix64 stream -> uniform (0,1) with 53-bit mantissa -> Box-Muller.
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64_uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n uniforms in (0,1): u_k = ((z_k >> 11) + 0.5) * 2**-53 with z_k the
    SplitMix64 output for state seed*2**32 + stream*2**48 + k."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * (1 << 32) + stream * (1 << 48)) & _M64)
        x = base + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def lognormal(seed: int, n: int, mu: float, sigma: float) -> np.ndarray:
    """exp(mu + sigma*N(0,1)) with N from Box-Muller on two SplitMix64 streams."""
    u1 = splitmix64_uniform(seed, n, 0)
    u2 = splitmix64_uniform(seed, n, 1)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return np.exp(mu + sigma * z)


# --------------------------------------------------------------------------
# Deck I/O (list-directed ParaFEM formats)
# --------------------------------------------------------------------------

def _tokens(path):
    with open(path) as fh:
        for line in fh:
            for t in shlex.split(line.replace(",", " "), posix=True):
                yield t


def read_deck(basename: str, program: str = "p121") -> dict:
    """Read a p121/p123/xx3-style deck: ``.dat``, ``.d``, ``.bnd``, ``.lds``
    (and ``.fix`` for p123).  Parsing follows READ_P121
    (``src/modules/mpi/input.f90:3234-3396``), READ_P123 (``:3620-3806``),
    READ_G_NUM_PP (``:935-1080``), READ_G_COORD_PP (``:288-445``),
    READ_REST (``:2570-2632``), READ_LOADS (``:2350-2411``), READ_FIXED
    (``:2483-2564``).  g_num is returned in S&G order (ABAQUS2SG applied when
    meshgen == 2)."""
    toks = list(_tokens(basename + ".dat"))
    out = {}
    if toks[0].lower() in ("gpu", "cpu"):
        out["xpu"] = toks.pop(0)
    if program == "p121":
        keys = ["element", "meshgen", "partitioner", "nels", "nn", "nr", "nip",
                "nod", "loaded_nodes", "e", "v", "tol", "limit"]
    elif program == "p123":
        keys = ["element", "meshgen", "partitioner", "nels", "nn", "nr", "nip",
                "nod", "loaded_nodes", "fixed_freedoms", "kx", "ky", "kz", "tol",
                "limit", "nres"]
    elif program == "p122":
        # READ_P122 (input.f90:3402-3582), then READ_QINC's incs values
        keys = ["element", "meshgen", "partitioner", "nels", "nn", "nr", "nip", "nod", "fixed_freedoms",
                "loaded_nodes", "phi", "c", "psi", "e", "v", "incs", "plasits", "cjits", "plastol", "cjtol"]
    elif program == "p124":
        # READ_P124 (input.f90:3997-4169)
        keys = ["element", "meshgen", "partitioner", "np_types", "nels", "nn", "nr", "nip",
                "nod", "loaded_nodes", "fixed_freedoms", "val0", "dtim", "nstep", "npri",
                "theta", "tol", "limit", "nres"]
    else:
        raise ValueError(program)
    for k, t in zip(keys, toks):
        out[k] = t
    ints = ("meshgen", "partitioner", "np_types", "nels", "nn", "nr", "nip", "nod", "loaded_nodes",
            "fixed_freedoms", "limit", "nres", "nstep", "npri", "incs", "plasits", "cjits")
    for k in keys[1:]:
        v = out[k]
        out[k] = int(float(v)) if k in ints else float(v)
    if program == "p122":
        out["qinc"] = np.array([float(t) for t in toks[len(keys):len(keys) + out["incs"]]])
    nod, nels, nn = out["nod"], out["nels"], out["nn"]
    coords = np.zeros((nn, 3))
    g_num = np.zeros((nels, nod), dtype=np.int32)
    with open(basename + ".d") as fh:
        lines = fh.read().splitlines()
    i = 0
    while not lines[i].strip().upper().startswith("*NODES"):
        i += 1
    i += 1
    arr = np.array(" ".join(lines[i:i + nn]).split(), dtype=np.float64).reshape(nn, 4)
    coords[arr[:, 0].astype(np.int64) - 1] = arr[:, 1:4]
    i += nn
    while not lines[i].strip().upper().startswith("*ELEMENTS"):
        i += 1
    i += 1
    arr = np.array(" ".join(lines[i:i + nels]).split(), dtype=np.int64).reshape(nels, nod + 5)
    g_num[:] = arr[:, 4:4 + nod]
    out["etype"] = arr[:, 4 + nod].astype(np.int32)
    if out["meshgen"] == 2:
        g_num = abaqus_to_sg(g_num)
    out["g_num"] = g_num
    out["coords"] = coords
    nodof = 3 if program in ("p121", "p122") else 1
    if out["nr"] > 0:
        rest = np.array(open(basename + ".bnd").read().split(), dtype=np.int64)
        out["rest"] = rest.reshape(out["nr"], nodof + 1).astype(np.int32)
    else:
        out["rest"] = np.zeros((0, nodof + 1), dtype=np.int32)
    nl = out["loaded_nodes"]
    if nl > 0:
        a = np.array(open(basename + ".lds").read().split(), dtype=np.float64)
        a = a[: nl * (nodof + 1)].reshape(nl, nodof + 1)
        out["load_node"] = a[:, 0].astype(np.int32)
        out["load_val"] = a[:, 1:]
    if program in ("p122", "p123", "p124") and out.get("fixed_freedoms", 0) > 0:
        a = np.array(open(basename + ".fix").read().split(), dtype=np.float64)
        a = a.reshape(out["fixed_freedoms"], 3)
        out["fix_node"] = a[:, 0].astype(np.int32)
        out["fix_sense"] = a[:, 1].astype(np.int32)
        out["fix_val"] = a[:, 2]
    if program == "p124":
        # READ_MATERIAL (input.f90:3067-3102): keyword nmats nvals, a skipped
        # line, then per type its number and nvals values
        with open(basename + ".mat") as fh:
            mlines = fh.read().splitlines()
        nmats, nvals = int(mlines[0].split()[1]), int(mlines[0].split()[2])
        vals = np.array(" ".join(mlines[2:]).split(), dtype=np.float64)[: nmats * (nvals + 1)]
        out["prop"] = vals.reshape(nmats, nvals + 1)[:, 1:]
    return out


def write_p121_deck(basename: str, mesh: Mesh, rest, load_node, load_val,
                    e: float, v: float, tol: float, limit: int, nip: int = 8,
                    abaqus: bool = True) -> None:
    """Write a p121 deck in p12meshgen's formats
    (``src/tools/preprocessing/p12meshgen/p12meshgen.f90:250-323``)."""
    os.makedirs(os.path.dirname(os.path.abspath(basename)), exist_ok=True)
    g = sg_to_abaqus(mesh.g_num) if abaqus else mesh.g_num
    with open(basename + ".d", "w") as fh:
        fh.write("*THREE_DIMENSIONAL\n*NODES\n")
        for i, (x, y, z) in enumerate(mesh.coords, 1):
            fh.write(f"{i:12d}{x:14.6E}{y:14.6E}{z:14.6E}\n")
        fh.write("*ELEMENTS\n")
        for i, row in enumerate(g, 1):
            fh.write(f"{i:12d} 3 {mesh.nod} 1 " + "".join(f"{n:12d}" for n in row) + " 1\n")
    with open(basename + ".bnd", "w") as fh:
        for r in rest:
            fh.write(f"{r[0]:15d}" + "".join(f"{x:6d}" for x in r[1:]) + "\n")
    with open(basename + ".lds", "w") as fh:
        for n, val in zip(load_node, np.asarray(load_val).reshape(len(load_node), -1)):
            fh.write(f"{n:12d}" + "".join(f"{x:16.8E}" for x in val) + "\n")
    with open(basename + ".dat", "w") as fh:
        fh.write("'hexahedron'\n")
        fh.write("2\n" if abaqus else "1\n")
        fh.write("1\n")
        fh.write(f"{mesh.nels:12d}{mesh.nn:12d}{len(rest):12d}{nip:5d}{mesh.nod:5d}{len(load_node):9d}\n")
        fh.write(f"{e:12.4E}{v:12.4E}{tol:12.4E}{limit:8d}\n")


def box_bc8_rest(nxe: int, nye: int, nze: int) -> np.ndarray:
    """BOX_BC8 (``src/modules/shared/geometry.f90:655-696``) for the
    GEOMETRY_8BXZ numbering: in each of the nye layers of (nxe+1)(nze+1)
    nodes, the first row (z = 0) and the last node of every row (x = max);
    then the whole last layer (y = max).  Rows (node, 0), ascending node."""
    face = (nxe + 1) * (nze + 1)
    nodes = []
    j = np.arange(1, face + 1)
    for i in range(nye):
        jj = i * face + j
        on = (jj % (nxe + 1) == 0) | (jj < i * face + nxe + 1)
        nodes.append(jj[on])
    nodes.append(nye * face + j)
    nodes = np.concatenate(nodes)
    return np.stack([nodes, np.zeros_like(nodes)], axis=1).astype(np.int32)


def write_p123_deck(basename: str, mesh: Mesh, rest, kx: float, ky: float, kz: float, tol: float,
                    limit: int, nres: int, loads=None, fix=None, nip: int = 8, abaqus: bool = True) -> None:
    """A p123 deck in the formats READ_P123, READ_G_NUM_PP, READ_G_COORD_PP,
    READ_REST, READ_LOADS and READ_FIXED read (input.f90:3620-3806, 935-1080,
    288-445, 2570-2632, 2350-2411, 2483-2564), laid out as p12meshgen's p123
    case writes them (p12meshgen.f90:706-800).  loads: (equation, value)
    rows; fix: (node, sense, value) rows."""
    os.makedirs(os.path.dirname(os.path.abspath(basename)), exist_ok=True)
    g = sg_to_abaqus(mesh.g_num) if abaqus else mesh.g_num
    with open(basename + ".d", "w") as fh:
        fh.write("*THREE_DIMENSIONAL\n*NODES\n")
        for i, (x, y, z) in enumerate(mesh.coords, 1):
            fh.write(f"{i:12d}{x:14.6E}{y:14.6E}{z:14.6E}\n")
        fh.write("*ELEMENTS\n")
        for i, row in enumerate(g, 1):
            fh.write(f"{i:12d} 3 8 1 " + "".join(f"{n:10d}" for n in row) + " 1\n")
    with open(basename + ".bnd", "w") as fh:
        for r in rest:
            fh.write(f"{r[0]:8d}" + "".join(f"{x:6d}" for x in r[1:]) + "\n")
    nl = 0 if loads is None else len(loads)
    nfix = 0 if fix is None else len(fix)
    if nl:
        with open(basename + ".lds", "w") as fh:
            for eq, val in loads:
                fh.write(f"{int(eq):10d}{val:16.8E}\n")
    if nfix:
        with open(basename + ".fix", "w") as fh:
            for node, sense, val in fix:
                fh.write(f"{int(node):10d}{int(sense):6d}{val:16.8E}\n")
    with open(basename + ".dat", "w") as fh:
        fh.write("'hexahedron'\n" + ("2" if abaqus else "1") + "\n1\n")
        fh.write(f"{mesh.nels:12d}{mesh.nn:12d}{len(rest):12d}{nip:6d}{8:6d}{nl:6d}{nfix:6d}\n")
        fh.write(f"{kx:12.4E}{ky:12.4E}{kz:12.4E}{tol:12.4E}{limit:8d}{nres:8d}\n")


def p124_box_rest(mesh: Mesh) -> np.ndarray:
    """The restraint of the p124 demo cube: temperature held (freedom 0) on
    the faces x = max, y = max and z = 0."""
    c = mesh.coords
    on = (np.isclose(c[:, 0], c[:, 0].max()) | np.isclose(c[:, 1], c[:, 1].max())
          | np.isclose(c[:, 2], 0.0))
    nodes = np.nonzero(on)[0] + 1
    return np.stack([nodes, np.zeros_like(nodes)], axis=1).astype(np.int32)


def write_p124_deck(basename: str, mesh: Mesh, rest, prop, etype, val0: float, dtim: float, nstep: int,
                    npri: int, theta: float, tol: float, limit: int, nres: int, nip: int = 8,
                    abaqus: bool = True, fix=None, loads=None) -> None:
    """A p124 deck in the formats READ_P124, READ_ELEMENTS, READ_G_COORD_PP,
    READ_REST, READ_MATERIAL, READ_FIXED and READ_LOADS read (input.f90:
    3997-4169, 1434-1583, 288-445, 2570-2632, 3067-3102, 2483-2564,
    2350-2411).  fix: (node, sense, value) rows; loads: (node, value) rows."""
    os.makedirs(os.path.dirname(os.path.abspath(basename)), exist_ok=True)
    g = sg_to_abaqus(mesh.g_num) if abaqus else mesh.g_num
    et = np.asarray(etype, dtype=np.int64)
    with open(basename + ".d", "w") as fh:
        fh.write("*THREE_DIMENSIONAL\n*NODES\n")
        for i, (x, y, z) in enumerate(mesh.coords, 1):
            fh.write(f"{i:12d}{x:14.6E}{y:14.6E}{z:14.6E}\n")
        fh.write("*ELEMENTS\n")
        for i, row in enumerate(g, 1):
            fh.write(f"{i:12d} 3 {mesh.nod} 1 " + "".join(f"{n:12d}" for n in row) + f" {et[i - 1]}\n")
    with open(basename + ".bnd", "w") as fh:
        for r in rest:
            fh.write(f"{r[0]:12d}" + "".join(f"{x:6d}" for x in r[1:]) + "\n")
    pr = np.atleast_2d(np.asarray(prop, dtype=np.float64))
    with open(basename + ".mat", "w") as fh:
        fh.write(f"*MATERIAL {pr.shape[0]:4d} {pr.shape[1]:4d}\n<material>\n")
        for i, row in enumerate(pr, 1):
            fh.write(f"{i:2d}" + "".join(f"{x:12.4E}" for x in row) + "\n")
    nfix = 0 if fix is None else len(fix)
    nld = 0 if loads is None else len(loads)
    if nfix:
        with open(basename + ".fix", "w") as fh:
            for n, sense, v in fix:
                fh.write(f"{int(n):12d}{int(sense):6d}{v:16.8E}\n")
    if nld:
        with open(basename + ".lds", "w") as fh:
            for n, v in loads:
                fh.write(f"{int(n):12d}{v:16.8E}\n")
    with open(basename + ".dat", "w") as fh:
        fh.write("'hexahedron'\n")
        fh.write("2\n" if abaqus else "1\n")
        fh.write("1\n")
        fh.write(f"{pr.shape[0]}\n")
        fh.write(f"{mesh.nels:12d}{mesh.nn:12d}{len(rest):12d}{nip:5d}{mesh.nod:5d}{nld:9d}{nfix:9d}\n")
        fh.write(f"{val0:12.4E}\n")
        fh.write(f"{dtim:12.4E}{nstep:8d}{npri:8d}{theta:12.4E}\n")
        fh.write(f"{tol:12.4E}{limit:8d}{nres:8d}\n")


def fortran_e(value: float, w: int = 12, d: int = 5) -> str:
    """``value`` as Fortran's Ew.d edit descriptor writes it (no scale
    factor): optional minus, ``0.``, d digits rounded to nearest, then
    ``E+xx`` (``+xxx`` for exponents of three digits), right-justified in w
    characters (w asterisks when it does not fit)."""
    v = float(value)
    if v != v or v in (float("inf"), float("-inf")):
        body = "NaN" if v != v else ("Infinity" if v > 0 else "-Infinity")
        return body.rjust(w) if len(body) <= w else "*" * w
    neg = np.signbit(v)
    if v == 0.0:
        digits, e10 = "0" * d, 0
    else:
        m, ex = f"{abs(v):.{d - 1}e}".split("e")
        digits, e10 = m.replace(".", ""), int(ex) + 1
    exp = f"E{e10:+03d}" if abs(e10) <= 99 else f"{e10:+04d}"
    body = ("-" if neg else "") + "0." + digits + exp
    if len(body) > w:
        body = body.replace("0.", ".", 1)
    return body.rjust(w) if len(body) <= w else "*" * w


ENSI_VECTOR_HEADER = ("Alya Ensight Gold --- Vector per-node variable file", "part", "     1", "coordinates")


def write_ensi_vector(path: str, values, d: int = 5) -> None:
    """A per-node vector in the EnSight Gold variable file p121 writes
    (``p121.f90:125-137`` with ``DISMSH_ENSI_P``, ``output.f90:2983-3111``):
    four header lines, then each component for all nodes, one value per line
    in Fortran's E12.d."""
    a = np.asarray(values, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write("\n".join(ENSI_VECTOR_HEADER) + "\n")
        for comp in range(a.shape[1]):
            fh.write("".join(fortran_e(x, 12, d) + "\n" for x in a[:, comp]))


def read_ensi_vector(path: str, ndim: int = 3) -> np.ndarray:
    """(nn, ndim) values of a file write_ensi_vector (or p121) wrote."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    vals = np.array([float(t) for t in lines[len(ENSI_VECTOR_HEADER):] if t.strip()])
    return vals.reshape(ndim, -1).T.copy()

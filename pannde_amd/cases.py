"""Synthetic RSG jobs: hexahedral meshes, materials, sources and sensors.

The conventions follow the reference's case makers and drivers (paths relative to the
reference tree):

* node numbering loops x outermost, then y, then z innermost, and cells are emitted in
  the same loop order with vertices in CGNS order [000,100,110,010,001,101,111,011]
  (py_utils/Case_Makers/Composite_Laminate_Transducer_Case_Maker.py:138-152;
  modules/Stubs/include/PlateBuilder.hpp:112-170);
* density and C11..C66 are CELL arrays (upper-triangle order C11,C12,..,C66, the
  ``combinations_with_replacement`` order of the case makers, :160-172);
* a ply stiffness is rotated about z as a fourth-order tensor (rotateStiffnessMatrix,
  :70-83);
* the source signal is the case makers' windowed tone burst: t = arange(0, steps*dt, dt),
  sin(2 pi f t + phase) * hann(steps), normalised to a maximum of 1 (:193-201);
* dt is a fraction of ``1/(c (1/dx+1/dy+1/dz))`` (py_utils/ParameterCalc.py:80);
* the transducer is HostNormalXDZ(center, radius, dz of cell 0, signal) driving the z
  force (applications/include/Boss.hpp:502-513,530-536;
  applications/source/CaseRunners/TransducerModel.cpp:185-198).

Nothing here touches a GPU; the arrays feed ``pannde_amd.rsg.DeviceSolver``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import List, Optional, Sequence

import numpy as np

FIELD_NODE, FIELD_CELL, FIELD_CONSTANT = 0, 1, 2

# (i, j) of the 21 stored stiffness entries, 0-based Voigt indices.
CIJ_PAIRS = [(i, j) for i in range(6) for j in range(i, 6)]
CIJ_NAMES = ["C%d%d" % (i + 1, j + 1) for i, j in CIJ_PAIRS]


@dataclasses.dataclass
class Disc:
    """HostNormalXDZ: circular window of ``radius`` around ``center`` and
    |z - zc| <= 0.75*ds, times a HostUnivariate signal (t, v)."""

    component: int
    center: Sequence[float]
    radius: float
    ds: float
    signal_t: np.ndarray
    signal_v: np.ndarray


@dataclasses.dataclass
class Field:
    type: int
    data: np.ndarray


@dataclasses.dataclass
class Case:
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    cells: np.ndarray  # int32 [nc, 8]
    rho: Field
    cij: List[Field]  # 21 entries
    dt: float
    nsteps: int
    force: List[Disc] = dataclasses.field(default_factory=list)
    moment: List[Disc] = dataclasses.field(default_factory=list)
    moment_form: int = 0
    sensors: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    seed: int = 0
    meta: dict = dataclasses.field(default_factory=dict)  # generator parameters (crack, ...)

    @property
    def n_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])


# ---------------------------------------------------------------- meshes

_NATIVE = None  # libpannde_cases.so (the native mesh builder), False when it is not built


def _native():
    global _NATIVE
    if _NATIVE is None:
        try:
            lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpannde_cases.so"))
        except OSError:
            _NATIVE = False
            return None
        i32, p = ctypes.c_int32, ctypes.c_void_p
        lib.cg_mesh_count.argtypes = [i32, i32, i32, p, p, p]
        lib.cg_mesh_fill.argtypes = [i32, i32, i32, p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     p, p, p, p, p, p, p]
        _NATIVE = lib
    return _NATIVE or None


def structured_mesh(nx, ny, nz, dx, dy, dz, keep=None, origin=(0.0, 0.0, 0.0), cell_grid=False):
    """Right-hex grid of nx*ny*nz cells in case-maker numbering.

    ``keep`` (bool [nx, ny, nz]) removes cells; nodes that no kept cell uses are dropped
    and the rest renumbered in the same loop order.  Coordinates are ``origin + k*d``.
    Returns (x, y, z, cells, cell_ijk), and with ``cell_grid`` also the [nx, ny, nz] int32
    grid of cell numbers (-1 where removed).  The arrays are built by the native builder
    (pannde_amd/csrc/case_gen.cpp) when it is built, else by structured_mesh_numpy; the two
    are equal array for array."""
    if keep is None:
        keep = np.ones((nx, ny, nz), dtype=bool)
    keep = np.ascontiguousarray(keep, dtype=bool)
    lib = _native()
    if lib is None:
        return structured_mesh_numpy(nx, ny, nz, dx, dy, dz, keep, origin, cell_grid)
    k8 = keep.view(np.uint8)
    nn, nc = ctypes.c_int64(0), ctypes.c_int64(0)
    if lib.cg_mesh_count(nx, ny, nz, k8.ctypes.data, ctypes.addressof(nn), ctypes.addressof(nc)) != 0:
        raise ValueError("grid %dx%dx%d is too large" % (nx, ny, nz))
    x, y, z = (np.empty(nn.value) for _ in range(3))
    cells = np.empty((nc.value, 8), dtype=np.int32)
    ijk = np.empty((nc.value, 3), dtype=np.int64)
    grid = np.empty((nx, ny, nz), dtype=np.int32) if cell_grid else None
    org = np.asarray(origin, dtype=np.float64)
    lib.cg_mesh_fill(nx, ny, nz, k8.ctypes.data, float(dx), float(dy), float(dz), org.ctypes.data, x.ctypes.data,
                     y.ctypes.data, z.ctypes.data, cells.ctypes.data, ijk.ctypes.data,
                     grid.ctypes.data if cell_grid else None)
    return (x, y, z, cells, ijk, grid) if cell_grid else (x, y, z, cells, ijk)


def structured_mesh_numpy(nx, ny, nz, dx, dy, dz, keep=None, origin=(0.0, 0.0, 0.0), cell_grid=False):
    """structured_mesh in numpy: the reference the native builder is tested against."""
    if keep is None:
        keep = np.ones((nx, ny, nz), dtype=bool)
    keep = np.asarray(keep, dtype=bool)
    used = np.zeros((nx + 1, ny + 1, nz + 1), dtype=bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                used[di:di + nx, dj:dj + ny, dk:dk + nz] |= keep
    # C order throughout (x outer, z inner): np.nonzero lists kept entries in that order
    ii, jj, kk = np.nonzero(used)
    nid = np.full(used.shape, -1, dtype=np.int32)
    nid[used] = np.arange(ii.shape[0], dtype=np.int32)
    x = origin[0] + ii * dx
    y = origin[1] + jj * dy
    z = origin[2] + kk * dz
    ci, cj, ck = np.nonzero(keep)
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    cells = np.empty((ci.shape[0], 8), dtype=np.int32)
    for v, (a, b, c) in enumerate(offs):
        cells[:, v] = nid[a:a + nx, b:b + ny, c:c + nz][keep]
    out = (np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64), np.asarray(z, dtype=np.float64),
           cells, np.stack([ci, cj, ck], axis=1))
    if not cell_grid:
        return out
    grid = np.full(keep.shape, -1, dtype=np.int32)
    grid[keep] = np.arange(ci.shape[0], dtype=np.int32)
    return out + (grid,)


def renumber_nodes(x, y, z, cells, new_of_old):
    """Applies a node permutation: node ``o`` becomes ``new_of_old[o]``."""
    inv = np.empty_like(new_of_old)
    inv[new_of_old] = np.arange(new_of_old.shape[0])
    return x[inv], y[inv], z[inv], new_of_old[cells].astype(np.int32)


def rcm_order(n_nodes, cells):
    """Reverse Cuthill-McKee order (scipy) of the graph whose edges are the 12 hex
    edges; returns new_of_old."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    e = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    a = np.concatenate([cells[:, i] for i, _ in e]).astype(np.int64)
    b = np.concatenate([cells[:, j] for _, j in e]).astype(np.int64)
    g = sp.coo_matrix((np.ones(2 * a.shape[0], dtype=np.int8), (np.concatenate([a, b]), np.concatenate([b, a]))),
                      shape=(n_nodes, n_nodes)).tocsr()
    perm = reverse_cuthill_mckee(g, symmetric_mode=True)  # perm[new] = old
    new_of_old = np.empty(n_nodes, dtype=np.int64)
    new_of_old[perm] = np.arange(n_nodes)
    return new_of_old


# ---------------------------------------------------------------- materials

def isotropic_c(lam, mu):
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    for i in range(3):
        c[i, i] = lam + 2 * mu
        c[i + 3, i + 3] = mu
    return c


_VOIGT = np.array([[0, 5, 4], [5, 1, 3], [4, 3, 2]])


def rotate_c_about_z(c6, degrees):
    """C'_ijkl = R_ia R_jb R_kc R_ld C_abcd with R a rotation by ``degrees`` about z."""
    c4 = c6[_VOIGT[:, :, None, None], _VOIGT[None, None, :, :]]
    a = math.radians(degrees)
    r = np.array([[math.cos(a), -math.sin(a), 0.0], [math.sin(a), math.cos(a), 0.0], [0.0, 0.0, 1.0]])
    c4r = np.einsum("ia,jb,kc,ld,abcd->ijkl", r, r, r, r, c4)
    out = np.zeros((6, 6))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    out[_VOIGT[i, j], _VOIGT[k, l]] = c4r[i, j, k, l]
    return out


def cij_cell_fields(c6_per_cell):
    """21 CELL fields from an [nc, 6, 6] stiffness array."""
    return [Field(FIELD_CELL, np.ascontiguousarray(c6_per_cell[:, i, j], dtype=np.float64)) for i, j in CIJ_PAIRS]


def cij_material_fields(c6_per_material, material_of_cell):
    """21 CELL fields from [nmat, 6, 6] stiffnesses and each cell's material index."""
    return [Field(FIELD_CELL, np.ascontiguousarray(c6_per_material[:, i, j], dtype=np.float64)[material_of_cell])
            for i, j in CIJ_PAIRS]


def max_wave_speed(c6, rho):
    return math.sqrt(float(np.max(np.diag(c6)[:3])) / rho)


def cfl_dt(c, dx, dy, dz, fraction=0.9):
    return fraction / (c * (1.0 / dx + 1.0 / dy + 1.0 / dz))


# ---------------------------------------------------------------- signals and sensors

def tone_burst(freq, cycles, dt, phase_deg=0.0):
    """The case makers' Hann-windowed sine burst, normalised to max 1."""
    from scipy.signal import windows

    steps = int(math.ceil((cycles / freq) / dt))
    t = np.arange(0, steps * dt, dt)[:steps]
    wf = np.sin(t * freq * 2 * np.pi + math.radians(phase_deg)) * windows.hann(steps)
    return t.astype(np.float64), (wf / wf.max()).astype(np.float64)


def top_face_nodes(z, tol=1e-12):
    return np.nonzero(np.abs(z - z.max()) <= tol * max(1.0, abs(z.max())))[0]


def seeded_sensors(rng, candidates, count):
    count = min(count, candidates.shape[0])
    return np.sort(rng.choice(candidates, size=count, replace=False)).astype(np.int32)


def cell0_dz(x, y, z, cells):
    """|z(v0) - z(v6)| of cell 0 (Boss.hpp:530-536)."""
    return abs(z[cells[0, 0]] - z[cells[0, 6]])


# ---------------------------------------------------------------- the workloads

def al_block(seed=1, nx=40, ny=40, nz=20, dx=1.0e-4, nsteps=500, n_sensors=16,
             rho_type=FIELD_CELL):
    """Isotropic aluminium block (rho 2700, lambda 51 GPa, mu 26 GPa), traction-free,
    5 MHz 3-cycle burst on a 2 mm disc centred on the top face, dt = 0.9 CFL."""
    rho, lam, mu = 2700.0, 51.0e9, 26.0e9
    x, y, z, cells, _ = structured_mesh(nx, ny, nz, dx, dx, dx)
    c6 = isotropic_c(lam, mu)
    dt = cfl_dt(math.sqrt((lam + 2 * mu) / rho), dx, dx, dx)
    t, v = tone_burst(5.0e6, 3, dt)
    center = (nx * dx / 2, ny * dx / 2, nz * dx)
    disc = Disc(2, center, 2.0e-3 * (nx / 40.0), cell0_dz(x, y, z, cells), t, v)
    rng = np.random.default_rng(seed)
    nc = cells.shape[0]
    if rho_type == FIELD_NODE:
        rho_f = Field(FIELD_NODE, np.full(x.shape[0], rho))
    elif rho_type == FIELD_CONSTANT:
        rho_f = Field(FIELD_CONSTANT, np.array([rho]))
    else:
        rho_f = Field(FIELD_CELL, np.full(nc, rho))
    return Case(name="al_block_%dx%dx%d" % (nx, ny, nz), x=x, y=y, z=z, cells=cells, rho=rho_f,
                cij=cij_cell_fields(np.broadcast_to(c6, (nc, 6, 6))), dt=dt, nsteps=nsteps,
                force=[disc], sensors=seeded_sensors(rng, top_face_nodes(z), n_sensors), seed=seed)


CFRP_PLY = dict(C11=135e9, C22=9.5e9, C33=9.5e9, C12=4.5e9, C13=4.5e9, C23=4.2e9, C44=2.6e9,
                C55=5.0e9, C66=5.0e9)
CFRP_LAYUP = [0, 45, 90, -45, 0, 45, 90, -45, -45, 90, 45, 0, -45, 90, 45, 0]


def cfrp_ply_c():
    p = CFRP_PLY
    c = np.zeros((6, 6))
    c[0, 0], c[1, 1], c[2, 2] = p["C11"], p["C22"], p["C33"]
    c[0, 1] = c[1, 0] = p["C12"]
    c[0, 2] = c[2, 0] = p["C13"]
    c[1, 2] = c[2, 1] = p["C23"]
    c[3, 3], c[4, 4], c[5, 5] = p["C44"], p["C55"], p["C66"]
    return c


def cfrp_plate(seed=2, nx=256, ny=256, cells_per_ply=4, dx=5.0e-5, nsteps=2000, n_holes=5,
               n_sensors=16, renumber=True, xd_radius=2.0e-3, crack=None, xd_center=None):
    """Orthotropic CFRP laminate, [0/45/90/-45]2s, 16 plies of ``cells_per_ply`` cells,
    rho 1580, seeded through-holes of radius 1-3 mm (cells removed, orphan nodes
    dropped), node ids seeded-permuted then RCM-renumbered, 2 MHz 5-cycle burst.
    ``crack`` (a Crack) is opened on the top face before renumbering; ``xd_center`` (x, y)
    moves the transducer from the plate centre (holes keep clear of it)."""
    rng = np.random.default_rng(seed)
    nplies = len(CFRP_LAYUP)
    nz = nplies * cells_per_ply
    lx, ly = nx * dx, ny * dx
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dx
    keep2d = np.ones((nx, ny), dtype=bool)
    center = (lx / 2, ly / 2, nz * dx) if xd_center is None else (xd_center[0], xd_center[1], nz * dx)
    holes = []
    scale = min(lx, ly) / 12.8e-3
    for _ in range(n_holes):
        for _attempt in range(200):
            r = rng.uniform(1.0e-3, 3.0e-3) * scale
            hx, hy = rng.uniform(r + dx, lx - r - dx), rng.uniform(r + dx, ly - r - dx)
            if math.hypot(hx - center[0], hy - center[1]) < r + xd_radius * scale + 4 * dx:
                continue
            if any(math.hypot(hx - a, hy - b) < r + q + 4 * dx for a, b, q in holes):
                continue
            holes.append((hx, hy, r))
            break
    for hx, hy, r in holes:
        keep2d &= (xc[:, None] - hx) ** 2 + (yc[None, :] - hy) ** 2 >= r * r
    keep = np.repeat(keep2d[:, :, None], nz, axis=2)
    x, y, z, cells, ijk, grid = structured_mesh(nx, ny, nz, dx, dx, dx, keep=keep, cell_grid=True)
    n_dup = 0
    if crack is not None:
        x, y, z, cells, n_dup = surface_crack(x, y, z, cells, ijk, (dx, dx, dx), crack, cell_grid=grid)
    ply_c = np.stack([rotate_c_about_z(cfrp_ply_c(), a) for a in CFRP_LAYUP])
    ply_of_cell = ijk[:, 2] // cells_per_ply
    rho = 1580.0
    if renumber:
        perm = rng.permutation(x.shape[0])
        x, y, z, cells = renumber_nodes(x, y, z, cells, perm)
        x, y, z, cells = renumber_nodes(x, y, z, cells, rcm_order(x.shape[0], cells))
    dt = cfl_dt(max_wave_speed(cfrp_ply_c(), rho), dx, dx, dx)
    t, v = tone_burst(2.0e6, 5, dt)
    disc = Disc(2, center, xd_radius * scale, dx, t, v)
    return Case(name="cfrp_plate_%dx%dx%d" % (nx, ny, nz), x=x, y=y, z=z, cells=cells,
                rho=Field(FIELD_CELL, np.full(cells.shape[0], rho)), cij=cij_material_fields(ply_c, ply_of_cell),
                dt=dt, nsteps=nsteps, force=[disc],
                sensors=seeded_sensors(rng, top_face_nodes(z), n_sensors), seed=seed,
                meta=dict(crack=crack, n_dup=n_dup, holes=holes))


@dataclasses.dataclass
class Crack:
    """A planar, surface-breaking crack: vertical plane through ``center`` (on the top
    face), normal at ``angle_deg`` from +x in the x-y plane, ``length`` along the surface
    and ``depth`` below it."""

    center: Sequence[float]
    angle_deg: float
    length: float
    depth: float


def surface_crack(x, y, z, cells, ijk, d, crack: Crack, origin=(0.0, 0.0, 0.0), cell_grid=None):
    """Opens ``crack`` in a right-hex mesh on a grid of spacing d = (dx, dy, dz) by
    duplicating nodes, as a mesh file describes a crack (coincident nodes on the two
    faces, each face's cells using their own copy).

    Every cell gets the side of the crack plane its centre lies on.  A face between two
    cells of opposite sides is cracked when its centre is within the crack rectangle.  A
    node is duplicated when every opposite-side face around it is cracked (so the crack
    tip stays bonded and the mouth at the surface opens); the negative side's cells take
    the copy, appended after the existing nodes.  Returns (x, y, z, cells, n_dup).

    ``cell_grid`` (structured_mesh's grid of cell numbers) finds the region's cells without a
    pass over the whole mesh; the result is the same."""
    dx, dy, dz = d
    th = math.radians(crack.angle_deg)
    nrm = np.array([math.cos(th), math.sin(th), 0.0])
    tan = np.array([-math.sin(th), math.cos(th), 0.0])
    cx, cy, ztop = (float(v) for v in crack.center)
    half = 0.5 * crack.length
    # cell region: the crack rectangle's bounding box plus two cells
    ext = np.abs(tan[:2]) * half
    lo = np.array([cx - ext[0] - 2 * dx, cy - ext[1] - 2 * dy, ztop - crack.depth - 2 * dz])
    hi = np.array([cx + ext[0] + 2 * dx, cy + ext[1] + 2 * dy, ztop + 2 * dz])
    g_lo = np.floor((lo - np.asarray(origin)) / np.array(d)).astype(np.int64)
    g_hi = np.ceil((hi - np.asarray(origin)) / np.array(d)).astype(np.int64)
    if cell_grid is not None:
        # the region clipped to the grid rather than to the kept cells: the extra rows hold
        # no cells, so the faces, and the duplicated nodes in C order, are the same
        g_lo = np.maximum(g_lo, 0)
        g_hi = np.minimum(g_hi, np.array(cell_grid.shape))
        if np.any(g_hi <= g_lo):
            return x, y, z, cells, 0
        shape = tuple(int(v) for v in g_hi - g_lo)
        cgrid = cell_grid[g_lo[0]:g_hi[0], g_lo[1]:g_hi[1], g_lo[2]:g_hi[2]].astype(np.int64)
        li, lj, lk = np.nonzero(cgrid >= 0)  # C order: ascending cell numbers
        loc = np.stack([li, lj, lk], axis=1)
        reg_cells = cgrid[li, lj, lk]
    else:
        g_lo = np.maximum(g_lo, ijk.min(axis=0))
        g_hi = np.minimum(g_hi, ijk.max(axis=0) + 1)
        if np.any(g_hi <= g_lo):
            return x, y, z, cells, 0
        shape = tuple(int(v) for v in g_hi - g_lo)
        inreg = np.all((ijk >= g_lo) & (ijk < g_hi), axis=1)
        reg_cells = np.nonzero(inreg)[0]
        loc = ijk[reg_cells] - g_lo
        cgrid = np.full(shape, -1, dtype=np.int64)
        cgrid[loc[:, 0], loc[:, 1], loc[:, 2]] = reg_cells
    # side of each region cell (+1 / -1) from its centre
    ii, jj, kk = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    gx = origin[0] + (g_lo[0] + ii + 0.5) * dx
    gy = origin[1] + (g_lo[1] + jj + 0.5) * dy
    side = np.where((gx - cx) * nrm[0] + (gy - cy) * nrm[1] >= 0.0, 1, -1)
    exists = cgrid >= 0
    ncrack = np.zeros(tuple(n + 1 for n in shape), dtype=np.int32)
    nbond = np.zeros_like(ncrack)
    for a in range(3):
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[a] = slice(0, shape[a] - 1)
        sl1[a] = slice(1, shape[a])
        sl0, sl1 = tuple(sl0), tuple(sl1)
        opp = exists[sl0] & exists[sl1] & (side[sl0] != side[sl1])
        fi, fj, fk = np.nonzero(opp)
        # face centre: the shared boundary between the two cells
        f = np.stack([fi, fj, fk], axis=1).astype(np.float64) + 0.5
        f[:, a] += 0.5
        fxyz = np.asarray(origin) + (f + g_lo) * np.array(d)
        along = (fxyz[:, 0] - cx) * tan[0] + (fxyz[:, 1] - cy) * tan[1]
        below = ztop - fxyz[:, 2]
        inside = (np.abs(along) <= half) & (below <= crack.depth)
        # the face's 4 nodes in region-node coordinates
        b = np.stack([fi, fj, fk], axis=1)
        b[:, a] += 1
        others = [q for q in range(3) if q != a]
        for d0 in (0, 1):
            for d1 in (0, 1):
                nidx = b.copy()
                nidx[:, others[0]] += d0
                nidx[:, others[1]] += d1
                np.add.at(ncrack, (nidx[inside, 0], nidx[inside, 1], nidx[inside, 2]), 1)
                np.add.at(nbond, (nidx[~inside, 0], nidx[~inside, 1], nidx[~inside, 2]), 1)
    dup_grid = (ncrack > 0) & (nbond == 0)
    if not dup_grid.any():
        return x, y, z, cells, 0
    # node ids of the duplicated grid nodes (through the region's cells' vertices)
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    node_at = np.full(ncrack.shape, -1, dtype=np.int64)
    for v, (a0, a1, a2) in enumerate(offs):
        node_at[loc[:, 0] + a0, loc[:, 1] + a1, loc[:, 2] + a2] = cells[reg_cells, v]
    old = node_at[dup_grid]
    assert np.all(old >= 0)
    nn = x.shape[0]
    new_of = np.full(nn, -1, dtype=np.int64)
    new_of[old] = nn + np.arange(old.shape[0])
    cells = cells.copy()
    neg = reg_cells[side[loc[:, 0], loc[:, 1], loc[:, 2]] < 0]
    sub = cells[neg]
    rep = new_of[sub]
    cells[neg] = np.where(rep >= 0, rep, sub).astype(np.int32)
    return (np.concatenate([x, x[old]]), np.concatenate([y, y[old]]), np.concatenate([z, z[old]]), cells,
            int(old.shape[0]))


def ti_bracket(seed=5, leg=512, width=96, height=96, dx=5.0e-5, nsteps=4000, crack_length=None,
               crack_depth=None, crack_angle=None, n_sensors=16, xd_radius=1.5e-3, xd_center=None):
    """L-shaped Ti-6Al-4V bracket (rho 4430, lambda 90 GPa, mu 44 GPa): leg A is
    leg x width x height cells along x, leg B width x leg x height cells along y from
    y = width, both ``height`` cells tall.  A seeded planar surface crack on leg A's top
    face (length U(0.5, 5) mm, depth U(0.2, 2) mm, orientation U(0, 180) deg unless
    given), modelled by duplicated nodes; a 10 MHz 3-cycle burst on a disc on the top face
    of leg A (at ``xd_center`` (x, y) if given); 16 seeded top-face sensors."""
    rng = np.random.default_rng(seed)
    nx, ny = leg, width + leg
    ci = np.arange(nx)[:, None]
    cj = np.arange(ny)[None, :]
    keep2d = (cj < width) | (ci < width)
    keep = np.repeat(keep2d[:, :, None], height, axis=2)
    x, y, z, cells, ijk, grid = structured_mesh(nx, ny, height, dx, dx, dx, keep=keep, cell_grid=True)
    L = leg * dx
    W = width * dx
    H = height * dx
    length = crack_length if crack_length is not None else rng.uniform(0.5e-3, 5.0e-3)
    depth = crack_depth if crack_depth is not None else rng.uniform(0.2e-3, 2.0e-3)
    depth = min(depth, 0.8 * H)
    angle = crack_angle if crack_angle is not None else rng.uniform(0.0, 180.0)
    ccx = rng.uniform(0.55 * L, 0.8 * L)
    ccy = rng.uniform(0.3 * W, 0.7 * W)
    crack = Crack((ccx, ccy, H), angle, length, depth)
    x, y, z, cells, n_dup = surface_crack(x, y, z, cells, ijk, (dx, dx, dx), crack, cell_grid=grid)
    mat = TI64
    c6 = isotropic_c(mat["lam"], mat["mu"])
    nc = cells.shape[0]
    dt = cfl_dt(math.sqrt((mat["lam"] + 2 * mat["mu"]) / mat["rho"]), dx, dx, dx)
    t, v = tone_burst(10.0e6, 3, dt)
    r = min(xd_radius, 0.4 * W)
    xd = (rng.uniform(0.25 * L, 0.4 * L), 0.5 * W) if xd_center is None else xd_center
    disc = Disc(2, (xd[0], xd[1], H), r, dx, t, v)
    top = top_face_nodes(z)
    case = Case(name="ti_bracket_%dx%dx%d" % (leg, width, height), x=x, y=y, z=z, cells=cells,
                rho=Field(FIELD_CELL, np.full(nc, mat["rho"])),
                cij=[Field(FIELD_CONSTANT, np.array([c6[i, j]])) for i, j in CIJ_PAIRS],
                dt=dt, nsteps=nsteps, force=[disc], sensors=seeded_sensors(rng, top, n_sensors), seed=seed)
    case.meta.update(crack=crack, n_dup=n_dup)
    return case


def water_column_block(seed=3, nx=12, ny=10, nz=6, dx=1.0e-4, nsteps=40):
    """HostUTModel form: isotropic block driven by an XX/YY/ZZ stress (moment) disc on
    the top face, as WaterColumnModel applies it (WaterColumnModel.cpp:127-180)."""
    case = al_block(seed=seed, nx=nx, ny=ny, nz=nz, dx=dx, nsteps=nsteps, n_sensors=8)
    t, v = case.force[0].signal_t, case.force[0].signal_v
    center = (nx * dx / 2, ny * dx / 2, nz * dx)
    r = 0.3 * min(nx, ny) * dx
    ds = case.force[0].ds
    case.moment = [Disc(0, center, r, ds, t, v), Disc(1, center, r, ds, t, 0.5 * v),
                   Disc(2, center, r, ds, t, -0.25 * v)]
    case.force = []
    case.moment_form = 1
    case.name = "water_column_%dx%dx%d" % (nx, ny, nz)
    return case


TI64 = dict(rho=4430.0, lam=90.0e9, mu=44.0e9)
AL = dict(rho=2700.0, lam=51.0e9, mu=26.0e9)


def batch_instances(seed=4, count=64, min_cells=32 ** 3, max_cells=128 ** 3, nsteps=1500, dx=1.0e-4,
                    n_sensors=8):
    """Independent parametric instances for one batched launch: cell counts log-uniform in
    [min_cells, max_cells] with seeded box shapes, isotropic Al or Ti-6Al-4V with a seeded
    +-5 % perturbation of rho, lambda and mu, a 5 MHz 3-cycle burst on a disc at a seeded
    position of the top face, each with its own CFL dt."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        target = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        a = np.exp(rng.uniform(-0.5, 0.5, size=3))
        a /= a.prod() ** (1 / 3)
        n = np.maximum(8, np.round(a * target ** (1 / 3))).astype(int)
        nx, ny, nz = (int(v) for v in n)
        base = AL if rng.uniform() < 0.5 else TI64
        rho, lam, mu = (base[k] * (1 + rng.uniform(-0.05, 0.05)) for k in ("rho", "lam", "mu"))
        x, y, z, cells, _ = structured_mesh(nx, ny, nz, dx, dx, dx)
        nc = cells.shape[0]
        dt = cfl_dt(math.sqrt((lam + 2 * mu) / rho), dx, dx, dx)
        t, v = tone_burst(5.0e6, 3, dt)
        r = min(2.0e-3, 0.3 * min(nx, ny) * dx)
        cx = rng.uniform(r, nx * dx - r)
        cy = rng.uniform(r, ny * dx - r)
        disc = Disc(2, (cx, cy, nz * dx), r, cell0_dz(x, y, z, cells), t, v)
        out.append(Case(name="batch%02d_%dx%dx%d" % (i, nx, ny, nz), x=x, y=y, z=z, cells=cells,
                        rho=Field(FIELD_CELL, np.full(nc, rho)),
                        cij=[Field(FIELD_CONSTANT, np.array([c])) for c in
                             (isotropic_c(lam, mu)[i_, j_] for i_, j_ in CIJ_PAIRS)],
                        dt=dt, nsteps=nsteps, force=[disc],
                        sensors=seeded_sensors(rng, top_face_nodes(z), n_sensors), seed=seed))
    return out

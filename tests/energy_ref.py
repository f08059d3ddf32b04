"""Numpy energy of an RSG state: the checker for rsg_energy (test infrastructure).

kinetic = 1/2 sum_n rho_n dxn dyn dzn |V_n|^2 and strain = 1/2 sum_c vol_c S_c . C_c^-1 S_c,
with rho_n the /8-rule node density (HostField.hpp:142-154), dxn..dzn the nodePrep spacings
(RightHex.hpp:258-309, mirrored bound at a face), vol_c = dxc dyc dzc (cellPrep,
RightHex.hpp:253-256) and C_c the cell's 6x6 stiffness (HostField::atCell for NODE fields,
HostField.hpp:121-133), solved with numpy.linalg.solve.
"""
from __future__ import annotations

import numpy as np

from pannde_amd import cases


def node_cells(n_nodes, cells):
    """HexMesh::buildLinks: slot = vertex ^ 6 (HexMesh.hpp:300-313)."""
    out = np.full((n_nodes, 8), -1, np.int64)
    for v in range(8):
        out[cells[:, v], v ^ 6] = np.arange(cells.shape[0])
    return out


def node_spacings(X, cells, ncl):
    lo, hi = X.copy(), X.copy()
    for kb in range(8):
        c = ncl[:, kb]
        ok = c >= 0
        for kbb in range(8):
            m = cells[c[ok], kbb]
            lo[:, ok] = np.minimum(lo[:, ok], X[:, m])
            hi[:, ok] = np.maximum(hi[:, ok], X[:, m])
    lo2 = np.where(lo == X, 2 * X - hi, lo)
    hi2 = np.where(hi == X, 2 * X - lo2, hi)
    return 0.5 * (hi2 - lo2)


def stiffness(case):
    cells = np.asarray(case.cells).reshape(-1, 8)
    nc = cells.shape[0]
    C = np.zeros((nc, 6, 6))
    for k, (i, j) in enumerate(cases.CIJ_PAIRS):
        f = case.cij[k]
        if f.type == cases.FIELD_CONSTANT:
            v = np.full(nc, f.data[0])
        elif f.type == cases.FIELD_CELL:
            v = np.asarray(f.data)
        else:
            v = np.asarray(f.data)[cells].sum(1) / 8
        C[:, i, j] = v
        C[:, j, i] = v
    return C


def weights(case):
    """(node masses, cell volumes, [nc,6,6] stiffness)."""
    X = np.stack([case.x, case.y, case.z])
    cells = np.asarray(case.cells).reshape(-1, 8)
    ncl = node_cells(X.shape[1], cells)
    dn = node_spacings(X, cells, ncl)
    rho = np.asarray(case.rho.data)
    if case.rho.type == cases.FIELD_NODE:
        rn = rho
    else:
        rc = np.full(cells.shape[0], rho[0]) if case.rho.type == cases.FIELD_CONSTANT else rho
        rn = np.where(ncl >= 0, rc[np.maximum(ncl, 0)], 0.0).sum(1) / 8
    mass = rn * dn[0] * dn[1] * dn[2]
    dc = X[:, cells[:, 6]] - X[:, cells[:, 0]]
    return mass, dc[0] * dc[1] * dc[2], stiffness(case)


def energy(case, V, S, w=None):
    """(kinetic, strain) of state V [3][nn], S [6][nc]."""
    mass, vol, C = weights(case) if w is None else w
    k = 0.5 * float((mass * (np.asarray(V) ** 2).sum(0)).sum())
    x = np.linalg.solve(C, np.asarray(S).T[:, :, None])[:, :, 0]
    u = 0.5 * float((vol * (np.asarray(S).T * x).sum(1)).sum())
    return k, u

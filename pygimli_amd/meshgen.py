"""Synthetic graded tri/tet meshes and survey layouts for the ERT path.

pyGIMLi builds ERT meshes with Triangle/TetGen (pygimli/meshtools,
createParaMesh); those libraries are not part of this build, so meshes are
generated here on graded tensor grids: every quad is split into two
triangles, every hexahedron into six tetrahedra sharing its main diagonal
(Kuhn subdivision, conforming across neighbours).  Markers follow pyGIMLi's
ERT conventions: surface boundary -1 (Neumann), all other outer boundaries -2
(mixed), electrode nodes -99, cell marker = parameter index.
"""
from __future__ import annotations

import numpy as np

from .mesh import (Mesh, MARKER_BOUND_HOMOGEN_NEUMANN, MARKER_BOUND_MIXED,
                   MARKER_NODE_ELECTRODE)


def graded_axis(fixed, h, growth, extent_lo, extent_hi, h_max=None):
    """Sorted 1D node coordinates.

    ``fixed`` points (e.g. electrode positions and interfaces) are kept, the
    span between min(fixed) and max(fixed) is filled at spacing <= h, and the
    spacing grows geometrically by ``growth`` out to extent_lo / extent_hi.
    """
    fixed = np.unique(np.asarray(fixed, np.float64))
    lo, hi = fixed[0], fixed[-1]
    pts = [fixed]
    for a, b in zip(fixed[:-1], fixed[1:]):
        n = int(np.ceil((b - a) / h - 1e-9))
        if n > 1:
            pts.append(np.linspace(a, b, n + 1)[1:-1])
    core = np.unique(np.concatenate(pts))
    out = [core]
    for direction, limit in ((-1, extent_lo), (1, extent_hi)):
        x = lo if direction < 0 else hi
        step = h
        extra = []
        while (x - limit) * direction < 0:
            step *= growth
            if h_max is not None:
                step = min(step, h_max)
            x = x + direction * step
            if (x - limit) * direction >= 0 or abs(x - limit) < 0.3 * step:
                x = limit
            extra.append(x)
        out.append(np.array(extra))
    return np.unique(np.concatenate(out))


def grid_mesh_2d(x, y, electrodes_x=(), cell_marker_fn=None):
    """Triangle mesh on the tensor grid x * y (y <= 0 is the subsurface).

    Boundary at y = max(y) gets marker -1, the other outer edges -2.  Nodes at
    (xe, max(y)) for xe in electrodes_x get marker -99.
    """
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    nx, ny = len(x), len(y)
    X, Y = np.meshgrid(x, y, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    nid = np.arange(nx * ny).reshape(nx, ny)
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    n00 = nid[i, j].ravel()
    n10 = nid[i + 1, j].ravel()
    n01 = nid[i, j + 1].ravel()
    n11 = nid[i + 1, j + 1].ravel()
    # alternate the diagonal so the mesh has no preferred direction
    flip = ((i + j) % 2 == 1).ravel()
    t1 = np.where(flip[:, None], np.stack([n00, n10, n01], 1), np.stack([n00, n10, n11], 1))
    t2 = np.where(flip[:, None], np.stack([n10, n11, n01], 1), np.stack([n00, n11, n01], 1))
    cells = np.concatenate([t1, t2]).astype(np.int32)
    # boundaries
    top = np.stack([nid[:-1, -1], nid[1:, -1]], 1)
    bot = np.stack([nid[:-1, 0], nid[1:, 0]], 1)
    lef = np.stack([nid[0, :-1], nid[0, 1:]], 1)
    rig = np.stack([nid[-1, :-1], nid[-1, 1:]], 1)
    bounds = np.concatenate([top, bot, lef, rig]).astype(np.int32)
    bm = np.concatenate([np.full(len(top), MARKER_BOUND_HOMOGEN_NEUMANN),
                         np.full(len(bot) + len(lef) + len(rig), MARKER_BOUND_MIXED)])
    nm = np.zeros(len(nodes), np.int32)
    for xe in electrodes_x:
        k = int(np.argmin(np.abs(x - xe)))
        if abs(x[k] - xe) > 1e-9:
            raise ValueError(f"electrode x={xe} is not a grid line")
        nm[nid[k, -1]] = MARKER_NODE_ELECTRODE
    cm = np.arange(len(cells), dtype=np.int32)
    if cell_marker_fn is not None:
        cm = cell_marker_fn(nodes[cells].mean(axis=1)).astype(np.int32)
    return Mesh(2, nodes, cells, cm, bounds, bm.astype(np.int32), nm)


# Kuhn subdivision of the unit cube: 6 tets along the (0,0,0)-(1,1,1) diagonal.
# Corner index = ix + 2*iy + 4*iz.
_KUHN = np.array([[0, 1, 3, 7], [0, 1, 5, 7], [0, 2, 3, 7],
                  [0, 2, 6, 7], [0, 4, 5, 7], [0, 4, 6, 7]], np.int64)


def grid_mesh_3d(x, y, z, electrodes_xy=(), cell_marker_fn=None):
    """Tetrahedron mesh on the tensor grid x * y * z (z <= 0 subsurface).

    Faces at z = max(z) get marker -1, other outer faces -2; nodes at
    (xe, ye, max(z)) get marker -99.
    """
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    nx, ny, nz = len(x), len(y), len(z)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    nid = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    i, j, k = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    corners = np.stack([nid[i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)]
                        for c in range(8)], axis=1)
    cells = corners[:, _KUHN].reshape(-1, 4).astype(np.int32)

    def quad_faces(a, b, c, d):
        # split quad (a, b, c, d in cyclic order) along a-c
        return np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])

    faces = []
    markers = []
    # z faces: the split diagonal must match the Kuhn tets, which use the
    # diagonal through the lower (ix, iy) corner of each face: (0,0)-(1,1)
    for kk, mk in ((nz - 1, MARKER_BOUND_HOMOGEN_NEUMANN), (0, MARKER_BOUND_MIXED)):
        a = nid[:-1, :-1, kk].ravel(); b = nid[1:, :-1, kk].ravel()
        c = nid[1:, 1:, kk].ravel(); d = nid[:-1, 1:, kk].ravel()
        f = quad_faces(a, b, c, d)
        faces.append(f); markers.append(np.full(len(f), mk))
    for ii in (0, nx - 1):
        a = nid[ii, :-1, :-1].ravel(); b = nid[ii, 1:, :-1].ravel()
        c = nid[ii, 1:, 1:].ravel(); d = nid[ii, :-1, 1:].ravel()
        f = quad_faces(a, b, c, d)
        faces.append(f); markers.append(np.full(len(f), MARKER_BOUND_MIXED))
    for jj in (0, ny - 1):
        a = nid[:-1, jj, :-1].ravel(); b = nid[1:, jj, :-1].ravel()
        c = nid[1:, jj, 1:].ravel(); d = nid[:-1, jj, 1:].ravel()
        f = quad_faces(a, b, c, d)
        faces.append(f); markers.append(np.full(len(f), MARKER_BOUND_MIXED))
    bounds = np.concatenate(faces).astype(np.int32)
    bm = np.concatenate(markers).astype(np.int32)
    nm = np.zeros(len(nodes), np.int32)
    for xe, ye in electrodes_xy:
        a = int(np.argmin(np.abs(x - xe)))
        b = int(np.argmin(np.abs(y - ye)))
        if abs(x[a] - xe) > 1e-9 or abs(y[b] - ye) > 1e-9:
            raise ValueError(f"electrode ({xe}, {ye}) is not a grid node")
        nm[nid[a, b, -1]] = MARKER_NODE_ELECTRODE
    cm = np.arange(len(cells), dtype=np.int32)
    if cell_marker_fn is not None:
        cm = cell_marker_fn(nodes[cells].mean(axis=1)).astype(np.int32)
    mesh = Mesh(3, nodes, cells, cm, bounds, bm, nm)
    lc = mesh.boundaryLeftCell()
    if (lc < 0).any():
        raise RuntimeError("boundary faces do not match the tetrahedra")
    return mesh

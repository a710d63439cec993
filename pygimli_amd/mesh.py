"""Unstructured P1 mesh container (triangles in 2D, tetrahedra in 3D).

Holds exactly the read-only arrays the ERT forward/Jacobian path consumes from
a pyGIMLi ``Mesh`` (core/src/mesh.h): node positions and markers, cell
connectivity and markers, boundary connectivity and markers, and each
boundary's "left" cell (core/src/meshentities.h Boundary::leftCell, set by
Mesh::createNeighborInfos).  Marker conventions follow the reference
(core/src/gimli.h): boundary -1 homogeneous Neumann (surface), -2 mixed
(Robin), -3 homogeneous Dirichlet; node -99 electrode, -1000 calibration.
"""
from __future__ import annotations

import numpy as np

MARKER_BOUND_HOMOGEN_NEUMANN = -1
MARKER_BOUND_MIXED = -2
MARKER_BOUND_HOMOGEN_DIRICHLET = -3
MARKER_BOUND_DIRICHLET = -4
MARKER_NODE_ELECTRODE = -99
MARKER_NODE_REFERENCEELECTRODE = -999
MARKER_NODE_CALIBRATION = -1000


class Mesh:
    """Minimal pyGIMLi-compatible mesh for the ERT hot path."""

    def __init__(self, dim, nodes, cells, cellMarkers=None, boundaries=None,
                 boundaryMarkers=None, nodeMarkers=None):
        self._dim = int(dim)
        if self._dim not in (2, 3):
            raise ValueError("only 2D triangle and 3D tetrahedron meshes are supported")
        nodes = np.asarray(nodes, dtype=np.float64)
        if nodes.ndim != 2 or nodes.shape[1] not in (2, 3):
            raise ValueError("nodes must be (n, 2) or (n, 3)")
        if nodes.shape[1] == 2:
            nodes = np.hstack([nodes, np.zeros((nodes.shape[0], 1))])
        self.nodes = np.ascontiguousarray(nodes)
        self.cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int32))
        if self.cells.ndim != 2 or self.cells.shape[1] != self._dim + 1:
            raise ValueError("cells must be (C, dim+1)")
        if self.cells.size and (self.cells.min() < 0 or self.cells.max() >= len(self.nodes)):
            raise ValueError("cell node index out of range")
        nc = self.cells.shape[0]
        self.cellMarkers_ = (np.zeros(nc, np.int32) if cellMarkers is None
                             else np.ascontiguousarray(np.asarray(cellMarkers, np.int32)))
        self.nodeMarkers_ = (np.zeros(len(self.nodes), np.int32) if nodeMarkers is None
                             else np.ascontiguousarray(np.asarray(nodeMarkers, np.int32)))
        if boundaries is None:
            boundaries = np.zeros((0, self._dim), np.int32)
            boundaryMarkers = np.zeros(0, np.int32)
        self.bounds = np.ascontiguousarray(np.asarray(boundaries, np.int32)).reshape(-1, self._dim)
        self.boundMarkers_ = np.ascontiguousarray(np.asarray(boundaryMarkers, np.int32))
        self._leftCell = None
        self._cellAttributes = np.ones(nc)
        self._data = {}

    # -- pyGIMLi Mesh API subset --------------------------------------------
    def dim(self):
        return self._dim

    def dimension(self):
        return self._dim

    def nodeCount(self):
        return self.nodes.shape[0]

    def cellCount(self):
        return self.cells.shape[0]

    def boundaryCount(self):
        return self.bounds.shape[0]

    def positions(self):
        return self.nodes

    def cellMarkers(self):
        return self.cellMarkers_

    def setCellMarkers(self, m):
        m = np.asarray(m, np.int32)
        if m.shape != (self.cellCount(),):
            raise ValueError("marker size mismatch")
        self.cellMarkers_ = np.ascontiguousarray(m)

    def nodeMarkers(self):
        return self.nodeMarkers_

    def boundaryMarkers(self):
        return self.boundMarkers_

    def findNodesIdxByMarker(self, marker):
        return np.nonzero(self.nodeMarkers_ == marker)[0]

    def cellAttributes(self):
        return self._cellAttributes

    def setCellAttributes(self, a):
        a = np.asarray(a, np.float64)
        if a.ndim == 0:
            a = np.full(self.cellCount(), float(a))
        if a.shape != (self.cellCount(),):
            raise ValueError("attribute size mismatch")
        self._cellAttributes = a.copy()

    def addData(self, name, v):
        """Named vector stored with the mesh (Mesh::addData, mesh.h)."""
        v = np.ascontiguousarray(np.asarray(v, np.float64).ravel())
        if not name or v.size == 0:
            raise ValueError("mesh data needs a name and at least one value")
        self._data[str(name)] = v

    def data(self, name):
        return self._data[name]

    def haveData(self, name):
        return name in self._data

    def dataMap(self):
        return dict(self._data)

    def save(self, fileName):
        """Binary mesh .bms (Mesh::save -> saveBinaryV2)."""
        from .fileio import save_mesh
        save_mesh(self, fileName)

    saveBinaryV2 = save

    def cellCenters(self):
        return self.nodes[self.cells].mean(axis=1)

    def boundarySizes(self):
        """Length (2D) or area (3D) of every boundary (Shape::domainSize)."""
        p = self.nodes[self.bounds]
        if self._dim == 2:
            return np.sqrt(((p[:, 1] - p[:, 0]) ** 2).sum(-1))
        return 0.5 * np.sqrt((np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]) ** 2).sum(-1))

    def boundaryCenters(self):
        return self.nodes[self.bounds].mean(axis=1)

    def cellSizes(self):
        p = self.nodes[self.cells]
        if self._dim == 2:
            a = p[:, 1, :2] - p[:, 0, :2]
            b = p[:, 2, :2] - p[:, 0, :2]
            return 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
        a = p[:, 1] - p[:, 0]
        b = p[:, 2] - p[:, 0]
        c = p[:, 3] - p[:, 0]
        return np.abs(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0

    def boundaryLeftCell(self):
        """Cell adjacent to each boundary (Mesh::createNeighborInfos).

        For an outer boundary the only adjacent cell; for an inner boundary
        the lower-indexed of the two.  -1 where no cell shares the face.
        """
        if self._leftCell is None:
            self._leftCell = _face_owner(self.cells, self.bounds, self.nodeCount())
        return self._leftCell

    def __repr__(self):
        return (f"Mesh: Nodes: {self.nodeCount()} Cells: {self.cellCount()} "
                f"Boundaries: {self.boundaryCount()} (dim={self._dim})")


def _face_keys(faces, n):
    f = np.sort(np.asarray(faces, np.int64), axis=1)
    key = np.zeros(len(f), np.int64)
    for j in range(f.shape[1]):
        key = key * np.int64(n) + f[:, j]
    return key


def _face_owner(cells, bounds, n):
    nv = cells.shape[1]
    d = nv - 1
    if len(bounds) == 0:
        return np.zeros(0, np.int32)
    # every cell face, i.e. every (d)-subset of the cell's nodes
    faces = []
    owners = []
    cid = np.arange(len(cells), dtype=np.int64)
    for skip in range(nv):
        cols = [j for j in range(nv) if j != skip]
        faces.append(cells[:, cols])
        owners.append(cid)
    faces = np.concatenate(faces)
    owners = np.concatenate(owners)
    fk = _face_keys(faces, n)
    order = np.lexsort((owners, fk))
    fk = fk[order]
    owners = owners[order]
    bk = _face_keys(bounds.reshape(-1, d), n)
    pos = np.searchsorted(fk, bk)
    pos = np.minimum(pos, len(fk) - 1)
    hit = fk[pos] == bk
    out = np.where(hit, owners[pos], -1).astype(np.int32)
    return out


# local edges of the P2 cells and faces: triangle (0,1), (1,2), (0,2);
# tetrahedron (0,1), (0,2), (0,3), (1,2), (1,3), (2,3); edge (0,1)
P2_LOCAL_EDGES = {1: ((0, 1),), 2: ((0, 1), (1, 2), (0, 2)),
                  3: ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))}


def p2_refinement(mesh):
    """Mesh::createP2 (core/src/mesh.cpp:1285-1330): the mesh with one more
    node at the midpoint of every edge (numbered after the vertices, which
    keep their ids), as quadratic cells and faces in the local order of
    P2_LOCAL_EDGES.  Returns nodes (N x 3), cells (C x 6 | 10), boundaries
    (F x 3 | 6) and node markers (vertices keep theirs; a midpoint takes
    markerT's rule, mesh.cpp:1291-1297)."""
    n, d = mesh.nodeCount(), mesh.dim()
    c = mesh.cells.astype(np.int64)
    pairs = np.sort(np.stack([c[:, [a, b]] for a, b in P2_LOCAL_EDGES[d]], axis=1), axis=2)
    key = pairs[..., 0] * n + pairs[..., 1]
    edges, inv = np.unique(key.ravel(), return_inverse=True)
    e0, e1 = edges // n, edges % n
    nodes = np.vstack([mesh.nodes, 0.5 * (mesh.nodes[e0] + mesh.nodes[e1])])
    cells = np.hstack([c, n + inv.reshape(key.shape)]).astype(np.int32)
    b = mesh.bounds.astype(np.int64)
    bpairs = np.sort(np.stack([b[:, [x, y]] for x, y in P2_LOCAL_EDGES[d - 1]], axis=1), axis=2)
    bkey = bpairs[..., 0] * n + bpairs[..., 1]
    pos = np.searchsorted(edges, bkey)
    if len(bkey) and (np.any(pos >= len(edges)) or np.any(edges[np.minimum(pos, len(edges) - 1)] != bkey)):
        raise ValueError("a boundary edge is no cell edge")
    bounds = np.hstack([b, n + pos]).astype(np.int32)
    nm = mesh.nodeMarkers()
    m0, m1 = nm[e0], nm[e1]
    mid = np.where((m0 == MARKER_NODE_ELECTRODE) & (m1 == MARKER_NODE_ELECTRODE), -1,
                   np.where(m0 == MARKER_NODE_ELECTRODE, m1,
                            np.where(m1 == MARKER_NODE_ELECTRODE, m0, np.where(m0 == m1, m0, 0))))
    return nodes, cells, bounds, np.concatenate([nm, mid]).astype(np.int32)

"""Host-side set-up of the ERT operator: wavenumbers, geometric factors and
electrode search (set-up cost O(nE^2 + D), done once per mesh/data set).

Restates core/src/bert/bertMisc.cpp and the electrode/mesh analysis of
DCMultiElectrodeModelling (core/src/bert/dcfemmodelling.cpp:711-1073).
"""
from __future__ import annotations

import math
import re
import warnings

import numpy as np

MAX_DOUBLE = float(np.finfo(np.float64).max)
TOLERANCE = 1e-12
MARKER_BOUND_ELECTRODE = -10000        # core/src/bert/bert.h:33
MARKER_FIXEDVALUE_REGION = -1000000    # core/src/gimli.h:241
MARKER_NODE_REFERENCEELECTRODE = -999  # core/src/bert/bert.h:31-32
MARKER_NODE_CALIBRATION = -1000
BYPASS_RESISTANCE = 1e-6               # dcfemmodelling.cpp:684-688 default


def _gauss_legendre_nr(n):
    """Gauss-Legendre nodes/weights on [0, 1] by Newton iteration on P_n with
    the reference's stopping rule |dz| <= 3e-6 (core/src/numericbase.cpp:102-156)."""
    x = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1.0) / 2.0
    i = 1
    while i <= m:
        z = math.cos(math.pi * (i - 0.25) / (n + 0.5))
        z1 = z + 6e-6
        pp = 1.0
        while abs(z - z1) > 3.0e-6:
            p_prev, p = 0.0, 1.0
            for j in range(1, n + 1):
                p_prev, p = p, ((2.0 * j - 1.0) * z * p - (j - 1.0) * p_prev) / j
            pp = n * (z * p - p_prev) / (z * z - 1.0)
            z1, z = z, z - p / pp
        x[i - 1] = 0.5 - 0.5 * z
        x[n - i] = 0.5 + 0.5 * z
        w[i - 1] = w[n - i] = 1.0 / ((1.0 - z * z) * pp * pp)
        i += 1
    return x, w


def _gauss_laguerre_nr(n):
    """Gauss-Laguerre (alpha = 0) by Newton iteration, |dz| <= 3e-11, at most
    20 steps (core/src/numericbase.cpp:51-100)."""
    x = np.zeros(n)
    w = np.zeros(n)
    z = 0.0
    for i in range(1, n + 1):
        if i == 1:
            z = 3.0 / (1.0 + 2.4 * n)
        elif i == 2:
            z += 15.0 / (1.0 + 2.5 * n)
        else:
            ai = i - 2
            z += (1.0 + 2.55 * ai) / (1.9 * ai) * (z - x[ai - 1])
        p = p_prev = pp = 0.0
        for _ in range(20):
            p_prev, p = 0.0, 1.0
            for j in range(1, n + 1):
                p_prev, p = p, ((2.0 * j - 1 - z) * p - (j - 1) * p_prev) / j
            pp = n * (p - p_prev) / z
            z1 = z
            z = z1 - p / pp
            if abs(z - z1) <= 3.0e-11:
                break
        x[i - 1] = z
        w[i - 1] = -1.0 / (pp * n * p_prev)
    return x, w


def kwave_from_range(rmin, rmax, n_leg, n_lag=4):
    """Wavenumbers/weights of the 2.5D inverse Fourier transform
    (initKWaveList, core/src/bert/bertMisc.cpp:89-106)."""
    k0 = 1.0 / (2.0 * rmin)
    x, w = _gauss_legendre_nr(n_leg)
    kl = k0 * x * x
    wl = 2.0 * k0 * x * w / math.pi
    x, w = _gauss_laguerre_nr(n_lag)
    kg = k0 * (x + 1.0)
    wg = k0 * np.exp(x) * w / math.pi
    return np.concatenate([kl, kg]), np.concatenate([wl, wg])


def kwave_list(dim, sensors, n_leg=None, n_lag=4):
    """initKWaveList (bertMisc.cpp:36-86): 3D -> k = {0}, w = {1}; 2.5D ->
    rMin = min pair distance / 2, rMax = max * 2, nLeg = max(floor(6 log10(rMax/rMin)), 4)."""
    if dim == 3:
        return np.array([0.0]), np.array([1.0])
    s = np.asarray(sensors, np.float64)
    diff = s[:, None, :] - s[None, :, :]
    d = np.sqrt((diff ** 2).sum(-1))[np.triu_indices(len(s), 1)]
    rmin = d.min() / 2.0
    rmax = d.max() * 2.0
    if n_leg is None:
        n_leg = max(int(math.floor(6.0 * math.log10(rmax / rmin))), 4)
    return kwave_from_range(rmin, rmax, n_leg, n_lag)


def halfspace_potential(points, src):
    """3D half-space unit potential with mirror plane z = 0:
    (1/r + 1/r') / (4 pi), 1 where r = 0 (bertMisc.cpp:178-181, :209-214)."""
    p = np.asarray(points, np.float64)
    s = np.asarray(src, np.float64)
    r = np.sqrt(((p - s) ** 2).sum(-1))
    sm = s * np.array([1.0, 1.0, -1.0])
    rm = np.sqrt(((p - sm) ** 2).sum(-1))
    with np.errstate(divide="ignore"):
        out = (1.0 / r + 1.0 / rm) / (4.0 * math.pi)
    return np.where(r < TOLERANCE, 1.0, out)


def geometric_factors(sensors, abmn, dim=3):
    """geometricFactors (bertMisc.cpp:131-176): k = 1 / (G_am - G_bm - G_an + G_bn).
    2D sensors with y != 0 are read as (x, z)."""
    s = np.array(sensors, np.float64)
    if dim == 2:
        sw = s[:, 1] != 0.0
        s[sw, 2] = s[sw, 1]
        s[sw, 1] = 0.0
    abmn = np.asarray(abmn)
    a, b, m, n = (abmn[:, i] for i in range(4))

    def G(i, j):
        ok = (i > -1) & (j > -1)
        out = np.zeros(len(i))
        if ok.any():
            out[ok] = np.array([halfspace_potential(s[jj][None], s[ii])[0]
                                for ii, jj in zip(i[ok], j[ok])])
        return out

    with np.errstate(divide="ignore"):
        return 1.0 / (G(a, m) - G(b, m) - G(a, n) + G(b, n))


def _find_cell(mesh, p, tol=1e-12):
    """Index of the cell holding position p (Mesh::findCell), -1 outside the
    mesh; on a face shared by two cells the lower cell index."""
    dim = mesh.dim()
    X = mesh.positions()[np.asarray(mesh.cells)][:, :, :dim]
    E = np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))
    b = np.asarray(p, np.float64)[:dim] - X[:, 0, :]
    lam = np.linalg.solve(E, b[..., None])[..., 0]
    inside = np.all(lam >= -tol, axis=1) & (lam.sum(1) <= 1.0 + tol)
    hit = np.nonzero(inside)[0]
    return int(hit[0]) if len(hit) else -1


class MeshAnalysis:
    """Boundary/topography analysis and electrode search
    (updateMeshDependency_ dcfemmodelling.cpp:711-773, searchElectrodes_ :790-1073)
    for node electrodes (marker -99), bypass electrodes (nodes sharing a
    marker <= -10000 form one electrode at the group's first node, joined
    node to node by bypass resistors, :815-826, :886-897) and free electrodes
    (a sensor matching neither lies in a cell: elecNode -1, elecCell, the
    position in elecPos, :979-985)."""

    def __init__(self, mesh, sensors):
        dim = mesh.dim()
        bm = mesh.boundaryMarkers()
        bc = mesh.boundaryCenters() if mesh.boundaryCount() else np.zeros((0, 3))
        self.neumannDomain = not bool(np.isin(bm, (-2, -3, -4)).any())
        self.topography = False
        self.surfaceZ = -MAX_DOUBLE
        neu = np.nonzero(bm == -1)[0]
        if len(neu):
            self.surfaceZ = float(bc[neu[0], dim - 1])
            if np.any(bc[neu, dim - 1] != self.surfaceZ):
                self.topography = True
        if self.neumannDomain:
            self.topography = True
            if dim == 2:
                self.neumannDomain = False
        pos = np.array(sensors, np.float64)
        if dim == 2:
            zvar = np.ptp(pos[:, 2]) > 0 or np.abs(pos[:, 2]).max() > 0
            ystill = not (np.ptp(pos[:, 1]) > 0) and np.abs(pos[:, 1]).max() < 1e-8
            if zvar and ystill:
                pos[:, [1, 2]] = pos[:, [2, 1]]
        cm = mesh.cellMarkers()
        if np.any((cm <= MARKER_BOUND_ELECTRODE) & (cm > MARKER_FIXEDVALUE_REGION)):
            raise NotImplementedError("complete electrode model: electrode cells (marker <= -10000) are not "
                                      "supported, electrode faces are")
        # reference electrode (the first node with marker -999) and calibration
        # nodes (marker -1000; node 0 in a Neumann domain without one; none
        # outside Neumann domains), dcfemmodelling.cpp:800-803, :1010-1071
        nm = np.asarray(mesh.nodeMarkers())
        ref = np.nonzero(nm == MARKER_NODE_REFERENCEELECTRODE)[0]
        self.refNode = int(ref[0]) if len(ref) else -1
        cal = [int(i) for i in np.nonzero(nm == MARKER_NODE_CALIBRATION)[0]]
        self.calibration = (cal if cal else [0]) if self.neumannDomain else []
        self.cemFaces = {}
        self.cemPassive = None
        if np.any(bm <= MARKER_BOUND_ELECTRODE + 1):
            if self.refNode >= 0:
                raise NotImplementedError("a reference electrode with the complete electrode model")
            self._search_cem(mesh, pos)
            return
        nmk = np.asarray(mesh.nodeMarkers())
        self.bypassGroups = {}             # marker -> node ids ascending, markers in order of appearance
        for i in np.nonzero(nmk <= MARKER_BOUND_ELECTRODE)[0]:
            self.bypassGroups.setdefault(int(nmk[i]), []).append(int(i))
        free = sorted(self.bypassGroups, key=lambda m: MARKER_BOUND_ELECTRODE - m)   # map order
        cand = list(mesh.findNodesIdxByMarker(-99))
        nodes = mesh.positions()
        self.elecNode = np.empty(len(pos), np.int32)
        self.elecCell = np.full(len(pos), -1, np.int32)
        self.freePos = np.zeros((len(pos), 3))
        for i, p in enumerate(pos):
            hit = None
            for m in free:
                first = self.bypassGroups[m][0]
                if np.sqrt(((nodes[first] - p) ** 2).sum()) < 0.01:
                    hit = m
                    break
            if hit is not None:
                free.remove(hit)
                self.elecNode[i] = self.bypassGroups[hit][0]
                continue
            if cand:
                d = np.sqrt(((nodes[cand] - p) ** 2).sum(-1))
                close = np.nonzero(d < 0.01)[0]
                if len(close):
                    hit = int(close[0])
            if hit is None:
                # a free electrode in the cell holding the position
                # (ElectrodeShapeEntity, searchElectrodes_ :979-985)
                c = _find_cell(mesh, p)
                if c < 0:
                    raise RuntimeError(f"sensor {i} at {p}: there is a requested electrode that does not match "
                                       "the given mesh")
                self.elecNode[i] = -1
                self.elecCell[i] = c
                self.freePos[i] = p
                continue
            self.elecNode[i] = cand.pop(hit)
        self.elecPos = np.where((self.elecNode >= 0)[:, None], nodes[np.maximum(self.elecNode, 0)], self.freePos)
        self.sourceCenter = self.elecPos.sum(0) / len(self.elecPos)

    def _search_cem(self, mesh, pos):
        """Complete-electrode-model electrodes (searchElectrodes_,
        dcfemmodelling.cpp:804-851, :868-884): boundary faces with marker
        -10000 - id grouped per id (ElectrodeShapeDomain: position = mean of
        the faces' nodes, size = their area); sensor i takes the first
        remaining electrode, in id order, within 2 sqrt(size) of it or with id
        i.  Every sensor must find one (the reference then gives node
        electrodes unknowns that no current reaches); the electrode unknowns
        follow the n mesh nodes: elecNode = n + i."""
        if mesh.dim() != 3:
            raise NotImplementedError("complete electrode model in 2.5D: the reference keeps only the last "
                                      "wavenumber's electrode potentials (dcfemmodelling.cpp:1910-1913)")
        bm = mesh.boundaryMarkers()
        ids = sorted({int(MARKER_BOUND_ELECTRODE - m) for m in bm[bm <= MARKER_BOUND_ELECTRODE + 1]})
        size = mesh.boundarySizes()
        nodes = mesh.positions()
        self.cemPassive = None
        if ids[0] < 0:
            # the passive body (faces with marker -9999): an unknown after the
            # electrodes' with its contact terms and no current (:838-847)
            self.cemPassive = np.nonzero(bm == MARKER_BOUND_ELECTRODE + 1)[0]
            self.cemPassivePos = nodes[np.unique(mesh.bounds[self.cemPassive])].mean(0)
            ids = ids[1:]
            if not ids:
                raise NotImplementedError("complete electrode model: a passive body without electrode faces (the "
                                          "reference then gives node electrodes unknowns no current reaches)")
        domains = []
        for eid in ids:
            faces = np.nonzero(bm == MARKER_BOUND_ELECTRODE - eid)[0]
            centre = nodes[np.unique(mesh.bounds[faces])].mean(0)
            domains.append((eid, faces, centre, float(size[faces].sum())))
        for i, p in enumerate(pos):
            for j, (eid, faces, centre, area) in enumerate(domains):
                if np.sqrt(((p - centre) ** 2).sum()) < math.sqrt(area) * 2.0 or eid == i:
                    self.cemFaces[i] = faces
                    domains.pop(j)
                    break
        if len(self.cemFaces) != len(pos):
            raise NotImplementedError("complete electrode model: every sensor needs an electrode face group "
                                      "(node or free electrodes beside it are not supported)")
        n = mesh.nodeCount()
        self.bypassGroups = {}
        self.elecNode = (n + np.arange(len(pos))).astype(np.int32)
        self.elecCell = np.full(len(pos), -1, np.int32)
        self.freePos = np.zeros((len(pos), 3))
        self.elecPos = pos.copy()
        self.sourceCenter = self.elecPos.sum(0) / len(self.elecPos)

    def cem_terms(self, mesh, impedances=None, resistances=None):
        """Matrix terms of the complete electrode model
        (assembleCompleteElectrodeModel_, dcfemmodelling.cpp:315-410, no
        reference electrode): electrode e with unknown v = n + e, faces f,
        area A = sum |f| and contact impedance z (impedances[e], default
        1 Ohm m^2; z = resistances[e] * A when contact resistances are given):
        (i, v), (v, i): -|f| / d / z per face node i (int_f N_i); (i, j):
        |f| (1 + d_ij) / (d (d + 1)) / z (int_f N_i N_j); (v, v): A / z.  The
        passive body follows the electrodes (unknown n + nE, impedance entry
        nE).  Returns rows, cols, vals and the node pairs (v, i) that need
        slots."""
        d = mesh.dim()
        n = mesh.nodeCount()
        size = mesh.boundarySizes()
        rows, cols, vals, pairs = [], [], [], []
        groups = dict(self.cemFaces)
        if self.cemPassive is not None:
            groups[len(self.cemFaces)] = self.cemPassive
        for e, faces in sorted(groups.items()):
            area = float(size[faces].sum())
            z = 1.0 if impedances is None else float(impedances[e])
            if resistances is not None:
                z = float(resistances[e]) * area
            v = n + e
            for f in faces:
                fn = mesh.bounds[f]
                u = size[f] / d
                for a in range(d):
                    rows += [fn[a], v]
                    cols += [v, fn[a]]
                    vals += [-u / z, -u / z]
                    pairs.append((v, int(fn[a])))
                    for b in range(d):
                        rows.append(fn[a])
                        cols.append(fn[b])
                        vals.append(size[f] / (d * (d + 1)) * (2.0 if a == b else 1.0) / z)
            rows.append(v)
            cols.append(v)
            vals.append(area / z)
        return (np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals, np.float64),
                sorted(set(pairs)))

    def bypass_pairs(self, resistances=None):
        """Node pairs (i, i+1) of every bypass group with their conductance
        1 / R, R from `resistances` (marker -> Ohm, the reference's
        electrodeBypassResistances.map) or 1e-6 Ohm
        (assembleStiffnessMatrixDCFEMByPass_, dcfemmodelling.cpp:671-707)."""
        a, b, g = [], [], []
        for m, ids in self.bypassGroups.items():
            r = BYPASS_RESISTANCE if resistances is None else float(resistances.get(m, BYPASS_RESISTANCE))
            for i in range(len(ids) - 1):
                a.append(ids[i])
                b.append(ids[i + 1])
                g.append(1.0 / r)
        return np.array(a, np.int32), np.array(b, np.int32), np.array(g, np.float64)


# ---------------------------------------------------------------- map files
def _map_rows(path):
    """Rows of a whitespace-separated text file with '#' comments, empty rows
    skipped (getNonEmptyRow / getRowSubstrings, core/src/gimli.cpp:245-259)."""
    with open(path) as f:
        for line in f:
            row = line.split("#", 1)[0].split()
            if row:
                yield row


def _atoi(tok):
    """std::atoi: the leading integer of the token, 0 if there is none
    (toInt, core/src/gimli.h:477)."""
    m = re.match(r"[+-]?\d+", tok)
    return int(m.group(0)) if m else 0


def _atof32(tok):
    """std::atof rounded to single precision (toFloat, core/src/gimli.h:478)."""
    m = re.match(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?", tok)
    return float(np.float32(float(m.group(0)))) if m else 0.0


def load_float_map(path):
    """loadFloatMap (core/src/gimli.cpp:312-334): rows "key value", both read
    as single-precision floats; a file with no two-column row before another
    row is an error."""
    out = {}
    for row in _map_rows(path):
        if len(row) == 2:
            out[_atof32(row[0])] = _atof32(row[1])
        elif not out:
            raise ValueError(f"{path}: no proper format found for map <float, float> ({len(row)} columns)")
    return out


def bypass_marker_resistances(path):
    """electrodeBypassResistances.map as marker -> Ohm: the reference looks a
    group up by float(marker) (dcfemmodelling.cpp:672-691), so only integral
    keys can match a marker."""
    return {int(k): v for k, v in load_float_map(path).items() if float(k).is_integer()}


def read_bypass_map(path, elecNode, nodeMarkers):
    """The resistor rows of a bypass.map file
    (assembleStiffnessMatrixDCFEMByPass_, dcfemmodelling.cpp:608-667), in file
    order, as node pairs and resistances (Ohm, single precision as the
    reference reads them):

    * "e1 e2 R": a resistor between the nodes of electrodes e1 and e2
      (counted from 1).  Like the reference, e1 must lie in 1 .. nE - 1 (its
      bound check is e1 < nE, so the last electrode is refused too) and rows
      failing that are skipped with a warning; the reference checks e1 again
      in place of e2 and reads past its electrode list for a bad e2, which
      raises here, as does an electrode without a node (a free electrode).
    * "marker R": resistors chaining the nodes with that node marker in index
      order (Mesh::findNodesIdxByMarker); a marker no node has is skipped with
      a warning.
    * rows of any other length are skipped with a warning.

    A pair of one node with itself adds nothing and is dropped; a resistance
    of zero or not finite raises."""
    elecNode = np.asarray(elecNode)
    nodeMarkers = np.asarray(nodeMarkers)
    nE = len(elecNode)
    a, b, r = [], [], []

    def add(n1, n2, res, row):
        if not (np.isfinite(res) and res != 0.0):
            raise ValueError(f"{path}: row {' '.join(row)}: resistance must be finite and non-zero")
        if n1 != n2:
            a.append(int(n1))
            b.append(int(n2))
            r.append(res)

    for row in _map_rows(path):
        if len(row) == 3:
            e1, e2, res = _atoi(row[0]), _atoi(row[1]), _atof32(row[2])
            if e1 < 1:
                warnings.warn(f"{path}: bypass electrode unknown: {e1}, electrode indices count from 1; ignoring")
                continue
            if e1 >= nE:
                warnings.warn(f"{path}: bypass electrode unknown {e1} (e-size = {nE}); ignoring")
                continue
            if not 1 <= e2 <= nE:
                raise ValueError(f"{path}: row {' '.join(row)}: electrode {e2} is not one of 1 .. {nE}")
            n1, n2 = int(elecNode[e1 - 1]), int(elecNode[e2 - 1])
            if n1 < 0 or n2 < 0:
                raise ValueError(f"{path}: row {' '.join(row)}: a free electrode (on no node) has no bypass node")
            add(n1, n2, res, row)
        elif len(row) == 2:
            m, res = _atoi(row[0]), _atof32(row[1])
            ids = np.nonzero(nodeMarkers == m)[0]
            if len(ids) == 0:
                warnings.warn(f"{path}: no node with marker {m}; ignoring")
                continue
            for i in range(len(ids) - 1):
                add(ids[i], ids[i + 1], res, row)
        else:
            warnings.warn(f"{path}: wrong format ({len(row)} columns), expected \"electrode electrode resistance\" "
                          "or \"node-marker resistance\"; ignoring")
    return np.array(a, np.int32), np.array(b, np.int32), np.array(r, np.float64)


def uncoupled_pairs(cells, a, b):
    """The node pairs (a[i], b[i]) that share no cell: they need their own
    matrix slots (pgert_set_mesh_coupled).  Returns a sorted list of
    (min, max) tuples without repeats."""
    if len(a) == 0:
        return []
    cells = np.asarray(cells)
    u = np.unique(np.concatenate([a, b]))
    sub = cells[np.isin(cells, u).any(1)]
    out = set()
    for x, y in zip(np.asarray(a).tolist(), np.asarray(b).tolist()):
        if not ((sub == x).any(1) & (sub == y).any(1)).any():
            out.add((min(x, y), max(x, y)))
    return sorted(out)


def load_electrode_vector(path, n, default):
    """checkIfMapFileExistAndLoadToVector (dcfemmodelling.cpp:114-137): a
    loadFloatMap file of electrode index -> value into a vector of n
    defaults; indices count from 1 unless the first key rounds to 0; key -1
    is the last electrode; other indices out of range are skipped."""
    v = np.full(n, float(default))
    m = load_float_map(path)
    if not m:
        return v
    from_one = int(np.rint(min(m))) != 0
    for key in sorted(m):
        if key == -1:
            v[-1] = m[key]
            continue
        idx = int(np.rint(key))
        if from_one:
            if 0 < idx <= n:
                v[idx - 1] = m[key]
        elif 0 <= idx < n:
            v[idx] = m[key]
    return v

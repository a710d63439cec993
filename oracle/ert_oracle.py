"""FP64 CPU restatement of pyGIMLi's ERT forward + Jacobian path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker the HIP path
is compared against; never the product path.

Each function cites the reference code it restates (paths relative to the
pyGIMLi repository).  The default forward is the singularity-removal (SR)
operator DCSRMultiElectrodeModelling that ERTModelling(sr=True) builds
(pygimli/physics/ert/ertModelling.py:87-106).  CHOLMOD (the reference's
direct solver, an external SuiteSparse library) is replaced by scipy's
SuperLU factorisation, an FP64 direct solver as well.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

PI = np.pi
TOLERANCE = 1e-12          # core/src/gimli.h:206
MAX_DOUBLE = np.finfo(np.float64).max


# --------------------------------------------------------------------------- a2
# Bessel functions: Abramowitz & Stegun polynomial approximations 9.8.1-9.8.8
# with the coefficients the reference uses (core/src/numericbase.h:80-180).

def bessel_i0(x):
    x = np.asarray(x, np.float64)
    ax = np.abs(x)
    out = np.empty_like(ax)
    s = ax < 3.75
    y = (x[s] / 3.75) ** 2
    out[s] = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (
        0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))))
    a = ax[~s]
    y = 3.75 / a
    out[~s] = (np.exp(a) / np.sqrt(a)) * (0.39894228 + y * (0.1328592e-1 + y * (
        0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (
            0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))))
    return out


def bessel_i1(x):
    x = np.asarray(x, np.float64)
    ax = np.abs(x)
    out = np.empty_like(ax)
    s = ax < 3.75
    y = (x[s] / 3.75) ** 2
    out[s] = ax[s] * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (
        0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))))
    a = ax[~s]
    y = 3.75 / a
    r = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2))
    r = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (
        0.163801e-2 + y * (-0.1031555e-1 + y * r))))
    out[~s] = r * (np.exp(a) / np.sqrt(a))
    return np.where(x < 0.0, -out, out)


def bessel_k0(x):
    # x = 0 (a point on the source): log(0) and 1/0 give the limits -inf / +inf
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.asarray(x, np.float64)
        out = np.empty_like(x)
        s = x <= 2.0
        xs = x[s]
        y = xs * xs / 4.0
        out[s] = (-np.log(xs / 2.0) * bessel_i0(xs)) + (-0.57721566 + y * (0.42278420 + y * (
            0.23069756 + y * (0.3488590e-1 + y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5))))))
        xl = x[~s]
        y = 2.0 / xl
        out[~s] = (np.exp(-xl) / np.sqrt(xl)) * (1.25331414 + y * (-0.7832358e-1 + y * (
            0.2189568e-1 + y * (-0.1062446e-1 + y * (0.587872e-2 + y * (
                -0.251540e-2 + y * 0.53208e-3))))))
        return out


def bessel_k1(x):
    # x = 0 (a point on the source): log(0) and 1/0 give the limits -inf / +inf
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.asarray(x, np.float64)
        out = np.empty_like(x)
        s = x <= 2.0
        xs = x[s]
        y = xs * xs / 4.0
        out[s] = (np.log(xs / 2.0) * bessel_i1(xs)) + (1.0 / xs) * (1.0 + y * (0.15443144 + y * (
            -0.67278579 + y * (-0.18156897 + y * (-0.1919402e-1 + y * (
                -0.110404e-2 + y * (-0.4686e-4)))))))
        xl = x[~s]
        y = 2.0 / xl
        out[~s] = (np.exp(-xl) / np.sqrt(xl)) * (1.25331414 + y * (0.23498619 + y * (
            -0.3655620e-1 + y * (0.1504268e-1 + y * (-0.780353e-2 + y * (
                0.325614e-2 + y * (-0.68245e-3)))))))
        return out


# --------------------------------------------------------------------------- a1
def gauss_legendre(x1, x2, n):
    """Newton iteration on P_n (core/src/numericbase.cpp:102-156, eps 3e-6)."""
    x = np.zeros(n)
    w = np.zeros(n)
    eps = 3.0e-6
    m = (n + 1.0) / 2.0
    xm = 0.5 * (x2 + x1)
    xl = 0.5 * (x2 - x1)
    i = 1
    while i <= m:
        z = np.cos(PI * (i - 0.25) / (n + 0.5))
        z1 = z + 2.0 * eps
        pp = 0.0
        while abs(z - z1) > eps:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
        x[i - 1] = xm - xl * z
        x[n - i] = xm + xl * z
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
        i += 1
    return x, w


def gauss_laguerre(n):
    """Newton iteration on L_n, alpha = 0 (core/src/numericbase.cpp:51-100,
    eps 3e-11, at most 20 iterations)."""
    x = np.zeros(n)
    w = np.zeros(n)
    eps = 3.0e-11
    z = 0.0
    for i in range(1, n + 1):
        if i == 1:
            z = 3.0 / (1.0 + 2.4 * n)
        elif i == 2:
            z = z + 15.0 / (1.0 + 2.5 * n)
        else:
            ai = i - 2
            z = z + (1.0 + 2.55 * ai) / (1.9 * ai) * (z - x[ai - 1])
        p1 = p2 = pp = 0.0
        for _ in range(20):
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2.0 * j - 1 - z) * p2 - (j - 1) * p3) / j
            pp = n * (p1 - p2) / z
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) <= eps:
                break
        x[i - 1] = z
        w[i - 1] = -1.0 / (pp * n * p2)
    return x, w


def kwave_from_range(rmin, rmax, n_leg, n_lag):
    """core/src/bert/bertMisc.cpp:89-106."""
    k0 = 1.0 / (2.0 * rmin)
    x, w = gauss_legendre(0.0, 1.0, n_leg)
    k_leg = k0 * x * x
    w_leg = 2.0 * k0 * x * w / PI
    x, w = gauss_laguerre(n_lag)
    k_lag = k0 * (x + 1.0)
    w_lag = k0 * np.exp(x) * w / PI
    return np.concatenate([k_leg, k_lag]), np.concatenate([w_leg, w_lag])


def kwave_list(dim, sensors):
    """initKWaveList (core/src/bert/bertMisc.cpp:36-86)."""
    if dim == 3:
        return np.array([0.0]), np.array([1.0])
    s = np.asarray(sensors, np.float64)
    d = np.sqrt(((s[:, None, :] - s[None, :, :]) ** 2).sum(-1))
    iu = np.triu_indices(len(s), 1)
    rmin = d[iu].min() / 2.0
    rmax = d[iu].max() * 2.0
    n_leg = max(int(np.floor(6.0 * np.log10(rmax / rmin))), 4)
    return kwave_from_range(rmin, rmax, n_leg, 4)


# --------------------------------------------------------------------------- a2
def exact_dc(points, src, k, surfaceZ, fallback):
    """exactDCSolution(v, source, k, surfaceZ, fallback)
    (core/src/bert/bertMisc.cpp:183-228), vectorised over points."""
    v = np.asarray(points, np.float64)
    src = np.asarray(src, np.float64)
    r = np.sqrt(((v - src) ** 2).sum(-1))
    dim = 2 if k > 0 else 3
    mir = src.copy()
    mir[dim - 1] = 2.0 * surfaceZ - src[dim - 1]
    rm = np.sqrt(((v - mir) ** 2).sum(-1))
    small = r < TOLERANCE
    rs = np.where(small, 1.0, r)
    if surfaceZ == -MAX_DOUBLE or np.isnan(surfaceZ):
        out = 1.0 / (4.0 * PI * rs) if k == 0.0 else bessel_k0(rs * k) / (2.0 * PI)
    else:
        if k == 0.0:
            out = (1.0 / rs + 1.0 / np.where(rm < TOLERANCE, 1.0, rm)) / (4.0 * PI)
        elif ((src - mir) ** 2).sum() < TOLERANCE:  # RVector3 == (pos.h:319-322)
            out = bessel_k0(rs * k) / PI
        else:
            out = (bessel_k0(rs * k) + bessel_k0(rm * k)) / (2.0 * PI)
    return np.where(small, fallback, out)


def geometric_factors(sensors, abmn, dim):
    """geometricFactors (core/src/bert/bertMisc.cpp:131-176): 2D sensors with
    y != 0 are taken as (x, z) and rotated into the y = 0 plane; then the
    3D half-space potential with mirror plane z = 0."""
    s = np.array(sensors, np.float64)
    if dim == 2:
        sw = s[:, 1] != 0.0
        s[sw, 2] = s[sw, 1]
        s[sw, 1] = 0.0
    a, b, m, n = [np.asarray(abmn[:, i]) for i in range(4)]

    def G(i, j):
        ok = (i > -1) & (j > -1)
        out = np.zeros(len(i))
        for t in np.nonzero(ok)[0]:
            out[t] = exact_dc(s[j[t]][None], s[i[t]], 0.0, 0.0, 1.0)[0]
        return out

    return 1.0 / (G(a, m) - G(b, m) - G(a, n) + G(b, n))


# --------------------------------------------------------------------------- mesh-side set-up
def find_cell(mesh, p, tol=1e-12):
    """The cell holding position p (Mesh::findCell, mesh.cpp) and the P1 shape
    functions there (barycentric coordinates, MeshEntity::N(rst(p)));
    (-1, None) outside the mesh.  A position on a face shared by two cells
    goes to the lower cell index."""
    dim = mesh.dim()
    P = mesh.nodes[mesh.cells][:, :, :dim]
    x = np.asarray(p, np.float64)[:dim]
    T = np.transpose(P[:, 1:, :] - P[:, :1, :], (0, 2, 1))      # (C, d, d): columns = edges
    lam = np.linalg.solve(T, (x - P[:, 0, :])[..., None])[..., 0]
    N = np.concatenate([1.0 - lam.sum(-1, keepdims=True), lam], axis=1)
    inside = np.nonzero(np.all(N >= -tol, axis=1))[0]
    if len(inside) == 0:
        return -1, None
    c = int(inside[0])
    return c, N[c]


class Setup:
    """Mesh/data dependent state of DCMultiElectrodeModelling
    (updateMeshDependency_ dcfemmodelling.cpp:711-773, searchElectrodes_
    :790-1073)."""

    def __init__(self, mesh, sensors):
        self.mesh = mesh
        dim = mesh.dim()
        self.dim = dim
        bm = mesh.boundaryMarkers()
        bc = mesh.boundaryCenters()
        self.neumannDomain = not np.any(np.isin(bm, (-2, -3, -4)))
        self.topography = False
        self.surfaceZ = -MAX_DOUBLE
        neu = np.nonzero(bm == -1)[0]
        if len(neu):
            self.surfaceZ = bc[neu[0], dim - 1]
            if np.any(bc[neu, dim - 1] != self.surfaceZ):
                self.topography = True
        if self.neumannDomain:
            self.topography = True
            if dim == 2:
                self.neumannDomain = False
        ePos = np.array(sensors, np.float64)
        if dim == 2:
            zvar = np.ptp(ePos[:, 2]) > 0 or np.abs(ePos[:, 2]).max() > 0
            yconst = not (np.ptp(ePos[:, 1]) > 0) and np.abs(ePos[:, 1]).max() < 1e-8
            if zvar and yconst:
                ePos[:, [1, 2]] = ePos[:, [2, 1]]
        # reference electrode (first node with marker -999, :1010-1015): every
        # current pattern becomes a dipole against it (createCurrentPattern
        # :1517-1523); calibration nodes (marker -1000, node 0 when a Neumann
        # domain has none, :1044-1071) are homogeneous Dirichlet nodes of the
        # total-field matrix (:1835), ignored outside Neumann domains
        nm = mesh.nodeMarkers()
        ref = np.nonzero(nm == -999)[0]
        self.refNode = int(ref[0]) if len(ref) else -1
        cal = [int(i) for i in np.nonzero(nm == -1000)[0]]
        if self.neumannDomain:
            self.calibration = cal if cal else [0]
            if self.refNode < 0:
                raise NotImplementedError("Neumann domain without a reference electrode: the reference takes the "
                                          "last electrode, whose Jacobian reads past its potential rows")
        else:
            self.calibration = []
        if self.refNode >= 0:
            # the reference electrode as an ElectrodeShapeNode: its cells,
            # nearest-neighbour distance and position (electrode.cpp:98-189)
            self.refCells = np.nonzero((mesh.cells == self.refNode).any(axis=1))[0]
            nb = np.unique(mesh.cells[self.refCells].ravel())
            nb = nb[nb != self.refNode]
            self.refRmin = np.sqrt(((mesh.nodes[nb] - mesh.nodes[self.refNode]) ** 2).sum(-1)).min()
        # bypass electrodes: nodes with one marker <= -10000 (searchElectrodes_
        # :815-826), matched to a sensor at the group's first node (:886-897)
        self.bypass = {}
        for nid in range(len(nm)):
            if nm[nid] <= -10000:
                self.bypass.setdefault(int(nm[nid]), []).append(nid)
        free = sorted(self.bypass, key=lambda m: -10000 - m)
        cand = list(mesh.findNodesIdxByMarker(-99))
        nodes = mesh.nodes
        self.elecNode = np.empty(len(ePos), np.int64)
        # complete electrode model (:804-851, :868-884): boundary faces with
        # marker <= -9999 grouped by id = -10000 - marker; a sensor takes the
        # first remaining electrode (by id) within 2 sqrt(area) or with its
        # index; every electrode then has an unknown after the n nodes
        bm = mesh.boundaryMarkers()
        cm = mesh.cellMarkers()
        if np.any((cm <= -10000) & (cm > -1000000)):
            raise NotImplementedError("electrode cells of the complete electrode model")
        self.cem = {}
        cemIds = sorted({int(-10000 - m) for m in bm[bm <= -9999]})
        self.cemPassive = None
        if cemIds:
            if self.refNode >= 0:
                raise NotImplementedError("a reference electrode with the complete electrode model")
            size, _, _ = boundary_geometry(mesh)
            if cemIds[0] < 0:
                # the passive body (marker -9999, id -1): an unknown after the
                # electrodes' with its contact terms and no current (:838-847)
                self.cemPassive = np.nonzero(bm == -9999)[0]
                cemIds = cemIds[1:]
            domains = []
            for eid in cemIds:
                faces = np.nonzero(bm == -10000 - eid)[0]
                un = np.unique(mesh.bounds[faces])
                domains.append((eid, faces, nodes[un].mean(0), size[faces].sum()))
            for i, p in enumerate(ePos):
                for j, (eid, faces, pos, area) in enumerate(domains):
                    if np.linalg.norm(p - pos) < np.sqrt(area) * 2.0 or eid == i:
                        self.cem[i] = faces
                        domains.pop(j)
                        break
            if len(self.cem) != len(ePos):
                raise NotImplementedError("complete electrode model with node or free electrodes beside it")
            self.elecNode[:] = mesh.nodeCount() + np.arange(len(ePos))
            self.elecCell = np.full(len(ePos), -1, np.int64)
            self.elecShape = np.zeros((len(ePos), dim + 1))
            self.elecPos = ePos
            self.sourceCenter = ePos.sum(0) / len(ePos)
            self.nodeCells = [np.zeros(0, np.int64)] * len(ePos)
            self.singNode = np.full(len(ePos), -1, np.int64)
            self.minRadius = np.zeros(len(ePos))
            return
        # free electrodes (no bypass group, no marker -99 node within 0.01):
        # ElectrodeShapeEntity of the cell holding the position (:979-985)
        self.elecCell = np.full(len(ePos), -1, np.int64)
        self.elecShape = np.zeros((len(ePos), dim + 1))
        for i, p in enumerate(ePos):
            hit = None
            for m in free:
                if np.linalg.norm(p - nodes[self.bypass[m][0]]) < 0.01:
                    hit = m
                    break
            if hit is not None:
                free.remove(hit)
                self.elecNode[i] = self.bypass[hit][0]
                continue
            for j, nid in enumerate(cand):
                if np.linalg.norm(p - nodes[nid]) < 0.01:
                    hit = j
                    break
            if hit is None:
                c, N = find_cell(mesh, p)
                if c < 0:
                    raise RuntimeError("There is a requested electrode that does not match the given mesh.")
                self.elecNode[i] = -1
                self.elecCell[i] = c
                self.elecShape[i] = N
                continue
            self.elecNode[i] = cand.pop(hit)
        self.elecPos = np.where((self.elecNode >= 0)[:, None], nodes[np.maximum(self.elecNode, 0)], ePos)
        self.sourceCenter = self.elecPos.sum(0) / len(self.elecPos)
        # cells around each electrode node, nearest neighbour distance; a free
        # electrode within 1e-4 of a node of its cell has its singular value
        # at that node (ElectrodeShapeEntity::setSingValue, electrode.cpp:207-249)
        self.nodeCells = []
        self.singNode = np.full(len(self.elecNode), -1, np.int64)
        self.minRadius = np.zeros(len(self.elecNode))
        for i, nid in enumerate(self.elecNode):
            if nid < 0:
                self.nodeCells.append(np.array([self.elecCell[i]]))
                cn = mesh.cells[self.elecCell[i]]
                d = np.sqrt(((nodes[cn] - self.elecPos[i]) ** 2).sum(-1))
                close = np.nonzero(d < 1e-4)[0]
                if len(close) == 0:
                    continue
                nid = int(cn[close[-1]])
            self.singNode[i] = nid
            cs = np.nonzero((mesh.cells == nid).any(axis=1))[0]
            if self.elecNode[i] >= 0:
                self.nodeCells.append(cs)
            nb = np.unique(mesh.cells[cs].ravel())
            nb = nb[nb != nid]
            self.minRadius[i] = np.sqrt(((nodes[nb] - nodes[nid]) ** 2).sum(-1)).min()

    def geom_mean_rho(self, i, rho):
        """ElectrodeShapeNode::geomMeanCellAttributes (electrode.cpp:102-127);
        a free electrode: the attribute of its cell
        (ElectrodeShapeEntity::geomMeanCellAttributes, electrode.cpp:252-266)."""
        if self.elecNode[i] < 0:
            return float(rho[self.elecCell[i]])
        v = rho[self.nodeCells[i]]
        return float(np.exp(np.log(v).sum() / len(v)))

    def primary(self, i, k):
        """exactDCSolution(mesh, elec, k, surfaceZ, setSingValue=true)
        (bertMisc.cpp:230-257) + ElectrodeShapeNode::setSingValue /
        ElectrodeShapeEntity::setSingValue with scale 0 (electrode.cpp:154-249)."""
        u = exact_dc(self.mesh.nodes, self.elecPos[i], k, self.surfaceZ, 0.0)
        sn = self.singNode[i]
        if sn < 0:
            return u
        r = self.minRadius[i]
        if k > 0.0:
            u[sn] = bessel_k0(np.array([r / 6.0 * k]))[0] / PI
        else:
            u[sn] = 1.0 / (2.0 * PI * r / 2.0)
        return u

    def primary_ref(self, k):
        """exactDCSolution of the reference electrode (eB of every pattern,
        checkPrimpotentials_ dcfemmodelling.cpp:2101-2102), singular value at
        its node as for a node electrode."""
        u = exact_dc(self.mesh.nodes, self.mesh.nodes[self.refNode], k, self.surfaceZ, 0.0)
        r = self.refRmin
        u[self.refNode] = bessel_k0(np.array([r / 6.0 * k]))[0] / PI if k > 0.0 else 1.0 / (2.0 * PI * r / 2.0)
        return u

    def rho_source(self, i, rho):
        """rhoSource of pattern i (DCSRMultiElectrodeModelling::calculateK,
        dcfemmodelling.cpp:2219-2244): the mean of the geometric means of eA
        and eB (the reference electrode) over those above TOLERANCE."""
        vals = [self.geom_mean_rho(i, rho)]
        if self.refNode >= 0:
            v = rho[self.refCells]
            vals.append(float(np.exp(np.log(v).sum() / len(v))))
        vals = [v for v in vals if v > TOLERANCE]
        return sum(vals) / len(vals)

    def source_rhs(self, i, n):
        """Point-source right-hand side of electrode i
        (ElectrodeShapeNode::assembleRHS, electrode.cpp:133-151;
        ElectrodeShapeEntity::assembleRHS: shape functions at the position,
        electrode.cpp:274-287)."""
        rhs = np.zeros(n)
        if self.elecNode[i] >= 0:
            rhs[self.elecNode[i]] = 1.0
        else:
            rhs[self.mesh.cells[self.elecCell[i]]] = self.elecShape[i]
        if self.refNode >= 0:
            rhs[self.refNode] -= 1.0    # dipole against the reference electrode (calculateK_ :1854-1855)
        return rhs

    def pot(self, i, sol):
        """Potential of the fields `sol` [.., n] at electrode i
        (ElectrodeShapeNode::pot, electrode.cpp:129-131;
        ElectrodeShapeEntity::pot -> MeshEntity::pot, meshentities.cpp:347-349)."""
        if self.elecNode[i] >= 0:
            return sol[..., self.elecNode[i]]
        return sol[..., self.mesh.cells[self.elecCell[i]]] @ self.elecShape[i]


# --------------------------------------------------------------------------- a4/a5
def element_matrices(mesh):
    """P1 element stiffness K_c = vol grad(N_i).grad(N_j) (ElementMatrix::
    ux2uy2uz2, elementmatrix.cpp:798-1122) and the mass factor: M_c =
    vol/((d+1)(d+2)) (1 + delta_ij) (ElementMatrix::u2, :682-796)."""
    dim = mesh.dim()
    p = mesh.nodes[mesh.cells][:, :, :dim]
    D = p[:, 1:, :] - p[:, :1, :]              # (C, d, d) rows = edges
    Dinv = np.linalg.inv(D)                     # columns = grads of N_1..N_d
    g = np.transpose(Dinv, (0, 2, 1))           # (C, d, d) row j = grad N_{j+1}
    g0 = -g.sum(axis=1, keepdims=True)
    grads = np.concatenate([g0, g], axis=1)     # (C, d+1, d)
    vol = np.abs(np.linalg.det(D)) / (2.0 if dim == 2 else 6.0)
    K = vol[:, None, None] * np.einsum("cid,cjd->cij", grads, grads)
    nv = dim + 1
    mfac = vol / (nv * (nv + 1))
    M = mfac[:, None, None] * (np.ones((nv, nv)) + np.eye(nv))[None]
    return K, M, vol, grads


# --------------------------------------------------------------- f4 topography
# Numerical primary potentials of the SR operator with topography
# (DCSRMultiElectrodeModelling::checkPrimpotentials_, dcfemmodelling.cpp:2009-2056):
# the total-field problem with rho = 1 on the P2 mesh (Mesh::createP2,
# mesh.cpp:1285-1330, quadratic triangles / tetrahedra), interpolated to the
# mesh nodes (the vertices of the P2 mesh, where the P2 field is its nodal value).

P2_EDGES = {1: [(0, 1)], 2: [(0, 1), (1, 2), (0, 2)], 3: [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]}


def p2_basis(d):
    """P2 Lagrange basis of a d-simplex as quadratic forms in the barycentric
    coordinates, phi_i = lam^T Q_i lam: vertex a, lam_a (2 lam_a - 1) =
    2 lam_a^2 - lam_a sum(lam); edge (p, q), 4 lam_p lam_q.  Nodes: the d + 1
    vertices, then the edges in P2_EDGES order."""
    nb = d + 1
    Q = []
    for a in range(nb):
        q = np.zeros((nb, nb))
        q[a, :] -= 0.5
        q[:, a] -= 0.5
        q[a, a] = 1.0
        Q.append(q)
    for p, r in P2_EDGES[d]:
        q = np.zeros((nb, nb))
        q[p, r] = q[r, p] = 2.0
        Q.append(q)
    return np.array(Q)


def _simplex_moment(d, idx):
    """Integral of prod(lam[idx]) over a d-simplex divided by its measure:
    d! prod(alpha!) / (d + |alpha|)!."""
    from math import factorial
    alpha = np.bincount(np.asarray(idx), minlength=d + 1)
    num = factorial(d)
    for a in alpha:
        num *= factorial(int(a))
    return num / factorial(d + len(idx))


def p2_tables(d):
    """Per-measure P2 element tables of a d-simplex: mass M[i, j] (exact, the
    reference's order-4 rules are exact for it, elementmatrix.cpp:761-782) and
    stiffness weights W[i, j, a, c] with K_ij = |T| sum_ac (grad lam_a .
    grad lam_c) W[i, j, a, c] (exact; order-2 rules, :872-1060)."""
    Q = p2_basis(d)
    nb, nn = d + 1, len(Q)
    I2 = np.array([[_simplex_moment(d, [b, e]) for e in range(nb)] for b in range(nb)])
    I4 = np.zeros((nb,) * 4)
    for a in range(nb):
        for b in range(nb):
            for c in range(nb):
                for e in range(nb):
                    I4[a, b, c, e] = _simplex_moment(d, [a, b, c, e])
    M = np.einsum("iab,jce,abce->ij", Q, Q, I4)
    W = 4.0 * np.einsum("iab,jce,be->ijac", Q, Q, I2)
    return M, W


def p2_mesh(mesh):
    """Mesh::createP2: the vertices keep their ids, one node per edge at its
    midpoint after them.  Returns nodes, cells (C x (d+1 + edges)) and
    boundaries (F x (d + face edges)) in P2_EDGES order."""
    n, d = mesh.nodeCount(), mesh.dim()
    c = mesh.cells.astype(np.int64)
    ce = np.sort(np.stack([c[:, [p, q]] for p, q in P2_EDGES[d]], 1), axis=2)
    keys = ce[..., 0] * n + ce[..., 1]
    uk, inv = np.unique(keys.ravel(), return_inverse=True)
    a, b = uk // n, uk % n
    nodes = np.vstack([mesh.nodes, 0.5 * (mesh.nodes[a] + mesh.nodes[b])])
    cells = np.hstack([c, n + inv.reshape(keys.shape)])
    bnd = mesh.bounds.astype(np.int64)
    be = np.sort(np.stack([bnd[:, [p, q]] for p, q in P2_EDGES[d - 1]], 1), axis=2)
    bk = be[..., 0] * n + be[..., 1]
    bounds = np.hstack([bnd, n + np.searchsorted(uk, bk)])
    return nodes, cells, bounds


def p2_primaries(mesh, setup, kvalues):
    """[nK * nE][n] primary potentials of unit point sources at the electrode
    nodes, total field on the P2 mesh with rho = 1 and the mixed and
    Dirichlet faces of the mesh (calculateK_ on the primary operator)."""
    if np.any(setup.elecNode < 0):
        raise NotImplementedError("numerical primaries: node electrodes only")
    if setup.bypass:
        raise NotImplementedError("numerical primaries with bypass electrodes")
    d, n = mesh.dim(), mesh.nodeCount()
    nodes, cells, bounds = p2_mesh(mesh)
    N = len(nodes)
    K1, _, vol, grads = element_matrices(mesh)
    M2, W2 = p2_tables(d)
    G = vol[:, None, None] * np.einsum("xad,xed->xae", grads, grads)    # |T| grad lam_a . grad lam_e
    Ke = np.einsum("xae,ijae->xij", G, W2)
    Me = vol[:, None, None] * M2[None]
    nl = cells.shape[1]
    rows = np.repeat(cells, nl, axis=1).ravel()
    cols = np.tile(cells, (1, nl)).ravel()
    bm = mesh.boundaryMarkers()
    mix = np.nonzero(bm == -2)[0]
    size, _, _ = boundary_geometry(mesh)
    Mf, _ = p2_tables(d - 1)
    fl = bounds.shape[1]
    hd = np.unique(bounds[bm == -3])
    P = np.empty((len(kvalues) * len(setup.elecNode), n))
    for ki, k in enumerate(kvalues):
        vals = [(Ke + (k * k) * Me).ravel() if k > 0.0 else Ke.ravel()]
        r, cc = [rows], [cols]
        if len(mix):
            alpha = mixed_bc_alpha(mesh, setup.sourceCenter, k)
            Be = (alpha[mix] * size[mix])[:, None, None] * Mf[None]
            bb = bounds[mix]
            r.append(np.repeat(bb, fl, axis=1).ravel())
            cc.append(np.tile(bb, (1, fl)).ravel())
            vals.append(Be.ravel())
        A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(r), np.concatenate(cc))), shape=(N, N)).tocsr()
        A.sum_duplicates()
        for dn in (hd, setup.calibration):
            if len(dn):
                A = A.tolil()
                A[dn, :] = 0.0
                A[:, dn] = 0.0
                A[dn, dn] = 1.0
                A = A.tocsr()
        lu = spla.splu(A.tocsc())
        for i, e in enumerate(setup.elecNode):
            rhs = np.zeros(N)
            rhs[e] = 1.0
            if setup.refNode >= 0:
                rhs[setup.refNode] -= 1.0    # dipole against the reference electrode
            P[i + ki * len(setup.elecNode)] = lu.solve(rhs)[:n]
    return P


def boundary_geometry(mesh):
    """Boundary measure, centre and unit normal (Boundary::norm)."""
    dim = mesh.dim()
    p = mesh.nodes[mesh.bounds]
    c = p.mean(axis=1)
    if dim == 2:
        t = p[:, 1, :] - p[:, 0, :]
        size = np.sqrt((t ** 2).sum(-1))
        nrm = np.stack([t[:, 1], -t[:, 0], np.zeros(len(t))], 1) / size[:, None]
    else:
        cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        a2 = np.sqrt((cr ** 2).sum(-1))
        size = a2 / 2.0
        nrm = cr / a2[:, None]
    return size, c, nrm


def mixed_bc_alpha(mesh, source, k):
    """mixedBoundaryCondition (dcfemmodelling.cpp:430-506), MIRRORSOURCE
    branch, mirror plane 0, fabs(r.n)."""
    dim = 2 if k > 0 else 3
    smir = np.array(source, np.float64)
    smir[dim - 1] = -source[dim - 1]
    _, c, nrm = boundary_geometry(mesh)
    r = source - c
    rm = smir - c
    ra = np.sqrt((r ** 2).sum(-1))
    rma = np.sqrt((rm ** 2).sum(-1))
    rn = np.abs((r * nrm).sum(-1))
    rmn = np.abs((rm * nrm).sum(-1))
    if k == 0:
        return (rma * rma * rn / ra + ra * ra * rmn / rma) / (rma * ra * (ra + rma))
    k0a = bessel_k0(ra * k)
    k0m = bessel_k0(rma * k)
    # faces far from the source at large k: both K0 underflow to 0 and the
    # quotient is 0/0; the reference sets those to 0 (the mask below)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = k * (rn / ra * bessel_k1(ra * k) + rmn / rma * bessel_k1(rma * k)) / (k0a + k0m)
    return np.where((np.abs(k0a) < TOLERANCE) | (np.abs(k0m) < TOLERANCE), 0.0, out)


def sparsity_pattern(mesh):
    """SparseMatrix::buildSparsityPattern (sparsematrix.h:966-1032): all node
    pairs sharing a cell, sorted column indices per row."""
    n = mesh.nodeCount()
    c = mesh.cells.astype(np.int64)
    nv = c.shape[1]
    rows = np.repeat(c, nv, axis=1).ravel()
    cols = np.tile(c, (1, nv)).ravel()
    A = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


def assemble(mesh, rho, k, source, geo=None):
    """System matrix of one wavenumber as DCSRMultiElectrodeModelling::
    calculateK assembles it (dcfemmodelling.cpp:2179-2182):
    dcfemDomainAssembleStiffnessMatrix (:163-242) + dcfemBoundaryAssemble-
    StiffnessMatrix (:243-292).  Returns scipy CSR with the reference's
    pattern (full symmetric storage)."""
    if geo is None:
        geo = element_matrices(mesh)
    K, M, vol, _ = geo
    n = mesh.nodeCount()
    rho = np.asarray(rho)
    rho = rho.astype(np.complex128 if np.iscomplexobj(rho) else np.float64)   # complex: 1 / rho complex
    use = np.abs(rho) > TOLERANCE
    sig = np.where(use, 1.0 / np.where(use, rho, 1.0), 0.0)
    Ae = (K + (k * k) * M) * sig[:, None, None] if k > 0.0 else K * sig[:, None, None]
    c = mesh.cells.astype(np.int64)
    nv = c.shape[1]
    rows = [np.repeat(c, nv, axis=1).ravel()]
    cols = [np.tile(c, (1, nv)).ravel()]
    vals = [Ae.ravel()]
    # mixed boundaries
    bm = mesh.boundaryMarkers()
    mix = np.nonzero(bm == -2)[0]
    if len(mix):
        size, _, _ = boundary_geometry(mesh)
        alpha = mixed_bc_alpha(mesh, source, k)
        lc = mesh.boundaryLeftCell()
        if np.any(lc[mix] < 0):
            raise RuntimeError("no cell found for a mixed boundary")
        d = mesh.dim()
        mb = size[mix] / (d * (d + 1))  # edge: len/6, triangle: area/12
        coef = alpha[mix] / rho[lc[mix]]
        Be = (mb * coef)[:, None, None] * (np.ones((d, d)) + np.eye(d))[None]
        bb = mesh.bounds[mix].astype(np.int64)
        rows.append(np.repeat(bb, d, axis=1).ravel())
        cols.append(np.tile(bb, (1, d)).ravel())
        vals.append(Be.ravel())
    # one COO -> CSR conversion keeps explicit zeros, i.e. the full pattern
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    if np.any(bm == -4):
        raise NotImplementedError("inhomogeneous Dirichlet boundaries (THROW_TO_IMPL in the reference)")
    A.sum_duplicates()
    A.sort_indices()
    # homogeneous Dirichlet faces (marker -3): rows and columns of their nodes
    # cleared, diagonal one (:280-298, assembleStiffnessMatrixHomogenDirichletBC
    # :141-153); pattern entries stay as explicit zeros
    hd = np.unique(mesh.bounds[bm == -3])
    if len(hd):
        flag = np.zeros(n, bool)
        flag[hd] = True
        rows_of = np.repeat(np.arange(n), np.diff(A.indptr))
        hit = flag[rows_of] | flag[A.indices]
        A.data[hit] = 0.0
        A.data[hit & (rows_of == A.indices)] = 1.0
    diag = A.diagonal().real
    if np.any(diag < TOLERANCE):
        raise RuntimeError("zero diagonal (all cells around a node have rho = 0)")
    return A.tocsr()


def stiffness_matrix(mesh, a):
    """RSparseMatrix::fillStiffnessMatrix(mesh, a) (sparsematrix.h:1043-1053)."""
    K, _, _, _ = element_matrices(mesh)
    return _scatter(mesh, K * np.asarray(a, np.float64)[:, None, None])


def mass_matrix(mesh, a):
    """RSparseMatrix::fillMassMatrix(mesh, a) (sparsematrix.h:1059-1069)."""
    _, M, _, _ = element_matrices(mesh)
    return _scatter(mesh, M * np.asarray(a, np.float64)[:, None, None])


def _scatter(mesh, Ae):
    n = mesh.nodeCount()
    c = mesh.cells.astype(np.int64)
    nv = c.shape[1]
    rows = np.repeat(c, nv, axis=1).ravel()
    cols = np.tile(c, (1, nv)).ravel()
    A = sp.coo_matrix((Ae.ravel(), (rows, cols)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


# --------------------------------------------------------------------------- a13
def map_model(mesh, model):
    """mapERTModel (dcfemmodelling.cpp:1211-1218) + ModellingBase::
    createMappedModel (modellingbase.cpp:401-470) with the default background
    -9e99 of response(model) (dcfemmodelling.h:164): parameter cells take
    model[marker], cells with a negative marker are prolongated."""
    model = np.asarray(model, np.float64)
    if len(model) == mesh.cellCount():
        return model.copy()
    cm = mesh.cellMarkers()
    if np.any(cm <= -1000000):
        raise NotImplementedError("fixed-value regions need the region manager")
    if cm.max() >= len(model):
        raise ValueError("marker >= model.size()")
    vals = np.where(cm >= 0, model[np.maximum(cm, 0)], 0.0)
    if np.all(cm < 0):
        raise ValueError("too many empty cells")
    return prolongate_empty_cells(mesh, vals)


def cell_neighbours(mesh):
    """Cell::findNeighborCell (meshentities.cpp:553-573): neighbour j of a cell
    is the unique other cell sharing the face opposite node j (Triangle /
    Tetrahedron::boundaryNodes, meshentities.cpp:1025-1033, :1153-1165), -1 if
    none.  Returns (nb [C, nv], face nodes [C, nv, dim])."""
    c = mesh.cells.astype(np.int64)
    C, nv = c.shape
    dim = nv - 1
    faces = np.stack([np.delete(c, j, axis=1) for j in range(nv)], axis=1)  # [C, nv, dim]
    key = np.sort(faces, axis=2).reshape(-1, dim)
    order = np.lexsort(key.T[::-1])
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    nb = -np.ones(C * nv, np.int64)
    for q in np.nonzero(same)[0]:
        nb[order[q]] = order[q + 1] // nv
        nb[order[q + 1]] = order[q] // nv
    return nb.reshape(C, nv), faces


def prolongate_empty_cells(mesh, vals):
    """Mesh::prolongateEmptyCellsValues with background -9e99 (mesh.cpp:2247-2318):
    sweeps in which every empty cell (|v| < 1e-12) with filled neighbours takes
    their mean weighted by |n * (1, 1, 0)| + 1e-6 (3D) or |n_x| + 1e-6 (2D), n the
    unit normal of the common face (RVector3::normXY for edges, pos.cpp:185-188;
    TriangleShape::norm, shape.cpp:535-540); values of a sweep use the state
    before it."""
    vals = np.array(vals, np.float64)
    nb, faces = cell_neighbours(mesh)
    xyz = mesh.nodes
    dim = mesh.dim()
    while True:
        empty = np.nonzero(np.abs(vals) < TOLERANCE)[0]
        if len(empty) == 0:
            return vals
        new = {}
        for i in empty:
            val = 0.0
            weight = 0.0
            for j in range(nb.shape[1]):
                o = nb[i, j]
                if o < 0 or abs(vals[o]) <= TOLERANCE:
                    continue
                p = xyz[faces[i, j]]
                if dim == 2:
                    d = p[0] - p[1]
                    w = abs(d[1]) / np.hypot(d[0], d[1]) + 1e-6
                else:
                    n = np.cross(p[1] - p[0], p[2] - p[0])
                    n = n / np.linalg.norm(n)
                    w = np.hypot(n[0], n[1]) + 1e-6
                val += vals[o] * w
                weight += w
            if weight > 1e-8:
                new[i] = val / weight
        if not new:
            raise RuntimeError("cannot fill the empty cells")
        for i, v in new.items():
            vals[i] = v


# --------------------------------------------------------------------------- a7/a9/a10
def bypass_map_rows(text, elecNode, nodeMarkers):
    """Resistors of a bypass.map file (assembleStiffnessMatrixDCFEMByPass_,
    dcfemmodelling.cpp:608-667) as (node, node, resistance) in file order:
    "e1 e2 R" joins the nodes of electrodes e1, e2 (from 1; rows with e1 < 1
    or e1 >= nE are skipped, the reference's bound check); "marker R" chains
    the nodes carrying the marker in index order; '#' starts a comment;
    integers as atoi, resistances as atof rounded to float (toInt / toFloat,
    gimli.h:477-478); other rows are skipped."""
    import re
    out = []
    nodeMarkers = np.asarray(nodeMarkers)
    for line in text.splitlines():
        tok = line.split("#")[0].split()
        ints = [int(m.group()) if (m := re.match(r"[+-]?[0-9]+", t)) else 0 for t in tok]
        res = float(np.float32(float(re.match(r"[+-]?[0-9.]+([eE][+-]?[0-9]+)?", tok[-1]).group()))) if tok else 0.0
        if len(tok) == 3 and 1 <= ints[0] < len(elecNode):
            out.append((int(elecNode[ints[0] - 1]), int(elecNode[ints[1] - 1]), res))
        elif len(tok) == 2:
            ids = np.flatnonzero(nodeMarkers == ints[0])
            out += [(int(i), int(j), res) for i, j in zip(ids[:-1], ids[1:])]
    return out


class Forward:
    """DCSRMultiElectrodeModelling restated: response() and createJacobian()."""

    def __init__(self, mesh, data, kvalues=None, weights=None, total_field=False):
        """total_field: the operator without singularity removal, which needs
        no primary potentials and so admits topography."""
        self.mesh = mesh
        self.data = data
        self.setup = Setup(mesh, data.sensorPositions())
        if kvalues is None:
            kvalues, weights = kwave_list(mesh.dim(), self.setup.elecPos)
        self.k = np.asarray(kvalues, np.float64)
        self.w = np.asarray(weights, np.float64)
        self.geo = element_matrices(mesh)
        self.nE = len(self.setup.elecNode)
        self.subpots = None
        self.prim = None
        self.bypassR = {}      # node marker -> bypass resistance (electrodeBypassResistances.map)
        self.bypassMap = []    # (node, node, resistance) rows of a bypass.map file (bypass_map_rows)
        self.contactImpedance = None    # per electrode, Ohm m^2 (setContactImpedances)
        self.contactResistance = None   # per electrode, Ohm (contactResistance.map)

    def primaries(self):
        """checkPrimpotentials_ (dcfemmodelling.cpp:2086-2110): rows
        source + nE * kIdx; numerical P2 primaries with topography
        (:2009-2056)."""
        if self.prim is None and self.setup.topography:
            self.prim = p2_primaries(self.mesh, self.setup, self.k)
        if self.prim is None:
            P = np.empty((len(self.k) * self.nE, self.mesh.nodeCount()))
            for ki, k in enumerate(self.k):
                pr = self.setup.primary_ref(k) if self.setup.refNode >= 0 else 0.0
                for i in range(self.nE):
                    P[i + ki * self.nE] = self.setup.primary(i, k) - pr
            self.prim = P
        return self.prim

    def solve_subpotentials(self, rho):
        """DCSRMultiElectrodeModelling::calculateK (dcfemmodelling.cpp:2152-2296)
        for every wavenumber."""
        rho = np.asarray(rho, np.float64)
        P = self.primaries()
        ones = np.ones_like(rho)
        U = np.empty_like(P)
        src = self.setup.sourceCenter
        cal = self.setup.calibration
        for ki, k in enumerate(self.k):
            S = assemble(self.mesh, rho, k, src, self.geo)
            S1 = assemble(self.mesh, ones, k, src, self.geo)
            if cal:
                # calibration nodes in both matrices (:2182, :2192), their
                # right-hand side zeroed (:2280-2283)
                S, S1 = S.tolil(), S1.tolil()
                for M in (S, S1):
                    M[cal, :] = 0.0
                    M[:, cal] = 0.0
                    M[cal, cal] = 1.0
                S, S1 = S.tocsr(), S1.tocsr()
            lu = spla.splu(S.tocsc())
            for i in range(self.nE):
                rs = self.setup.rho_source(i, rho)
                prim = P[i + ki * self.nE] * rs
                rhs = S1 @ prim / rs - S @ prim
                rhs[cal] = 0.0
                U[i + ki * self.nE] = lu.solve(rhs) + prim
        return U

    def solve_point_source(self, rho):
        """Non-SR total-field potentials: unit point source at the electrode
        node (DCMultiElectrodeModelling::calculateK_, dcfemmodelling.cpp:1755-1928
        with ElectrodeShapeNode::assembleRHS, electrode.cpp:133-151)."""
        rho = np.asarray(rho)
        n = self.mesh.nodeCount()
        dt = np.complex128 if np.iscomplexobj(rho) else np.float64
        cem = self.setup.cem
        if cem:
            if dt != np.float64:
                raise NotImplementedError("the complete electrode model is real only (dcfemmodelling.cpp:412-417)")
            n = n + len(self.setup.elecNode) + (self.setup.cemPassive is not None)   # unknowns after the nodes
        U = np.empty((len(self.k) * self.nE, n), dt)
        for ki, k in enumerate(self.k):
            S = assemble(self.mesh, rho, k, self.setup.sourceCenter, self.geo)
            if cem:
                S = self.cem_matrix(S)
            if dt == np.float64 and (self.setup.bypass or self.bypassMap):
                # assembleStiffnessMatrixDCFEMByPass_ (:601-707): the bypass.map
                # rows, then 1e-6 Ohm (or the group's resistance) between
                # consecutive nodes of a group (real systems only, :597-599)
                pairs = list(self.bypassMap)
                for m, ids in self.setup.bypass.items():
                    pairs += [(a1, a2, self.bypassR.get(m, 1e-6)) for a1, a2 in zip(ids[:-1], ids[1:])]
                S = S.tolil()
                for a1, a2, r in pairs:
                    g = 1.0 / r
                    S[a1, a1] += g
                    S[a2, a2] += g
                    S[a1, a2] -= g
                    S[a2, a1] -= g
                S = S.tocsr()
            if self.setup.calibration:
                # calibration nodes: rows and columns cleared, diagonal one,
                # after the bypass resistors (:1835); the right-hand side stays
                S = S.tolil()
                for c in self.setup.calibration:
                    S[c, :] = 0.0
                    S[:, c] = 0.0
                    S[c, c] = 1.0
                S = S.tocsr()
            lu = spla.splu(S.tocsc())
            for i in range(self.nE):
                rhs = self.setup.source_rhs(i, n).astype(dt)
                U[i + ki * self.nE] = lu.solve(rhs)
        return U

    def cem_matrix(self, S):
        """assembleCompleteElectrodeModel_ (dcfemmodelling.cpp:315-410) without a
        reference electrode: for electrode e (unknown v = n + e, contact
        impedance z_e, faces f): S[i, v] = S[v, i] -= int_f N_i / z,
        S[i, j] += int_f N_i N_j / z, S[v, v] = area / z; z = contactImpedance
        (default 1 Ohm m^2), or contactResistance * area when one is given."""
        n = self.mesh.nodeCount()
        groups = dict(self.setup.cem)
        if self.setup.cemPassive is not None:
            groups[len(self.setup.elecNode)] = self.setup.cemPassive
        nE = len(groups)
        d = self.mesh.dim()
        size, _, _ = boundary_geometry(self.mesh)
        S = sp.bmat([[S, None], [None, sp.csr_matrix((nE, nE))]]).tolil()
        for e, faces in groups.items():
            area = size[faces].sum()
            z = self.contactImpedance[e] if self.contactImpedance is not None else 1.0
            if self.contactResistance is not None:
                z = self.contactResistance[e] * area
            v = n + e
            for f in faces:
                nodes = self.mesh.bounds[f]
                u = size[f] / d                              # int_f N_i (P1 face)
                Mf = size[f] / (d * (d + 1)) * (np.ones((d, d)) + np.eye(d))
                for a, na in enumerate(nodes):
                    S[na, v] -= u / z
                    S[v, na] -= u / z
                    for b, nb in enumerate(nodes):
                        S[na, nb] += Mf[a, b] / z
            S[v, v] = area / z
        return S.tocsr()

    def potentials(self, U=None):
        """Wavenumber sum (dcfemmodelling.cpp:1707-1729)."""
        U = self.subpots if U is None else U
        nK = len(self.k)
        return sum(self.w[ki] * U[ki * self.nE:(ki + 1) * self.nE] for ki in range(nK))

    def electrode_map(self, sol):
        """DataMap::collect (datamap.cpp:57-98): map[source][electrode]."""
        return np.stack([self.setup.pot(f, sol) for f in range(self.nE)], axis=1)

    def geometric_factors(self):
        return geometric_factors(self.data.sensorPositions(), self.data.abmn(), self.mesh.dim())

    def response(self, model):
        """DCMultiElectrodeModelling::response (dcfemmodelling.cpp:1085-1197)."""
        model = np.asarray(model, np.float64)
        if model.min() < TOLERANCE:
            raise ValueError("response for model with negative or zero resistivity is not defined")
        kf = self.data["k"]
        if not self.data.haveData("k") or np.min(np.abs(kf)) < TOLERANCE:
            kf = self.geometric_factors()
            self.data.set("k", kf)
        rho = map_model(self.mesh, model)
        self.subpots = self.solve_subpotentials(rho)
        emap = self.electrode_map(self.potentials())
        u = collect_data(emap, self.data.abmn(), False)
        ur = collect_data(emap, self.data.abmn(), True)
        resp = np.rint(u / 1e-10) * 1e-10 * kf
        respRez = np.rint(ur / 1e-10) * 1e-10 * kf
        return np.sqrt(np.abs(resp * respRez))

    def jacobian(self, model):
        """createJacobian_ (dcfemmodelling.cpp:1318-1408) with
        createSensitivityCol / calc2 (bertJacobian.cpp:170-219, :367-521)."""
        model = np.asarray(model, np.float64)
        if self.subpots is None:
            self.response(model)
        J = sensitivity_calc2(self.mesh, self.geo, self.data.abmn(), self.subpots,
                              self.k, self.w, self.data.sensorCount())
        if len(model) == J.shape[1]:
            J *= self.data["k"][:, None] / (model * model)[None, :]
        return J


def response_total(fwd, model):
    """DCMultiElectrodeModelling::response (dcfemmodelling.cpp:1085-1197) with
    the total-field calculateK_ (:1755-1928): unit point sources, no
    singularity removal; the response rounding and reciprocity of the SR path."""
    model = np.asarray(model, np.float64)
    kf = fwd.data["k"]
    if not fwd.data.haveData("k") or np.min(np.abs(kf)) < TOLERANCE:
        kf = fwd.geometric_factors()
        fwd.data.set("k", kf)
    fwd.subpots = fwd.solve_point_source(map_model(fwd.mesh, model))
    emap = fwd.electrode_map(fwd.potentials())
    u = collect_data(emap, fwd.data.abmn(), False)
    ur = collect_data(emap, fwd.data.abmn(), True)
    return np.sqrt(np.abs(np.rint(u / 1e-10) * 1e-10 * kf * (np.rint(ur / 1e-10) * 1e-10 * kf)))


def response_complex(fwd, model_re, model_im):
    """DCMultiElectrodeModelling::response, complex branch
    (dcfemmodelling.cpp:1103-1120): complex cell resistivities (mapERTModel of
    both halves, :1199-1209), total-field potentials of the complex system
    (calculateK_, :1755-1928), k * ((uAM - uAN) - (uBM - uBN)) of the real and
    the imaginary potential maps, no rounding, no reciprocity.  Returns (re, im)."""
    re = map_model(fwd.mesh, np.asarray(model_re, np.float64))
    im = map_model(fwd.mesh, np.asarray(model_im, np.float64))
    kf = fwd.data["k"]
    if not fwd.data.haveData("k") or np.min(np.abs(kf)) < TOLERANCE:
        kf = fwd.geometric_factors()
        fwd.data.set("k", kf)
    fwd.subpots = fwd.solve_point_source(re + 1j * im)
    emap = fwd.electrode_map(fwd.potentials())
    u = collect_data(emap, fwd.data.abmn(), False) * kf
    return u.real, u.imag


def jacobian_complex(fwd, model_re, model_im):
    """createJacobian_ complex (dcfemmodelling.cpp:1410-1444): calc2 on the
    complex potentials without conjugation, divided by m^2 / k (m complex)."""
    J = sensitivity_calc2(fwd.mesh, fwd.geo, fwd.data.abmn(), fwd.subpots, fwd.k, fwd.w,
                          fwd.data.sensorCount())
    m = np.asarray(model_re, np.float64) + 1j * np.asarray(model_im, np.float64)
    if len(m) == J.shape[1]:
        J = J / ((m * m)[None, :] / fwd.data["k"][:, None])
    return J.real, J.imag


def collect_data(emap, abmn, reciprocity):
    """DataMap::data (datamap.cpp:159-231)."""
    a, b, m, n = [abmn[:, i].copy() for i in range(4)]
    if reciprocity:
        a, m = m, a
        b, n = n, b

    def g(i, j):
        ok = (i > -1) & (j > -1)
        return np.where(ok, emap[np.maximum(i, 0), np.maximum(j, 0)], 0.0)

    return (g(a, m) - g(a, n)) - (g(b, m) - g(b, n))


def sensitivity_calc2(mesh, geo, abmn, U, kvals, weights, nE):
    """CreateSensitivityColMT::calc2 (bertJacobian.cpp:170-219), vectorised
    over cells: J[i][marker_c] += w_k (m-n)^T (K_c + k^2 M_c) (a-b) over the
    cell's nodes, potentials row = electrode + nE * kIdx."""
    K, M, _, _ = geo
    cm = mesh.cellMarkers()
    use = cm >= 0
    nModel = int(cm.max()) + 1
    cells = mesh.cells[use]
    mk = cm[use]
    K = K[use]
    M = M[use]
    D = len(abmn)
    J = np.zeros((D, nModel), np.result_type(U.dtype, np.float64))
    nn = U.shape[1]
    zero = np.zeros(nn)
    for ki, k in enumerate(kvals):
        S = M * (k * k) + K
        Uk = U[ki * nE:(ki + 1) * nE]
        for i in range(D):
            a, b, m, n = abmn[i]
            va = Uk[a] if a > -1 else zero
            vb = Uk[b] if b > -1 else zero
            vm = Uk[m] if m > -1 else zero
            vn = Uk[n] if n > -1 else zero
            dab = (va - vb)[cells]
            dmn = (vm - vn)[cells]
            val = np.einsum("ci,cij,cj->c", dmn, S, dab)
            if np.iscomplexobj(val):
                J[i] += weights[ki] * (np.bincount(mk, weights=val.real, minlength=nModel)
                                       + 1j * np.bincount(mk, weights=val.imag, minlength=nModel))
            else:
                J[i] += weights[ki] * np.bincount(mk, weights=val, minlength=nModel)
    return J

"""Homogeneous Dirichlet boundaries on the GPU: faces with marker -3 clear the
rows and columns of their nodes and set the diagonal to one after the
boundary terms (core/src/bert/dcfemmodelling.cpp:280-298 with
assembleStiffnessMatrixHomogenDirichletBC :141-153), in both SR matrices,
the total-field matrix and the complex one; and 2.5D meshes with Neumann
faces only, which the reference solves with the total-field operator and no
calibration node.  Against the oracle's restatement (equal to the reference
build in tests/test_oracle.py) and the reference build itself."""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads
from oracle import ert_oracle as eo

pytestmark = pytest.mark.gpu


def _fop(mesh, data, sr, method, cplx=False):
    from pygimli_amd.ertModelling import ERTModelling
    fop = ERTModelling(sr=sr)
    if cplx:
        fop.setComplex(True)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method=method, rtol=1e-12, coarse=0)
    return fop


@pytest.mark.parametrize("name", ["small_2d", "small_3d"])
@pytest.mark.parametrize("sr", [True, False])
@pytest.mark.parametrize("method", ["band", "block64", "mixed"])
def test_dirichlet_vs_oracle(name, sr, method):
    mesh, data, rho = workloads.small_dirichlet(name)
    if method == "band" and mesh.dim() == 3:
        pytest.skip("direct band solver: 2D meshes")
    fop = _fop(mesh, data, sr, method)
    if method == "mixed":
        fop.solverOptions.update(rtol=1e-10)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=not sr)
    o = f.response(rho) if sr else eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print(name, "sr", sr, method, e1, e2, fop.lastStats)
    tol = 1e-7 if method == "mixed" else 1e-9
    assert e1 < tol and e2 < tol
    # the potentials at the Dirichlet nodes are the ones the cleared rows fix
    hd = np.unique(mesh.bounds[mesh.boundaryMarkers() == -3])
    U = fop.subPotentials().cpu().numpy()
    if sr:
        P = f.primaries()
        assert np.abs(U[:, hd] - P[:, hd]).max() <= tol * np.abs(P).max()
    else:
        assert np.abs(U[:, hd]).max() <= tol * np.abs(U).max()


@pytest.mark.parametrize("sr", [True, False])
def test_dirichlet_vs_reference_build(sr):
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build oracle/_ref not present")
    mesh, data, rho = workloads.small_dirichlet("small_3d")
    s = ref.Session(mesh, copy.deepcopy(data), nthreads=1, sr=sr)
    r_ref = s.response(rho)
    J_ref = s.jacobian(rho)
    fop = _fop(mesh, data, sr, "block64")
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    e1 = np.abs(r - r_ref).max() / np.abs(r_ref).max()
    e2 = np.linalg.norm(J - J_ref) / np.linalg.norm(J_ref)
    print("Dirichlet vs reference build, sr", sr, e1, e2)
    assert e1 < 1e-9 and e2 < 1e-9


def test_dirichlet_complex_vs_oracle():
    mesh, data, rho = workloads.small_dirichlet("small_2d")
    rng = np.random.default_rng(1)
    rim = -rho * rng.uniform(0.005, 0.05, len(rho))
    fop = _fop(mesh, data, False, "band", cplx=True)
    resp = fop.response(np.concatenate([rho, rim]))
    D = data.size()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    orr, ori = eo.response_complex(f, rho, rim)
    e = max(np.abs(resp[:D] - orr).max() / np.abs(orr).max(), np.abs(resp[D:] - ori).max() / np.abs(ori).max())
    print("complex Dirichlet", e)
    assert e < 1e-9


@pytest.mark.parametrize("method", ["band", "block64"])
def test_neumann_2d_total_field(method):
    """Neumann faces only in 2.5D: the total field against the oracle and
    the reference build; singularity removal too, with numerical primaries
    (the reference counts such a mesh as topography).  Without a Robin or
    Dirichlet term the
    matrix of the smallest wavenumber is held away from singular only by
    k^2 times the mass matrix, so its FP64 direct solve leaves a relative
    residual near 3e-10: the solves run to rtol 1e-9 and the comparison
    allows 1e-7."""
    from oracle import ref
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_2d_neumann()
    fop = _fop(mesh, data, False, method)
    fop.solverOptions.update(rtol=1e-9)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print("Neumann 2D", method, e1, e2, fop.lastStats)
    assert e1 < 1e-7 and e2 < 1e-7
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), nthreads=1, sr=False)
        rr = s.response(rho)
        print("Neumann 2D vs reference build", np.abs(r - rr).max() / np.abs(rr).max())
        assert np.abs(r - rr).max() <= 1e-7 * np.abs(rr).max()
    g = ERTModelling(sr=True)
    g.setMesh(mesh)
    g.setData(copy.deepcopy(data))
    g.solverOptions.update(method=method, rtol=1e-9, coarse=0)
    rs = g.response(rho)
    fs = eo.Forward(mesh, copy.deepcopy(data), g.kValues(), g.weights())
    os_ = fs.response(rho)
    e3 = np.abs(rs - os_).max() / np.abs(os_).max()
    print("Neumann 2D SR", method, e3)
    assert e3 < 1e-7


def test_neumann_3d_and_inhomogeneous_dirichlet_raise():
    """3D Neumann-only domains without a reference electrode node (the
    reference takes the last electrode, and its Jacobian then reads past its
    potential rows) and inhomogeneous Dirichlet faces (marker -4, not
    implemented in the reference either) are refused with an error."""
    from pygimli_amd.ertModelling import ERTModelling
    from pygimli_amd.mesh import Mesh
    mesh, data, _ = workloads.small_3d()
    bm = mesh.boundaryMarkers().copy()
    bm[bm == -2] = -1
    neu = Mesh(3, mesh.nodes, mesh.cells, mesh.cellMarkers(), mesh.bounds, bm, mesh.nodeMarkers())
    fop = ERTModelling(sr=False)
    fop.setMesh(neu)
    with pytest.raises(NotImplementedError):
        fop.setData(copy.deepcopy(data))
    bm = mesh.boundaryMarkers().copy()
    bm[bm == -2] = -4
    dm = Mesh(3, mesh.nodes, mesh.cells, mesh.cellMarkers(), mesh.bounds, bm, mesh.nodeMarkers())
    with pytest.raises(RuntimeError):
        ERTModelling(sr=False).setMesh(dm)


@pytest.mark.parametrize("kw", [dict(neumann=True, calibration=True), dict(neumann=True, calibration=False),
                                dict(neumann=False)], ids=["neumann-cal", "neumann-node0", "mixed"])
@pytest.mark.parametrize("method", ["block64", "pcg64"])
def test_reference_electrode_and_calibration(kw, method):
    """A reference electrode node (marker -999) makes every source a dipole
    against it; in a 3D Neumann-only domain the calibration node (marker
    -1000, else node 0) is a homogeneous Dirichlet node after the bypass
    resistors (dcfemmodelling.cpp:1010-1071, :1517-1523, :1835).  Against the
    oracle (equal to the reference build in tests/test_oracle.py), potentials
    included, and against the reference build; SR refuses it."""
    from oracle import ref
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_3d_refnode(**kw)
    fop = _fop(mesh, data, False, method)
    fop.solverOptions.update(rtol=1e-11)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    U = fop.subPotentials().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    e3 = np.abs(U - f.subpots).max() / np.abs(f.subpots).max()
    print("reference electrode", kw, method, e1, e2, e3, fop.lastStats)
    assert e1 < 1e-8 and e2 < 1e-8 and e3 < 1e-8
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), nthreads=1, sr=False)
        rr = s.response(rho)
        assert np.abs(r - rr).max() <= 1e-8 * np.abs(rr).max()
    # singularity removal with dipole primaries (the reference electrode's
    # analytic potential subtracted, rho_s the mean of both geometric means;
    # a Neumann domain: numerical P2 dipole primaries, calibration nodes in
    # both matrices and their right-hand side zeroed)
    g = ERTModelling()
    g.setMesh(mesh)
    g.setData(copy.deepcopy(data))
    g.solverOptions.update(method=method, rtol=1e-11, coarse=0)
    rs = g.response(rho)
    g.createJacobian(rho)
    Js = g.jacobian().cpu().numpy()
    Us = g.subPotentials().cpu().numpy()
    fs = eo.Forward(mesh, copy.deepcopy(data), g.kValues(), g.weights())
    os_ = fs.response(rho)
    oJs = fs.jacobian(rho)
    e4 = np.abs(rs - os_).max() / np.abs(os_).max()
    e5 = np.linalg.norm(Js - oJs) / np.linalg.norm(oJs)
    e6 = np.abs(Us - fs.subpots).max() / np.abs(fs.subpots).max()
    print("reference electrode SR", method, e4, e5, e6)
    assert e4 < 1e-8 and e5 < 1e-8 and e6 < 1e-8
    if ref.available():
        s2 = ref.Session(mesh, copy.deepcopy(data), nthreads=1)
        if kw["neumann"]:
            s2.primary_p2()
        s2.response(rho)
        U2 = s2.subpotentials()
        assert np.abs(Us - U2).max() <= 1e-8 * np.abs(U2).max()


@pytest.mark.parametrize("neumann", [False, True])
def test_reference_electrode_complex(neumann):
    """Complex resistivities with a reference electrode node (and, in a 3D
    Neumann domain, the calibration node with a complex unit diagonal):
    against the oracle and the reference build's complex branch."""
    from oracle import ref
    mesh, data, rho = workloads.small_3d_refnode(neumann=neumann)
    rng = np.random.default_rng(5)
    rim = -rho * rng.uniform(0.005, 0.03, len(rho))
    fop = _fop(mesh, data, False, "block64", cplx=True)
    resp = fop.response(np.concatenate([rho, rim]))
    D = data.size()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    orr, ori = eo.response_complex(f, rho, rim)
    e = max(np.abs(resp[:D] - orr).max() / np.abs(orr).max(), np.abs(resp[D:] - ori).max() / np.abs(ori).max())
    print("complex reference electrode, neumann", neumann, e)
    assert e < 1e-8
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), nthreads=1, complex_=True)
        rr, ri = s.response_complex(rho, rim)
        assert np.abs(resp[:D] - rr).max() <= 1e-8 * np.abs(rr).max()
        assert np.abs(resp[D:] - ri).max() <= 1e-8 * np.abs(ri).max()


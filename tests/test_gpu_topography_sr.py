"""Singularity removal with topography on the GPU: the primary potentials come
from the total-field problem with rho = 1 on the P2 mesh (Mesh::createP2;
DCSRMultiElectrodeModelling::checkPrimpotentials_,
core/src/bert/dcfemmodelling.cpp:2009-2056), solved by the library on a
quadratic-mesh context (pgert_set_mesh_p2) and handed to the SR operator
(pgert_set_primary_potentials).  Against the oracle's P2 restatement (equal
to the reference build's own primaries to 6e-13 in tests/test_oracle.py)
and the reference build itself."""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads
from oracle import ert_oracle as eo

pytestmark = pytest.mark.gpu


def _fop(mesh, data, method, rtol=1e-12):
    from pygimli_amd.ertModelling import ERTModelling
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method=method, rtol=rtol, coarse=0)
    return fop


@pytest.mark.parametrize("name", ["small_2d_topo", "small_3d_topo"])
@pytest.mark.parametrize("method", ["band", "block64", "mixed"])
def test_sr_topography_vs_oracle(name, method):
    mesh, data, rho = getattr(workloads, name)()
    if method == "band" and mesh.dim() == 3:
        pytest.skip("direct band solver: 2D meshes")
    fop = _fop(mesh, data, method, 1e-10 if method == "mixed" else 1e-12)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights())
    assert f.setup.topography
    P = f.primaries()
    nK, nE = len(fop.kValues()), len(f.setup.elecNode)
    Pg = fop._prim.cpu().numpy().transpose(0, 2, 1).reshape(nK * nE, -1)
    o = f.response(rho)
    oJ = f.jacobian(rho)
    e0 = np.abs(Pg - P).max() / np.abs(P).max()
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print(name, method, "primaries", e0, "rhoa", e1, "J", e2, fop.primaryStats, fop.lastStats)
    tol = 1e-7 if method == "mixed" else 1e-9
    assert e0 < tol and e1 < tol and e2 < tol


def test_sr_topography_vs_reference_build():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build oracle/_ref not present")
    mesh, data, rho = workloads.small_3d_topo()
    s = ref.Session(mesh, copy.deepcopy(data), nthreads=1)
    P = s.primary_p2()
    r_ref = s.response(rho)
    J_ref = s.jacobian(rho)
    fop = _fop(mesh, data, "block64")
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    nE = data.sensorCount()
    Pg = fop._prim.cpu().numpy().transpose(0, 2, 1).reshape(-1, mesh.nodeCount())
    e0 = np.abs(Pg - P).max() / np.abs(P).max()
    e1 = np.abs(r - r_ref).max() / np.abs(r_ref).max()
    e2 = np.linalg.norm(J - J_ref) / np.linalg.norm(J_ref)
    print("SR topography vs reference build", nE, e0, e1, e2)
    assert e0 < 1e-9 and e1 < 1e-9 and e2 < 1e-9


def test_sr_topography_primaries_follow_the_set_up():
    """The numerical primaries are computed once per mesh / electrode /
    wavenumber set; a second response reuses them, new wavenumbers make new
    ones, and the library refuses the SR forward when they are stale."""
    import ctypes

    import torch
    from pygimli_amd.ertModelling import _dptr
    mesh, data, rho = workloads.small_2d_topo()
    fop = _fop(mesh, data, "band")
    r1 = fop.response(rho)
    p1 = fop._prim
    r2 = fop.response(rho)
    assert fop._prim is p1 and np.array_equal(r1, r2)
    k, w = fop.kValues()[:-1].copy(), fop.weights()[:-1].copy()
    fop.setkValues(k)
    fop.setWeights(w)
    r3 = fop.response(rho)
    assert fop._prim is not p1 and fop._prim.shape[0] == len(k)
    assert np.abs(r3 - r1).max() > 0.0
    # stale primaries: new wavenumbers set on the library directly
    L, h = fop._ctx.L, fop._ctx.h
    kk = (ctypes.c_double * len(k))(*k)
    ww = (ctypes.c_double * len(w))(*w)
    assert L.pgert_set_wavenumbers(h, len(k), kk, ww) == 0
    out = torch.empty(data.size(), dtype=torch.float64, device=fop.device)
    rt = torch.as_tensor(rho, device=fop.device)
    opts = fop._opts()
    assert L.pgert_forward(h, _dptr(rt), ctypes.byref(opts), _dptr(out), None) != 0
    assert b"primary" in L.pgert_last_error(h)


def test_sr_topography_with_reference_electrode():
    """Topography, singularity removal and a reference electrode node: the P2
    primary problem is solved with the dipole sources as well (the
    reference's inner operator searches the same -999 node); against the
    oracle (equal to the reference build, 6e-15 for the primaries) and the
    reference build."""
    from oracle import ref
    mesh, data, rho = workloads.small_3d_topo_ref()
    fop = _fop(mesh, data, "block64")
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights())
    assert f.setup.refNode >= 0 and f.setup.topography
    o = f.response(rho)
    oJ = f.jacobian(rho)
    U = fop.subPotentials().cpu().numpy()
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    e3 = np.abs(U - f.subpots).max() / np.abs(f.subpots).max()
    print("SR topography + reference electrode", e1, e2, e3)
    assert e1 < 1e-9 and e2 < 1e-9 and e3 < 1e-9
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), nthreads=1)
        s.primary_p2()
        rr = s.response(rho)
        assert np.abs(r - rr).max() <= 1e-9 * np.abs(rr).max()


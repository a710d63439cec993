"""Complete electrode model on the GPU (assembleCompleteElectrodeModel_,
core/src/bert/dcfemmodelling.cpp:315-410, applied in calculateK_ :1790-1829):
the electrode unknowns are library nodes of no cell after the mesh nodes,
coupled to their faces (pgert_set_mesh_coupled), with the constant contact
terms added to every matrix (pgert_set_cem); the unit current enters at the
unknown and the electrode potential is its value.  Against the oracle
(equal to the reference build in tests/test_oracle.py) and the reference
build itself, with the default contact impedance, a contactImpedance.map and
setContactImpedances."""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads
from oracle import ert_oracle as eo

pytestmark = pytest.mark.gpu


def _fop(mesh, data, method):
    from pygimli_amd.ertModelling import ERTModelling
    fop = ERTModelling(sr=False)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method=method, rtol=1e-10 if method == "mixed" else 1e-12, coarse=0)
    return fop


@pytest.mark.parametrize("method", ["block64", "mixed", "pcg64"])
def test_cem_vs_oracle(method):
    mesh, data, rho = workloads.small_3d_cem()
    fop = _fop(mesh, data, method)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print("CEM", method, e1, e2, fop.lastStats)
    tol = 1e-7 if method == "mixed" else 1e-9
    assert e1 < tol and e2 < tol
    U = fop.subPotentials().cpu().numpy()
    assert U.shape[1] == mesh.nodeCount() + data.sensorCount()


def test_cem_contact_impedances_vs_reference_build(tmp_path, monkeypatch):
    from oracle import ref
    from pygimli_amd.bert import load_electrode_vector
    mesh, data, rho = workloads.small_3d_cem()
    monkeypatch.chdir(tmp_path)
    (tmp_path / "contactImpedance.map").write_text("1 0.5\n3 2.0\n9 0.25\n")
    fop = _fop(mesh, data, "block64")
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    zi = load_electrode_vector("contactImpedance.map", data.sensorCount(), 1.0)
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    f.contactImpedance = zi
    o = eo.response_total(f, rho)
    assert np.abs(r - o).max() <= 1e-9 * np.abs(o).max()
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), nthreads=1, sr=False)
        rr = s.response(rho)
        Jr = s.jacobian(rho)
        e1 = np.abs(r - rr).max() / np.abs(rr).max()
        e2 = np.linalg.norm(J - Jr) / np.linalg.norm(Jr)
        print("CEM vs reference build", e1, e2)
        assert e1 < 1e-9 and e2 < 1e-9
    # the same impedances through the API, the file gone
    (tmp_path / "contactImpedance.map").unlink()
    g = _fop(mesh, data, "block64")
    g.setContactImpedances(zi)
    assert np.abs(g.response(rho) - r).max() <= 1e-9 * np.abs(r).max()


def test_cem_refusals():
    """SR and complex resistivities with the complete electrode model are
    refused, as the reference cannot compute them either."""
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_3d_cem()
    sr = ERTModelling()
    sr.setMesh(mesh)
    with pytest.raises(NotImplementedError):
        sr.setData(copy.deepcopy(data))
    c = ERTModelling(sr=False)
    c.setComplex(True)
    c.setMesh(mesh)
    c.setData(copy.deepcopy(data))
    with pytest.raises(RuntimeError):
        c.response(np.concatenate([rho, -0.01 * rho]))


def test_cem_passive_body():
    """A passive body of the complete electrode model (faces with marker
    -9999: an unknown with contact terms and no current, after the
    electrodes' unknowns, dcfemmodelling.cpp:838-847, :1811): against the
    oracle and the reference build; the plate changes the result."""
    from oracle import ref
    mesh, data, rho = workloads.small_3d_cem_passive()
    fop = _fop(mesh, data, "block64")
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print("CEM passive body", e1, e2)
    assert e1 < 1e-9 and e2 < 1e-9
    assert fop.subPotentials().shape[1] == mesh.nodeCount() + data.sensorCount() + 1
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), nthreads=1, sr=False)
        rr = s.response(rho)
        assert np.abs(r - rr).max() <= 1e-9 * np.abs(rr).max()
    m0, _, _ = workloads.small_3d_cem()
    g = _fop(m0, data, "block64")
    assert np.abs(g.response(rho) - r).max() > 0.1 * np.abs(r).max()


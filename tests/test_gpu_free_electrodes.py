"""Free (in-cell) electrodes on the GPU: sensors off every mesh node, the
reference's ElectrodeShapeEntity (core/src/electrode.cpp:198-287, placed by
searchElectrodes_, core/src/bert/dcfemmodelling.cpp:979-985): point source
spread over the cell's nodes by the shape functions at the position,
potential interpolated the same way, rho_s the cell's resistivity, primary
potential from the position.  Against the oracle's restatement (equal to the
reference build to 1e-11 in tests/test_oracle.py) and the reference build
itself, with and without singularity removal, real and complex."""
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
@pytest.mark.parametrize("method", ["band", "block64"])
def test_free_electrodes_vs_oracle(name, sr, method):
    mesh, data, rho = workloads.small_free(name)
    if method == "band" and mesh.dim() == 3:
        pytest.skip("direct band solver: 2D meshes")
    fop = _fop(mesh, data, sr, method)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=not sr)
    assert np.all(f.setup.elecNode < 0)
    o = f.response(rho) if sr else eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print(name, "sr", sr, method, e1, e2, fop.lastStats)
    assert e1 < 1e-9 and e2 < 1e-9


@pytest.mark.parametrize("sr", [True, False])
def test_free_electrodes_vs_reference_build(sr):
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build oracle/_ref not present")
    mesh, data, rho = workloads.small_free("small_3d")
    s = ref.Session(mesh, copy.deepcopy(data), nthreads=1, sr=sr)
    r_ref = s.response(rho)
    J_ref = s.jacobian(rho)
    fop = _fop(mesh, data, sr, "block64")
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    e1 = np.abs(r - r_ref).max() / np.abs(r_ref).max()
    e2 = np.linalg.norm(J - J_ref) / np.linalg.norm(J_ref)
    print("free electrodes vs reference build, sr", sr, e1, e2)
    assert e1 < 1e-9 and e2 < 1e-9


def test_free_electrodes_complex_vs_oracle():
    mesh, data, rho = workloads.small_free("small_2d")
    rng = np.random.default_rng(1)
    rim = -rho * rng.uniform(0.005, 0.05, len(rho))
    fop = _fop(mesh, data, False, "band", cplx=True)
    resp = fop.response(np.concatenate([rho, rim]))
    D = data.size()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    orr, ori = eo.response_complex(f, rho, rim)
    fop.createJacobian(np.concatenate([rho, rim]))
    Jr, Ji = [x.cpu().numpy() for x in fop.jacobianComplex()]
    oJr, oJi = eo.jacobian_complex(f, rho, rim)
    errs = (np.abs(resp[:D] - orr).max() / np.abs(orr).max(), np.abs(resp[D:] - ori).max() / np.abs(ori).max(),
            np.linalg.norm(Jr - oJr) / np.linalg.norm(oJr), np.linalg.norm(Ji - oJi) / np.linalg.norm(oJi))
    print("complex free electrodes", errs)
    assert max(errs) < 1e-9


def test_mixed_node_and_free_electrodes():
    """Half the sensors on nodes, half in cells, in one survey (block FP64)."""
    mesh, data, rho = workloads.small_3d()
    s = data.sensorPositions().copy()
    s[1::2] += np.array(workloads.FREE_SHIFT["small_3d"])
    d2 = workloads.make_data(s, data.abmn())
    fop = _fop(mesh, d2, True, "block64")
    r = fop.response(rho)
    f = eo.Forward(mesh, copy.deepcopy(d2), fop.kValues(), fop.weights())
    assert np.all(f.setup.elecNode[1::2] < 0) and np.all(f.setup.elecNode[0::2] >= 0)
    o = f.response(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    print("mixed node / free electrodes", e1)
    assert e1 < 1e-9

"""Complex resistivity (induced polarisation) on the GPU: ERTModelling(sr=False)
with setComplex(True), the reference's DCMultiElectrodeModelling complex branch
(core/src/bert/dcfemmodelling.cpp:1103-1120 response, :1410-1476 Jacobian).

Checked against the oracle's complex restatement (oracle/ert_oracle.py
response_complex / jacobian_complex, itself equal to the reference's own C++
core to 1e-14 in tests/test_oracle.py) and against the uniform-phase identity:
scaling every resistivity by one complex z scales the response by z and leaves
the Jacobian real and equal to the real-model one.
"""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads
from oracle import ert_oracle as eo

pytestmark = pytest.mark.gpu

CASES = ["small_2d", "small_3d"]


def _phases(rho, seed=1):
    rng = np.random.default_rng(seed)
    return -rho * rng.uniform(0.005, 0.05, len(rho))     # -5 .. -50 mrad


def _fop(mesh, data, method, k=None, w=None):
    from pygimli_amd.ertModelling import ERTModelling
    fop = ERTModelling(sr=False)
    fop.setComplex(True)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    if k is not None:
        fop.setkValues(k)
        fop.setWeights(w)
    fop.solverOptions.update(method=method, rtol=1e-12, coarse=0)
    return fop


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("method", ["band", "block64", "mixed"])
def test_complex_forward_and_jacobian_vs_oracle(name, method):
    mesh, data, rho = getattr(workloads, name)()
    if method == "band" and mesh.dim() == 3:
        pytest.skip("direct band solver: 2D meshes")
    rim = _phases(rho)
    fop = _fop(mesh, data, method)
    resp = fop.response(np.concatenate([rho, rim]))
    D = data.size()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights())
    orr, ori = eo.response_complex(f, rho, rim)
    e_re = np.abs(resp[:D] - orr).max() / np.abs(orr).max()
    e_im = np.abs(resp[D:] - ori).max() / np.abs(ori).max()
    fop.createJacobian(np.concatenate([rho, rim]))
    Jr, Ji = [x.cpu().numpy() for x in fop.jacobianComplex()]
    oJr, oJi = eo.jacobian_complex(f, rho, rim)
    j_re = np.linalg.norm(Jr - oJr) / np.linalg.norm(oJr)
    j_im = np.linalg.norm(Ji - oJi) / np.linalg.norm(oJi)
    print(name, method, "resp", e_re, e_im, "J", j_re, j_im, fop.lastStats)
    tol = 1e-9 if method != "mixed" else 1e-8
    assert e_re < tol and e_im < tol and j_re < tol and j_im < tol
    # real block form [[Re, -Im], [Im, Re]] of the reference's toRealMatrix
    J = fop.jacobian().cpu().numpy()
    P = Jr.shape[1]
    assert J.shape == (2 * D, 2 * P)
    assert np.array_equal(J[:D, P:], -Ji) and np.array_equal(J[D:, :P], Ji) and np.array_equal(J[D:, P:], Jr)
    assert fop.parameterCount == 2 * P


@pytest.mark.parametrize("name", CASES)
def test_complex_uniform_phase(name):
    mesh, data, rho = getattr(workloads, name)()
    z = np.exp(-0.03j)
    fop = _fop(mesh, data, "block64")
    r0 = fop.response(np.concatenate([rho, 0 * rho]))
    D = data.size()
    fop.createJacobian(np.concatenate([rho, 0 * rho]))
    J0 = fop.jacobianComplex()[0].cpu().numpy()
    rz = rho * z
    r1 = fop.response(np.concatenate([rz.real, rz.imag]))
    u0 = r0[:D] + 1j * r0[D:]
    u1 = r1[:D] + 1j * r1[D:]
    assert np.abs(r0[D:]).max() == 0.0
    assert np.abs(u1 - z * u0).max() <= 1e-9 * np.abs(u0).max()
    fop.createJacobian(np.concatenate([rz.real, rz.imag]))
    Jr, Ji = [x.cpu().numpy() for x in fop.jacobianComplex()]
    assert np.linalg.norm(Ji) <= 1e-9 * np.linalg.norm(J0)
    assert np.linalg.norm(Jr - J0) <= 1e-9 * np.linalg.norm(J0)


def test_complex_vs_reference_build():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build oracle/_ref not present")
    mesh, data, rho = workloads.small_2d()
    rim = _phases(rho, seed=5)
    ses = ref.Session(mesh, copy.deepcopy(data), complex_=True)
    rr, ri = ses.response_complex(rho, rim)
    Jr_r, Ji_r = ses.jacobian_complex(rho, rim)
    k, w = ses.kw()
    fop = _fop(mesh, data, "band", k, w)
    resp = fop.response(np.concatenate([rho, rim]))
    fop.createJacobian(np.concatenate([rho, rim]))
    Jr, Ji = [x.cpu().numpy() for x in fop.jacobianComplex()]
    D = data.size()
    assert np.abs(resp[:D] - rr).max() <= 1e-10 * np.abs(rr).max()
    assert np.abs(resp[D:] - ri).max() <= 1e-10 * np.abs(ri).max()
    assert np.linalg.norm(Jr - Jr_r) <= 1e-10 * np.linalg.norm(Jr_r)
    assert np.linalg.norm(Ji - Ji_r) <= 1e-10 * np.linalg.norm(Ji_r)


def test_complex_errors():
    from pygimli_amd.ertModelling import ERTModelling
    with pytest.raises(NotImplementedError):
        ERTModelling().setComplex(True)
    mesh, data, rho = workloads.small_2d()
    fop = ERTModelling(sr=False)
    fop.setComplex(True)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    with pytest.raises(ValueError):
        fop.response(rho[:-1])                     # not [re, im] of one length
    with pytest.raises(ValueError):
        fop.response(np.concatenate([rho * 0, rho * 0]))   # |rho| = 0


def test_complex_gauss_newton_step():
    """The complex inversion step as the reference sets it up: the real block
    Jacobian [[Re, -Im], [Im, Re]] (toRealMatrix), constraints repeated for
    both halves (RepeatDMatrix(C, 2)), model / data = [re, im]; the GPU CGLS
    (pgert_cgls) against the oracle's CGLS on the same inputs."""
    from oracle import cgls as cg
    from pygimli_amd.inversion import complex_constraints, smoothness_constraints, solveCGLSCDWWhtrans
    mesh, data, rho = workloads.small_2d()
    rim = _phases(rho)
    fop = _fop(mesh, data, "band")
    m = np.concatenate([rho, rim])
    resp = fop.response(m)
    fop.createJacobian(m)
    J = fop.jacobian()
    D2, P2 = J.shape
    C = complex_constraints(smoothness_constraints(mesh))
    assert C.shape[1] == P2 == fop.parameterCount
    rng = np.random.default_rng(2)
    b = 0.01 * resp * rng.standard_normal(D2)
    dW = 1.0 / (0.03 * np.abs(resp) + 1e-3)
    tm, td = np.ones(P2), np.ones(D2)
    wc, wm = np.ones(C.shape[0]), np.ones(P2)
    rough = C @ m
    x, it = solveCGLSCDWWhtrans(J, C, dW, b, np.zeros(P2), wc, wm, tm, td, 10.0, rough, maxIter=5, tol=0.0)
    xo, ito, _ = cg.cgls(J.cpu().numpy(), C, dW, b, np.zeros(P2), wc, wm, tm, td, 10.0, rough, max_iter=5, tol=0.0)
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    print("complex GN step: 5-iteration rel err", err)
    assert it == ito == 5 and err < 1e-9

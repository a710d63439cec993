"""GPU parity: libpgert (through the C-ABI) against the FP64 oracle.

Tolerances (relative): system matrix 1e-13; potentials 1e-6 (the solver
stops at a residual of 1e-10 here); rhoa 1e-7; Jacobian 1e-5 Frobenius.
"""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads
from oracle import ert_oracle as eo

pytestmark = pytest.mark.gpu

CASES = ["small_2d", "small_3d"]


def _case(name):
    mesh, data, rho = getattr(workloads, name)()
    return mesh, data, rho


@pytest.fixture(scope="module")
def oracle_results():
    out = {}
    for name in CASES:
        mesh, data, rho = _case(name)
        fo = eo.Forward(mesh, copy.deepcopy(data))
        rhoa = fo.response(rho)
        J = fo.jacobian(rho)
        out[name] = (fo, rhoa, J)
    return out


def _fop(name, method, rtol=1e-10):
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = _case(name)
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method=method, rtol=rtol)
    return fop, mesh, data, rho


@pytest.mark.parametrize("name", CASES)
def test_system_matrix(name, oracle_results):
    from pygimli_amd.ertModelling import system_matrix
    fo = oracle_results[name][0]
    fop, mesh, data, rho = _fop(name, "pcg64")
    for k in fo.k[:3]:
        A = system_matrix(fop, rho, k)
        Ao = eo.assemble(mesh, rho, k, fo.setup.sourceCenter)
        assert (A.indptr == Ao.indptr).all() and (A.indices == Ao.indices).all()
        err = abs(A - Ao).max() / abs(Ao).max()
        print(name, "k", k, "rel", err)
        assert err < 1e-13


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("method", ["pcg64", "block64", "mixed"])
def test_forward_and_jacobian(name, method, oracle_results):
    fo, rhoa_o, J_o = oracle_results[name]
    fop, mesh, data, rho = _fop(name, method)
    rhoa = fop.response(rho)
    st = fop.lastStats
    print(name, method, st, fop.timings())
    assert st["converged"] == 1
    U = fop.subPotentials().cpu().numpy()
    pot_err = np.linalg.norm(U - fo.subpots) / np.linalg.norm(fo.subpots)
    rho_err = np.abs(rhoa - rhoa_o).max() / np.abs(rhoa_o).max()
    print("potential rel", pot_err, "rhoa rel", rho_err)
    assert pot_err < 1e-6
    assert rho_err < 1e-7
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    jerr = np.linalg.norm(J - J_o) / np.linalg.norm(J_o)
    print("J rel fro", jerr)
    assert jerr < 1e-5

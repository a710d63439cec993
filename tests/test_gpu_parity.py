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
    coarse, smooth = 0, 2
    if method.endswith("+coarse"):      # multiplicative two-level cycle (the default with a coarse space)
        method, coarse = method.split("+")[0], 64
    elif method.endswith("+additive"):  # additive two-level preconditioner
        method, coarse, smooth = method.split("+")[0], 64, 0
    fop.solverOptions.update(method=method, rtol=rtol, coarse=coarse, smooth=smooth)
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
@pytest.mark.parametrize("method", ["pcg64", "block64", "mixed", "block64+coarse", "mixed+coarse", "block64+additive",
                                    "mixed+additive"])
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


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("method", ["block64", "mixed"])
def test_two_level_cycle_cuts_iterations(name, method, oracle_results):
    """The multiplicative two-level cycle with Chebyshev smoothing (smooth=2)
    and the additive preconditioner (smooth=0) reach the same potentials (the
    oracle's, to the solve tolerance), the cycle in fewer iterations; a solve
    with too few iterations for the cycle reports non-convergence."""
    fo = oracle_results[name][0]
    its = {}
    for smooth in (0, 2):
        fop, mesh, data, rho = _fop(name, method + "+coarse", rtol=1e-10)
        fop.solverOptions.update(smooth=smooth)
        fop.response(rho)
        its[smooth] = fop.lastStats["iterations"]
        U = fop.subPotentials().cpu().numpy()
        err = np.linalg.norm(U - fo.subpots) / np.linalg.norm(fo.subpots)
        print(name, method, "smooth", smooth, fop.lastStats, "potential rel", err)
        assert fop.lastStats["converged"] == 1 and err < 1e-6
    assert its[2] <= 0.6 * its[0]
    fop.solverOptions.update(smooth=2, max_iter=2, report_only=True)
    fop.response(rho)
    assert fop.lastStats["converged"] == 0


@pytest.mark.parametrize("name", CASES)
def test_jacobian_fp32_and_regions(name, oracle_results):
    """FP32 Jacobian (FP32 features in LDS) against the oracle, and a region
    model: 13 parameters shared by many cells (segmented sums over the
    marker-sorted cells, markers larger than a tile summed from chunk
    partials, scaling k_i / m_j^2 after the sum), FP64 and FP32."""
    import torch
    fo, rhoa_o, J_o = oracle_results[name]
    fop, mesh, data, rho = _fop(name, "block64")
    fop.response(rho)
    fop.jacobianDtype = torch.float32
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy().astype(np.float64)
    err = np.linalg.norm(J - J_o) / np.linalg.norm(J_o)
    print(name, "fp32 J rel fro", err)
    assert err < 1e-5
    # regions
    mesh = copy.deepcopy(mesh)
    cm = (np.arange(mesh.cellCount()) // 5) % 13
    mesh.setCellMarkers(cm)
    model = 20.0 + 10.0 * np.arange(13)
    fo2 = eo.Forward(mesh, copy.deepcopy(data))
    fo2.response(model)
    J_o2 = fo2.jacobian(model)
    from pygimli_amd.ertModelling import ERTModelling
    fop2 = ERTModelling()
    fop2.setMesh(mesh)
    fop2.setData(copy.deepcopy(data))
    fop2.solverOptions.update(method="block64", rtol=1e-11, coarse=0)
    for dt, tol in ((torch.float64, 1e-9), (torch.float32, 1e-5)):
        fop2.jacobianDtype = dt
        fop2.createJacobian(model)
        J2 = fop2.jacobian().cpu().numpy().astype(np.float64)
        assert J2.shape == J_o2.shape
        err = np.linalg.norm(J2 - J_o2) / np.linalg.norm(J_o2)
        print(name, dt, "regions J rel fro", err)
        assert err < tol


def test_band_direct_small_2d(oracle_results):
    """Direct banded Cholesky: the direct solution agrees with the oracle's
    SuperLU solution to rounding."""
    fo, rhoa_o, J_o = oracle_results["small_2d"]
    fop, mesh, data, rho = _fop("small_2d", "band")
    rhoa = fop.response(rho)
    print("band", fop.lastStats, fop.timings())
    assert fop.lastStats["converged"] == 1 and fop.lastStats["max_rel_residual"] < 1e-12
    U = fop.subPotentials().cpu().numpy()
    assert np.all(np.isfinite(U))
    pot_err = np.linalg.norm(U - fo.subpots) / np.linalg.norm(fo.subpots)
    rho_err = np.abs(rhoa - rhoa_o).max() / np.abs(rhoa_o).max()
    print("potential rel", pot_err, "rhoa rel", rho_err)
    assert pot_err < 1e-10 and rho_err < 1e-10
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    assert np.linalg.norm(J - J_o) / np.linalg.norm(J_o) < 1e-10


def test_config1_2d_vs_oracle():
    """2.5D workload at full size (5203 nodes, 10080 triangles, 41 electrodes,
    888 dipole-dipole data, 11 wavenumbers set explicitly) against the oracle."""
    from pygimli_amd import bert
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.ert2d_layers_block()
    an = bert.MeshAnalysis(mesh, data.sensorPositions())
    d = np.sqrt(((an.elecPos[:, None] - an.elecPos[None]) ** 2).sum(-1))
    k, w = bert.kwave_from_range(d[d > 0].min() / 2, d.max() * 2, 7, 4)
    fo = eo.Forward(mesh, copy.deepcopy(data), k, w)
    rhoa_o = fo.response(rho)
    J_o = fo.jacobian(rho)
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.setkValues(k)
    fop.setWeights(w)
    fop.solverOptions.update(method="mixed", rtol=1e-8, coarse=1024)
    rhoa = fop.response(rho)
    print(fop.lastStats, fop.timings())
    assert fop.lastStats["converged"] == 1
    assert np.abs(rhoa - rhoa_o).max() / np.abs(rhoa_o).max() < 1e-6
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    assert np.linalg.norm(J - J_o) / np.linalg.norm(J_o) < 1e-5
    # the direct band solver on the same problem
    fop.solverOptions.update(method="band", rtol=1e-10)
    rhoa_b = fop.response(rho)
    print("band", fop.lastStats, fop.timings())
    assert fop.lastStats["converged"] == 1
    U = fop.subPotentials().cpu().numpy()
    assert np.linalg.norm(U - fo.subpots) / np.linalg.norm(fo.subpots) < 1e-10
    assert np.abs(rhoa_b - rhoa_o).max() / np.abs(rhoa_o).max() < 1e-10
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    assert np.linalg.norm(J - J_o) / np.linalg.norm(J_o) < 1e-10


def test_config2_3d_invariants():
    """3D workload at full size (201898 nodes, 1150488 tets, 64 sources,
    2000 data): homogeneous model gives rhoa == rho; the GRF model converges
    to the FP64 residual target with every method, the mixed solution agrees
    with the FP64 block solution, and rhoa is homogeneous of degree one."""
    import torch
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.ert3d_grid_grf()
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    r = fop.response(np.full(mesh.cellCount(), 30.0))
    # potentials are rounded to 1e-10 before the ratio (vector.h:992-1000)
    np.testing.assert_allclose(r, 30.0, rtol=1e-7)
    fop.solverOptions.update(method="block64", rtol=1e-10)
    r64 = fop.response(rho)
    U64 = fop.subPotentials(node_major=True).clone()
    print("block64", fop.lastStats, fop.timings())
    assert fop.lastStats["converged"] == 1
    fop.solverOptions.update(method="mixed", rtol=1e-6)
    r32 = fop.response(rho)
    U32 = fop.subPotentials(node_major=True)
    print("mixed", fop.lastStats, fop.timings())
    assert fop.lastStats["converged"] == 1 and fop.lastStats["max_rel_residual"] <= 1e-6
    err = float(torch.linalg.norm(U32 - U64) / torch.linalg.norm(U64))
    print("mixed vs fp64 potentials", err)
    assert err < 1e-6
    assert np.abs(r32 - r64).max() / np.abs(r64).max() < 1e-6
    fop.solverOptions.update(method="mixed", rtol=1e-8)
    r2 = fop.response(2.0 * rho)
    fop.response(rho)
    np.testing.assert_allclose(r2, 2.0 * fop.response(rho), rtol=1e-6)


def test_config1_vs_reference_build():
    """The full 2.5D workload (5203 nodes, 41 electrodes, 888 data) against
    the reference's own C++ core built from its sources (oracle/_ref): the
    default wavenumber list (17), direct band solver, FP64 Jacobian."""
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build oracle/_ref not present")
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.ert2d_layers_block()
    ses = ref.Session(mesh, copy.deepcopy(data), nthreads=4)
    rhoa_r = ses.response(rho)
    J_r = ses.jacobian(rho)
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="band", rtol=1e-10)
    rhoa = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    assert len(fop._kv) == len(ses.kw()[0]) == 17
    e1 = np.abs(rhoa - rhoa_r).max() / np.abs(rhoa_r).max()
    e2 = np.linalg.norm(J - J_r) / np.linalg.norm(J_r)
    print("vs reference build: rhoa", e1, "J", e2)
    assert e1 < 1e-10 and e2 < 1e-10


def test_forward_from_files(tmp_path):
    """A survey written to the unified data format and a .bms mesh, read back
    and modelled: the same apparent resistivities and Jacobian as the
    in-memory survey (sensor positions pass through 14-digit text, so to
    1e-11 rather than bit for bit)."""
    from pygimli_amd import fileio
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_3d()
    mesh.save(tmp_path / "mesh.bms")
    data.save(tmp_path / "survey.dat", "a b m n k")
    out = []
    for m, d in ((mesh, data), (fileio.load_mesh(tmp_path / "mesh.bms"), fileio.load_data(tmp_path / "survey.dat"))):
        fop = ERTModelling()
        fop.setMesh(m)
        fop.setData(copy.deepcopy(d))
        fop.solverOptions.update(method="block64", rtol=1e-12)
        r = fop.response(rho)
        fop.createJacobian(rho)
        out.append((r, fop.jacobian().cpu().numpy()))
    (r0, J0), (r1, J1) = out
    print("from files: rhoa", np.abs(r1 - r0).max() / np.abs(r0).max())
    assert np.abs(r1 - r0).max() <= 1e-11 * np.abs(r0).max()
    assert np.linalg.norm(J1 - J0) <= 1e-11 * np.linalg.norm(J0)


@pytest.mark.parametrize("dim", [2, 3])
def test_shared_markers_vs_reference_deterministic(dim):
    """Parameters shared by groups of 4 (2D) / 8 (3D) scattered cells, plus one
    marker over 150 cells (more than a Jacobian tile): rhoa and J against the
    reference's own C++ core (oracle/_ref) and the oracle, FP64 and FP32, and
    two Jacobian calls give identical bits (fixed-order sums, no atomics)."""
    import torch
    from oracle import ref
    from pygimli_amd.ertModelling import ERTModelling
    if dim == 2:
        mesh, data, _ = workloads.small_2d(n_elec=6, seed=7)
        grp = 4
    else:
        mesh, data, _ = workloads.small_3d()
        grp = 8
    rng = np.random.default_rng(3)
    C = mesh.cellCount()
    mk = (rng.permutation(C) // grp).astype(np.int32)
    big = int(mk.max()) + 1
    mk[rng.choice(C, 150, replace=False)] = big
    mesh.setCellMarkers(mk)
    model = np.exp(rng.normal(3.0, 0.5, big + 1))
    fo = eo.Forward(mesh, copy.deepcopy(data))
    rhoa_o = fo.response(model)
    J_o = fo.jacobian(model)
    if ref.available():
        # one thread: the reference's multi-threaded calc2 adds the cells of a
        # shared marker into one column without synchronisation
        # (bertJacobian.cpp:197), which makes its sums vary between runs
        s = ref.Session(mesh, copy.deepcopy(data), nthreads=1)
        J_r = s.jacobian(model)
        assert np.linalg.norm(J_r - J_o) / np.linalg.norm(J_r) < 1e-11
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="block64", rtol=1e-11, coarse=0)
    rhoa = fop.response(model)
    assert np.abs(rhoa - rhoa_o).max() / np.abs(rhoa_o).max() < 1e-9
    for dt, tol in ((torch.float64, 1e-9), (torch.float32, 1e-5)):
        fop.jacobianDtype = dt
        fop.createJacobian(model)
        J1 = fop.jacobian().clone()
        fop.createJacobian(model)
        J2 = fop.jacobian()
        assert torch.equal(J1, J2), "shared-marker Jacobian differs between two calls"
        J = J1.cpu().numpy().astype(np.float64)
        err = np.linalg.norm(J - J_o) / np.linalg.norm(J_o)
        print(dim, dt, "shared-marker J rel fro", err)
        assert J.shape == J_o.shape and err < tol


@pytest.mark.parametrize("dim", [2, 3])
def test_pg_solver_stiffness_and_mass(dim):
    """pg.solver.createStiffnessMatrix / createMassMatrix (solver.py:1908-2038)
    -> RSparseMatrix::fillStiffnessMatrix / fillMassMatrix (sparsematrix.h:1039-1065)
    on the GPU against the reference build (oracle/_ref) and the oracle: identical
    pattern (row pointer, sorted column indices), values within 1e-13."""
    from oracle import ref
    from pygimli_amd import solver
    mesh = workloads.small_2d()[0] if dim == 2 else workloads.small_3d()[0]
    rng = np.random.default_rng(5 + dim)
    a = rng.uniform(0.1, 5.0, mesh.cellCount())
    s = ref.Session(mesh) if ref.available() else None
    for mass, fn, ofn in ((False, solver.createStiffnessMatrix, eo.stiffness_matrix),
                          (True, solver.createMassMatrix, eo.mass_matrix)):
        for coef in (a, None, 2.5):
            A = fn(mesh, coef)
            cf = np.broadcast_to(1.0 if coef is None else np.asarray(coef, float), (mesh.cellCount(),))
            B = ofn(mesh, cf)
            refs = [B] + ([s.fill_matrix(np.ascontiguousarray(cf), mass=mass)] if s is not None else [])
            for R in refs:
                np.testing.assert_array_equal(A.indptr, R.indptr)
                np.testing.assert_array_equal(A.indices, R.indices)
                err = np.abs(A.data - R.data).max() / np.abs(R.data).max()
                print(dim, "mass" if mass else "stiffness", err)
                assert err < 1e-13
    with pytest.raises(Exception):
        solver.createStiffnessMatrix(mesh, np.ones(mesh.cellCount() + 1))


@pytest.mark.parametrize("dim", [2, 3])
def test_background_prolongation(dim):
    """Background cells (marker -1) around a half-ellipse of parameter cells:
    mapERTModel's prolongation sweeps (Mesh::prolongateEmptyCellsValues,
    mesh.cpp:2247-2318) on the GPU against the oracle (pinned to the reference
    build), then rhoa and J (no background columns) against the oracle."""
    from pygimli_amd.ertModelling import ERTModelling
    if dim == 2:
        mesh, data, _ = workloads.small_2d(n_elec=6, seed=7)
    else:
        mesh, data, _ = workloads.small_3d()
    cc = mesh.cellCenters()
    sp = data.sensorPositions()
    ctr = sp.mean(axis=0)
    half = np.maximum(0.5 * (sp.max(axis=0) - sp.min(axis=0)), 1.0)
    r2 = ((cc[:, 0] - ctr[0]) / (1.1 * half[0])) ** 2 + (cc[:, dim - 1] / (0.8 * half[0])) ** 2
    if dim == 3:
        r2 += ((cc[:, 1] - ctr[1]) / (1.1 * half[1])) ** 2
    inside = r2 < 1.0
    mk = -np.ones(mesh.cellCount(), np.int32)
    mk[inside] = np.arange(inside.sum())
    mesh.setCellMarkers(mk)
    model = np.exp(np.random.default_rng(1).normal(3.0, 0.5, int(inside.sum())))
    fo = eo.Forward(mesh, copy.deepcopy(data))
    rho_o = eo.map_model(mesh, model)
    rhoa_o = fo.response(model)
    J_o = fo.jacobian(model)
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="block64", rtol=1e-11, coarse=0)
    rho = fop.mapERTModel(model).cpu().numpy()
    err = np.abs(rho - rho_o).max() / rho_o.max()
    print(dim, "prolongated rho rel", err)
    assert err < 1e-14
    rhoa = fop.response(model)
    assert np.abs(rhoa - rhoa_o).max() / np.abs(rhoa_o).max() < 1e-9
    fop.createJacobian(model)
    J = fop.jacobian().cpu().numpy()
    assert J.shape == J_o.shape
    assert np.linalg.norm(J - J_o) / np.linalg.norm(J_o) < 1e-9

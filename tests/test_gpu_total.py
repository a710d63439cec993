"""Total-field operator on the GPU: ERTModelling(sr=False), the reference's
DCMultiElectrodeModelling (core/src/bert/dcfemmodelling.cpp:1755-1928
calculateK_ with unit point sources, :1085-1197 response, :1318-1408
Jacobian), against the oracle's restatement (oracle/ert_oracle.py
response_total, equal to the reference's own C++ core in tests/test_oracle.py)
and against the reference build directly."""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads
from oracle import ert_oracle as eo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["small_2d", "small_3d"])
@pytest.mark.parametrize("method", ["band", "block64", "mixed"])
def test_total_field_vs_oracle(name, method):
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = getattr(workloads, name)()
    if method == "band" and mesh.dim() == 3:
        pytest.skip("direct band solver: 2D meshes")
    fop = ERTModelling(sr=False)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method=method, rtol=1e-12, coarse=0)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights())
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print(name, method, e1, e2, fop.lastStats)
    tol = 1e-9 if method != "mixed" else 1e-8
    assert e1 < tol and e2 < tol


def test_total_field_vs_reference_build():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build oracle/_ref not present")
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.ert2d_layers_block()
    ses = ref.Session(mesh, copy.deepcopy(data), nthreads=4, sr=False)
    r_ref = ses.response(rho)
    J_ref = ses.jacobian(rho)
    fop = ERTModelling(sr=False)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="band", rtol=1e-10)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    e1 = np.abs(r - r_ref).max() / np.abs(r_ref).max()
    e2 = np.linalg.norm(J - J_ref) / np.linalg.norm(J_ref)
    print("total field vs reference build:", e1, e2)
    assert e1 < 1e-10 and e2 < 1e-10


def test_bypass_vs_oracle_and_reference(tmp_path, monkeypatch):
    """Bypass electrodes (node groups with marker <= -10000) in the
    total-field operator: against the oracle with the same resistances, and
    against the reference build, whose resistors act as R / 2 (each added once
    per electrode search, two searches)."""
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_2d_bypass()
    R = {-10000: 50.0, -10001: 100.0}
    out = {}
    for method in ("band", "block64"):
        fop = ERTModelling(sr=False)
        fop.setBypassResistances(R)
        fop.setMesh(mesh)
        fop.setData(copy.deepcopy(data))
        fop.solverOptions.update(method=method, rtol=1e-12, coarse=0)
        r = fop.response(rho)
        fop.createJacobian(rho)
        out[method] = (r, fop.jacobian().cpu().numpy(), fop.kValues(), fop.weights())
    r, J, k, w = out["band"]
    f = eo.Forward(mesh, copy.deepcopy(data), k, w)
    f.bypassR = R
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    for method, (rm, Jm, _, _) in out.items():
        e1 = np.abs(rm - o).max() / np.abs(o).max()
        e2 = np.linalg.norm(Jm - oJ) / np.linalg.norm(oJ)
        print("bypass", method, e1, e2)
        assert e1 < 1e-9 and e2 < 1e-9
    from oracle import ref
    if ref.available():
        (tmp_path / "electrodeBypassResistances.map").write_text("-10000 100\n-10001 200\n")
        monkeypatch.chdir(tmp_path)
        s = ref.Session(mesh, copy.deepcopy(data), k, w, sr=False)
        rr = s.response(rho)
        assert np.abs(r - rr).max() <= 1e-9 * np.abs(rr).max()
    # a sensor whose bypass group lost its first node matches neither a group
    # nor a node electrode: like the reference (searchElectrodes_,
    # dcfemmodelling.cpp:979-985) it becomes a free electrode in a cell at
    # the same position, which the oracle models the same way
    from pygimli_amd.mesh import Mesh
    nm = mesh.nodeMarkers().copy()
    g = np.nonzero(nm == -10000)[0]
    nm[g[0]] = 0
    bad = Mesh(2, mesh.nodes, mesh.cells, mesh.cellMarkers(), mesh.bounds, mesh.boundaryMarkers(), nm)
    fop = ERTModelling(sr=False)
    fop.setBypassResistances(R)
    fop.setMesh(bad)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="band", rtol=1e-10, coarse=0)
    rb = fop.response(rho)
    fb = eo.Forward(bad, copy.deepcopy(data), k, w, total_field=True)
    fb.bypassR = R
    assert (fb.setup.elecNode < 0).sum() == 1
    ob = eo.response_total(fb, rho)
    assert np.abs(rb - ob).max() <= 1e-9 * np.abs(ob).max()


def test_bypass_map_vs_oracle_and_reference(tmp_path, monkeypatch):
    """bypass.map in the working directory (dcfemmodelling.cpp:565, :608-667):
    electrode pairs and a node-marker chain whose nodes share no cell, so
    the matrix pattern gains slots for them (pgert_set_mesh_coupled).  Against
    the oracle given the same rows (rhoa and J 1e-9, band and block64), and
    against the reference build reading the same file.  A changed file is
    read again at the next response."""
    from oracle import ref
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho, text = workloads.small_2d_bypass_map()
    monkeypatch.chdir(tmp_path)
    (tmp_path / "bypass.map").write_text(text)
    R = {-10000: 50.0, -10001: 100.0}
    out = {}
    for method in ("band", "block64"):
        fop = ERTModelling(sr=False)
        fop.setBypassResistances(R)
        fop.setMesh(mesh)
        nnz0 = fop._nnz
        with pytest.warns(UserWarning):
            fop.setData(copy.deepcopy(data))
        assert fop._nnz == nnz0 + 8          # four uncoupled pairs, both directions
        fop.solverOptions.update(method=method, rtol=1e-12, coarse=0)
        r = fop.response(rho)
        fop.createJacobian(rho)
        out[method] = (r, fop.jacobian().cpu().numpy(), fop.kValues(), fop.weights(), fop)
    r, J, k, w, fop = out["band"]
    f = eo.Forward(mesh, copy.deepcopy(data), k, w)
    f.bypassR = R
    f.bypassMap = eo.bypass_map_rows(text, f.setup.elecNode, mesh.nodeMarkers())
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    for method, (rm, Jm, *_) in out.items():
        e1 = np.abs(rm - o).max() / np.abs(o).max()
        e2 = np.linalg.norm(Jm - oJ) / np.linalg.norm(oJ)
        print("bypass.map", method, e1, e2)
        assert e1 < 1e-9 and e2 < 1e-9
    if ref.available():
        (tmp_path / "electrodeBypassResistances.map").write_text("-10000 100\n-10001 200\n")
        s = ref.Session(mesh, copy.deepcopy(data), k, w, sr=False)
        rr = s.response(rho)
        print("bypass.map vs reference build", np.abs(r - rr).max() / np.abs(rr).max())
        assert np.abs(r - rr).max() <= 1e-9 * np.abs(rr).max()
    # the file changes: the next response reads it again
    (tmp_path / "bypass.map").write_text("1 6 150\n")
    r2 = fop.response(rho)
    f.bypassMap = [f.bypassMap[0]]
    o2 = eo.response_total(f, rho)
    assert np.abs(r2 - o2).max() <= 1e-9 * np.abs(o2).max()
    assert np.abs(r2 - r).max() > 1e-3 * np.abs(r).max()


def test_topography_total_field():
    """Topography with the total-field operator against the oracle and the
    reference build; SR with topography and a free electrode, and missing
    geometric factors, raise."""
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_2d_topo()
    fop = ERTModelling(sr=False)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="band", rtol=1e-12)
    r = fop.response(rho)
    fop.createJacobian(rho)
    J = fop.jacobian().cpu().numpy()
    f = eo.Forward(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), total_field=True)
    o = eo.response_total(f, rho)
    oJ = f.jacobian(rho)
    e1 = np.abs(r - o).max() / np.abs(o).max()
    e2 = np.linalg.norm(J - oJ) / np.linalg.norm(oJ)
    print("topography", e1, e2)
    assert e1 < 1e-10 and e2 < 1e-10
    from oracle import ref
    if ref.available():
        s = ref.Session(mesh, copy.deepcopy(data), fop.kValues(), fop.weights(), sr=False)
        assert np.abs(r - s.response(rho)).max() <= 1e-10 * np.abs(r).max()
    sr = ERTModelling()
    sr.setMesh(mesh)
    free = copy.deepcopy(data)
    pos = free.sensorPositions().copy()
    pos[0, 0] += 0.25
    free.setSensorPositions(pos)
    with pytest.raises(NotImplementedError):
        sr.setData(free)
    nok = copy.deepcopy(data)
    nok.set("k", np.zeros(nok.size()))
    f2 = ERTModelling(sr=False)
    f2.setMesh(mesh)
    f2.setData(nok)
    with pytest.raises(ValueError):
        f2.response(rho)


def test_batched_api_refuses_total_field_and_complex():
    from pygimli_amd.batch import batch_jacobian, batch_response
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.small_2d()
    f = ERTModelling(sr=False)
    f.setMesh(mesh)
    f.setData(copy.deepcopy(data))
    with pytest.raises(NotImplementedError):
        batch_response([f], [rho])
    f.setComplex(True)
    with pytest.raises(NotImplementedError):
        batch_jacobian([f], [np.concatenate([rho, 0 * rho])])


def test_learned_warm_start_refuses_total_field():
    """The learned warm start predicts SR secondary potentials: a
    total-field operator is refused on the host, and collecting training
    samples after a total-field forward is refused by the library."""
    import ctypes
    import torch
    from pygimli_amd.ertModelling import ERTModelling
    from pygimli_amd.learned import LearnedWarmStart
    mesh, data, rho = workloads.small_2d()
    f = ERTModelling(sr=False)
    f.setMesh(mesh)
    f.setData(copy.deepcopy(data))
    with pytest.raises(NotImplementedError):
        LearnedWarmStart(f)
    f.response(rho)
    idx = torch.zeros(4, dtype=torch.int64, device=f.device)
    feat = torch.empty((4, 64), dtype=torch.float32, device=f.device)
    tgt = torch.empty(4, dtype=torch.float32, device=f.device)
    rc = f._ctx.L.pgert_mlp_collect(f._ctx.h, 0, ctypes.c_void_p(idx.data_ptr()), 4, ctypes.c_void_p(feat.data_ptr()),
                                    ctypes.c_void_p(tgt.data_ptr()))
    assert rc != 0 and b"singularity-removal" in f._ctx.L.pgert_last_error(f._ctx.h)


def test_start_model_constraints_and_imaginary_flip():
    """createStartModel / createConstraints / flipImagPart of the reference's
    ERTModelling (ertModelling.py:152-202): the median start model, the
    smoothness constraints (repeated for a complex model), and the imaginary
    part flipped to positive and back."""
    from pygimli_amd.ertModelling import ERTModelling
    from pygimli_amd.inversion import smoothness_constraints
    mesh, data, rho = workloads.small_2d()
    fop = ERTModelling(sr=False)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    r = fop.response(rho)
    sm = fop.createStartModel(r)
    assert sm.shape == (fop.parameterCount,) and np.all(sm == np.median(r))
    C = fop.createConstraints()
    assert (C != smoothness_constraints(mesh)).nnz == 0
    fop.setComplex(True)
    d = np.concatenate([r, -0.01 * r])
    sm = fop.createStartModel(d)
    P = fop.parameterCount // 2
    assert sm.shape == (2 * P,) and np.all(sm[:P] == np.median(r))
    np.testing.assert_allclose(sm[P:], 0.01 * np.median(r), rtol=1e-14)
    v = np.concatenate([np.ones(P), 2 * np.ones(P)])
    assert np.array_equal(fop.flipImagPart(v), np.concatenate([np.ones(P), -2 * np.ones(P)]))
    C2 = fop.createConstraints()
    assert C2.shape == (2 * C.shape[0], 2 * C.shape[1])


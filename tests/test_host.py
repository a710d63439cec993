"""CPU tests of the host side: the C-ABI library loads and exports every
entry point include/pgert.h declares (no compute call without a GPU), the
host set-up logic (electrode search, wavenumbers, geometric factors) agrees
with the oracle, and the synthetic meshes are conforming."""
import os
import re

import numpy as np
import pytest

from oracle import ert_oracle as eo
from pygimli_amd import bert, workloads
from pygimli_amd.data import DataContainerERT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pgert.h")).read()
    return sorted(set(re.findall(r"\b(pgert_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from pygimli_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpgert.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 18
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_create_without_gpu_fails_cleanly():
    import torch
    from pygimli_amd import _lib
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpgert.so not built")
    with pytest.raises(RuntimeError):
        _lib.Context(0)


def test_ertmodelling_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from pygimli_amd.ertModelling import ERTModelling
    with pytest.raises(RuntimeError):
        ERTModelling()


@pytest.mark.parametrize("name", ["small_2d", "small_3d"])
def test_electrode_search_and_wavenumbers_match_oracle(name):
    mesh, data, rho = getattr(workloads, name)()
    an = bert.MeshAnalysis(mesh, data.sensorPositions())
    st = eo.Setup(mesh, data.sensorPositions())
    np.testing.assert_array_equal(an.elecNode, st.elecNode)
    np.testing.assert_allclose(an.sourceCenter, st.sourceCenter, rtol=1e-15)
    assert an.surfaceZ == st.surfaceZ and an.topography == st.topography
    k1, w1 = bert.kwave_list(mesh.dim(), an.elecPos)
    k2, w2 = eo.kwave_list(mesh.dim(), st.elecPos)
    np.testing.assert_array_equal(k1, k2)
    np.testing.assert_array_equal(w1, w2)
    g1 = bert.geometric_factors(data.sensorPositions(), data.abmn(), mesh.dim())
    g2 = eo.geometric_factors(data.sensorPositions(), data.abmn(), mesh.dim())
    np.testing.assert_allclose(g1, g2, rtol=1e-14)


def test_electrode_search_2d_swap_and_errors():
    """2D sensors given as (x, 0, z) are read as (x, z) (dcfemmodelling.cpp:859-867);
    a sensor off every electrode node is a free electrode in the cell holding
    it (same cell as the oracle's search); a sensor outside the mesh raises."""
    mesh, data, _ = workloads.small_2d()
    s = data.sensorPositions().copy()
    s2 = np.stack([s[:, 0], np.zeros(len(s)), s[:, 1]], 1)
    an = bert.MeshAnalysis(mesh, s2)
    an0 = bert.MeshAnalysis(mesh, s)
    np.testing.assert_array_equal(an.elecNode, an0.elecNode)
    free = s.copy()
    free[0, 0] += 0.3
    free[3, 1] -= 0.2
    an = bert.MeshAnalysis(mesh, free)
    st = eo.Setup(mesh, free)
    np.testing.assert_array_equal(an.elecNode, st.elecNode)
    assert an.elecNode[0] == -1 and an.elecNode[3] == -1
    np.testing.assert_array_equal(an.elecCell[[0, 3]], st.elecCell[[0, 3]])
    np.testing.assert_allclose(an.elecPos, st.elecPos, rtol=0, atol=0)
    np.testing.assert_allclose(an.sourceCenter, st.sourceCenter, rtol=1e-15)
    out = s.copy()
    out[0, 1] += 1.0   # above the surface
    with pytest.raises(RuntimeError):
        bert.MeshAnalysis(mesh, out)


def test_bypass_map_files(tmp_path):
    """bypass.map rows as the reference reads them (dcfemmodelling.cpp:608-667):
    the product's parser gives the oracle's rows (which the reference build
    pins), warns for each row the reference skips, raises where the reference
    would read past its electrode list or use a free electrode's missing
    node, and finds the pairs that need matrix slots of their own.
    electrodeBypassResistances.map is read as loadFloatMap does."""
    mesh, data, _, text = workloads.small_2d_bypass_map()
    an = bert.MeshAnalysis(mesh, data.sensorPositions())
    f = tmp_path / "bypass.map"
    f.write_text(text)
    with pytest.warns(UserWarning) as rec:
        a, b, r = bert.read_bypass_map(str(f), an.elecNode, mesh.nodeMarkers())
    assert len(rec) == 4
    rows = eo.bypass_map_rows(text, an.elecNode, mesh.nodeMarkers())
    assert list(zip(a.tolist(), b.tolist(), r.tolist())) == [tuple(x) for x in rows]
    # electrode 4 is the bypass electrode: its node is the group's first node
    assert a[1] == an.bypassGroups[-10000][0] and r[1] == float(np.float32(80.3))
    assert bert.uncoupled_pairs(mesh.cells, a, b) == sorted({(min(x, y), max(x, y)) for x, y in zip(a, b)})
    ga, gb, _ = an.bypass_pairs()
    assert bert.uncoupled_pairs(mesh.cells, ga, gb) == []
    for bad in ("1 9 10\n", "2 0 10\n", "1 2 0\n"):
        f.write_text(bad)
        with pytest.raises(ValueError):
            bert.read_bypass_map(str(f), an.elecNode, mesh.nodeMarkers())
    free = an.elecNode.copy()
    free[1] = -1
    f.write_text("1 2 10\n")
    with pytest.raises(ValueError):
        bert.read_bypass_map(str(f), free, mesh.nodeMarkers())
    m = tmp_path / "electrodeBypassResistances.map"
    m.write_text("# marker resistance\n-10000 100.1\n-10001 2e2\n-10000.5 7\n")
    assert bert.bypass_marker_resistances(str(m)) == {-10000: float(np.float32(100.1)), -10001: 200.0}
    m.write_text("1 2 3\n")
    with pytest.raises(ValueError):
        bert.load_float_map(str(m))


def test_cem_electrode_search_and_map_files(tmp_path):
    """Complete-electrode-model electrodes: the same face groups and matching
    as the oracle's search (which the reference build pins), electrode
    unknowns after the nodes; the contact map files read as
    checkIfMapFileExistAndLoadToVector does (from 1, from 0 when the first key
    is 0, -1 = last); 2.5D, passive bodies and sensors without faces are
    refused."""
    from pygimli_amd.mesh import Mesh
    mesh, data, _ = workloads.small_3d_cem()
    an = bert.MeshAnalysis(mesh, data.sensorPositions())
    st = eo.Setup(mesh, data.sensorPositions())
    assert sorted(an.cemFaces) == sorted(st.cem)
    for i in an.cemFaces:
        np.testing.assert_array_equal(an.cemFaces[i], st.cem[i])
    np.testing.assert_array_equal(an.elecNode, mesh.nodeCount() + np.arange(data.sensorCount()))
    rows, cols, vals, pairs = an.cem_terms(mesh)
    areas = [mesh.boundarySizes()[f].sum() for _, f in sorted(an.cemFaces.items())]
    diag = vals[rows == cols][np.asarray(rows[rows == cols]) >= mesh.nodeCount()]
    np.testing.assert_allclose(diag, areas, rtol=1e-15)
    # each electrode row sums to zero: area/z - sum_f int N / z
    n = mesh.nodeCount()
    for e in range(data.sensorCount()):
        assert abs(vals[rows == n + e].sum()) < 1e-14
    f = tmp_path / "z.map"
    f.write_text("1 0.5\n3 2.0\n")
    np.testing.assert_array_equal(bert.load_electrode_vector(str(f), 4, 1.0), [0.5, 1.0, 2.0, 1.0])
    f.write_text("0 0.5\n3 2.0\n")
    np.testing.assert_array_equal(bert.load_electrode_vector(str(f), 4, 1.0), [0.5, 1.0, 1.0, 2.0])
    f.write_text("-1 7\n2 3\n9 4\n")
    np.testing.assert_array_equal(bert.load_electrode_vector(str(f), 4, 1.0), [1.0, 3.0, 1.0, 7.0])
    bm = mesh.boundaryMarkers().copy()
    bm[bm == -10000] = -9999
    with pytest.raises(NotImplementedError):
        bert.MeshAnalysis(Mesh(3, mesh.nodes, mesh.cells, mesh.cellMarkers(), mesh.bounds, bm,
                               mesh.nodeMarkers()), data.sensorPositions())
    bm = mesh.boundaryMarkers().copy()
    bm[bm == -10008] = -1
    with pytest.raises(NotImplementedError):
        bert.MeshAnalysis(Mesh(3, mesh.nodes, mesh.cells, mesh.cellMarkers(), mesh.bounds, bm,
                               mesh.nodeMarkers()), data.sensorPositions())
    m2, d2, _ = workloads.small_2d()
    bm2 = m2.boundaryMarkers().copy()
    bm2[np.nonzero(bm2 == -1)[0][:1]] = -10000
    with pytest.raises(NotImplementedError):
        bert.MeshAnalysis(Mesh(2, m2.nodes, m2.cells, m2.cellMarkers(), m2.bounds, bm2, m2.nodeMarkers()),
                          d2.sensorPositions())


@pytest.mark.parametrize("name", ["small_2d", "small_3d"])
def test_meshes_are_conforming(name):
    mesh, _, _ = getattr(workloads, name)()
    nv = mesh.cells.shape[1]
    faces = np.concatenate([np.sort(np.delete(mesh.cells, j, axis=1), axis=1) for j in range(nv)])
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2
    outer = (cnt == 1).sum()
    assert outer == mesh.boundaryCount()
    assert (mesh.boundaryLeftCell() >= 0).all()
    assert np.all(mesh.cellSizes() > 0)


def test_config_sizes():
    """The benchmarked workloads have the stated sizes."""
    m1, d1, r1 = workloads.ert2d_layers_block()
    assert 4500 < m1.nodeCount() < 5600 and 9000 < m1.cellCount() < 10500
    assert len(m1.findNodesIdxByMarker(-99)) == 41
    assert set(np.unique(r1)) == {10.0, 20.0, 100.0, 500.0}


def test_datacontainer_api():
    d = DataContainerERT()
    for i in range(4):
        d.createSensor([i, 0, 0])
    assert d.createSensor([1.0004, 0, 0]) == 1  # within 1e-3: same sensor
    d.resize(2)
    d.createFourPointData(0, 0, 1, 2, 3)
    d.createFourPointData(1, 1, -1, 2, -1)
    assert d.size() == 2 and d.sensorCount() == 4
    np.testing.assert_array_equal(d.abmn()[1], [1, -1, 2, -1])


def test_batched_2d_workload():
    """Batch generator: sizes within range, Wenner-Schlumberger counts,
    11 wavenumbers, positive resistivities, deterministic."""
    from pygimli_amd import workloads
    a = workloads.batched_2d(count=4, seed=11, n_range=(2000, 6000))
    b = workloads.batched_2d(count=4, seed=11, n_range=(2000, 6000))
    for (m, d, rho, k, w), (m2, d2, rho2, k2, w2) in zip(a, b):
        ne = len(d.sensorPositions())
        assert 24 <= ne <= 96
        assert 0.8 * 2000 <= m.nodeCount() <= 1.1 * 6000
        assert d.size() == sum(max(0, ne - 1 - 2 * sep) for sep in range(1, ne - 1))
        assert d.abmn().max() < ne
        assert len(k) == 11 and len(w) == 11 and np.all(k > 0)
        assert np.all(rho >= 10.0 - 1e-9) and np.all(rho <= 1000.0 + 1e-9)
        np.testing.assert_array_equal(rho, rho2)


def test_large_survey_schemes():
    """Grid dipole-dipole and gradient schemes: indices on the grid, four
    distinct electrodes, every potential dipole of neighbouring electrodes;
    the large-survey generator keeps only quadrupoles with a non-vanishing
    half-space signal and is deterministic."""
    from pygimli_amd import workloads
    ns = 6
    dd = workloads.line_dipole_dipole_grid(ns, spacings=(1, 2), seps=(1, 2))
    per_line = len(workloads.dipole_dipole(ns, (1, 2), (1, 2)))
    assert len(dd) == 2 * ns * per_line
    gr = workloads.gradient_grid(ns, insets=(0,), reach=1)
    for q in (dd, gr):
        assert q.min() >= 0 and q.max() < ns * ns
        assert all(len(set(r)) == 4 for r in q.tolist())
    # potential dipoles of the gradient arrays are neighbours on one grid line
    m, n = gr[:, 2], gr[:, 3]
    assert np.all((np.abs(m - n) == 1) | (np.abs(m - n) == ns))
    mesh, data, rho = workloads.ert3d_survey_large(n_side=4, spacing=1.0, n_data=50, h=1.0, growth=2.0, pad=40.0,
                                                   z_h=1.0, z_growth=2.0)
    s = data.sensorPositions()
    ab = data.abmn()
    d = np.sqrt(((s[:, None] - s[None]) ** 2).sum(-1))
    g = 1 / d[ab[:, 0], ab[:, 2]] - 1 / d[ab[:, 0], ab[:, 3]] - 1 / d[ab[:, 1], ab[:, 2]] + 1 / d[ab[:, 1], ab[:, 3]]
    assert np.all(np.abs(g) * d[d > 0].min() > 1e-3)
    assert len(s) == 16 and data.size() <= 50 and len(rho) == mesh.cellCount()
    _, data2, rho2 = workloads.ert3d_survey_large(n_side=4, spacing=1.0, n_data=50, h=1.0, growth=2.0, pad=40.0,
                                                  z_h=1.0, z_growth=2.0)
    np.testing.assert_array_equal(ab, data2.abmn())
    np.testing.assert_array_equal(rho, rho2)


def test_warm_start_checkpoint_round_trip(tmp_path):
    """Trained warm-start parameters survive save/load bit for bit; foreign
    files and other layer sizes are refused."""
    from pygimli_amd import learned
    p = learned.init_params(3)
    path = tmp_path / "ws.npz"
    learned.save_params(path, p, {"epochs": 20, "loss": 0.125})
    q, meta = learned.load_params(path)
    np.testing.assert_array_equal(p, q)
    assert int(meta["epochs"]) == 20 and float(meta["loss"]) == 0.125
    with pytest.raises(ValueError):
        learned.save_params(tmp_path / "bad.npz", p[:-1])
    np.savez(tmp_path / "other.npz", params=p)
    with pytest.raises(ValueError):
        learned.load_params(tmp_path / "other.npz")
    np.savez(tmp_path / "sizes.npz", format=np.asarray(learned.FORMAT), sizes=np.asarray([32, 64, 1]), params=p)
    with pytest.raises(ValueError):
        learned.load_params(tmp_path / "sizes.npz")

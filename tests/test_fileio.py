"""Unified data format and binary mesh files (pygimli_amd.fileio).

Parity is bytes: a file read and written again here must equal the file the
reference's own DataContainerERT::load + DataContainer::save
(core/src/datacontainer.cpp:209-633) and Mesh::load + Mesh::save
(core/src/mesh_io.cpp:391-630) make of the same input, through
oracle/_ref/libgimli_ref.so.  The hand-checked tests pin the alias table, unit
scales, index shift and validity rules by values fixed by the file text.
"""
import copy
import struct

import numpy as np
import pytest

from pygimli_amd import fileio, workloads
from oracle import ref

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (make -C oracle ref)")

# Comments, a sensor format line with two columns, a duplicated sensor (1e-13
# from sensor 2), aliases with unit scales, a custom column, a NaN, a zero
# rhoa, equal electrodes, an index past the sensor count, a short row and a
# long row.
SURVEY = b"""# a comment line before the count
6   # number of electrodes
# x z
0 0
1 0     # second
2 0
3 0
1 0.0000000000001
5 0
9
# C1 c2 P1 p2 rhoa/Ohmm u/mV i/mA err/% ip/mrad custom
1 2 3 4 100.5 12.5 100 3 1.5 7
1 2 4 5 101 13 100 3 nan 7
2 3 4 5 0 13 100 3 1 7
2 2 4 5 99 13 100 3 1 7
1 2 3 9 99 13 100 3 1 7
1 2 5 6 98.25 11 100 2.5 1 7
1 3 4 5 97 10 100 2 1 7 extra
2 3 4 6 96 10 100 2
1 2 3 4 95 10 100 2 1 7
"""

# Pole-dipole rows (b = 0 on file = no electrode), no sensor format line, a
# format line with r computed from u / i, millimetre sensor coordinates.
POLE = b"""4
# x/mm y z
0 0 0
1500 0 0
3000 0 0
4500 0 0
3
# a b m n u i k
1 0 2 3 0.5 0.1 6.28
1 0 3 4 0.2 0.1 12.5
2 0 3 4 0.7 0.1 6.28
"""


# a topography section after the data, with a format line and hex / exponent forms
TOPO = b"""3
# x y z
0 0 0x1p-2
1.5 0 2e-1
3.0 0 -.5
"""


def _write(tmp_path, name, raw):
    p = tmp_path / name
    p.write_bytes(raw)
    return p


def test_aliases_scales_and_validity(tmp_path):
    src = _write(tmp_path, "s.dat", SURVEY)
    d = fileio.load_data(src, removeInvalid=False)
    # six rows of sensors; the fifth lies 1e-13 from the second, so it is
    # mapped onto sensor 1 and dropped: five sensors, x = 0 1 2 3 5 (z column
    # only, y = 0)
    assert np.array_equal(d.sensorPositions(), [[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0], [5, 0, 0]])
    assert d.size() == 9
    assert np.array_equal(d["a"], [0, 0, 1, 1, 0, 0, 0, 1, 0])          # C1, counted from one on file
    assert np.array_equal(d["m"], [2, 3, 3, 3, 2, 1, 3, 3, 2])          # P1; 5 on file -> duplicate -> 1
    assert np.array_equal(d["n"], [3, 1, 1, 1, 8, 4, 1, 4, 3])          # 6 on file -> sensor 4; 9 -> 8 stays
    assert np.allclose(d["u"][:2], [0.0125, 0.013], rtol=1e-15)         # mV -> V
    assert np.allclose(d["i"], 0.1, rtol=1e-15)                          # mA -> A
    assert np.allclose(d["err"][5:7], [0.025, 0.02], rtol=1e-15)        # % -> 1
    assert d["rhoa"][5] == 98.25                                         # rhoa/Ohmm, unscaled
    assert np.array_equal(d["custom"], [7, 7, 7, 7, 7, 7, 7, 0, 7])     # no alias; short row -> 0
    assert d["ip"][7] == 0.0 and np.isnan(d["ip"][1])
    # rows 1-5 invalid: NaN (and b = n), zero rhoa (and a = n), a = b, index 8
    # past the sensors, b = m after the duplicate merge
    assert np.array_equal(d["valid"], [1, 0, 0, 0, 0, 0, 1, 1, 1])
    kept = fileio.load_data(src)
    assert kept.size() == 4 and np.all(kept["valid"] == 1)
    assert np.array_equal(kept["rhoa"], [100.5, 97, 96, 95])


def test_sensor_units_pole_dipole_and_r(tmp_path):
    p = fileio.load_data(_write(tmp_path, "p.dat", POLE))
    assert np.array_equal(p.sensorPositions()[:, 0], [0, 1.5, 3, 4.5])     # mm -> m
    assert np.array_equal(p["b"], [-1, -1, -1])                          # 0 on file: no electrode
    assert np.allclose(p["r"], [5, 2, 7], rtol=1e-15)                    # r = u / i, no rhoa and r
    assert np.array_equal(p["k"], [6.28, 12.5, 6.28])
    # indices from zero on file: 4 is past the last sensor
    d = fileio.load_data(_write(tmp_path, "z.dat", POLE), sensorIndicesFromOne=False, removeInvalid=False)
    assert np.array_equal(d["a"], [1, 1, 2]) and np.array_equal(d["b"], [0, 0, 0])
    assert np.array_equal(d["valid"], [1, 0, 0])
    d.save(tmp_path / "z2.dat", "a b m n")
    assert (tmp_path / "z2.dat").read_bytes().split(b"\n")[8] == b"1\t0\t2\t3"   # written from zero again


def test_save_load_round_trip(tmp_path):
    mesh, data, rho = workloads.small_3d()
    data.set("rhoa", np.linspace(10, 20, data.size()))
    data.set("err", np.full(data.size(), 0.03))
    data.save(tmp_path / "d.dat")
    back = fileio.load_data(tmp_path / "d.dat")
    assert back.size() == data.size()
    for t in "abmn":
        assert np.array_equal(back[t], data[t])
    for t in ("k", "rhoa", "err"):
        assert np.allclose(back[t], data[t], rtol=1e-13, atol=0)
    assert np.allclose(back.sensorPositions(), data.sensorPositions(), rtol=0, atol=1e-12)
    back.save(tmp_path / "e.dat")
    again = fileio.load_data(tmp_path / "e.dat")
    back.save(tmp_path / "f.dat")
    again.save(tmp_path / "g.dat")
    assert (tmp_path / "f.dat").read_bytes() == (tmp_path / "g.dat").read_bytes()


def test_topography_points_and_number_forms(tmp_path):
    raw = (b"2\n0 0 0\n1 0 0\n1\n# a b m n u i\n1 0 2 0 0x1p-1 -1e-1\n"
           b"2\n# x z\n0.5 0x10\n1.5 -2.25   # second\n")
    d = fileio.load_data(_write(tmp_path, "t.dat", raw), removeInvalid=False)
    assert d["u"][0] == 0.5 and d["i"][0] == -0.1           # hexadecimal and signed exponent forms
    assert np.array_equal(d.additionalPoints(), [[0.5, 0, 16], [1.5, 0, -2.25]])
    d.save(tmp_path / "t2.dat")
    assert (tmp_path / "t2.dat").read_bytes().endswith(b"\n0\n")   # not written back
    with pytest.raises(ValueError, match="topo electrode format"):
        fileio.load_data(_write(tmp_path, "u.dat", raw.replace(b"# x z", b"# x q")))
    with pytest.raises(ValueError, match="too few topo"):
        fileio.load_data(_write(tmp_path, "v.dat", raw.replace(b"1.5 -2.25   # second\n", b"")))


def test_data_errors(tmp_path):
    with pytest.raises(ValueError, match="sensor count"):
        fileio.load_data(_write(tmp_path, "a.dat", b"0\n0\n"))
    with pytest.raises(ValueError, match="too few sensor"):
        fileio.load_data(_write(tmp_path, "b.dat", b"3\n0 0 0\n1 0 0\n"))
    with pytest.raises(ValueError, match="too few data"):
        fileio.load_data(_write(tmp_path, "c.dat", b"2\n0 0 0\n1 0 0\n3\n# a m\n1 2\n"))
    with pytest.raises(ValueError, match="data format"):
        fileio.load_data(_write(tmp_path, "d.dat", b"1 2\n"))
    d = fileio.load_data(_write(tmp_path, "e.dat", b"2\n0 0 0\n1 0 0\n0\n"))
    assert d.size() == 0 and d.sensorCount() == 2
    with pytest.raises(ValueError, match="no such data"):
        d.save(tmp_path / "f.dat", "a b nothing")


@needs_ref
@pytest.mark.parametrize("remove", [True, False])
@pytest.mark.parametrize("fmt,no_filter", [("all", False), ("a b m n rhoa u i err ip valid", False),
                                           ("a b m n rhoa u i err ip valid", True),
                                           ("c1 c2 p1 p2 rhoa/Ohmm", False)])
@pytest.mark.parametrize("fmt_sensor", ["x y z", "x z", "X Y q"])
@pytest.mark.parametrize("name", ["survey", "pole", "topo", "crlf"])
def test_data_bytes_match_reference(tmp_path, remove, fmt, no_filter, fmt_sensor, name):
    if name in ("pole", "topo") and fmt.startswith("a b m n rhoa u i err ip"):
        fmt = "a b m n u i r k valid"
    src = _write(tmp_path, "in.dat", {"survey": SURVEY, "pole": POLE, "topo": POLE + TOPO,
                                      "crlf": SURVEY.replace(b"\n", b"\r\n")}[name])
    d = fileio.load_data(src, removeInvalid=remove)
    fileio.save_data(d, tmp_path / "ours.dat", fmt, fmt_sensor, no_filter)
    ref.data_convert(src, tmp_path / "theirs.dat", fmt, fmt_sensor, remove_invalid=remove, no_filter=no_filter)
    assert (tmp_path / "ours.dat").read_bytes() == (tmp_path / "theirs.dat").read_bytes()


@needs_ref
def test_scheme_written_here_reads_in_reference(tmp_path):
    mesh, data, rho = workloads.small_2d()
    data.set("rhoa", np.linspace(10, 20, data.size()))
    data.save(tmp_path / "ours.dat")
    ref.data_convert(tmp_path / "ours.dat", tmp_path / "theirs.dat")
    fileio.load_data(tmp_path / "ours.dat").save(tmp_path / "again.dat")
    assert (tmp_path / "again.dat").read_bytes() == (tmp_path / "theirs.dat").read_bytes()


def _mesh_arrays_equal(a, b):
    return (np.array_equal(a.nodes, b.nodes) and np.array_equal(a.cells, b.cells)
            and np.array_equal(a.bounds, b.bounds) and np.array_equal(a.nodeMarkers(), b.nodeMarkers())
            and np.array_equal(a.cellMarkers(), b.cellMarkers())
            and np.array_equal(a.boundaryMarkers(), b.boundaryMarkers()))


@pytest.mark.parametrize("name", ["small_2d", "small_3d"])
def test_mesh_round_trip_and_v2(tmp_path, name):
    mesh, _, _ = getattr(workloads, name)()
    mesh.addData("res", np.arange(mesh.cellCount()) * 0.5)
    mesh.save(tmp_path / "m")                     # .bms appended
    raw = (tmp_path / "m.bms").read_bytes()
    assert raw[:2] == bytes([mesh.dim(), 3]) and raw[2:130] == bytes(128)
    back = fileio.load_mesh(tmp_path / "m.bms")
    assert _mesh_arrays_equal(back, mesh)
    assert np.array_equal(back.data("res"), mesh.data("res"))
    # version 2: the same without the 128-byte header
    (tmp_path / "v2.bms").write_bytes(bytes([mesh.dim(), 2]) + raw[130:])
    assert _mesh_arrays_equal(fileio.load_mesh(tmp_path / "v2.bms"), mesh)
    (tmp_path / "v9.bms").write_bytes(bytes([mesh.dim(), 9]) + raw[130:])
    with pytest.raises(ValueError, match="wrong version"):
        fileio.load_mesh(tmp_path / "v9.bms")
    with pytest.raises(ValueError, match="end of file"):
        (tmp_path / "cut.bms").write_bytes(raw[:200])
        fileio.load_mesh(tmp_path / "cut.bms")


def test_mesh_rejects_non_simplex(tmp_path):
    # one quadrilateral cell in 2D
    raw = struct.pack("<BB", 2, 2) + struct.pack("<I", 4) + np.array(
        [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "<f8").tobytes() + np.zeros(4, "<i4").tobytes()
    raw += struct.pack("<I", 1) + bytes([4]) + np.arange(4, dtype="<u4").tobytes() + np.zeros(1, "<i4").tobytes()
    raw += struct.pack("<I", 0) + struct.pack("<Q", 0)
    (tmp_path / "q.bms").write_bytes(raw)
    with pytest.raises(ValueError, match="4 nodes"):
        fileio.load_mesh(tmp_path / "q.bms")


@needs_ref
@pytest.mark.parametrize("name", ["small_2d", "small_3d"])
def test_mesh_bytes_match_reference(tmp_path, name):
    mesh, data, _ = getattr(workloads, name)()
    mesh.addData("res", np.arange(mesh.cellCount()) * 0.5)
    mesh.addData("Attr", np.ones(3))
    fileio.save_mesh(mesh, tmp_path / "ours.bms")
    ref.mesh_convert(tmp_path / "ours.bms", tmp_path / "theirs.bms")
    assert (tmp_path / "ours.bms").read_bytes() == (tmp_path / "theirs.bms").read_bytes()
    # a mesh written by the reference after Mesh::createNeighborInfos (every
    # cell face a boundary, with the reference's left / right cells)
    s = ref.Session(mesh, copy.deepcopy(data))
    s.save_mesh(tmp_path / "session.bms")
    m = fileio.load_mesh(tmp_path / "session.bms")
    fileio.save_mesh(m, tmp_path / "again.bms")
    assert (tmp_path / "session.bms").read_bytes() == (tmp_path / "again.bms").read_bytes()
    nb = mesh.boundaryCount()
    assert np.array_equal(m.bounds[:nb], mesh.bounds) and np.array_equal(m.nodes, mesh.nodes)
    left, right = m._fileNeighbors
    # outer boundaries: exactly one neighbour, the cell this build assembles
    # the Robin term on
    assert np.all((left[:nb] < 0) != (right[:nb] < 0))
    assert np.array_equal(np.where(left[:nb] >= 0, left[:nb], right[:nb]), mesh.boundaryLeftCell())
    # every face: the same unordered neighbour pair as computed here
    m._fileNeighbors = None
    l2, r2 = fileio._neighbors(m)
    assert np.array_equal(np.sort(np.stack([left, right]), 0), np.sort(np.stack([l2, r2]), 0))


@needs_ref
def test_topography_without_format_line_refused_like_reference(tmp_path):
    """A topography section with no '#' line keeps the data section's format in
    the reference (datacontainer.cpp:431-441), whose first token is no
    coordinate: both readers refuse the file."""
    src = _write(tmp_path, "in.dat", POLE + TOPO.replace(b"# x y z\n", b""))
    with pytest.raises(ValueError, match="topo electrode format"):
        fileio.load_data(src)
    with pytest.raises(RuntimeError):
        ref.data_convert(src, tmp_path / "theirs.dat")

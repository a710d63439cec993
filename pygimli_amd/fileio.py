"""On-disk formats of real surveys: the unified data format and binary meshes.

Unified data format (DataContainer::load / DataContainer::write,
core/src/datacontainer.cpp:209-419 and :527-633), with the ERT token aliases of
DataContainerERT (core/src/bert/bertDataContainer.cpp:45-89) and its data
validity rules (datacontainer.cpp:480-517, bertDataContainer.cpp:91-146).  A
file reads

    <sensor count>
    # x y z                      (optional sensor format line)
    <one row per sensor>
    <data count>
    # a b m n rhoa ...           (data format line)
    <one row per datum>
    <topography point count>     (written as 0)

with '#' starting a comment anywhere, sensor indices counted from one on
file, token names translated through the alias table (c1 -> a, rhoa/Ohmm ->
rhoa, u/mV -> u scaled by 1/1000, err/% -> err scaled by 1/100, ...).

Binary mesh v2/v3 ``.bms`` (Mesh::saveBinaryV2 / Mesh::loadBinaryV2,
core/src/mesh_io.cpp:391-630): little-endian
    uint8 dim, uint8 version, [uint8[128] v3 header, byte 0 = isGeometry],
    uint32 nNodes, float64[3 nNodes] xyz, int32[nNodes] node markers,
    uint32 nCells, uint8[nCells] node counts, uint32[sum] node ids, int32[nCells] markers,
    uint32 nBounds, uint8[nBounds] node counts, uint32[sum] node ids,
        int32[nBounds] markers, int32[nBounds] left cell, int32[nBounds] right cell,
    uint64 nData, per entry: uint64 name length, name, uint64 length, float64[length],
with each block after a zero count left out.  Only P1 simplex meshes (the
cells the ERT path assembles: triangles in 2D, tetrahedra in 3D) load here.

Text is handled as bytes (C locale), so token matching, whitespace and the
'#' comment rule are those of the reference's stream reader (gimli.cpp:245-259).
"""
from __future__ import annotations

import os
import re
import struct

import numpy as np

from .data import DataContainerERT
from .mesh import Mesh

TOLERANCE = 1e-12          # gimli.h TOLERANCE

# DataContainer::initTokenTranslator (datacontainer.cpp:84-96) then the ERT
# table (bertDataContainer.cpp:64-89): token -> aliases, entered key by key in
# sorted order; the first entry of an alias wins.
_BASE_ALIASES = {"err": "err std error err/%", "t": "t t/s t/ms t/(ms)"}
_ERT_ALIASES = {
    "a": "a c1", "b": "b c2", "m": "m p1", "n": "n p2",
    "rhoa": "rhoa rho_a ra rs rhos rhoa/Ohmm rhoa(Ohmm)",
    "eca": "ECa EC_a",
    "r": "r rho r(Ohm) imp z u/i",
    "ip": "ip ip/mrad ip/° phase phase/mrad phase/° phi phi/mrad phi/°",
    "iperr": "iperr iperr/mrad iperr/° phierr phierr/mrad phierr/°",
    "u": "u u/V u(V) u/mV u(mV) v v/V v(V) v/mV v(mV)",
    "i": "i i/A i(A) i/mA i(mA)",
    "k": "k",
}
SENSOR_TOKENS = (b"a", b"b", b"m", b"n")                   # registerSensorIndex
ERT_TOKENS = (b"i", b"u", b"r", b"err", b"k", b"rhoa", b"ip", b"iperr")   # init()


def _translator():
    tt = {}
    for table in (_BASE_ALIASES, _ERT_ALIASES):
        for tok in sorted(table):
            for alias in table[tok].split():
                tt.setdefault(alias.encode("utf-8"), tok.encode())
    return tt


_TT = _translator()


def _alias_key(tok):
    """First byte lowered as C tolower does in the C locale."""
    c = tok[:1]
    if b"A" <= c <= b"Z":
        c = c.lower()
    return c + tok[1:]


def translate_alias(tok):
    """Token name for a column name of a data file, or None (no alias)."""
    return _TT.get(_alias_key(tok))


_NUM = re.compile(rb"\s*[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.I)
_HEX = re.compile(rb"\s*([+-]?)0[xX]((?:[0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)(?:[pP][+-]?\d+)?)")
_INT = re.compile(rb"\s*[+-]?\d+")


def _strtod(tok):
    """std::strtod: the longest numeric prefix, 0.0 without one."""
    try:
        if b"_" not in tok:
            return float(tok)
    except ValueError:
        pass
    h = _HEX.match(tok)
    if h:
        return float.fromhex((h.group(1) + b"0x" + h.group(2)).decode())
    m = _NUM.match(tok)
    return float(m.group(0)) if m else 0.0


def _atoi(tok):
    m = _INT.match(tok)
    return int(m.group(0)) if m else 0


class _Rows:
    """Line reader with the reference's row rules: text after '#' dropped,
    whitespace-separated tokens, empty rows skipped by `nonempty`."""

    def __init__(self, raw):
        self.lines = raw.split(b"\n")
        self.i = 0

    def row(self):
        if self.i >= len(self.lines):
            return []
        s = self.lines[self.i]
        self.i += 1
        c = s.find(b"#")
        return (s if c < 0 else s[:c]).split()

    def nonempty(self):
        while self.i < len(self.lines):
            r = self.row()
            if r:
                return r
        return []

    def format_line(self):
        """A '#' as the first byte of the next line makes it a format line."""
        if self.i < len(self.lines) and self.lines[self.i][:1] == b"#":
            self.lines[self.i] = self.lines[self.i][1:]
            return self.row()
        return None


def _round(v, tol):
    return np.rint(v / tol) * tol


def _checked_validity(tok, valid, nSensors):
    """checkDataValidity (datacontainer.cpp:480-517) with the ERT rules
    (bertDataContainer.cpp:91-146); updates `valid` (and r) in place."""
    for v in tok.values():
        valid[~np.isfinite(v)] = 0.0
    for name in SENSOR_TOKENS:
        v = tok[name]
        valid[(v < -1) | (v >= float(nSensors))] = 0.0
    if len(valid) == 0:
        return

    def have(name):
        v = tok.get(name)
        return v is not None and len(v) > 0 and not np.max(np.abs(v)) < TOLERANCE

    def all_nonzero(name):
        v = tok.get(name)
        return v is not None and len(v) > 0 and np.min(np.abs(v)) > TOLERANCE

    if have(b"rhoa") or have(b"r") or have(b"u"):
        if not all_nonzero(b"rhoa") and not all_nonzero(b"r"):
            if all_nonzero(b"i") and all_nonzero(b"u"):
                tok[b"r"] = tok[b"u"] / tok[b"i"]
    if have(b"rhoa"):
        valid[tok[b"rhoa"] < TOLERANCE] = 0.0
    if have(b"r"):
        valid[np.abs(tok[b"r"]) < TOLERANCE] = 0.0
    a, b, m, n = (tok[x] for x in SENSOR_TOKENS)
    bad = np.zeros(len(valid), bool)
    for p, q in ((a, b), (a, m), (a, n), (b, m), (b, n), (m, n)):
        bad |= (np.abs(p - q) < TOLERANCE) & (p > 0)
    valid[bad] = 0.0


def load_data(fileName, sensorIndicesFromOne=True, removeInvalid=True):
    """DataContainerERT(fileName, removeInvalid) (bertDataContainer.cpp:29-33,
    datacontainer.cpp:209-419).  Raises ValueError where the reference throws."""
    with open(fileName, "rb") as f:
        rows = _Rows(f.read())
    row = rows.nonempty()
    if len(row) != 1:
        raise ValueError(f"{fileName}: cannot determine data format. {len(row)}")
    nSensors = _atoi(row[0])
    if nSensors < 1:
        raise ValueError(f"{fileName}: cannot determine sensor count {row[0]!r}")
    fmt = rows.format_line()
    if fmt is None:
        fmt = [b"x", b"y", b"z"]
    xyz = np.zeros((nSensors, 3))
    axis = {}
    for t in fmt:
        for ax, c in enumerate(b"xyz"):
            lo, up = bytes([c]), bytes([c]).upper()
            if t in (lo, up, lo + b"/m", up + b"/m"):
                axis[t] = (ax, 1.0)
            elif t in (lo + b"/mm", up + b"/mm") and not (c == ord("z") and t == b"Z/mm"):
                axis[t] = (ax, 1000.0)
    for i in range(nSensors):
        row = rows.nonempty()
        if not row:
            raise ValueError(f"{fileName}: too few sensor data. {nSensors} sensors expected but {i} found.")
        for j, t in enumerate(row):
            if j == len(fmt):
                break
            ax = axis.get(fmt[j])
            if ax is not None:
                xyz[i, ax[0]] = _strtod(t) / ax[1] if ax[1] != 1.0 else _strtod(t)

    # createSensor(pos.round(1e-12)) with the 1e-3 tolerance (datacontainer.cpp:170-184):
    # a position within reach of an earlier one becomes a duplicate mapped to the
    # last such sensor and is dropped again once the indices are remapped.
    sensors = _round(xyz, 1e-12)
    dup = []
    for i in range(1, nSensors):
        d = sensors[:i] - sensors[i]
        hit = np.nonzero(np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) < 1e-3)[0]
        if len(hit):
            dup.append((i, int(hit[-1])))

    row = rows.nonempty()
    if len(row) != 1:
        raise ValueError(f"{fileName}: cannot determine data size. {len(row)}")
    nData = _atoi(row[0])
    size = max(nData, 0)
    tok = {name: np.full(size, -1.0) for name in SENSOR_TOKENS}
    tok.update({name: np.zeros(size) for name in ERT_TOKENS})
    if nData > 0:
        f2 = rows.format_line()
        if f2 is not None:
            if not f2:
                raise ValueError(f"{fileName}: can not determine data format.")
            fmt = f2
    cols = {}
    for d in range(max(nData, 0)):
        row = rows.nonempty()
        if not row:
            raise ValueError(f"{fileName}: too few data. {nData} data expected and {d} data found.")
        for j, t in enumerate(row):
            if j == len(fmt):
                break
            col = cols.get(fmt[j])
            if col is None:
                col = cols[fmt[j]] = np.zeros(nData)
            col[d] = _strtod(t)
    for name in sorted(cols):
        target = translate_alias(name)
        if target is not None:
            scale = 1.0
            if b"mV" in name or b"mA" in name or b"ms" in name:
                scale = 1.0 / 1000.0
            if b"%" in name:
                scale = 1.0 / 100.0
            tok[target] = cols[name] * scale
        else:
            tok[name] = cols[name]
    if sensorIndicesFromOne and size > 0:
        for name in SENSOR_TOKENS:
            tok[name] = tok[name] - 1.0
    tok[b"valid"] = np.ones(size)

    nS = len(sensors)
    if dup:
        for old, new in dup:
            for name in SENSOR_TOKENS:
                v = tok[name]
                v[v == old] = new
        used = np.zeros(nS, bool)
        for name in SENSOR_TOKENS:
            v = tok[name]
            ids = v.astype(np.int64)
            ok = (ids > -1) & (ids < nS)
            used[ids[ok]] = True
        perm = np.cumsum(used) - 1
        sensors = sensors[used]
        for name in SENSOR_TOKENS:
            v = tok[name]
            ids = v.astype(np.int64)
            ok = (ids > -1) & (ids < nS)
            v[ok] = perm[ids[ok]]
        nS = len(sensors)

    valid = tok[b"valid"]
    _checked_validity(tok, valid, nS)
    if removeInvalid:
        keep = valid == 1
        tok = {k: v[keep] for k, v in tok.items()}

    # optional topography points after the data (datacontainer.cpp:421-474)
    topo = np.zeros((0, 3))
    row = rows.nonempty()
    if len(row) == 1:
        nTopo = _atoi(row[0])
        if nTopo > 0:
            # without a '#' line the reference keeps the format of the section before
            # (the data or sensor format) and refuses its non-coordinate tokens
            tf = rows.format_line()
            if tf is None:
                tf = fmt
            elif not tf or tf[0] not in (b"x", b"X"):
                tf = [b"x", b"y", b"z"]
            topo = np.zeros((nTopo, 3))
            for i in range(nTopo):
                row = rows.nonempty()
                if not row:
                    raise ValueError(f"{fileName}: too few topo data. {nTopo} topopoints expected and {i} found.")
                for j, t in enumerate(row):
                    if j == len(tf):
                        break
                    ax = {b"x": 0, b"X": 0, b"y": 1, b"Y": 1, b"z": 2, b"Z": 2}.get(tf[j])
                    if ax is None:
                        raise ValueError(f"{fileName}: format description unknown: topo electrode "
                                         f"format[{j}] = {tf[j].decode('latin-1')}")
                    topo[i, ax] = _strtod(t)
            topo = _round(topo, 1e-12)

    data = DataContainerERT()
    data.setSensorPositions(sensors)
    data._fromOne = bool(sensorIndicesFromOne)
    data._topoPoints = topo
    n = len(tok[b"valid"])
    data.resize(n)
    for name, v in tok.items():
        data.set(name.decode("latin-1"), v)
    return data


def _fmt_g(v):
    return "%.14g" % v


def _fmt_e(v):
    return "%.14e" % v


def save_data(data, fileName, formatData="all", formatSensor="x y z", noFilter=False):
    """DataContainer::save / write (datacontainer.cpp:519-633): 14 significant
    digits, sensor indices written from one (unless the container was loaded
    with indices from zero), '# ' format lines, tab-separated columns.  The
    "all" format writes every token, invalid data included."""
    fromOne = 1 if getattr(data, "_fromOne", True) else 0
    n = data.size()
    tok = {k.encode("latin-1"): np.asarray(v) for k, v in data._tok.items()}
    for name in ERT_TOKENS:
        tok.setdefault(name, np.zeros(n))
    if b"valid" not in tok:
        tok[b"valid"] = np.ones(n)
    valid = tok[b"valid"]
    if formatData.lower() == "all":
        idx = np.nonzero(valid > -1)[0]
        names = sorted(tok)
        fmt = b"".join(t + b" " for t in names if t in SENSOR_TOKENS) + \
            b"".join(t + b" " for t in names if t not in SENSOR_TOKENS)
    else:
        idx = np.nonzero(valid > -1 if noFilter else valid == 1)[0]
        fmt = formatData.encode("latin-1")
    out = [b"%d\n" % data.sensorCount(), b"# " + formatSensor.encode("latin-1") + b"\n"]
    sfmt = formatSensor.encode("latin-1").split()
    pos = data.sensorPositions()
    for p in pos:
        cols = []
        for t in sfmt:
            c = b""
            if t in (b"x", b"X"):
                c = _fmt_g(p[0]).encode()
            elif t in (b"y", b"Y"):
                c = _fmt_g(p[1]).encode()
            elif t in (b"z", b"Z"):
                c = _fmt_g(p[2]).encode()
            cols.append(c)
        out.append(b"\t".join(cols) + b"\n")
    out.append(b"%d\n" % len(idx))
    out.append(b"# " + fmt + b"\n")
    writers = []
    for t in fmt.split():
        name = translate_alias(t) or t
        if name not in tok:
            raise ValueError(f"no such data: {name.decode('latin-1')}")
        v = tok[name]
        if name in SENSOR_TOKENS:
            writers.append(lambda i, v=v: b"%d" % (int(v[i]) + fromOne))
        elif t == b"valid":
            writers.append(lambda i, v=v: b"%d" % int(v[i]))
        else:
            writers.append(lambda i, v=v: _fmt_e(float(v[i])).encode())
    for i in idx:
        out.append(b"\t".join(w(i) for w in writers) + b"\n")
    out.append(b"0\n")
    with open(fileName, "wb") as f:
        f.write(b"".join(out))
    return 1


# -- binary mesh -------------------------------------------------------------
def _bms_name(fileName):
    """Mesh::saveBinaryV2 / loadBinaryV2 append .bms unless present."""
    s = str(fileName)
    i = s.rfind(".bms")
    return (s if i < 0 else s[:i]) + ".bms"


class _Reader:
    def __init__(self, raw):
        self.raw = raw
        self.o = 0

    def take(self, dtype, count):
        dt = np.dtype(dtype).newbyteorder("<")
        nbytes = dt.itemsize * count
        if self.o + nbytes > len(self.raw):
            raise ValueError("unexpected end of file")
        a = np.frombuffer(self.raw, dt, count, self.o)
        self.o += nbytes
        return a

    def one(self, dtype):
        return int(self.take(dtype, 1)[0])


def load_mesh(fileName):
    """Mesh::loadBinaryV2 (mesh_io.cpp:540-630) of a P1 simplex mesh.  The
    file's data map becomes Mesh.dataMap(); its left/right cell ids are kept
    for saving the mesh unchanged."""
    with open(_bms_name(fileName), "rb") as f:
        r = _Reader(f.read())
    dim = r.one(np.uint8)
    if dim not in (2, 3):
        raise ValueError(f"cannot determine dimension {dim}")
    version = r.one(np.uint8)
    geometry = False
    if version == 3:
        geometry = bool(r.take(np.uint8, 128)[0])
    elif version != 2:
        raise ValueError(f"wrong version {version}")
    nn = r.one(np.uint32)
    if nn > 1e9:
        raise ValueError(f"probably something wrong: nVerts > 1e9 {nn}")
    xyz = np.zeros((0, 3))
    nmark = np.zeros(0, np.int32)
    if nn:
        xyz = r.take(np.float64, 3 * nn).reshape(nn, 3).copy()
        nmark = r.take(np.int32, nn).copy()
    nc = r.one(np.uint32)
    cells = np.zeros((0, dim + 1), np.int32)
    cmark = np.zeros(0, np.int32)
    if nc:
        cv = r.take(np.uint8, nc)
        if np.any(cv != dim + 1):
            k = int(cv[np.argmax(cv != dim + 1)])
            raise ValueError(f"cell with {k} nodes: only {'triangle' if dim == 2 else 'tetrahedron'} "
                             "meshes (P1) are supported")
        cells = r.take(np.uint32, int(cv.sum(dtype=np.int64))).astype(np.int32).reshape(nc, dim + 1)
        cmark = r.take(np.int32, nc).copy()
    nb = r.one(np.uint32)
    bounds = np.zeros((0, dim), np.int32)
    bmark = np.zeros(0, np.int32)
    lr = None
    if nb:
        bv = r.take(np.uint8, nb)
        if np.any(bv != dim):
            raise ValueError(f"boundary with {int(bv[np.argmax(bv != dim)])} nodes: P1 simplex meshes only")
        bounds = r.take(np.uint32, int(bv.sum(dtype=np.int64))).astype(np.int32).reshape(nb, dim)
        bmark = r.take(np.int32, nb).copy()
        lr = (r.take(np.int32, nb).copy(), r.take(np.int32, nb).copy())
    nd = r.one(np.uint64)
    data = {}
    for _ in range(nd):
        ln = r.one(np.uint64)
        name = bytes(r.take(np.uint8, ln)).decode("latin-1")
        cnt = r.one(np.uint64)
        data[name] = r.take(np.float64, cnt).copy()
    if (cells.size and cells.max() >= nn) or (bounds.size and bounds.max() >= nn):
        raise ValueError("node index out of range")
    mesh = Mesh(dim, xyz, cells, cellMarkers=cmark, boundaries=bounds, boundaryMarkers=bmark,
                nodeMarkers=nmark)
    mesh._geometry = geometry
    mesh._fileNeighbors = lr
    for k, v in data.items():
        mesh.addData(k, v)
    return mesh


def _neighbors(mesh):
    """Left / right cell of each boundary: the adjacent cell (the lower id of
    two) and the other one, -1 for none."""
    left = np.asarray(mesh.boundaryLeftCell(), np.int32)
    right = np.full(len(left), -1, np.int32)
    nb = mesh.boundaryCount()
    if nb:
        cells = mesh.cells
        nv = cells.shape[1]
        n = mesh.nodeCount()
        faces = np.sort(np.concatenate([np.delete(cells, v, axis=1) for v in range(nv)]), axis=1)
        owner = np.tile(np.arange(len(cells), dtype=np.int64), nv)
        key = np.zeros(len(faces), np.int64)
        for j in range(faces.shape[1]):
            key = key * n + faces[:, j]
        bkey = np.zeros(nb, np.int64)
        sb = np.sort(mesh.bounds, axis=1)
        for j in range(sb.shape[1]):
            bkey = bkey * n + sb[:, j]
        order = np.lexsort((owner, key))
        key, owner = key[order], owner[order]
        lo = np.searchsorted(key, bkey, "left")
        hi = np.searchsorted(key, bkey, "right")
        two = hi - lo >= 2
        right[two] = owner[lo[two] + 1]
    return left, right


def save_mesh(mesh, fileName):
    """Mesh::saveBinaryV2 (mesh_io.cpp:391-538): binary version 3."""
    dim = mesh.dim()
    out = [struct.pack("<BB", dim, 3)]
    head = bytearray(128)
    head[0] = 1 if getattr(mesh, "_geometry", False) else 0
    out.append(bytes(head))
    nn = mesh.nodeCount()
    out.append(struct.pack("<I", nn))
    if nn:
        out.append(np.ascontiguousarray(mesh.nodes, "<f8").tobytes())
        out.append(np.ascontiguousarray(mesh.nodeMarkers(), "<i4").tobytes())
    nc = mesh.cellCount()
    out.append(struct.pack("<I", nc))
    if nc:
        out.append(np.full(nc, dim + 1, np.uint8).tobytes())
        out.append(np.ascontiguousarray(mesh.cells, "<u4").tobytes())
        out.append(np.ascontiguousarray(mesh.cellMarkers(), "<i4").tobytes())
    nb = mesh.boundaryCount()
    out.append(struct.pack("<I", nb))
    if nb:
        lr = getattr(mesh, "_fileNeighbors", None)
        left, right = lr if lr is not None else _neighbors(mesh)
        out.append(np.full(nb, dim, np.uint8).tobytes())
        out.append(np.ascontiguousarray(mesh.bounds, "<u4").tobytes())
        out.append(np.ascontiguousarray(mesh.boundaryMarkers(), "<i4").tobytes())
        out.append(np.ascontiguousarray(left, "<i4").tobytes())
        out.append(np.ascontiguousarray(right, "<i4").tobytes())
    dm = mesh.dataMap()
    names = sorted(dm, key=lambda k: k.encode("latin-1"))
    out.append(struct.pack("<Q", len(names)))
    for k in names:
        kb = k.encode("latin-1")
        v = np.ascontiguousarray(dm[k], "<f8")
        out.append(struct.pack("<Q", len(kb)) + kb + struct.pack("<Q", len(v)) + v.tobytes())
    with open(_bms_name(fileName), "wb") as f:
        f.write(b"".join(out))


def load(fileName, **kw):
    """A .bms mesh or a unified-format data file, by file suffix."""
    if str(fileName).endswith(".bms") or os.path.exists(str(fileName) + ".bms"):
        return load_mesh(fileName)
    return load_data(fileName, **kw)

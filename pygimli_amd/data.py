"""ERT data container (subset of pyGIMLi's DataContainerERT).

Mirrors core/src/bert/bertDataContainer.h + core/src/datacontainer.h for the
fields the forward/Jacobian path reads: sensor positions and the per-datum
electrode indices a, b, m, n (-1 = none), geometric factor k and rhoa.
"""
from __future__ import annotations

import numpy as np


class DataContainerERT:
    """Sensor positions plus a/b/m/n/k/rhoa token vectors."""

    def __init__(self, sensors=None, a=None, b=None, m=None, n=None, k=None):
        self._sensors = np.zeros((0, 3))
        self._tok = {}
        if sensors is not None:
            self.setSensorPositions(sensors)
        if a is not None:
            self.resize(len(a))
            for name, v in (("a", a), ("b", b), ("m", m), ("n", n)):
                self.set(name, v if v is not None else np.full(len(a), -1))
        if k is not None:
            self.set("k", k)

    # -- sensors -------------------------------------------------------------
    def setSensorPositions(self, pos):
        pos = np.asarray(pos, np.float64)
        if pos.ndim == 1:
            pos = pos[:, None]
        if pos.shape[1] < 3:
            pos = np.hstack([pos, np.zeros((len(pos), 3 - pos.shape[1]))])
        self._sensors = np.ascontiguousarray(pos[:, :3])

    def sensorPositions(self):
        return self._sensors

    def sensors(self):
        return self._sensors

    def sensorCount(self):
        return self._sensors.shape[0]

    def createSensor(self, pos, tolerance=1e-3):
        """Append a sensor unless one lies within ``tolerance`` (datacontainer.h:172)."""
        p = np.zeros(3)
        pos = np.asarray(pos, np.float64).ravel()
        p[:len(pos)] = pos
        if self.sensorCount():
            d = np.linalg.norm(self._sensors - p, axis=1)
            i = int(np.argmin(d))
            if d[i] < tolerance:
                return i
        self._sensors = np.vstack([self._sensors, p])
        return self.sensorCount() - 1

    # -- data tokens -----------------------------------------------------------
    def size(self):
        return len(self._tok.get("a", ()))

    def __len__(self):
        return self.size()

    def resize(self, n):
        for name in ("a", "b", "m", "n"):
            v = self._tok.get(name, np.zeros(0, np.int64))
            w = np.full(n, -1, np.int64)
            w[:min(n, len(v))] = v[:n]
            self._tok[name] = w
        for name, v in list(self._tok.items()):
            if name not in ("a", "b", "m", "n"):
                w = np.zeros(n)
                w[:min(n, len(v))] = v[:n]
                self._tok[name] = w

    def set(self, name, v):
        v = np.asarray(v)
        if self.size() and len(v) != self.size():
            raise ValueError(f"token {name}: size {len(v)} != {self.size()}")
        if name in ("a", "b", "m", "n"):
            v = v.astype(np.int64)
        else:
            v = v.astype(np.float64)
        self._tok[name] = v.copy()

    def get(self, name):
        if name not in self._tok:
            return np.zeros(self.size())
        return self._tok[name]

    def __getitem__(self, name):
        return self.get(name)

    def __setitem__(self, name, v):
        self.set(name, v)

    def haveData(self, name):
        return name in self._tok and (name in "abmn" or np.any(self._tok[name] != 0))

    def allNonZero(self, name):
        return name in self._tok and bool(np.all(self._tok[name] != 0))

    def createFourPointData(self, i, a, b, m, n):
        if i >= self.size():
            self.resize(i + 1)
        for name, v in (("a", a), ("b", b), ("m", m), ("n", n)):
            self._tok[name][i] = v
        return i

    def save(self, fileName, formatData="all", formatSensor="x y z", noFilter=False):
        """Unified data format file (DataContainer::save)."""
        from .fileio import save_data
        return save_data(self, fileName, formatData, formatSensor, noFilter)

    @staticmethod
    def load(fileName, sensorIndicesFromOne=True, removeInvalid=True):
        """DataContainerERT(fileName, removeInvalid) of the unified data format."""
        from .fileio import load_data
        return load_data(fileName, sensorIndicesFromOne, removeInvalid)

    def additionalPoints(self):
        """Topography points read with the data (DataContainer::additionalPoints);
        they are not written back (DataContainer::write writes a zero count)."""
        return getattr(self, "_topoPoints", np.zeros((0, 3)))

    def abmn(self):
        return np.stack([self._tok[x] for x in "abmn"], axis=1).astype(np.int32)

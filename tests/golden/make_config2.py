"""Write tests/golden/config2_3d.npz: an independent FP64 answer for the 3D
workload (201,898 nodes / 1,150,488 tetrahedra, 8 x 8 electrode grid at 2 m,
2000 random quadrupoles, log10 rho Gaussian random field) from the oracle's
direct path:

  * singularity-removal potentials from SuperLU factorisations (FP64) of the
    assembled system (the oracle's stand-in for CHOLMOD,
    dcfemmodelling.cpp:2152-2296 -> linSolver.cpp:96-178),
  * rhoa through the reference's response rules (round to 1e-10, x k,
    reciprocal geometric mean; dcfemmodelling.cpp:1085-1197),
  * the 64 x 64 electrode potential map (DataMap::collect, datamap.cpp:57-98),
  * Jacobian columns of 256 sampled cells (calc2, bertJacobian.cpp:170-219,
    scaled by k_i / rho_c^2, dcfemmodelling.cpp:1375-1384), stored as FP32.

    python tests/golden/make_config2.py        (30-40 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ert_oracle as eo  # noqa: E402
from pygimli_amd import workloads  # noqa: E402

N_SAMPLE = 256


class _CellSubset:
    """The cells of `idx` of a mesh, each its own parameter (marker = its
    position in idx), for sensitivity_calc2."""

    def __init__(self, mesh, idx):
        self.cells = mesh.cells[idx]
        self._m = np.arange(len(idx), dtype=np.int32)

    def cellMarkers(self):
        return self._m


def sample_cells(mesh, sensors, seed=7):
    """128 cells uniformly over the mesh and 128 among the cells within 6 m of
    an electrode (where the sensitivity and the potential gradients are
    largest)."""
    rng = np.random.default_rng(seed)
    C = mesh.cellCount()
    cc = mesh.cellCenters()
    d = np.min(np.linalg.norm(cc[:, None, :] - sensors[None, :, :], axis=2), axis=1)
    near = np.flatnonzero(d < 6.0)
    a = rng.choice(C, N_SAMPLE // 2, replace=False)
    b = rng.choice(np.setdiff1d(near, a), N_SAMPLE // 2, replace=False)
    return np.sort(np.concatenate([a, b])).astype(np.int32)


def fingerprint(mesh, rho, data):
    return np.array([mesh.nodeCount(), mesh.cellCount(), mesh.positions().sum(),
                     np.abs(mesh.positions()).sum(), mesh.cells.astype(np.float64).sum(),
                     rho.sum(), np.log(rho).sum(), data.abmn().astype(np.float64).sum()], np.float64)


def main():
    t0 = time.time()
    mesh, data, rho = workloads.ert3d_grid_grf()
    print(f"mesh {mesh.nodeCount()} nodes / {mesh.cellCount()} cells", flush=True)
    fo = eo.Forward(mesh, data)
    rhoa = fo.response(rho)
    print(f"response {time.time() - t0:.0f} s", flush=True)
    emap = fo.electrode_map(fo.potentials())
    idx = sample_cells(mesh, data.sensorPositions())
    K, M, vol, grad = fo.geo
    geo = (K[idx], M[idx], None, None)
    Jc = eo.sensitivity_calc2(_CellSubset(mesh, idx), geo, data.abmn(), fo.subpots, fo.k, fo.w,
                              data.sensorCount())
    Jc *= data["k"][:, None] / (rho[idx] * rho[idx])[None, :]
    np.savez_compressed(os.path.join(HERE, "config2_3d.npz"), fingerprint=fingerprint(mesh, rho, data),
                        rhoa=rhoa, k=np.asarray(data["k"], np.float64), emap=emap, cells=idx,
                        J=Jc.astype(np.float32))
    print(f"done {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()

"""The full-size 3D fixture (tests/golden/config2_3d.npz) belongs to the
workload the generator builds today: its fingerprint (node/cell counts,
coordinate, connectivity, model and data checksums) matches, and its arrays
have the documented shapes."""
import os
import sys

import numpy as np

from pygimli_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_config2  # noqa: E402


def test_config2_fixture_matches_workload():
    mesh, data, rho = workloads.ert3d_grid_grf()
    with np.load(os.path.join(HERE, "golden", "config2_3d.npz")) as z:
        fp, rhoa, emap, cells, J, k = (z[n] for n in ("fingerprint", "rhoa", "emap", "cells", "J", "k"))
    np.testing.assert_array_equal(make_config2.fingerprint(mesh, rho, data), fp)
    assert rhoa.shape == (2000,) and emap.shape == (64, 64) and J.shape == (2000, 256)
    assert cells.shape == (256,) and cells.min() >= 0 and cells.max() < mesh.cellCount()
    assert np.all(rhoa > 0) and np.all(np.isfinite(J))
    # the map is symmetric to the solver's accuracy only through reciprocity of
    # the total potentials; its diagonal (source electrode) dominates each row
    assert np.all(np.argmax(np.abs(emap), axis=1) == np.arange(64))

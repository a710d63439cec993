"""GPU: the batched forward (one block solve over several independent 2.5D
problems) against each problem solved on its own, captured-graph replay
against eager launches, and the Jacobian after a batched forward."""
import copy

import numpy as np
import pytest

from pygimli_amd import workloads

pytestmark = pytest.mark.gpu


def _fops(probs, coarse):
    from pygimli_amd.ertModelling import ERTModelling
    fops = []
    for mesh, data, rho, k, w in probs:
        f = ERTModelling()
        f.setMesh(mesh)
        f.setData(copy.deepcopy(data))
        f.setkValues(k)
        f.setWeights(w)
        f.solverOptions.update(method="mixed", rtol=1e-8, coarse=coarse)
        fops.append(f)
    return fops


@pytest.mark.parametrize("coarse", [0, 32])
def test_batch_matches_single(coarse):
    import torch
    from pygimli_amd.batch import batch_response
    probs = workloads.batched_2d(count=5, seed=7, n_range=(800, 3000), ne_range=(12, 70))
    print("problems", [(m.nodeCount(), d.size(), len(k)) for m, d, _, k, _ in probs], flush=True)
    fops = _fops(probs, coarse)
    single = []
    for f, (_, _, rho, _, _) in zip(fops, probs):
        single.append(f.response(rho))
        assert f.lastStats["converged"] == 1
    batched = batch_response(fops, [p[2] for p in probs])
    for f, r1, r0 in zip(fops, batched, single):
        assert f.lastStats["converged"] == 1, f.lastStats
        assert np.all(np.isfinite(r1)) and np.all(np.isfinite(r0))
        np.testing.assert_allclose(r1, r0, rtol=1e-6)
    # eager launches give the same bits as the graph replays
    for f in fops:
        f.solverOptions["graph"] = False
    eager = batch_response(fops, [p[2] for p in probs])
    for r1, r2 in zip(batched, eager):
        np.testing.assert_array_equal(r1, r2)
    # the Jacobian of a problem of the batch from its own potentials; problem
    # 0 has 67 electrodes, so its 2.5D tile is 32 cells (the LDS-size rule)
    f = fops[0]
    f.createJacobian(probs[0][2])
    J_b = f.jacobian().cpu().numpy().copy()
    f.response(probs[0][2])
    f.createJacobian(probs[0][2])
    J_s = f.jacobian().cpu().numpy()
    assert np.linalg.norm(J_b - J_s) <= 1e-6 * np.linalg.norm(J_s)
    if coarse == 0:
        from oracle import ert_oracle as eo
        mesh, data, rho, k, w = probs[0]
        J_o = eo.Forward(mesh, copy.deepcopy(data), k, w).jacobian(rho)
        err = np.linalg.norm(J_s - J_o) / np.linalg.norm(J_o)
        print("J vs oracle", err)
        assert err < 1e-6
    torch.cuda.synchronize()


def test_batch_band_direct():
    """The band solver over a batch (one factor launch, one solve launch for
    every wavenumber of every problem) equals the single-problem band solves
    bit for bit and the block solver to its tolerance."""
    from pygimli_amd.batch import batch_response
    probs = workloads.batched_2d(count=4, seed=9, n_range=(800, 4000), ne_range=(20, 100))
    fops = _fops(probs, 0)
    for f in fops:
        f.solverOptions.update(method="band", rtol=1e-10)
    single = [f.response(p[2]) for f, p in zip(fops, probs)]
    batched = batch_response(fops, [p[2] for p in probs])
    for f, r1, r0 in zip(fops, batched, single):
        assert f.lastStats["converged"] == 1 and f.lastStats["max_rel_residual"] < 1e-11, f.lastStats
        assert np.all(np.isfinite(r1))
        np.testing.assert_array_equal(r1, r0)
    for f in fops:
        f.solverOptions.update(method="block64", rtol=1e-10, coarse=0)
    ref = batch_response(fops, [p[2] for p in probs])
    for r1, r2 in zip(batched, ref):
        np.testing.assert_allclose(r1, r2, rtol=1e-8)


def test_batch_jacobian():
    """One Jacobian launch over several problems (pgert_batch_jacobian)
    against each problem's own createJacobian, FP64 and FP32, and problem 0
    against the oracle."""
    import torch
    from pygimli_amd.batch import batch_jacobian, batch_response
    probs = workloads.batched_2d(count=4, seed=11, n_range=(800, 3000), ne_range=(12, 100))
    print("problems", [(m.nodeCount(), d.size(), len(d.sensorPositions())) for m, d, *_ in probs], flush=True)
    fops = _fops(probs, 0)
    models = [p[2] for p in probs]
    for f in fops:
        f.solverOptions.update(method="band", rtol=1e-10)
    batch_response(fops, models)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 1e-6)):
        for f in fops:
            f.jacobianDtype = dt
        Js = [J.cpu().numpy().astype(np.float64) for J in batch_jacobian(fops, models)]
        for f, m, Jb in zip(fops, models, Js):
            f.createJacobian(m)
            Js_ = f.jacobian().cpu().numpy().astype(np.float64)
            assert np.all(np.isfinite(Jb))
            err = np.linalg.norm(Jb - Js_) / np.linalg.norm(Js_)
            assert err <= tol, (dt, err)
        if dt == torch.float32:
            from oracle import ert_oracle as eo
            mesh, data, rho, k, w = probs[0]
            J_o = eo.Forward(mesh, copy.deepcopy(data), k, w).jacobian(rho)
            err = np.linalg.norm(Js[0] - J_o) / np.linalg.norm(J_o)
            print("batched fp32 J vs oracle", err)
            assert err < 1e-5

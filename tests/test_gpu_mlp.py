"""GPU: the learned warm start's MFMA kernels against the numpy oracle
(oracle/mlp_oracle.py, same bf16 rounding), and what the trained network
changes in the solver (iterations only; accuracy is fixed by the residual)."""
import copy

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from pygimli_amd import workloads

pytestmark = pytest.mark.gpu


def _fop(case="small_3d"):
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = getattr(workloads, case)()
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    return fop, mesh, data, rho


def _params(seed):
    rng = np.random.default_rng(seed)
    W, b = mo.init_params(seed)
    b = [(0.05 * rng.standard_normal(x.shape)).astype(np.float32) for x in b]
    return W, b


def test_mlp_forward_matches_oracle():
    import torch
    from pygimli_amd.learned import LearnedWarmStart
    fop, *_ = _fop()
    W, b = _params(7)
    lw = LearnedWarmStart(fop, mo.pack(W, b))
    rng = np.random.default_rng(0)
    X = rng.standard_normal((200, 32)).astype(np.float32)
    y = lw.predict(torch.as_tensor(X, device=fop.device)).cpu().numpy()
    yo, _, _ = mo.forward(X, W, b)
    err = np.abs(y - yo).max() / np.abs(yo).max()
    print("forward rel err", err)
    assert err < 2e-3


def test_mlp_gradients_match_oracle():
    import torch
    from pygimli_amd.learned import LearnedWarmStart
    fop, *_ = _fop()
    W, b = _params(11)
    lw = LearnedWarmStart(fop, mo.pack(W, b))
    rng = np.random.default_rng(1)
    X = rng.standard_normal((256, 32)).astype(np.float32)
    t = rng.standard_normal(256).astype(np.float32)
    loss, g = lw.step(torch.as_tensor(X, device=fop.device), torch.as_tensor(t, device=fop.device), 0.0,
                      apply=False, want_grad=True)
    lo, gW, gb = mo.backward(X, t, W, b)
    go = mo.pack([x.astype(np.float32) for x in gW], [x.astype(np.float32) for x in gb])
    print("loss", loss, lo)
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    # per block relative error
    o = 0
    for fi, fo in zip(mo.SIZES[:-1], mo.SIZES[1:]):
        for size in (fi * fo, fo):
            a, r = g[o:o + size], go[o:o + size]
            err = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            print("block", o, size, err)
            assert err < 2e-2
            o += size


def test_learned_warm_start_gain():
    """Train on the method's own solves of random log-normal fields, then
    compare iterations on held-out fields with and without the warm start at
    the same FP64 residual target."""
    import torch
    from pygimli_amd.learned import LearnedWarmStart
    fop, mesh, data, rho = _fop()
    rng = np.random.default_rng(5)
    c = mesh.cellCenters()

    def field(seed):
        lr = workloads.gaussian_random_field(c, 2.0, 0.5, 2.0 + seed % 3, seed, (-6, -6, -8), (6, 6, 0))
        return 10.0 ** lr

    fop.solverOptions.update(method="mixed", rtol=1e-8)
    lw = LearnedWarmStart(fop, seed=0)
    losses = lw.train([field(s) for s in range(10, 22)], samples_per_model=8192, epochs=40, batch=1024, lr=2e-3)
    print("loss first/last", losses[0], losses[-1])
    assert losses[-1] < losses[0]
    it0, it1 = [], []
    for s in (101, 102, 103):
        m = field(s)
        fop.solverOptions["learned"] = False
        r0 = fop.response(m)
        it0.append(fop.lastStats["iterations"])
        fop.solverOptions["learned"] = True
        r1 = fop.response(m)
        it1.append(fop.lastStats["iterations"])
        assert fop.lastStats["converged"] == 1
        np.testing.assert_allclose(r1, r0, rtol=1e-6)
    print("iterations without / with learned warm start", it0, it1)
    # the network must earn its place: at least a 20 % cut in iterations
    assert np.mean(it1) <= 0.8 * np.mean(it0), (it0, it1)
    # a checkpoint restores the same predictions on a fresh context
    import os
    import tempfile
    from pygimli_amd.ertModelling import ERTModelling
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "warmstart.npz")
        lw.save(path, epochs=40)
        fop2 = ERTModelling()
        fop2.setMesh(mesh)
        fop2.setData(copy.deepcopy(data))
        fop2.solverOptions.update(method="mixed", rtol=1e-8, learned=True)
        lw2 = LearnedWarmStart(fop2, seed=1)
        meta = lw2.load(path)
        assert int(meta["epochs"]) == 40
        np.testing.assert_array_equal(lw2.params(), lw.params())
        fop2.response(field(101))
        assert fop2.lastStats["iterations"] == it1[0]

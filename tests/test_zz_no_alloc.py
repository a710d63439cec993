"""GPU: pgert_cgls_noalloc (work vectors in the library's fixed per-device
scratch array, no device allocation) gives the results of pgert_cgls, through
the C-ABI, on two calls in a row (the second smaller, so stale workspace
contents from the first would show), and rejects a problem whose workspace
exceeds the array before touching the GPU."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, D, P, nC):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    S = rng.standard_normal((D, P))
    C = (sp.random(nC, P, density=4.0 / P, random_state=seed, format="csr") + sp.eye(nC, P, format="csr")).tocsr()
    vec = dict(dWeight=rng.uniform(0.5, 2.0, D), b=rng.standard_normal(D), wc=rng.uniform(0.5, 1.5, nC),
               wm=rng.uniform(0.5, 1.5, P), tm=rng.uniform(0.5, 2.0, P), td=rng.uniform(0.5, 2.0, D),
               roughness=0.1 * rng.standard_normal(nC), x0=0.01 * rng.standard_normal(P))
    return S, C, vec


def _run(ctx, entry, S, C, v, dtype, lam, max_iter, tol):
    import torch
    D, P = S.shape
    nC = C.shape[0]
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    rp = np.ascontiguousarray(C.indptr, np.int32)
    ci = np.ascontiguousarray(C.indices, np.int32)
    va = np.ascontiguousarray(C.data, np.float64)
    ctx.check(ctx.L.pgert_set_constraints(ctx.h, nC, P, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip),
                                          va.ctypes.data_as(dp)), "set_constraints")
    dev = torch.device("cuda", 0)
    Sd = torch.as_tensor(S.astype(dtype), device=dev)
    t = {k: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev) for k, a in v.items()}
    x = t["x0"].clone()
    vp = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731
    it, nr2 = ctypes.c_int(), ctypes.c_double()
    torch.cuda.synchronize()
    fn = getattr(ctx.L, entry)
    ctx.check(fn(ctx.h, vp(Sd), 0 if dtype == np.float64 else 1, D, P, vp(t["dWeight"]), vp(t["b"]), vp(x),
                 vp(t["wc"]), vp(t["wm"]), vp(t["tm"]), vp(t["td"]), float(lam), vp(t["roughness"]), int(max_iter),
                 float(tol), ctypes.byref(it), ctypes.byref(nr2)), entry)
    torch.cuda.synchronize()
    return x.cpu().numpy(), it.value, nr2.value


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cgls_noalloc_matches_cgls(dtype):
    from pygimli_amd import _lib
    ctx = _lib.Context(0)
    # two calls in a row, the second with a smaller problem
    for seed, D, P, nC, its in ((1, 257, 1031, 700, 25), (2, 120, 60, 80, 40)):
        S, C, v = _problem(seed, D, P, nC)
        xa, ita, ra = _run(ctx, "pgert_cgls", S, C, v, dtype, 2.0, its, 0.0)
        xb, itb, rb = _run(ctx, "pgert_cgls_noalloc", S, C, v, dtype, 2.0, its, 0.0)
        print(D, P, nC, ita, itb, ra, rb)
        assert ita == itb == its
        np.testing.assert_array_equal(xb, xa)
        assert ra == rb
        # default stopping rule
        xa, ita, ra = _run(ctx, "pgert_cgls", S, C, v, dtype, 0.5, 1000, -1.0)
        xb, itb, rb = _run(ctx, "pgert_cgls_noalloc", S, C, v, dtype, 0.5, 1000, -1.0)
        assert ita == itb
        np.testing.assert_array_equal(xb, xa)


def test_cgls_noalloc_rejects_oversized_problem():
    from pygimli_amd import _lib
    ctx = _lib.Context(0)
    ctx.check(ctx.L.pgert_set_constraints(ctx.h, 0, 2_000_000, (ctypes.c_int * 1)(0), None, None), "set_constraints")
    d = (ctypes.c_double * 1)()
    p = ctypes.cast(d, ctypes.c_void_p)
    rc = ctx.L.pgert_cgls_noalloc(ctx.h, p, 0, 100, 2_000_000, p, p, p, None, p, p, p, 1.0, None, 10, 0.0, None,
                                  None)
    assert rc != 0
    assert "scratch array" in ctx.L.pgert_last_error(ctx.h).decode()

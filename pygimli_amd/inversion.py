"""Gauss-Newton step on the GPU.

`solveCGLSCDWWhtrans` mirrors pg.core.solveCGLSCDWWhtrans
(core/src/solver.cpp:27-224), the regularised CGLS that RInversion::oneStep
(core/src/inversion.cpp:308) runs on the Jacobian and the constraint matrix of
the forward operator: here the Jacobian is the device tensor that
ERTModelling.createJacobian() left in HBM and every product runs in libpgert
(pgert_cgls).  `smoothness_constraints` builds the first-order constraint
matrix between cells sharing a face, one row per such pair (+1 / -1 on the
two cells' parameters), the matrix RegionManager makes for a single region
with constraint type 1.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

_ctx = {}


def _context(device):
    d = torch.device(device)
    idx = d.index if d.index is not None else torch.cuda.current_device()
    if idx not in _ctx:
        _ctx[idx] = {"ctx": _lib.Context(idx), "C": None}
    return _ctx[idx]


def _dev(v, n, device, name):
    t = torch.as_tensor(v, dtype=torch.float64, device=device).contiguous()
    if t.numel() != n:
        raise ValueError(f"{name}: size {t.numel()} != {n}")
    return t


def smoothness_constraints(mesh):
    """First-order smoothness: a row (+1, -1) per pair of cells with a common
    face (both with parameter marker >= 0; columns = markers).  Returns scipy CSR."""
    import scipy.sparse as sp
    cells = np.asarray(mesh.cells)
    nv = cells.shape[1]
    faces = np.concatenate([np.delete(cells, v, axis=1) for v in range(nv)])
    owner = np.tile(np.arange(len(cells)), nv)
    faces = np.sort(faces, axis=1)
    order = np.lexsort(faces.T[::-1])
    f, o = faces[order], owner[order]
    same = np.all(f[1:] == f[:-1], axis=1)
    a, b = o[:-1][same], o[1:][same]
    mk = np.asarray(mesh.cellMarkers())
    keep = (mk[a] >= 0) & (mk[b] >= 0) & (mk[a] != mk[b])
    a, b = mk[a[keep]], mk[b[keep]]
    n = len(a)
    rows = np.repeat(np.arange(n), 2)
    cols = np.stack([a, b], 1).ravel()
    vals = np.tile([1.0, -1.0], n)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, int(mk.max()) + 1))


def complex_constraints(C):
    """Constraints of a complex-resistivity inversion: the real-model matrix
    repeated for the real and the imaginary parameter halves
    (ERTModelling.createConstraints, ertModelling.py:150-162 ->
    pg.matrix.RepeatDMatrix(C1, 2)).  Returns scipy CSR."""
    import scipy.sparse as sp
    return sp.block_diag([C, C], format="csr")


def solveCGLSCDWWhtrans(S, C, dWeight, b, x, wc, wm, tm, td, lam, roughness, maxIter=200, tol=-1.0,
                        verbose=False):
    """Regularised CGLS of one Gauss-Newton step (solver.cpp:27-224):
    minimise |dW (b - td * S (x / tm))|^2 + lam |wc (C (wm x)) + wc roughness|^2
    starting from x.  S: device tensor [nData, nModel] (FP32 or FP64, e.g.
    ERTModelling.jacobian()); C: scipy sparse [nConst, nModel] or None;
    vectors: array-likes.  x is updated in place when it is an FP64 device
    tensor.  Returns (x as FP64 device tensor, iterations)."""
    if not isinstance(S, torch.Tensor) or S.device.type != "cuda" or S.dim() != 2:
        raise ValueError("S must be a 2D device tensor (the Jacobian)")
    if S.dtype not in (torch.float32, torch.float64):
        raise ValueError("S must be FP32 or FP64")
    S = S.contiguous()
    D, P = S.shape
    dev = S.device
    st = _context(dev)
    ctx = st["ctx"]
    with torch.cuda.device(dev):
        torch.cuda.current_stream().synchronize()
        nC = 0 if C is None else C.shape[0]
        if C is not None:
            if C.shape[1] != P:
                raise ValueError("C has another number of columns than S")
            key = (id(C), C.shape, C.nnz)
            if st["C"] != key:
                Cc = C.tocsr()
                rp = np.ascontiguousarray(Cc.indptr, np.int32)
                ci = np.ascontiguousarray(Cc.indices, np.int32)
                va = np.ascontiguousarray(Cc.data, np.float64)
                ip = ctypes.POINTER(ctypes.c_int)
                dp = ctypes.POINTER(ctypes.c_double)
                ctx.check(ctx.L.pgert_set_constraints(ctx.h, nC, P, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip),
                                                      va.ctypes.data_as(dp)), "set_constraints")
                st["C"] = key
        elif st["C"] is not None:
            ctx.check(ctx.L.pgert_set_constraints(ctx.h, 0, P, (ctypes.c_int * 1)(0), None, None), "set_constraints")
            st["C"] = None
        inplace = isinstance(x, torch.Tensor) and x.dtype == torch.float64 and x.device == dev and x.is_contiguous()
        xt = x if inplace else _dev(x, P, dev, "x").clone()
        if xt.numel() != P:
            raise ValueError("x has another size than the model")
        vec = {"dWeight": _dev(dWeight, D, dev, "dWeight"), "b": _dev(b, D, dev, "b"), "td": _dev(td, D, dev, "td"),
               "wm": _dev(wm, P, dev, "wm"), "tm": _dev(tm, P, dev, "tm")}
        if nC:
            vec["wc"] = _dev(wc, nC, dev, "wc")
            vec["roughness"] = _dev(roughness, nC, dev, "roughness")
        vp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        it = ctypes.c_int()
        nr2 = ctypes.c_double()
        ctx.check(ctx.L.pgert_set_stream(ctx.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                  "set_stream")
        ctx.check(ctx.L.pgert_cgls(ctx.h, vp(S), 0 if S.dtype == torch.float64 else 1, D, P, vp(vec["dWeight"]),
                                   vp(vec["b"]), vp(xt), vp(vec.get("wc")), vp(vec["wm"]), vp(vec["tm"]),
                                   vp(vec["td"]), float(lam), vp(vec.get("roughness")), int(maxIter), float(tol),
                                   ctypes.byref(it), ctypes.byref(nr2)), "cgls")
        if verbose:
            print(f"[ {it.value}/{nr2.value}]")
    return xt, it.value

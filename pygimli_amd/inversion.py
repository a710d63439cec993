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


def _set_constraints(st, C, nModel=None):
    """Hand C (scipy sparse or None) to the context's library state once per
    matrix (pgert_set_constraints)."""
    ctx = st["ctx"]
    if C is not None:
        key = (id(C), C.shape, C.nnz)
        if st["C"] != key:
            Cc = C.tocsr()
            rp = np.ascontiguousarray(Cc.indptr, np.int32)
            ci = np.ascontiguousarray(Cc.indices, np.int32)
            va = np.ascontiguousarray(Cc.data, np.float64)
            ip = ctypes.POINTER(ctypes.c_int)
            dp = ctypes.POINTER(ctypes.c_double)
            ctx.check(ctx.L.pgert_set_constraints(ctx.h, C.shape[0], C.shape[1], rp.ctypes.data_as(ip),
                                                  ci.ctypes.data_as(ip), va.ctypes.data_as(dp)), "set_constraints")
            st["C"] = key
    elif st["C"] is not None:
        ctx.check(ctx.L.pgert_set_constraints(ctx.h, 0, nModel or 1, (ctypes.c_int * 1)(0), None, None),
                  "set_constraints")
        st["C"] = None


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
                        verbose=False, stats=None):
    """Regularised CGLS of one Gauss-Newton step (solver.cpp:27-224):
    minimise |dW (b - td * S (x / tm))|^2 + lam |wc (C (wm x)) + wc roughness|^2
    starting from x.  S: device tensor [nData, nModel] (FP32 or FP64, e.g.
    ERTModelling.jacobian()); C: scipy sparse [nConst, nModel] or None;
    vectors: array-likes.  x is updated in place when it is an FP64 device
    tensor.  Returns (x as FP64 device tensor, iterations); a dict passed as
    `stats` receives the final squared residual of the normal equations
    ("normR2", the value the stopping rule tests)."""
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
        if C is not None and C.shape[1] != P:
            raise ValueError("C has another number of columns than S")
        _set_constraints(st, C, P)
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
        if stats is not None:
            stats["normR2"] = nr2.value
    return xt, it.value


class Inversion:
    """The reference's Gauss-Newton inversion loop, RInversion::run / oneStep /
    linesearch (core/src/inversion.cpp:135-370, core/src/inversion.h:554-700),
    for an ERTModelling on the GPU: log data (TransLog) and log model
    (TransLogLU(modelLower, modelUpper); upper 0 = TransLog) transforms,
    constraint matrix C (first-order smoothness by default) with weights 1,
    fixed lambda, no reference model, Jacobian recomputed every iteration.
    Forward, Jacobian, CGLS, transforms and the line-search objective run in
    libpgert; the Jacobian stays in HBM.

    warmStart: a trained LearnedWarmStart of `fop`; when given, the forward
    solves start from its prediction and it is fine-tuned on every accepted
    forward solution of the loop (fineTuneSteps Adam steps on fineTuneSamples
    fresh samples), so the network follows the model as it changes.
    """

    def __init__(self, fop, data, error, C=None, lam=50.0, maxIter=20, lineSearch=True, maxCGLSIter=200,
                 CGLStol=-1.0, stopAtChi1=True, dPhiAbortPercent=2.0, modelLower=0.0, modelUpper=0.0,
                 warmStart=None, fineTuneSamples=65536, fineTuneSteps=8, fineTuneBatch=4096, fineTuneLR=5e-4,
                 verbose=False):
        self.fop = fop
        self.dev = fop.device
        self.data = _dev(data, len(data), self.dev, "data")
        self.error = _dev(error, len(data), self.dev, "error")
        self.C = C if C is not None else smoothness_constraints(fop._mesh)
        self.lam = float(lam)
        self.maxIter = int(maxIter)
        self.lineSearch = bool(lineSearch)
        self.maxCGLSIter = int(maxCGLSIter)
        self.CGLStol = float(CGLStol)
        self.stopAtChi1 = bool(stopAtChi1)
        self.dPhiAbortPercent = float(dPhiAbortPercent)
        self.mLower, self.mUpper = float(modelLower), float(modelUpper)
        self.warmStart = warmStart
        self.fineTune = dict(samples=int(fineTuneSamples), steps=int(fineTuneSteps), batch=int(fineTuneBatch),
                             lr=float(fineTuneLR))
        self.verbose = verbose
        nC, nM = self.C.shape
        self.mw = torch.ones(nM, dtype=torch.float64, device=self.dev)
        self.cw = torch.ones(nC, dtype=torch.float64, device=self.dev)
        self._rng = np.random.default_rng(0)
        self.model = self.response = None
        self.chi2History, self.tauHistory, self.timings = [], [], []

    # ------------------------------------------------------------ device helpers
    def _st(self):
        st = _context(self.dev)
        _set_constraints(st, self.C)
        return st

    def _trans(self, op, a, b=None, tau=0.0, data=False):
        """pgert_trans: 0 trans, 1 invTrans, 2 deriv, 3 update(a, tau b), 4 error(a, b), 5 trans(a) - trans(b)."""
        st = self._st()
        ctx = st["ctx"]
        lo, up = (0.0, 0.0) if data else (self.mLower, self.mUpper)
        out = torch.empty_like(a)
        vp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        ctx.check(ctx.L.pgert_set_stream(ctx.h, ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)),
                  "set_stream")
        ctx.check(ctx.L.pgert_trans(ctx.h, op, a.numel(), vp(a), vp(b), float(tau), lo, up, vp(out)), "trans")
        return out

    def _phi(self, taus, model, dm, response, dd):
        """(phiD, phiM) at the step lengths taus (pgert_gn_phi)."""
        st = self._st()
        ctx = st["ctx"]
        taus = np.ascontiguousarray(taus, np.float64)
        pd = np.empty(len(taus))
        pm = np.empty(len(taus))
        vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        dp = ctypes.POINTER(ctypes.c_double)
        ctx.check(ctx.L.pgert_set_stream(ctx.h, ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)),
                  "set_stream")
        ctx.check(ctx.L.pgert_gn_phi(ctx.h, len(taus), taus.ctypes.data_as(dp), model.numel(), vp(model), vp(dm),
                                     vp(self.mw), vp(self.cw), self.mLower, self.mUpper, self.data.numel(),
                                     vp(self.data), vp(response), vp(dd), vp(self._derr), pd.ctypes.data_as(dp),
                                     pm.ctypes.data_as(dp)), "gn_phi")
        return pd, pm

    def phi(self, model, response):
        """getPhi (inversion.h:581-583) and its parts (phiD, phiM)."""
        z = torch.zeros_like(model)
        pd, pm = self._phi([0.0], model, z, response, torch.zeros_like(response))
        return pd[0] + self.lam * pm[0], pd[0], pm[0]

    def _forward(self, model):
        r = self.fop.response_device(model)
        if self.warmStart is not None:
            self._fine_tune()
        return r

    def _fine_tune(self):
        """Adam steps on samples of the solution the forward just produced."""
        ft = self.fineTune
        if ft["steps"] <= 0:
            return
        f, t = self.warmStart.collect(ft["samples"], self._rng)
        t = torch.clamp(t, -5.0, 5.0)
        nb = max(1, f.shape[0] // ft["batch"])
        for s in range(ft["steps"]):
            b = s % nb
            self.warmStart.step(f[b * ft["batch"]:(b + 1) * ft["batch"]].contiguous(),
                                t[b * ft["batch"]:(b + 1) * ft["batch"]].contiguous(), ft["lr"])

    # ------------------------------------------------------------ the loop
    def linesearch(self, modelNew, responseNew):
        """RInversion::linesearch (inversion.h:615-675)."""
        dModel = self._trans(5, modelNew, self.model)
        dData = self._trans(5, responseNew, self.response, data=True)
        taus = np.concatenate([[0.0], 0.01 * np.arange(1, 101)])
        pd, pm = self._phi(taus, self.model, dModel, self.response, dData)
        phi = pd + self.lam * pm
        minPhi, tau = phi[0], 0.0
        for i in range(1, 101):
            if phi[i] < minPhi:
                minPhi, tau = phi[i], taus[i]
        if tau < 0.03:
            # parabola through tau = 0, 0.3, 1 (linesearchQuad, inversion.h:677-691)
            tauquad = 0.3
            modelQuad = self._trans(3, self.model, dModel, tauquad)
            responseQuad = self._forward(modelQuad)
            phi0 = phi[0]
            phi10 = self.phi(modelNew, responseNew)[0] - phi0
            phit0 = self.phi(modelQuad, responseQuad)[0] - phi0
            dphit = phit0 - phi10 * tauquad
            tau = 0.0 if abs(dphit) < 1e-12 else (phit0 - phi10 * tauquad * tauquad) / dphit / 2.0
            tau = min(tau, 1.0)
            tau = max(tau, 0.03)
        return tau

    def oneStep(self):
        """RInversion::oneStep (inversion.cpp:249-370)."""
        import time
        t0 = time.perf_counter()
        self.iter += 1
        dData = self._trans(5, self.data, self.response, data=True)
        if float(torch.sum(torch.abs(dData))) < 1e-12:
            return False
        tJ = 0.0
        if self.iter > 1:
            t = time.perf_counter()
            self.fop.createJacobian(self.model)
            torch.cuda.synchronize(self.dev)
            tJ = time.perf_counter() - t
        tm = self._trans(2, self.model)
        td = self._trans(2, self.response, data=True)
        rough = (self.C @ (self._trans(0, self.model) * self.mw).cpu().numpy()) * self.cw.cpu().numpy()
        t = time.perf_counter()
        dm, itC = solveCGLSCDWWhtrans(self.fop.jacobian(), self.C, 1.0 / self._derr, dData,
                                      torch.zeros_like(self.model), self.cw, self.mw, tm, td, self.lam, rough,
                                      maxIter=self.maxCGLSIter, tol=self.CGLStol)
        torch.cuda.synchronize(self.dev)
        tC = time.perf_counter() - t
        modelNew = self._trans(3, self.model, dm, 1.0)
        t = time.perf_counter()
        responseNew = self._forward(modelNew)
        tau = self.linesearch(modelNew, responseNew) if self.lineSearch else 1.0
        if tau >= 0.95:
            self.response = responseNew
        else:
            modelNew = self._trans(3, self.model, dm, tau)
            self.response = self._forward(modelNew)
        torch.cuda.synchronize(self.dev)
        tF = time.perf_counter() - t
        self.model = modelNew
        self.tauHistory.append(tau)
        self.timings.append({"jacobian_s": tJ, "cgls_s": tC, "cgls_iterations": itC, "forward_linesearch_s": tF,
                             "step_s": time.perf_counter() - t0})
        return True

    def run(self, startModel):
        """RInversion::run (inversion.cpp:135-246); returns the model (device tensor)."""
        self.model = _dev(startModel, self.C.shape[1], self.dev, "startModel")
        # checkError: dataWeight = 1 / tD.error(fixZero(data), error) (inversion.h:241-263)
        fz = torch.where(torch.abs(self.data) < 1e-12, torch.full_like(self.data, 1e-12), self.data)
        self._derr = self._trans(4, fz, self.error, data=True)
        self.response = self._forward(self.model)
        self.fop.createJacobian(self.model)
        self.iter = 0
        self.chi2History, self.tauHistory, self.timings = [], [], []
        _, phiD, _ = self.phi(self.model, self.response)
        self.chi2History.append(phiD / self.data.numel())
        oldPhi = phiD
        while self.iter < self.maxIter:
            if not self.oneStep():
                break
            phi, phiD, _ = self.phi(self.model, self.response)
            self.chi2History.append(phiD / self.data.numel())
            if self.verbose:
                print(f"GN {self.iter}: chi2 {self.chi2History[-1]:.4f}, tau {self.tauHistory[-1]:.2f}, "
                      f"{self.timings[-1]}")
            if self.stopAtChi1 and phiD < self.data.numel():
                break
            if phi / oldPhi > (1.0 - self.dPhiAbortPercent / 100.0) and self.iter > 2:
                break
            oldPhi = phi
        return self.model

"""Test infrastructure only (tests/, bench.py's CPU baseline): numpy
restatement of the reference's Gauss-Newton inversion loop for ERT,
RInversion::run / oneStep / linesearch (core/src/inversion.cpp:135-370,
core/src/inversion.h:554-700) with the logarithmic transforms of
core/src/trans.h:241-300 / trans.cpp:24-63 (TransLog, TransLogLU), on the
oracle's forward operator and the oracle's CGLS (oracle/cgls.py).

Pinned against the reference's own RInversion (oracle/_ref, ref.Session.invert)
on the small test meshes: same model and response after the same iterations.
"""
from __future__ import annotations

import copy

import numpy as np

from . import cgls as cg
from . import ert_oracle as eo

TOLERANCE = 1e-12  # gimli.h TOLERANCE
TRANSTOL = 1e-8    # trans.h:27


class TransLogLU:
    """TransLogLU(lower, upper) (trans.h:292-322, trans.cpp:24-63); upper = 0
    is TransLog(lower) (trans.h:241-287).  Elementwise forms of the
    reference's capping (it caps every element below / above the bound when
    any element is)."""

    def __init__(self, lower=0.0, upper=0.0):
        self.lower = float(lower)
        self.upper = float(upper)

    def _rangify(self, a):
        v = np.maximum(a, self.lower * (1.0 + TRANSTOL))
        if abs(self.upper) >= TOLERANCE:
            v = np.minimum(v, self.upper * (1.0 - TRANSTOL))
        return v

    def trans(self, a):
        v = self._rangify(np.asarray(a, np.float64))
        if abs(self.upper) < TOLERANCE:
            return np.log(v - self.lower)
        return np.log(v - self.lower) - np.log(self.upper - v)

    def invTrans(self, f):
        f = np.asarray(f, np.float64)
        if abs(self.upper) < TOLERANCE:
            return np.exp(f) + self.lower
        e = np.exp(np.minimum(f, 50.0))
        return (e * self.upper + self.lower) / (e + 1.0)

    def deriv(self, a):
        v = self._rangify(np.asarray(a, np.float64))
        if abs(self.upper) < TOLERANCE:
            return 1.0 / (v - self.lower)
        return 1.0 / (v - self.lower) + 1.0 / (self.upper - v)

    def update(self, a, b):
        """Trans::update (trans.h:76-79): invTrans(trans(a) + b)."""
        return self.invTrans(self.trans(a) + b)

    def error(self, a, daBya):
        """Trans::error (trans.h:88-93): |a daBya deriv(a)|."""
        a = np.asarray(a, np.float64)
        if np.all(np.asarray(daBya) == 0.0):
            return np.ones_like(a)
        return np.abs(a * daBya * self.deriv(a))


def fix_zero(v, tol=TOLERANCE):
    """fixZero (vector.h): values with |v| < tol replaced by tol (sign kept)."""
    v = np.array(v, np.float64, copy=True)
    small = np.abs(v) < tol
    v[small] = np.where(v[small] < 0, -tol, tol)
    return v


class Inversion:
    """RInversion for an ERT operator with log data / log(-LU) model
    transforms, constraint matrix C (weights 1), fixed lambda, no reference
    model, no local regularisation, Jacobian recomputed every iteration.
    fop: oracle Forward (response(model), jacobian(model))."""

    def __init__(self, fop, data, err, C, lam=50.0, max_iter=20, line_search=True, max_cgls=200, cgls_tol=-1.0,
                 stop_at_chi1=True, dphi_abort_percent=2.0, m_lower=0.0, m_upper=0.0):
        self.fop = fop
        self.data = np.asarray(data, np.float64)
        self.err = np.asarray(err, np.float64)
        self.C = C
        self.lam = float(lam)
        self.max_iter = int(max_iter)
        self.line_search = bool(line_search)
        self.max_cgls = int(max_cgls)
        self.cgls_tol = float(cgls_tol)
        self.stop_at_chi1 = bool(stop_at_chi1)
        self.dphi = float(dphi_abort_percent)
        self.tD = TransLogLU(0.0, 0.0)
        self.tM = TransLogLU(m_lower, m_upper)
        self.mw = np.ones(C.shape[1])
        self.cw = np.ones(C.shape[0])
        # checkError (inversion.h:241-263)
        self.dweight = 1.0 / self.tD.error(fix_zero(self.data), self.err)
        self.chi2_history = []

    # getPhiD / getPhiM / getPhi (inversion.cpp:84-117, inversion.h:581-595)
    def roughness(self, model):
        return (self.C @ (self.tM.trans(model) * self.mw)) * self.cw

    def phiD(self, response):
        d = (self.tD.trans(self.data) - self.tD.trans(response)) / self.tD.error(fix_zero(self.data), self.err)
        return float(d @ d)

    def phiM(self, model):
        r = self.roughness(model)
        return float(r @ r)

    def phi(self, model, response):
        return self.phiD(response) + self.phiM(model) * self.lam

    def linesearch(self, model_new, response_new):
        """inversion.h:615-675: the objective at tau = 0.01 .. 1 along the
        transformed step (linearised response), the first minimum; a
        parabola through tau = 0, 0.3, 1 (one more forward run) when that is
        below 0.03; clamped to [0.03, 1]."""
        phi0 = self.phi(self.model, self.response)
        dModel = self.tM.trans(model_new) - self.tM.trans(self.model)
        dData = self.tD.trans(response_new) - self.tD.trans(self.response)
        min_phi, min_tau = phi0, 0.0
        for i in range(1, 101):
            tau = 0.01 * i
            app_model = self.tM.update(self.model, dModel * tau)
            app_resp = self.tD.update(self.response, dData * tau)
            p = self.phi(app_model, app_resp)
            if p < min_phi:
                min_phi, min_tau = p, tau
        tau = min_tau
        if tau < 0.03:
            tauquad = 0.3
            model_quad = self.tM.update(self.model, dModel * tauquad)
            resp_quad = self.fop.response(model_quad)
            phi10 = self.phi(model_new, response_new) - phi0
            phit0 = self.phi(model_quad, resp_quad) - phi0
            dphit = phit0 - phi10 * tauquad
            tau = 0.0 if abs(dphit) < TOLERANCE else (phit0 - phi10 * tauquad * tauquad) / dphit / 2.0
            if tau > 1.0:
                tau = 1.0
            if tau < 0.03:
                tau = 0.03
        return tau

    def one_step(self):
        """RInversion::oneStep (inversion.cpp:249-370)."""
        self.iter += 1
        ddata = self.tD.trans(self.data) - self.tD.trans(self.response)
        if np.sum(np.abs(ddata)) < TOLERANCE:
            return False
        if self.iter > 1:
            self.J = self.fop.jacobian(self.model)          # checkJacobian, recalcJacobian_
        dm, _, _ = cg.cgls(self.J, self.C, self.dweight, ddata, np.zeros(len(self.model)), self.cw, self.mw,
                           self.tM.deriv(self.model), self.tD.deriv(self.response), self.lam,
                           self.roughness(self.model), self.max_cgls, self.cgls_tol)
        model_new = self.tM.update(self.model, dm)
        response_new = self.fop.response(model_new)
        tau = self.linesearch(model_new, response_new) if self.line_search else 1.0
        if tau >= 0.95:
            self.response = response_new
        else:
            model_new = self.tM.update(self.model, dm * tau)
            self.response = self.fop.response(model_new)
        self.model = model_new
        self.tau_history.append(tau)
        return True

    def run(self, start_model):
        """RInversion::run (inversion.cpp:135-246)."""
        self.model = np.array(start_model, np.float64, copy=True)
        self.response = self.fop.response(self.model)
        self.J = self.fop.jacobian(self.model)
        self.iter = 0
        self.tau_history = []
        self.chi2_history = [self.phiD(self.response) / len(self.data)]
        old_phi = self.phiD(self.response)
        while self.iter < self.max_iter:
            if not self.one_step():
                break
            phiD = self.phiD(self.response)
            self.chi2_history.append(phiD / len(self.data))
            if self.stop_at_chi1 and phiD < len(self.data):
                break
            phi = self.phi(self.model, self.response)
            if phi / old_phi > (1.0 - self.dphi / 100.0) and self.iter > 2:
                break
            old_phi = phi
        return self.model


def forward(mesh, data):
    """The oracle's SR operator with response / jacobian for Inversion."""
    return eo.Forward(mesh, copy.deepcopy(data))

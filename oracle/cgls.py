"""Test infrastructure only (tests/, bench.py's CPU baseline): numpy
restatement of the Gauss-Newton step's regularised CGLS,
solveCGLSCDWWhtrans (core/src/solver.cpp:27-224), as called by
RInversion::oneStep (core/src/inversion.cpp:308).

Minimises, by conjugate gradients on the normal equations,
    || dW (b - td * S (x / tm)) ||^2 + lambda || wc (C (wm x)) + wc roughness ||^2
(x: model update, S: Jacobian, C: constraints; tm, td: model / data
transformation derivatives; dW: data weights; wc, wm: constraint / model
weights).  Pinned by `normal_equations`: the fixed point of the iteration is
the solution of
    (T S' Dt W Dt S T + lambda M C' Wc^2 C M) x = T S' Dt W b - lambda M C' Wc r
with T = diag(1/tm), Dt = diag(td), W = diag(dW^2), M = diag(wm), Wc = diag(wc).
"""
from __future__ import annotations

import numpy as np

TOLERANCE = 1e-12  # gimli.h TOLERANCE


def _round(v, tol):
    """RVector::round (vector.h:992-1000): multiples of tol."""
    return np.rint(v / tol) * tol


def cgls(S, C, dW, b, x, wc, wm, tm, td, lam, roughness, max_iter, tol):
    """solveCGLSCDWWhtrans (solver.cpp:27-224).  S: dense [nData, nModel]
    (or anything with @ and .T), C: scipy sparse [nConst, nModel].  Returns
    (x, iterations, normR2 history)."""
    x = np.array(x, np.float64, copy=True)
    b = np.asarray(b, np.float64)
    cdx = (C.T @ (wc * roughness)) * wm * lam                                   # :62
    z = (b - (S @ (x / tm)) * td) * dW                                          # :63
    p = (S.T @ (z * dW * td)) / tm - (C.T @ (wc * wc * (C @ (wm * x)))) * wm * lam - cdx   # :64-66
    r = (S.T @ (b * dW * dW * td)) / tm - cdx                                   # :67
    accuracy = tol
    if accuracy < 0.0:
        accuracy = max(TOLERANCE, 1e-8 * float(r @ r))                          # :98
    r = p.copy()                                                                 # :99
    normR2 = float(r @ r)
    hist = [normR2]
    count = 0
    while count < max_iter and normR2 > accuracy:                                # :121
        count += 1
        q = (S @ (p / tm)) * (dW * td)                                           # :123-124
        Cpwm = _round(C @ (p * wm), 1e-10)                                       # :131-135
        wcp = wc * Cpwm                                                          # :138
        aQ = float(q @ q) + lam * float(wcp @ wcp)                               # :159
        alpha = normR2 / aQ                                                      # :165
        x += p * alpha                                                           # :168
        z -= q * alpha                                                           # :175
        r = (S.T @ (z * dW * td)) / tm - (C.T @ (wc * wc * (C @ (wm * x)))) * wm * lam - cdx   # :179-185
        old = normR2
        normR2 = float(r @ r)                                                    # :190
        p = r + p * (normR2 / old)                                               # :191-194
        hist.append(normR2)
    return x, count, hist


def normal_equations(S, C, dW, b, wc, wm, tm, td, lam, roughness):
    """Direct solution of the normal equations the iteration converges to
    (dense, small problems only)."""
    S = np.asarray(S, np.float64)
    Cd = C.toarray() if hasattr(C, "toarray") else np.asarray(C, np.float64)
    A = (S / tm[None, :]) * td[:, None] * dW[:, None]          # Dw Dt S T
    Cm = (Cd * wm[None, :]) * wc[:, None]                       # Wc C M
    lhs = A.T @ A + lam * Cm.T @ Cm
    rhs = A.T @ (dW * b) - lam * (wm * (Cd.T @ (wc * roughness)))
    return np.linalg.solve(lhs, rhs)

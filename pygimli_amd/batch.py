"""Batched forward modelling of independent problems.

The reference evaluates problems one after another (each an
ERTModelling.response, ertModelling.py:204-219); here the block solves of
all problems run as one solve whose kernels cover every wavenumber block of
every problem (pgert_batch_forward).  Each ERTModelling keeps its own
potentials, so createJacobian() works per problem afterwards.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .ertModelling import _dptr


def batch_response_device(fops, models):
    """Apparent resistivities of every problem (float64 device tensors).
    All operators must use the same solver method/tolerance (block methods);
    fops[0]'s solverOptions are the batch's."""
    if len(fops) != len(models) or not fops:
        raise ValueError("one model per operator")
    if any(not f._sr or f.complex() for f in fops):
        raise NotImplementedError("batched forward: singularity-removal operators with real resistivities only")
    rhos, rhoas = [], []
    for f, m in zip(fops, models):
        f._sync_setup()
        t = f._model_tensor(m)
        if float(t.min()) < 1e-12:
            raise ValueError("response for model with negative or zero resistivity is not defined")
        rhos.append(f.mapERTModel(t))
        rhoas.append(torch.empty(f._data.size(), dtype=torch.float64, device=f.device))
        f._model = t
    opts = None
    for f in fops:                       # per-problem coarse set-up
        o = f._opts()
        opts = o if opts is None else opts
    fops[0]._set_stream()
    n = len(fops)
    vp = ctypes.c_void_p
    ctxs = (vp * n)(*[f._ctx.h for f in fops])
    rp = (vp * n)(*[_dptr(r) for r in rhos])
    ap = (vp * n)(*[_dptr(a) for a in rhoas])
    stats = (_lib.SolveStats * n)()
    L = fops[0]._ctx.L
    fops[0]._ctx.check(L.pgert_batch_forward(ctxs, n, rp, ctypes.byref(opts), ap, stats), "batch_response")
    for f, st in zip(fops, stats):
        f.lastStats = st.as_dict()
    return rhoas


def batch_jacobian(fops, models):
    """Jacobians of every problem after a (batched) forward of the same
    models, one launch per wavenumber pass over the cells of all problems
    (pgert_batch_jacobian).  Each fop.jacobian() holds its own afterwards;
    returns the list of Jacobians."""
    if len(fops) != len(models) or not fops:
        raise ValueError("one model per operator")
    if any(f.complex() for f in fops):
        raise NotImplementedError("batched Jacobian: real resistivities only")
    ts, Js = [], []
    for f, m in zip(fops, models):
        t = f._model_tensor(m)
        if f._model is None or f._model.shape != t.shape or not torch.equal(f._model, t):
            raise ValueError("batch_jacobian needs the forward of the same model first (batch_response)")
        ts.append(t)
        Js.append(torch.empty((f._data.size(), f._nParam), dtype=f.jacobianDtype, device=f.device))
    dt = {f.jacobianDtype for f in fops}
    if len(dt) != 1:
        raise ValueError("all operators need the same jacobianDtype")
    code = 0 if fops[0].jacobianDtype == torch.float64 else 1
    fops[0]._set_stream()
    n = len(fops)
    vp = ctypes.c_void_p
    ctxs = (vp * n)(*[f._ctx.h for f in fops])
    mp = (vp * n)(*[_dptr(t) for t in ts])
    nm = (ctypes.c_int * n)(*[t.numel() for t in ts])
    jp = (vp * n)(*[_dptr(J) for J in Js])
    L = fops[0]._ctx.L
    fops[0]._ctx.check(L.pgert_batch_jacobian(ctxs, n, mp, nm, jp, code), "batch_jacobian")
    for f, J in zip(fops, Js):
        f._J = J
    return Js


def batch_response(fops, models):
    """Numpy apparent resistivities of every problem."""
    return [r.cpu().numpy() for r in batch_response_device(fops, models)]

"""Learned warm start: training driver for the per-(node, source) MLP whose
forward, backward and Adam kernels live in libpgert (csrc/mlp.hip).

The network predicts the relative secondary potential u_s / (rho_s p) from 32
local features (position relative to the source, distances, the analytic
half-space primary, local log-resistivity averages at three smoothing scales,
the source resistivity).  Training data are the method's own converged solves
on random resistivity fields; the solver stops on the FP64 residual either
way, so the network changes iteration counts only.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

SIZES = (32, 128, 128, 128, 1)


def init_params(seed=0):
    """He-initialised weights, zero biases, packed FP32 (pgert_mlp_set_params layout)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fi, fo in zip(SIZES[:-1], SIZES[1:]):
        parts.append((rng.standard_normal((fi, fo)) * np.sqrt(2.0 / fi)).astype(np.float32).ravel())
        parts.append(np.zeros(fo, np.float32))
    p = np.concatenate(parts)
    p[-(128 + 1):-1] *= 0.1  # small head: start close to "no correction"
    return p


FORMAT = "pgert-mlp-warmstart-1"


def save_params(path, params, meta=None):
    """Write trained parameters (packed FP32, pgert_mlp_set_params layout) with
    the layer sizes and a format tag to an .npz file; `meta` adds scalar
    entries (training fields, epochs, loss, ...)."""
    params = np.ascontiguousarray(params, np.float32)
    n = sum(fi * fo + fo for fi, fo in zip(SIZES[:-1], SIZES[1:]))
    if params.size != n:
        raise ValueError(f"expected {n} parameters, got {params.size}")
    extra = {f"meta_{k}": np.asarray(v) for k, v in (meta or {}).items()}
    np.savez(path, format=np.asarray(FORMAT), sizes=np.asarray(SIZES, np.int32), params=params, **extra)


def load_params(path):
    """Read a file written by save_params: (params, meta dict).  Raises on a
    foreign format or other layer sizes."""
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z or str(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not a warm-start parameter file ({FORMAT})")
        if tuple(int(v) for v in z["sizes"]) != SIZES:
            raise ValueError(f"{path}: layer sizes {tuple(z['sizes'])} differ from {SIZES}")
        params = np.array(z["params"], np.float32)
        meta = {k[5:]: z[k][()] for k in z.files if k.startswith("meta_")}
    return params, meta


class LearnedWarmStart:
    """Owns the parameters on one ERTModelling's context."""

    def __init__(self, fop, params=None, seed=0):
        if not fop._sr or fop.complex():
            raise NotImplementedError("the learned warm start predicts singularity-removal (secondary) "
                                      "potentials: ERTModelling(sr=True) with real resistivities only")
        self.fop = fop
        L = fop._ctx.L
        self.n = L.pgert_mlp_param_count()
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if p.size != self.n:
            raise ValueError(f"expected {self.n} parameters")
        self.set_params(p)

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        ctx = self.fop._ctx
        ctx.check(ctx.L.pgert_mlp_set_params(ctx.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size),
                  "mlp_set_params")

    def params(self):
        p = np.empty(self.n, np.float32)
        ctx = self.fop._ctx
        ctx.check(ctx.L.pgert_mlp_get_params(ctx.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size),
                  "mlp_get_params")
        return p

    def save(self, path, **meta):
        """Checkpoint the current parameters (save_params)."""
        save_params(path, self.params(), meta)

    def load(self, path):
        """Restore parameters written by save(); returns the stored meta entries."""
        p, meta = load_params(path)
        self.set_params(p)
        return meta

    def collect(self, n_samples, rng, k_index=0):
        """Features/targets of random (node, source) samples of the last response()."""
        fop = self.fop
        nE = len(fop._analysis.elecNode)
        total = fop._mesh.nodeCount() * nE
        idx = torch.as_tensor(rng.choice(total, size=min(n_samples, total), replace=False).astype(np.int64),
                              device=fop.device)
        feat = torch.empty((idx.numel(), 32), dtype=torch.float32, device=fop.device)
        tgt = torch.empty(idx.numel(), dtype=torch.float32, device=fop.device)
        ctx = fop._ctx
        ctx.check(ctx.L.pgert_mlp_collect(ctx.h, k_index, ctypes.c_void_p(idx.data_ptr()), idx.numel(),
                                          ctypes.c_void_p(feat.data_ptr()), ctypes.c_void_p(tgt.data_ptr())),
                  "mlp_collect")
        return feat, tgt

    def step(self, feat, tgt, lr, apply=True, want_grad=False):
        ctx = self.fop._ctx
        loss = ctypes.c_float()
        g = np.empty(self.n, np.float32) if want_grad else None
        ctx.check(ctx.L.pgert_mlp_train_step(ctx.h, ctypes.c_void_p(feat.data_ptr()),
                                             ctypes.c_void_p(tgt.data_ptr()), feat.shape[0], float(lr),
                                             1 if apply else 0, ctypes.byref(loss),
                                             g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if want_grad else None),
                  "mlp_train_step")
        return loss.value, g

    def predict(self, feat):
        out = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
        ctx = self.fop._ctx
        ctx.check(ctx.L.pgert_mlp_predict(ctx.h, ctypes.c_void_p(feat.data_ptr()), feat.shape[0],
                                          ctypes.c_void_p(out.data_ptr())), "mlp_predict")
        return out

    def train(self, models, samples_per_model=65536, epochs=30, batch=4096, lr=1e-3, seed=0, log=None):
        """Solve every training model (no warm start), collect samples, then Adam."""
        rng = np.random.default_rng(seed)
        fop = self.fop
        learned = fop.solverOptions.get("learned", False)
        fop.solverOptions["learned"] = False
        feats, tgts = [], []
        for m in models:
            fop.response_device(m)
            f, t = self.collect(samples_per_model, rng)
            feats.append(f)
            tgts.append(t)
        fop.solverOptions["learned"] = learned
        F = torch.cat(feats)
        T = torch.cat(tgts)
        # clip extreme relative corrections (singular-node neighbourhoods)
        T = torch.clamp(T, -5.0, 5.0)
        nb = F.shape[0] // batch
        losses = []
        for ep in range(epochs):
            perm = torch.as_tensor(rng.permutation(F.shape[0]), device=F.device)
            tot = 0.0
            for b in range(nb):
                sel = perm[b * batch:(b + 1) * batch]
                loss, _ = self.step(F[sel].contiguous(), T[sel].contiguous(), lr)
                tot += loss
            losses.append(tot / max(nb, 1))
            if log:
                log(f"[mlp] epoch {ep} loss {losses[-1]:.4e}")
        return losses

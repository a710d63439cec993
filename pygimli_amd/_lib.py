"""ctypes binding of libpgert.so (C-ABI in include/pgert.h).

The product path has no CPU fallback: if the HIP library is missing or no GPU
is visible, every compute call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpgert.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)

PGERT_PCG_FP64 = 0
PGERT_BLOCK_FP64 = 1
PGERT_BLOCK_MIXED = 2

PGERT_BAND_FP64 = 3
METHODS = {"pcg64": PGERT_PCG_FP64, "block64": PGERT_BLOCK_FP64, "mixed": PGERT_BLOCK_MIXED,
           "band": PGERT_BAND_FP64}

# every symbol include/pgert.h declares
EXPORTS = (
    "pgert_create", "pgert_destroy", "pgert_last_error", "pgert_set_stream", "pgert_set_mesh",
    "pgert_set_electrodes", "pgert_set_wavenumbers", "pgert_set_data", "pgert_info",
    "pgert_pattern", "pgert_map_model", "pgert_assemble", "pgert_fill_matrix", "pgert_forward",
    "pgert_potentials", "pgert_potentials_node_major", "pgert_jacobian", "pgert_last_timings",
    "pgert_kernel_timing", "pgert_mlp_param_count", "pgert_mlp_set_params", "pgert_mlp_get_params",
    "pgert_mlp_collect", "pgert_mlp_predict", "pgert_mlp_train_step", "pgert_set_coarse",
    "pgert_batch_forward", "pgert_batch_jacobian", "pgert_set_constraints", "pgert_cgls", "pgert_cgls_noalloc",
    "pgert_forward_complex", "pgert_jacobian_complex", "pgert_forward_total", "pgert_set_bypass",
    "pgert_trans", "pgert_gn_phi", "pgert_set_electrodes_free", "pgert_set_mesh_coupled", "pgert_set_mesh_p2",
    "pgert_set_primary_potentials", "pgert_set_cem", "pgert_set_reference",
)


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("rtol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("check_every", ctypes.c_int), ("inner_rtol", ctypes.c_double),
                ("warm_start", ctypes.c_void_p), ("learned_warm_start", ctypes.c_int), ("coarse", ctypes.c_int),
                ("no_graph", ctypes.c_int), ("report_only", ctypes.c_int), ("smooth", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("max_iterations", ctypes.c_int),
                ("refinements", ctypes.c_int), ("converged", ctypes.c_int),
                ("max_rel_residual", ctypes.c_double), ("fallback", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def load():
    """Load libpgert.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libpgert.so not built ({LIB_PATH}); run python -m pygimli_amd.build")
    # torch's HIP runtime first: the library then binds to the runtime torch
    # loaded (loaded the other way round, torch finds no device)
    import torch
    torch.cuda.is_available()
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.pgert_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.pgert_destroy.argtypes = [vp]
    L.pgert_destroy.restype = None
    L.pgert_last_error.argtypes = [vp]
    L.pgert_last_error.restype = ctypes.c_char_p
    L.pgert_set_stream.argtypes = [vp, vp]
    L.pgert_set_mesh.argtypes = [vp, ctypes.c_int, ctypes.c_int, _c_double_p, ctypes.c_int, _c_int_p,
                                 _c_int_p, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]
    L.pgert_set_mesh_coupled.argtypes = L.pgert_set_mesh.argtypes + [ctypes.c_int, _c_int_p, _c_int_p]
    L.pgert_set_mesh_p2.argtypes = L.pgert_set_mesh.argtypes
    L.pgert_set_primary_potentials.argtypes = [vp, vp]
    L.pgert_set_cem.argtypes = [vp, ctypes.c_int, _c_int_p, _c_int_p, _c_double_p]
    L.pgert_set_reference.argtypes = [vp, ctypes.c_int, ctypes.c_int, _c_int_p]
    L.pgert_set_electrodes.argtypes = [vp, ctypes.c_int, _c_int_p, _c_double_p, ctypes.c_double]
    L.pgert_set_electrodes_free.argtypes = [vp, ctypes.c_int, _c_int_p, _c_int_p, _c_double_p, _c_double_p,
                                            ctypes.c_double]
    L.pgert_set_wavenumbers.argtypes = [vp, ctypes.c_int, _c_double_p, _c_double_p]
    L.pgert_set_data.argtypes = [vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_double_p]
    L.pgert_info.argtypes = [vp, _c_int_p, _c_int_p]
    L.pgert_pattern.argtypes = [vp, _c_int_p, _c_int_p]
    L.pgert_map_model.argtypes = [vp, vp, ctypes.c_int, vp]
    L.pgert_assemble.argtypes = [vp, vp, ctypes.c_double, vp]
    L.pgert_fill_matrix.argtypes = [vp, ctypes.c_int, vp, vp]
    L.pgert_forward.argtypes = [vp, vp, ctypes.POINTER(SolveOpts), vp, ctypes.POINTER(SolveStats)]
    L.pgert_batch_forward.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp),
                                      ctypes.POINTER(SolveOpts), ctypes.POINTER(vp), ctypes.POINTER(SolveStats)]
    L.pgert_batch_jacobian.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp), _c_int_p,
                                       ctypes.POINTER(vp), ctypes.c_int]
    L.pgert_set_constraints.argtypes = [vp, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_double_p]
    L.pgert_cgls.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp,
                             ctypes.c_double, vp, ctypes.c_int, ctypes.c_double, _c_int_p, _c_double_p]
    L.pgert_cgls_noalloc.argtypes = L.pgert_cgls.argtypes
    L.pgert_trans.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, vp]
    L.pgert_gn_phi.argtypes = [vp, ctypes.c_int, _c_double_p, ctypes.c_int, vp, vp, vp, vp, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int, vp, vp, vp, vp, _c_double_p, _c_double_p]
    L.pgert_set_bypass.argtypes = [vp, ctypes.c_int, _c_int_p, _c_int_p, _c_double_p]
    L.pgert_forward_total.argtypes = [vp, vp, ctypes.POINTER(SolveOpts), vp, ctypes.POINTER(SolveStats)]
    L.pgert_forward_complex.argtypes = [vp, vp, vp, ctypes.POINTER(SolveOpts), vp, vp, ctypes.POINTER(SolveStats)]
    L.pgert_jacobian_complex.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp]
    L.pgert_potentials.argtypes = [vp, vp]
    L.pgert_potentials_node_major.argtypes = [vp, vp]
    L.pgert_jacobian.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int]
    L.pgert_last_timings.argtypes = [vp, _c_double_p]
    L.pgert_kernel_timing.argtypes = [vp, ctypes.c_int, ctypes.c_int, _c_double_p,
                                      ctypes.POINTER(ctypes.c_long), _c_double_p]
    fp = ctypes.POINTER(ctypes.c_float)
    L.pgert_set_coarse.argtypes = [vp, ctypes.c_int]
    L.pgert_mlp_param_count.argtypes = []
    L.pgert_mlp_set_params.argtypes = [vp, fp, ctypes.c_int]
    L.pgert_mlp_get_params.argtypes = [vp, fp, ctypes.c_int]
    L.pgert_mlp_collect.argtypes = [vp, ctypes.c_int, vp, ctypes.c_long, vp, vp]
    L.pgert_mlp_predict.argtypes = [vp, vp, ctypes.c_int, vp]
    L.pgert_mlp_train_step.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, fp, fp]
    for name in EXPORTS:
        if name not in ("pgert_destroy", "pgert_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


class Context:
    """Owns one pgert_ctx on one device."""

    def __init__(self, device=0):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.pgert_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise RuntimeError(f"pgert_create failed ({rc}): no usable HIP device {device}")
        self.h = h
        self.device = device

    def check(self, rc, what=""):
        if rc != 0:
            msg = self.L.pgert_last_error(self.h).decode(errors="replace")
            raise RuntimeError(f"{what}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.pgert_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

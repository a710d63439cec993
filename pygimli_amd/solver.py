"""pg.solver assembly entry points on the GPU.

createStiffnessMatrix / createMassMatrix mirror pygimli/solver/solver.py:1908-2038
(RSparseMatrix::fillStiffnessMatrix / fillMassMatrix, core/src/sparsematrix.h:1039-1065):
P1 element matrices scaled by a per-cell coefficient, summed into the CSR
pattern of all node pairs sharing a cell (full symmetric storage, sorted
column indices).  The values are computed by libpgert's atomics-free
assembly kernel; the result is a scipy CSR matrix with the reference's
layout (indptr = vecColPtr, indices = vecRowIdx).
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib

_cache = {}


def _context_for(mesh, device=0):
    key = (id(mesh), device)
    ent = _cache.get(key)
    if ent is not None and ent[0] is mesh:
        return ent[1], ent[2], ent[3]
    from .ertModelling import _host_d, _host_i
    ctx = _lib.Context(device)
    xyz, pxyz = _host_d(mesh.positions())
    cells, pc = _host_i(mesh.cells)
    cm, pcm = _host_i(mesh.cellMarkers())
    b, pb = _host_i(mesh.bounds)
    bm, pbm = _host_i(mesh.boundaryMarkers())
    lc, plc = _host_i(mesh.boundaryLeftCell())
    # the fill matrices ignore boundary conditions; pass no boundaries
    ctx.check(ctx.L.pgert_set_mesh(ctx.h, mesh.dim(), mesh.nodeCount(), pxyz, mesh.cellCount(), pc, pcm,
                                   0, pb, pbm, plc), "createStiffnessMatrix")
    nnz = ctypes.c_int()
    ctx.L.pgert_info(ctx.h, ctypes.byref(nnz), None)
    n = mesh.nodeCount()
    rowptr = np.empty(n + 1, np.int32)
    colidx = np.empty(nnz.value, np.int32)
    ctx.check(ctx.L.pgert_pattern(ctx.h, rowptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  colidx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "pattern")
    _cache.clear()
    _cache[key] = (mesh, ctx, rowptr, colidx)
    return ctx, rowptr, colidx


def _fill(mesh, a, mass, device=0):
    if mesh.cellCount() == 0:
        raise Exception("Mesh invalid")
    if a is None:
        a = 1.0
    a = np.asarray(a)
    if np.iscomplexobj(a):
        raise NotImplementedError("complex cell values: the GPU fill is real-valued")
    if a.ndim == 1 and len(a) != mesh.cellCount():
        raise ValueError("Number of cell values need to match cell count")
    a = np.broadcast_to(a.astype(np.float64), (mesh.cellCount(),))
    ctx, rowptr, colidx = _context_for(mesh, device)
    dev = torch.device("cuda", device)
    at = torch.as_tensor(np.array(a, np.float64), device=dev)
    vals = torch.empty(len(colidx), dtype=torch.float64, device=dev)
    ctx.check(ctx.L.pgert_fill_matrix(ctx.h, int(mass), ctypes.c_void_p(at.data_ptr()),
                                      ctypes.c_void_p(vals.data_ptr())), "fill_matrix")
    n = mesh.nodeCount()
    return sp.csr_matrix((vals.cpu().numpy(), colidx.copy(), rowptr.copy()), shape=(n, n))


def createStiffnessMatrix(mesh, a=None):
    """Stiffness matrix sum_c a_c int grad N_i . grad N_j (solver.py:1908-1997)."""
    return _fill(mesh, a, mass=False)


def createMassMatrix(mesh, b=None):
    """Mass matrix sum_c b_c int N_i N_j (solver.py:2000-2038)."""
    return _fill(mesh, b, mass=True)

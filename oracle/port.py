"""Test infrastructure only: ctypes access to oracle/_build/liboracle.so (the
C calc2 port, CPU baseline of bench.py)."""
import ctypes
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")


def calc2(mesh, U, kvals, weights, abmn, nParam, nthreads=1, cell_range=None, J=None, markers=None):
    L = ctypes.CDLL(LIB)
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    xyz = np.ascontiguousarray(mesh.nodes, np.float64)
    cells = np.ascontiguousarray(mesh.cells, np.int32)
    mk = np.ascontiguousarray(mesh.cellMarkers() if markers is None else markers, np.int32)
    U = np.ascontiguousarray(U, np.float64)
    kv = np.ascontiguousarray(kvals, np.float64)
    wv = np.ascontiguousarray(weights, np.float64)
    ab = np.ascontiguousarray(abmn, np.int32)
    if J is None:
        J = np.zeros((len(ab), nParam))
    c0, c1 = (0, mesh.cellCount()) if cell_range is None else cell_range
    nE = U.shape[0] // len(kv)
    L.oracle_calc2.argtypes = [ctypes.c_int, ctypes.c_int, D, ctypes.c_long, I, I, ctypes.c_int, ctypes.c_int, D, D,
                               D, ctypes.c_int, I, ctypes.c_int, D, ctypes.c_int, ctypes.c_long, ctypes.c_long]
    L.oracle_calc2(mesh.dim(), mesh.nodeCount(), xyz.ctypes.data_as(D), mesh.cellCount(), cells.ctypes.data_as(I),
                   mk.ctypes.data_as(I), nE, len(kv), kv.ctypes.data_as(D), wv.ctypes.data_as(D),
                   U.ctypes.data_as(D), len(ab), ab.ctypes.data_as(I), nParam, J.ctypes.data_as(D), int(nthreads),
                   int(c0), int(c1))
    return J


def pcg(A, B, maxit, rtol=0.0, nthreads=1):
    """Plain FP64 Jacobi-PCG per column of B (C port, threads over columns).
    A: scipy CSR; B: [n, nrhs].  Returns (X, iterations run)."""
    L = ctypes.CDLL(LIB)
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    A = A.tocsr()
    rp = np.ascontiguousarray(A.indptr, np.int32)
    ci = np.ascontiguousarray(A.indices, np.int32)
    va = np.ascontiguousarray(A.data, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    X = np.zeros_like(B)
    L.oracle_pcg.argtypes = [ctypes.c_int, I, I, D, ctypes.c_int, D, D, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    it = L.oracle_pcg(A.shape[0], rp.ctypes.data_as(I), ci.ctypes.data_as(I), va.ctypes.data_as(D), B.shape[1],
                      B.ctypes.data_as(D), X.ctypes.data_as(D), int(maxit), float(rtol), int(nthreads))
    return X, it

"""Test infrastructure only: ctypes access to oracle/_ref/libgimli_ref.so, the
reference's own C++ ERT core built in place from the pyGIMLi sources by
oracle/Makefile (see oracle/ref_driver.cpp for the entry points).

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this,
and only as a checker / baseline.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libgimli_ref.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"reference library not built: {LIB_PATH} (make -C oracle ref)")
        L = ctypes.CDLL(LIB_PATH)
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_bessel.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int]
        L.ref_gauss_legendre.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _D, _D]
        L.ref_gauss_laguerre.argtypes = [ctypes.c_int, _D, _D]
        L.ref_kwave.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D, _D]
        L.ref_exact_dc.argtypes = [_D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ref_exact_dc.restype = ctypes.c_double
        L.ref_session_create.restype = ctypes.c_void_p
        L.ref_session_create.argtypes = [ctypes.c_int, ctypes.c_int, _D, _I, ctypes.c_int, _I, _I,
                                         ctypes.c_int, _I, _I]
        L.ref_session_destroy.argtypes = [ctypes.c_void_p]
        L.ref_session_set_data.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int,
                                           _I, _I, _I, _I, _D]
        L.ref_session_init_fop.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.c_int]
        L.ref_session_nk.argtypes = [ctypes.c_void_p]
        L.ref_session_primary_p2.argtypes = [ctypes.c_void_p, _D]
        L.ref_session_kw.argtypes = [ctypes.c_void_p, _D, _D]
        L.ref_session_geometric_factors.argtypes = [ctypes.c_void_p, _D]
        L.ref_session_response.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D]
        L.ref_session_subpot_shape.argtypes = [ctypes.c_void_p, _I, _I]
        L.ref_session_subpots.argtypes = [ctypes.c_void_p, _D]
        L.ref_session_data_k.argtypes = [ctypes.c_void_p, _D]
        L.ref_session_jacobian.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D, _I, _I]
        L.ref_session_set_subpots.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]
        L.ref_session_invert.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.c_int, _D, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _I, _I, _D, _D, _D, _D, _I]
        L.ref_assemble.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, _D, _I, _I, _I, _D]
        L.ref_fill_matrix.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _I, _I, _I, _D]
        L.ref_mixed_bc.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, _D]
        L.ref_boundary_count.argtypes = [ctypes.c_void_p]
        L.ref_boundary_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _I]
        L.ref_session_init_fop_total.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int]
        L.ref_session_response_complex.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_int, _D, _D]
        L.ref_session_jacobian_complex.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_int, _D, _D, _I, _I]
        L.ref_mesh_convert.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.ref_session_save_mesh.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.ref_data_convert.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                       ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(_D)


def _i(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(_I)


def _check(rc):
    if rc != 0:
        raise RuntimeError("reference: " + lib().ref_last_error().decode())


def bessel(which, x):
    """which in {'i0','i1','k0','k1'} -> reference A&S polynomial values."""
    code = {"i0": 0, "i1": 1, "k0": 2, "k1": 3}[which]
    x, px = _d(np.atleast_1d(x))
    y = np.empty_like(x)
    lib().ref_bessel(code, px, y.ctypes.data_as(_D), len(x))
    return y


def gauss_legendre(x1, x2, n):
    x = np.empty(n); w = np.empty(n)
    lib().ref_gauss_legendre(x1, x2, n, x.ctypes.data_as(_D), w.ctypes.data_as(_D))
    return x, w


def gauss_laguerre(n):
    x = np.empty(n); w = np.empty(n)
    lib().ref_gauss_laguerre(n, x.ctypes.data_as(_D), w.ctypes.data_as(_D))
    return x, w


def kwave(rmin, rmax, nleg, nlag):
    k = np.empty(nleg + nlag); w = np.empty(nleg + nlag)
    lib().ref_kwave(rmin, rmax, nleg, nlag, k.ctypes.data_as(_D), w.ctypes.data_as(_D))
    return k, w


def exact_dc(v, src, k=0.0, surfaceZ=0.0, fallback=1.0):
    v, pv = _d(v); s, ps = _d(src)
    return lib().ref_exact_dc(pv, ps, k, surfaceZ, fallback)


def mesh_convert(src, dst):
    """Mesh::load(src, no neighbour infos) then Mesh::save(dst) (binary v3)."""
    _check(lib().ref_mesh_convert(str(src).encode(), str(dst).encode()))


def data_convert(src, dst, fmt_data="all", fmt_sensor="x y z", remove_invalid=True,
                 no_filter=False):
    """DataContainerERT(src, removeInvalid) then save(dst, fmt_data, fmt_sensor,
    noFilter); the load runs in dst's folder (it may write invalid.data there)."""
    import os
    work = os.path.dirname(os.path.abspath(str(dst)))
    _check(lib().ref_data_convert(str(src).encode(), str(dst).encode(), fmt_data.encode(),
                                  fmt_sensor.encode(), int(remove_invalid), int(no_filter),
                                  work.encode()))


class Session:
    """The reference DCSRMultiElectrodeModelling on one mesh + data set."""

    def __init__(self, mesh, data=None, kvalues=None, weights=None, nthreads=0, complex_=False, sr=True):
        """sr=False: the total-field DCMultiElectrodeModelling (complex_=True
        implies it)."""
        L = lib()
        self._keep = []
        nodes, pn = _d(mesh.nodes)
        nm, pnm = _i(mesh.nodeMarkers())
        cells, pc = _i(mesh.cells)
        cm, pcm = _i(mesh.cellMarkers())
        b, pb = _i(mesh.bounds)
        bm, pbm = _i(mesh.boundaryMarkers())
        self.h = L.ref_session_create(mesh.dim(), mesh.nodeCount(), pn, pnm, mesh.cellCount(),
                                      pc, pcm, mesh.boundaryCount(), pb, pbm)
        if not self.h:
            _check(-1)
        self.mesh = mesh
        self.data = data
        if data is not None:
            s, ps = _d(data.sensorPositions())
            a, pa = _i(data["a"]); bb, pbb = _i(data["b"])
            m, pm = _i(data["m"]); n, pnn = _i(data["n"])
            k = data["k"] if data.haveData("k") else None
            pk = None
            if k is not None:
                k, pk = _d(k)
            _check(L.ref_session_set_data(self.h, data.sensorCount(), ps, data.size(),
                                          pa, pbb, pm, pnn, pk))
            if complex_ or not sr:
                def init(h, nk, pk, pw, nt):
                    return L.ref_session_init_fop_total(h, nk, pk, pw, nt, 1 if complex_ else 0)
            else:
                init = L.ref_session_init_fop
            if kvalues is not None:
                kv, pkv = _d(kvalues); wv, pwv = _d(weights)
                _check(init(self.h, len(kv), pkv, pwv, nthreads))
            else:
                _check(init(self.h, -1, None, None, nthreads))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ref_session_destroy(self.h)
        except Exception:
            pass

    def kw(self):
        nk = lib().ref_session_nk(self.h)
        k = np.empty(nk); w = np.empty(nk)
        lib().ref_session_kw(self.h, k.ctypes.data_as(_D), w.ctypes.data_as(_D))
        return k, w

    def primary_p2(self):
        """Numerical primary potentials of the SR operator with topography
        (checkPrimpotentials_, dcfemmodelling.cpp:2009-2056), installed in the
        session; returns them, [nK * nE][n]."""
        nk = lib().ref_session_nk(self.h)
        out = np.empty((nk * self.data.sensorCount(), self.mesh.nodeCount()))
        _check(lib().ref_session_primary_p2(self.h, out.ctypes.data_as(_D)))
        return out

    def geometric_factors(self):
        out = np.empty(self.data.size())
        _check(lib().ref_session_geometric_factors(self.h, out.ctypes.data_as(_D)))
        return out

    def response(self, model):
        model, pm = _d(model)
        out = np.empty(self.data.size())
        _check(lib().ref_session_response(self.h, pm, len(model), out.ctypes.data_as(_D)))
        return out

    def response_complex(self, re, im):
        """Complex response (setComplex(True) session): (real, imag) of k * U."""
        re, pr = _d(re); im, pi = _d(im)
        out_re = np.empty(self.data.size()); out_im = np.empty(self.data.size())
        _check(lib().ref_session_response_complex(self.h, pr, pi, len(re), out_re.ctypes.data_as(_D),
                                                  out_im.ctypes.data_as(_D)))
        return out_re, out_im

    def jacobian_complex(self, re, im):
        re, pr = _d(re); im, pi = _d(im)
        r = ctypes.c_int(); c = ctypes.c_int()
        _check(lib().ref_session_jacobian_complex(self.h, pr, pi, len(re), None, None,
                                                  ctypes.byref(r), ctypes.byref(c)))
        Jr = np.empty((r.value, c.value)); Ji = np.empty((r.value, c.value))
        _check(lib().ref_session_jacobian_complex(self.h, pr, pi, len(re), Jr.ctypes.data_as(_D),
                                                  Ji.ctypes.data_as(_D), ctypes.byref(r), ctypes.byref(c)))
        return Jr, Ji

    def data_k(self):
        out = np.empty(self.data.size())
        lib().ref_session_data_k(self.h, out.ctypes.data_as(_D))
        return out

    def subpotentials(self):
        r = ctypes.c_int(); c = ctypes.c_int()
        lib().ref_session_subpot_shape(self.h, ctypes.byref(r), ctypes.byref(c))
        out = np.empty((r.value, c.value))
        lib().ref_session_subpots(self.h, out.ctypes.data_as(_D))
        return out

    def invert(self, data, err, start, C, lam=20.0, max_iter=3, line_search=True, max_cgls=200, cgls_tol=-1.0,
               m_lower=0.0, m_upper=0.0):
        """The reference's RInversion::run on this operator (see ref_driver.cpp):
        returns (model, response, chi2, iterations)."""
        import scipy.sparse as sp
        C = sp.csr_matrix(C)
        d, pd = _d(data); e, pe = _d(err); m0, pm = _d(start)
        rp, prp = _i(C.indptr); ci, pci = _i(C.indices); cv, pcv = _d(C.data)
        out_m = np.empty(len(m0)); out_r = np.empty(len(d)); chi = np.empty(1)
        it = ctypes.c_int()
        _check(lib().ref_session_invert(self.h, len(d), pd, pe, len(m0), pm, lam, max_iter, 1 if line_search else 0,
                                        max_cgls, cgls_tol, m_lower, m_upper, C.shape[0], prp, pci, pcv,
                                        out_m.ctypes.data_as(_D), out_r.ctypes.data_as(_D),
                                        chi.ctypes.data_as(_D), ctypes.byref(it)))
        return out_m, out_r, float(chi[0]), it.value

    def set_subpotentials(self, U):
        """Hand the core per-k potentials [nK * nE, n] (rows source + nE * kIdx)
        for createJacobian, as a previous response() would have left them."""
        U = np.ascontiguousarray(U, np.float64)
        _check(lib().ref_session_set_subpots(self.h, U.shape[0], U.shape[1], U.ctypes.data_as(_D)))

    def jacobian_only(self, model):
        """createJacobian(model) without copying J out; returns its shape."""
        model, pm = _d(model)
        r = ctypes.c_int(); c = ctypes.c_int()
        _check(lib().ref_session_jacobian(self.h, pm, len(model), None, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def jacobian(self, model):
        model, pm = _d(model)
        r = ctypes.c_int(); c = ctypes.c_int()
        _check(lib().ref_session_jacobian(self.h, pm, len(model), None,
                                          ctypes.byref(r), ctypes.byref(c)))
        out = np.empty((r.value, c.value))
        _check(lib().ref_session_jacobian(self.h, pm, len(model), out.ctypes.data_as(_D),
                                          ctypes.byref(r), ctypes.byref(c)))
        return out

    def save_mesh(self, path):
        """The reference mesh (with every cell face as a boundary) as .bms."""
        _check(lib().ref_session_save_mesh(self.h, str(path).encode()))

    def _csr(self, fn, *args):
        nnz = ctypes.c_int()
        _check(fn(self.h, *args, ctypes.byref(nnz), None, None, None))
        n = self.mesh.nodeCount()
        rp = np.empty(n + 1, np.int32); ci = np.empty(nnz.value, np.int32)
        v = np.empty(nnz.value)
        _check(fn(self.h, *args, ctypes.byref(nnz), rp.ctypes.data_as(_I),
                  ci.ctypes.data_as(_I), v.ctypes.data_as(_D)))
        import scipy.sparse as sp
        return sp.csr_matrix((v, ci, rp), shape=(n, n))

    def assemble(self, rho, k, source_pos):
        rho, pr = _d(rho); s, ps = _d(source_pos)
        return self._csr(lib().ref_assemble, pr, ctypes.c_double(k), ps)

    def fill_matrix(self, a, mass=False):
        a, pa = _d(a)
        return self._csr(lib().ref_fill_matrix, int(mass), pa)

    def mixed_bc(self, source_pos, k):
        nb = lib().ref_boundary_count(self.h)
        s, ps = _d(source_pos)
        out = np.empty(nb)
        _check(lib().ref_mixed_bc(self.h, ps, k, out.ctypes.data_as(_D)))
        return out

    def boundaries(self):
        nb = lib().ref_boundary_count(self.h)
        d = self.mesh.dim()
        ids = np.empty((nb, d), np.int32); mk = np.empty(nb, np.int32)
        tmp = np.empty(d, np.int32); m = ctypes.c_int()
        for i in range(nb):
            lib().ref_boundary_nodes(self.h, i, tmp.ctypes.data_as(_I), ctypes.byref(m))
            ids[i] = tmp; mk[i] = m.value
        return ids, mk

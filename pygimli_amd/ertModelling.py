"""ERT forward operator with pyGIMLi's ERTModelling API on MI355X.

Mirrors pygimli/physics/ert/ertModelling.py:73-246 (ERTModelling): the
reference wraps the C++ DCSRMultiElectrodeModelling core
(core/src/bert/dcfemmodelling.cpp); here the core is libpgert.so (HIP kernels
behind the C-ABI of include/pgert.h).  Device memory is held in torch
tensors; all arithmetic runs in the library's kernels.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _lib
from .bert import (MeshAnalysis, bypass_marker_resistances, geometric_factors, kwave_list, load_electrode_vector,
                   read_bypass_map, uncoupled_pairs)
from .mesh import p2_refinement
from .data import DataContainerERT

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _host_d(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(_D)


def _host_i(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(_I)


class ERTModelling:
    """Forward operator for Electrical Resistivity Tomography.

    Parameters
    ----------
    sr : bool
        Singularity removal (the reference default).  sr=False gives the
        total-field operator (DCMultiElectrodeModelling), also the one for
        complex resistivities (setComplex(True)).
    verbose : bool
        Print solver statistics.
    device : int
        HIP device index.
    """

    def __init__(self, sr=True, verbose=False, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("ERTModelling needs a HIP device (no GPU visible)")
        self.verbose = verbose
        self._sr = bool(sr)
        self._bypassR = None
        # files the total-field operator reads from the working directory at
        # every assembly, like the reference (byPassFile_, dcfemmodelling.cpp:565;
        # :672-674); read here whenever they change
        self.byPassFile = "bypass.map"
        self.bypassResistanceFile = "electrodeBypassResistances.map"
        self.contactImpedanceFile = "contactImpedance.map"
        self.contactResistanceFile = "contactResistance.map"
        self._contactZ = None
        self._bypassSig = None
        self._couplings = ()
        self._extra = None         # positions of the electrode unknowns (complete electrode model)
        self._nLib = 0             # nodes the library holds: mesh nodes + electrode unknowns
        self._prim = None          # numerical primary potentials (topography with SR), [nK][n][nE]
        self._primKey = None
        self._complex = False
        self._Jc = None
        self.device = torch.device("cuda", device)
        self._ctx = _lib.Context(device)
        self._mesh = None
        self._data = None
        self._kv = None
        self._wv = None
        self._kUser = False
        self._analysis = None
        self._J = None
        self._model = None
        self.jacobianDtype = torch.float64
        # report_only=False: a solve that misses rtol is repeated from the analytic
        # start, then in FP64, and raises if that also misses; True: lastStats
        # carries converged = 0 instead
        self.solverOptions = dict(method="mixed", rtol=1e-6, max_iter=20000, check_every=8,
                                  inner_rtol=1e-3, learned=False, coarse=0, smooth=2, graph=True,
                                  report_only=False)
        self._coarseSet = None
        self.lastStats = None
        self.warmStart = None
        self._dirty = True

    # ------------------------------------------------------------------ set-up
    def setComplex(self, c):
        """Complex resistivity (DCMultiElectrodeModelling::setComplex,
        dcfemmodelling.cpp:573-582); needs the total-field operator (sr=False),
        as in the reference, whose SR operator has no complex branch."""
        c = bool(c)
        if c and self._sr:
            raise NotImplementedError("complex resistivity needs ERTModelling(sr=False)")
        if c != self._complex:
            self._model = None
            self._J = None
            self._Jc = None
        self._complex = c

    def complex(self):
        return self._complex

    def setBypassResistances(self, resistances):
        """Resistance (Ohm) of the bypass resistors per node marker <= -10000
        (default 1e-6), in place of the electrodeBypassResistances.map file
        that is read otherwise.  Bypass resistors, these and the rows of a
        bypass.map file, act in the total-field operator (sr=False), as in the
        reference."""
        self._bypassR = {int(k): float(v) for k, v in dict(resistances).items()}
        if self._mesh is not None and self._data is not None:
            self._setup_electrodes()

    def setContactImpedances(self, zi):
        """Contact impedance (Ohm m^2) of every electrode of the complete
        electrode model (DCMultiElectrodeModelling::setContactImpedances,
        dcfemmodelling.cpp:776-778), in place of the contactImpedance.map file
        that is read otherwise (default 1 Ohm m^2)."""
        self._contactZ = np.asarray(zi, np.float64).copy()
        if self._mesh is not None and self._data is not None:
            self._setup_electrodes()

    def setVerbose(self, v):
        self.verbose = bool(v)

    def setMesh(self, mesh, ignoreRegionManager=False):
        """ModellingBase::setMesh (core/src/modellingbase.cpp:179-211)."""
        self._mesh = mesh
        self._upload_mesh(())
        if self._data is not None:
            self._setup_electrodes()

    def _upload_mesh(self, couplings, extra=None):
        """Mesh to the library; `couplings`: node pairs sharing no cell that
        need matrix slots (bypass resistors between distant nodes, electrode
        unknowns); `extra`: positions of nodes of no cell appended after the
        mesh nodes (the electrode unknowns of the complete electrode model)."""
        mesh = self._mesh
        ctx = self._ctx
        pos = mesh.positions() if extra is None else np.vstack([mesh.positions(), extra])
        xyz, pxyz = _host_d(pos)
        cells, pc = _host_i(mesh.cells)
        cm, pcm = _host_i(mesh.cellMarkers())
        b, pb = _host_i(mesh.bounds)
        bm, pbm = _host_i(mesh.boundaryMarkers())
        lc, plc = _host_i(mesh.boundaryLeftCell())
        self._set_stream()
        if couplings:
            ca, pca = _host_i([p[0] for p in couplings])
            cb, pcb = _host_i([p[1] for p in couplings])
            ctx.check(ctx.L.pgert_set_mesh_coupled(ctx.h, mesh.dim(), len(pos), pxyz, mesh.cellCount(), pc,
                                                   pcm, mesh.boundaryCount(), pb, pbm, plc, len(ca), pca, pcb),
                      "setMesh")
        else:
            ctx.check(ctx.L.pgert_set_mesh(ctx.h, mesh.dim(), mesh.nodeCount(), pxyz, mesh.cellCount(), pc, pcm,
                                           mesh.boundaryCount(), pb, pbm, plc), "setMesh")
        self._couplings = tuple(couplings)
        self._extra = None if extra is None else np.array(extra, np.float64)
        self._nLib = len(pos)
        nnz = ctypes.c_int()
        npar = ctypes.c_int()
        ctx.L.pgert_info(ctx.h, ctypes.byref(nnz), ctypes.byref(npar))
        self._nnz = nnz.value
        self._nParam = npar.value
        self._coarseSet = None
        self._analysis = None
        self._model = None
        self._dirty = True

    def mesh(self):
        return self._mesh

    def setData(self, data):
        self._data = data
        self._dirty = True
        if self._mesh is not None:
            self._setup_electrodes()

    def data(self):
        return self._data

    def setkValues(self, k):
        self._kv = np.asarray(k, np.float64)
        self._kUser = True
        self._dirty = True

    def setWeights(self, w):
        self._wv = np.asarray(w, np.float64)
        self._kUser = True
        self._dirty = True

    def kValues(self):
        return self._kv

    def weights(self):
        return self._wv

    def _set_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        self._ctx.check(self._ctx.L.pgert_set_stream(self._ctx.h, ctypes.c_void_p(s)), "set_stream")

    def _setup_electrodes(self):
        mesh, data = self._mesh, self._data
        an = MeshAnalysis(mesh, data.sensorPositions())
        if an.neumannDomain and an.refNode < 0:
            raise NotImplementedError("a 3D Neumann-only domain without a reference electrode node (marker -999): "
                                      "the reference then takes the last electrode, and its Jacobian reads past its "
                                      "potential rows")
        if an.topography and self._sr:
            # numerical primaries from the P2 total-field problem; its
            # operator would also carry free electrodes (P2 shape functions)
            # and bypass resistors (on the P2 nodes), which are not built
            if np.any(an.elecNode < 0) or an.bypassGroups or os.path.isfile(self.byPassFile):
                raise NotImplementedError("singularity removal with topography: numerical primary potentials are "
                                          "built for node electrodes without bypass resistors")
        # bypass resistors: bypass.map rows, then the node groups
        # (assembleStiffnessMatrixDCFEMByPass_, dcfemmodelling.cpp:601-707);
        # total-field operator only
        self._bypassSig = self._bypass_file_sig()
        ba, bb, bg = an.bypass_pairs(self._bypass_resistances())
        if not self._sr and os.path.isfile(self.byPassFile):
            fa, fb, fr = read_bypass_map(self.byPassFile, an.elecNode, mesh.nodeMarkers())
            ba, bb, bg = np.concatenate([fa, ba]), np.concatenate([fb, bb]), np.concatenate([1.0 / fr, bg])
        cem = None
        if an.cemFaces:
            # complete electrode model (total field only, as in the reference,
            # whose SR operator cannot read electrode potentials from it)
            if self._sr:
                raise NotImplementedError("the complete electrode model needs ERTModelling(sr=False)")
            nz = len(an.elecNode) + (an.cemPassive is not None)   # the passive body after the electrodes
            zi = self._contactZ
            if zi is not None and len(zi) < nz:
                zi = np.concatenate([zi, np.ones(nz - len(zi))])
            if zi is None and os.path.isfile(self.contactImpedanceFile):
                zi = load_electrode_vector(self.contactImpedanceFile, nz, 1.0)
            zr = None
            if os.path.isfile(self.contactResistanceFile):
                zr = load_electrode_vector(self.contactResistanceFile, nz, 1.0)
            cem = an.cem_terms(mesh, zi, zr)
        if not self._sr:
            need = set(uncoupled_pairs(mesh.cells, ba, bb)) | set(cem[3] if cem else ())
            extra = None
            if cem:
                extra = an.elecPos if an.cemPassive is None else np.vstack([an.elecPos, an.cemPassivePos])
            same_extra = (extra is None and self._extra is None) or (
                extra is not None and self._extra is not None and np.array_equal(extra, self._extra))
            if not need <= set(self._couplings) or not same_extra:
                keep = set(self._couplings) if same_extra else set()
                self._upload_mesh(tuple(sorted(need | keep)), extra)
        self._analysis = an
        en, pen = _host_i(an.elecNode)
        sc, psc = _host_d(an.sourceCenter)
        ctx = self._ctx
        if np.any(an.elecNode < 0):
            ec, pec = _host_i(an.elecCell)
            ep, pep = _host_d(an.elecPos)
            ctx.check(ctx.L.pgert_set_electrodes_free(ctx.h, len(en), pen, pec, pep, psc, an.surfaceZ),
                      "setElectrodes")
        else:
            ctx.check(ctx.L.pgert_set_electrodes(ctx.h, len(en), pen, psc, an.surfaceZ), "setElectrodes")
        ba, pba = _host_i(ba)
        bb, pbb = _host_i(bb)
        bg, pbg = _host_d(bg)
        ctx.check(ctx.L.pgert_set_bypass(ctx.h, len(ba), pba, pbb, pbg), "bypass")
        cal, pcal = _host_i(an.calibration)
        ctx.check(ctx.L.pgert_set_reference(ctx.h, an.refNode, len(cal), pcal), "reference electrode")
        if cem:
            cr, pcr = _host_i(cem[0])
            cc, pcc = _host_i(cem[1])
            cv, pcv = _host_d(cem[2])
            ctx.check(ctx.L.pgert_set_cem(ctx.h, len(cr), pcr, pcc, pcv), "complete electrode model")
        if not self._kUser or self._kv is None:
            self._kv, self._wv = kwave_list(mesh.dim(), an.elecPos)
        self._dirty = True

    def _bypass_file_sig(self):
        """State of the two bypass files the total-field operator reads."""
        if self._sr:
            return None
        sig = []
        for f in (self.byPassFile, self.bypassResistanceFile, self.contactImpedanceFile, self.contactResistanceFile):
            try:
                st = os.stat(f)
                sig.append((os.path.abspath(f), st.st_mtime_ns, st.st_size))
            except OSError:
                sig.append(None)
        return tuple(sig)

    def _bypass_resistances(self):
        if self._bypassR is not None:
            return self._bypassR
        if not self._sr and os.path.isfile(self.bypassResistanceFile):
            return bypass_marker_resistances(self.bypassResistanceFile)
        return None

    def _sync_setup(self):
        if self._mesh is None or self._data is None:
            raise RuntimeError("mesh and data must be set")
        if self._analysis is not None and self._bypass_file_sig() != self._bypassSig:
            self._analysis = None
            self._model = None
            self._dirty = True
        if not self._dirty:
            return
        if self._analysis is None:
            self._setup_electrodes()
        ctx = self._ctx
        k, pk = _host_d(self._kv)
        w, pw = _host_d(self._wv)
        if len(k) != len(w):
            raise ValueError("kValues and weights differ in length")
        ctx.check(ctx.L.pgert_set_wavenumbers(ctx.h, len(k), pk, pw), "setWavenumbers")
        self._ensure_k()
        self._upload_data()
        if self._sr and self._analysis.topography:
            self._set_numerical_primaries()
        self._dirty = False

    def _set_numerical_primaries(self):
        """DCSRMultiElectrodeModelling::checkPrimpotentials_ with topography
        (dcfemmodelling.cpp:2009-2056): the total-field problem with rho = 1
        on the P2 mesh (Mesh::createP2) for every electrode and wavenumber,
        its values at the mesh nodes (the P2 mesh's vertices) as the primary
        potentials.  Solved with this operator's solver options on a second
        library context; kept while mesh, electrodes and wavenumbers stay.
        The reference's inner operator takes its own default wavenumbers,
        equal to these unless setkValues changed them."""
        an = self._analysis
        key = (id(self._mesh), self._mesh.nodeCount(), tuple(an.elecNode.tolist()), tuple(self._kv.tolist()),
               tuple(an.sourceCenter.tolist()), an.refNode, tuple(an.calibration))
        if self._prim is None or self._primKey != key:
            self._prim = None
            self._prim = self._p2_primaries()
            self._primKey = key
        self._ctx.check(self._ctx.L.pgert_set_primary_potentials(self._ctx.h, _dptr(self._prim)), "primaries")

    def _p2_primaries(self):
        mesh, an, d = self._mesh, self._analysis, self._data
        nodes, cells, bounds, _ = p2_refinement(mesh)
        ctx = _lib.Context(self._ctx.device)
        try:
            s = torch.cuda.current_stream(self.device).cuda_stream
            ctx.check(ctx.L.pgert_set_stream(ctx.h, ctypes.c_void_p(s)), "set_stream")
            xyz, pxyz = _host_d(nodes)
            c2, pc = _host_i(cells)
            cm, pcm = _host_i(np.zeros(mesh.cellCount()))
            b2, pb = _host_i(bounds)
            bm, pbm = _host_i(mesh.boundaryMarkers())
            lc, plc = _host_i(mesh.boundaryLeftCell())
            ctx.check(ctx.L.pgert_set_mesh_p2(ctx.h, mesh.dim(), len(nodes), pxyz, mesh.cellCount(), pc, pcm,
                                              mesh.boundaryCount(), pb, pbm, plc), "P2 mesh")
            en, pen = _host_i(an.elecNode)
            sc, psc = _host_d(an.sourceCenter)
            ctx.check(ctx.L.pgert_set_electrodes(ctx.h, len(en), pen, psc, an.surfaceZ), "P2 electrodes")
            # a reference electrode keeps its node id on the P2 mesh: dipole primaries
            cal, pcal = _host_i(an.calibration)
            ctx.check(ctx.L.pgert_set_reference(ctx.h, an.refNode, len(cal), pcal), "P2 reference electrode")
            k, pk = _host_d(self._kv)
            w, pw = _host_d(self._wv)
            ctx.check(ctx.L.pgert_set_wavenumbers(ctx.h, len(k), pk, pw), "P2 wavenumbers")
            a, pa = _host_i(d["a"])
            bb, pbb = _host_i(d["b"])
            m, pm = _host_i(d["m"])
            nn, pn = _host_i(d["n"])
            kf, pkf = _host_d(d["k"])
            ctx.check(ctx.L.pgert_set_data(ctx.h, d.size(), pa, pbb, pm, pn, pkf), "P2 data")
            o = self.solverOptions
            opts = _lib.SolveOpts()
            opts.method = _lib.METHODS[o["method"]] if isinstance(o["method"], str) else int(o["method"])
            opts.rtol = float(o["rtol"])
            opts.max_iter = int(o["max_iter"])
            opts.check_every = int(o["check_every"])
            opts.inner_rtol = float(o["inner_rtol"])
            opts.no_graph = 0 if o.get("graph", True) else 1
            rho = torch.ones(mesh.cellCount(), dtype=torch.float64, device=self.device)
            rhoa = torch.empty(d.size(), dtype=torch.float64, device=self.device)
            stats = _lib.SolveStats()
            ctx.check(ctx.L.pgert_forward_total(ctx.h, _dptr(rho), ctypes.byref(opts), _dptr(rhoa),
                                                ctypes.byref(stats)), "P2 primary potentials")
            U = torch.empty((len(k), len(nodes), len(en)), dtype=torch.float64, device=self.device)
            ctx.check(ctx.L.pgert_potentials_node_major(ctx.h, _dptr(U)), "P2 potentials")
            self.primaryStats = stats.as_dict()
            return U[:, :mesh.nodeCount(), :].contiguous()
        finally:
            ctx.close()

    def _ensure_k(self):
        """response() fills missing geometric factors analytically
        (dcfemmodelling.cpp:1088-1098)."""
        d = self._data
        kf = d["k"]
        if not d.haveData("k") or np.min(np.abs(kf)) < 1e-12:
            if self._analysis is not None and (self._analysis.topography or self._analysis.cemFaces):
                raise ValueError("data contains no K-factors (needed with topography or the complete electrode "
                                 "model, dcfemmodelling.cpp:1088-1098)")
            d.set("k", self.calcGeometricFactor(d))

    def _upload_data(self):
        d = self._data
        a, pa = _host_i(d["a"])
        b, pb = _host_i(d["b"])
        m, pm = _host_i(d["m"])
        n, pn = _host_i(d["n"])
        k, pk = _host_d(d["k"])
        ctx = self._ctx
        ctx.check(ctx.L.pgert_set_data(ctx.h, d.size(), pa, pb, pm, pn, pk), "setData")

    def calcGeometricFactor(self, data=None, nModel=0):
        """Analytic geometric factors (geometricFactors, bertMisc.cpp:131-176)."""
        data = self._data if data is None else data
        return geometric_factors(data.sensorPositions(), data.abmn(), self._mesh.dim())

    # ------------------------------------------------------------------ model
    def _model_tensor(self, model):
        if isinstance(model, torch.Tensor):
            t = model.to(device=self.device, dtype=torch.float64).contiguous()
        else:
            t = torch.as_tensor(np.ascontiguousarray(model, np.float64), device=self.device)
        return t

    def mapERTModel(self, model, background=-9e99):
        """Cell resistivities from a model (dcfemmodelling.cpp:1211-1218,
        modellingbase.cpp:401-470): a cell-length model is copied, otherwise
        rho_c = model[marker_c] and cells with a negative marker are
        prolongated from their neighbours (Mesh::prolongateEmptyCellsValues,
        mesh.cpp:2247-2318; only the default background -9e99 of
        response(model) is supported).  Returns a device tensor."""
        if background > -9e99:
            raise NotImplementedError("only the prolongating background -9e99 is supported")
        t = self._model_tensor(model)
        nm = t.numel()
        C = self._mesh.cellCount()
        if nm != C:
            if self._mesh.cellMarkers().max() >= nm:
                raise IndexError("marker >= model.size()")
        rho = torch.empty(C, dtype=torch.float64, device=self.device)
        self._set_stream()
        self._ctx.check(self._ctx.L.pgert_map_model(self._ctx.h, _dptr(t), nm, _dptr(rho)), "mapERTModel")
        return rho

    # ------------------------------------------------------------------ forward
    def _opts(self):
        o = self.solverOptions
        opts = _lib.SolveOpts()
        opts.method = _lib.METHODS[o["method"]] if isinstance(o["method"], str) else int(o["method"])
        opts.rtol = float(o["rtol"])
        opts.max_iter = int(o["max_iter"])
        opts.check_every = int(o["check_every"])
        opts.inner_rtol = float(o["inner_rtol"])
        opts.warm_start = None
        opts.learned_warm_start = 1 if o.get("learned", False) else 0
        nagg = int(o.get("coarse", 0) or 0)
        if nagg != self._coarseSet:
            self._ctx.check(self._ctx.L.pgert_set_coarse(self._ctx.h, nagg), "set_coarse")
            self._coarseSet = nagg
        opts.coarse = 1 if nagg > 0 else 0
        opts.smooth = int(o.get("smooth", 0) or 0) if nagg > 0 else 0
        opts.no_graph = 0 if o.get("graph", True) else 1
        opts.report_only = 1 if o.get("report_only", False) else 0
        if self.warmStart is not None:
            ws = self.warmStart
            if ws.dtype != torch.float64 or not ws.is_contiguous():
                raise ValueError("warm start must be a contiguous float64 device tensor [nK][n][nE]")
            opts.warm_start = ctypes.c_void_p(ws.data_ptr())
        return opts

    def response_device(self, model):
        """Apparent resistivities as a float64 device tensor; for complex
        resistivities model = [re, im] and the result [re, im] of the complex
        apparent resistivities."""
        self._sync_setup()
        t = self._model_tensor(model)
        if self._complex:
            return self._response_complex(t)
        if float(t.min()) < 1e-12:
            raise ValueError("response for model with negative or zero resistivity is not defined")
        rho = self.mapERTModel(t)
        rhoa = torch.empty(self._data.size(), dtype=torch.float64, device=self.device)
        stats = _lib.SolveStats()
        opts = self._opts()
        self._set_stream()
        if self._sr:
            self._ctx.check(self._ctx.L.pgert_forward(self._ctx.h, _dptr(rho), ctypes.byref(opts), _dptr(rhoa),
                                                      ctypes.byref(stats)), "response")
        else:
            opts.learned_warm_start = 0
            self._ctx.check(self._ctx.L.pgert_forward_total(self._ctx.h, _dptr(rho), ctypes.byref(opts),
                                                            _dptr(rhoa), ctypes.byref(stats)), "response")
        self.lastStats = stats.as_dict()
        self._model = t
        if self.verbose:
            print("ERTModelling.response:", self.lastStats, self.timings())
        return rhoa

    def _response_complex(self, t):
        """DCMultiElectrodeModelling::response, complex branch
        (dcfemmodelling.cpp:1103-1120)."""
        if t.numel() % 2:
            raise ValueError("complex model: [real part, imaginary part] expected")
        h = t.numel() // 2
        re, im = t[:h].contiguous(), t[h:].contiguous()
        if float(torch.sqrt(re * re + im * im).min()) < 1e-12:
            raise ValueError("complex response for abs model with negative or zero resistivity is not defined")
        rre, rim = self.mapERTModel(re), self.mapERTModel(im)
        D = self._data.size()
        out = torch.empty(2 * D, dtype=torch.float64, device=self.device)
        stats = _lib.SolveStats()
        opts = self._opts()
        opts.learned_warm_start = 0
        self._set_stream()
        self._ctx.check(self._ctx.L.pgert_forward_complex(self._ctx.h, _dptr(rre), _dptr(rim), ctypes.byref(opts),
                                                          _dptr(out), ctypes.c_void_p(out.data_ptr() + 8 * D),
                                                          ctypes.byref(stats)), "response")
        self.lastStats = stats.as_dict()
        self._model = t
        return out

    def response(self, mod):
        """Forward response (apparent resistivity), numpy float64 [nData]
        (ERTModelling.response, ertModelling.py:204-219)."""
        return self.response_device(mod).cpu().numpy()

    def timings(self):
        ms = (ctypes.c_double * 5)()
        self._ctx.L.pgert_last_timings(self._ctx.h, ms)
        return dict(zip(("assembly", "primary_rhs", "solve", "response", "jacobian"), list(ms)))

    def subPotentials(self, node_major=False):
        """Per-wavenumber unweighted total potentials of the last response:
        rows electrode + nE * kIdx, columns nodes (subSolutions_,
        dcfemmodelling.cpp:1681-1682); node_major=True gives [nK, n, nE]."""
        nK, nE, n = len(self._kv), len(self._analysis.elecNode), self._nLib
        if node_major:
            out = torch.empty((nK, n, nE), dtype=torch.float64, device=self.device)
            self._ctx.check(self._ctx.L.pgert_potentials_node_major(self._ctx.h, _dptr(out)), "potentials")
        else:
            out = torch.empty((nK * nE, n), dtype=torch.float64, device=self.device)
            self._ctx.check(self._ctx.L.pgert_potentials(self._ctx.h, _dptr(out)), "potentials")
        return out

    def solution(self):
        """k-weighted potentials of every current electrode, [nE, n]
        (ModellingBase::solution)."""
        P = self.subPotentials(node_major=True)
        w = torch.as_tensor(self._wv, device=self.device)
        return torch.einsum("k,kne->en", w, P).contiguous()

    # ------------------------------------------------------------------ Jacobian
    def createJacobian(self, mod):
        """Compute and store the Jacobian (ERTModelling.createJacobian,
        ertModelling.py:221-238 -> DCMultiElectrodeModelling::createJacobian,
        dcfemmodelling.cpp:1446-1476).  Reuses the potentials of the last
        response() when it was computed for the same model."""
        self._sync_setup()
        t = self._model_tensor(mod)
        if self._model is None or self._model.shape != t.shape or not torch.equal(self._model, t):
            self.response_device(t)
        if self._complex:
            return self._jacobian_complex(t)
        dtype = self.jacobianDtype
        J = torch.empty((self._data.size(), self._nParam), dtype=dtype, device=self.device)
        self._Jc = None
        code = 0 if dtype == torch.float64 else 1
        self._set_stream()
        self._ctx.check(self._ctx.L.pgert_jacobian(self._ctx.h, _dptr(t), t.numel(), _dptr(J), code),
                        "createJacobian")
        self._J = J
        return None

    def _jacobian_complex(self, t):
        """createJacobian complex branch (dcfemmodelling.cpp:1446-1463,
        :1410-1444) and its real block form (ERTModelling.createJacobian,
        ertModelling.py:224-236 -> pg.utils.toRealMatrix, utils/complex.py:79-110):
        [[Re J, -Im J], [Im J, Re J]]."""
        h = t.numel() // 2
        re, im = t[:h].contiguous(), t[h:].contiguous()
        D, P = self._data.size(), self._nParam
        Jr = torch.empty((D, P), dtype=torch.float64, device=self.device)
        Ji = torch.empty((D, P), dtype=torch.float64, device=self.device)
        self._set_stream()
        self._ctx.check(self._ctx.L.pgert_jacobian_complex(self._ctx.h, _dptr(re), _dptr(im), h, _dptr(Jr),
                                                           _dptr(Ji)), "createJacobian")
        self._Jc = (Jr, Ji)
        self._J = None   # the real block form is built on demand (4 D x P)
        return None

    def jacobian(self):
        """The Jacobian of the last createJacobian, [nData, nParam] device
        tensor ([2 nData, 2 nParam] real block form for complex resistivities,
        materialised on the first call: it holds four D x P FP64 blocks, while
        jacobianComplex() returns the two parts without a copy)."""
        if self._J is None and self._Jc is not None:
            Jr, Ji = self._Jc
            self._J = torch.cat([torch.cat([Jr, -Ji], 1), torch.cat([Ji, Jr], 1)], 0)
        return self._J

    def jacobianComplex(self):
        """(real, imaginary) parts of the complex Jacobian, [nData, nParam] each."""
        return self._Jc

    @property
    def parameterCount(self):
        return self._nParam * (1 + int(self._complex))

    def createStartModel(self, dataVals):
        """Homogeneous start model (ERTModelling.createStartModel,
        ertModelling.py:168-186): the median of the data (real); for complex
        data [re, im] the medians of both parts, the imaginary one flipped to
        positive (and remembered for flipImagPart) when negative."""
        v = np.asarray(dataVals, np.float64)
        if not self._complex:
            return np.full(self.parameterCount, float(np.median(v)))
        h = len(v) // 2
        n = self.parameterCount // 2
        re = np.full(n, float(np.median(v[:h])))
        im = np.full(n, float(np.median(v[h:])))
        self._conjImag = bool(im.min() < 0)
        return np.concatenate([re, -im if self._conjImag else im])

    def flipImagPart(self, v):
        """[re, im] with the imaginary part negated when createStartModel
        flipped it (ertModelling.py:188-202)."""
        v = np.asarray(v, np.float64)
        if not getattr(self, "_conjImag", False):
            return v.copy()
        h = len(v) // 2
        return np.concatenate([v[:h], -v[h:]])

    def createConstraints(self, C=None):
        """Constraint matrix for the inversion (ERTModelling.createConstraints,
        ertModelling.py:152-166): first-order smoothness between neighbouring
        parameter cells, or the given C; repeated for both halves of a complex
        model (RepeatDMatrix).  Returns scipy CSR."""
        from .inversion import complex_constraints, smoothness_constraints
        C1 = smoothness_constraints(self._mesh) if C is None else C
        return complex_constraints(C1) if self._complex else C1

    def close(self):
        self._ctx.close()


def system_matrix(fop, model, k):
    """System matrix of one wavenumber for a model, as scipy CSR (the matrix
    DCSRMultiElectrodeModelling::calculateK factorises, dcfemmodelling.cpp:2179-2182)."""
    import scipy.sparse as sp
    fop._sync_setup()
    rho = fop.mapERTModel(model)
    ctx = fop._ctx
    n = fop._mesh.nodeCount()
    rowptr = np.empty(n + 1, np.int32)
    colidx = np.empty(fop._nnz, np.int32)
    ctx.check(ctx.L.pgert_pattern(ctx.h, rowptr.ctypes.data_as(_I), colidx.ctypes.data_as(_I)), "pattern")
    vals = torch.empty(fop._nnz, dtype=torch.float64, device=fop.device)
    fop._set_stream()
    ctx.check(ctx.L.pgert_assemble(ctx.h, _dptr(rho), float(k), _dptr(vals)), "assemble")
    return sp.csr_matrix((vals.cpu().numpy(), colidx, rowptr), shape=(n, n))

"""ERT forward + Jacobian throughput on one MI355X.

Workload (default): 3D ERT, tetrahedral mesh of 201,898 nodes / 1,150,488
tetrahedra (graded tensor grid, 6 tets per hexahedron), 64 electrodes on an
8 x 8 surface grid at 2 m, 2000 random 4-point quadrupoles (seed 0),
log10(rho) Gaussian random field (mean 2.0, std 0.5, exponential covariance,
5 m correlation, seed 1); 64-RHS block-PCG with FP32 SpMM and FP64 residual
replacement; Jacobian 2000 x 1,150,488 in FP32.

A step = one forward response (assembly, SR right-hand sides, solve of all
64 sources, rhoa) + one Jacobian, from model resident in HBM.  The plain FP64
independent-column Jacobi-PCG is timed beside it at the same tolerance.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ERT forward+Jacobian/s at 1e-6 rel err: plain FP64 PCG vs mixed+learned, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_worker(out_path, nthreads, sample_frac, pcg_iters, workload="ert3d_grid_grf", ref_data=0):
    """Runs in a child process (started before the GPU work): CPU work on the
    host cores, on a bounded sample of the same workload.
    Forward = FP64 assembly of the system (oracle) + the plain FP64 Jacobi-PCG
    C port over all sources, `pcg_iters` iterations timed (per-iteration cost
    does not depend on the values), scaled by the caller to the iteration
    count the same algorithm needs on this workload (the reference's own
    direct solver cannot run here: CHOLMOD is absent).
    Jacobian = with ref_data > 0 the reference's own createJacobian (calc2 on
    its thread pool, oracle/_ref) for the first ref_data quadrupoles on the
    whole mesh, scaled by D / ref_data (its cost is linear in the data count;
    the potentials handed to it are the analytic primaries, which cost calc2
    the same as any other values); otherwise the C restatement of calc2 on a
    contiguous sample of cells, scaled to all cells.
    Threads: min(8, nCPU-2) (the reference's default thread count,
    modellingbase.cpp:77), the PCG split over right-hand sides."""
    import faulthandler
    faulthandler.enable()  # a crash in native code prints the Python stack to stderr
    from oracle import ert_oracle as eo
    from oracle import port
    from pygimli_amd import workloads
    mesh, data, rho = getattr(workloads, workload)()
    fo = eo.Forward(mesh, copy.deepcopy(data))
    kv, wv = fo.k, fo.w
    t = time.time()
    S = eo.assemble(mesh, rho, kv[0], fo.setup.sourceCenter, fo.geo)
    ta = time.time() - t
    B = np.random.default_rng(0).standard_normal((S.shape[0], fo.nE * len(kv)))
    t = time.time()
    _, it = port.pcg(S, B, maxit=pcg_iters, rtol=0.0, nthreads=nthreads)
    t_it = (time.time() - t) / max(it, 1)
    C = mesh.cellCount()
    c1 = max(1, int(C * sample_frac))
    U = fo.primaries()
    res = {"ta": ta, "t_it": t_it, "it_sampled": it, "c1": c1, "C": C, "nrhs": B.shape[1], "D": data.size()}
    if ref_data > 0:
        from oracle import ref
        if ref.available():
            # the reference fills its shape-function caches lazily
            # (ShapeFunctionCache, shape.cpp:253-317) with no lock, so its
            # threaded createJacobian can crash when the threads race on the
            # first fill: fill them first on a small mesh of the same cell type,
            # single-threaded
            ms, ds, rs = (workloads.small_3d if mesh.dim() == 3 else workloads.small_2d)()
            warm = ref.Session(ms, ds, nthreads=1)
            warm.response(rs)
            warm.jacobian_only(rs)
            sub = workloads.make_data(data.sensorPositions(), data.abmn()[:ref_data])
            sub.set("k", fo.geometric_factors()[:ref_data])
            t = time.time()
            sess = ref.Session(mesh, sub, nthreads=nthreads)
            res["ref_setup_s"] = time.time() - t
            sess.set_subpotentials(U)
            t = time.time()
            sess.jacobian_only(rho)
            res["tj_sample"] = time.time() - t
            res["tj"] = res["tj_sample"] * data.size() / ref_data
            res["tj_kind"] = "reference"
            res["ref_data"] = ref_data
    if "tj" not in res:
        # the sampled cells get their own columns 0..c1-1 (J of the sample only)
        mk = np.full(C, -1, np.int32)
        mk[:c1] = np.arange(c1)
        t = time.time()
        port.calc2(mesh, U, kv, wv, data.abmn(), c1, nthreads=nthreads, cell_range=(0, c1),
                   J=np.zeros((data.size(), c1)), markers=mk)
        res["tj"] = (time.time() - t) * C / c1
        res["tj_kind"] = "port"
    with open(out_path, "w") as f:
        json.dump(res, f)


def cpu_baseline_result(path, nthreads, pcg_iters_needed):
    try:
        with open(path) as f:
            r = json.load(f)
    except Exception as e:
        return {"error": f"cpu baseline produced no result ({e!r})"}
    if not pcg_iters_needed:
        return {"error": "no plain-PCG iteration count to scale the CPU PCG sample with (run without --no-plain)",
                "partial": r, "cores": nthreads}
    tfwd = r["ta"] + r["t_it"] * pcg_iters_needed
    if r.get("tj_kind") == "reference":
        jac = (f"Jacobian: the reference's own createJacobian (calc2, oracle/_ref, {nthreads} threads) on the whole "
               f"mesh for the first {r['ref_data']} of {r['D']} quadrupoles = {r['tj_sample']:.2f} s, x "
               f"{r['D'] / r['ref_data']:.0f} = {r['tj']:.1f} s")
    else:
        jac = f"Jacobian: C calc2 port on {r['c1']} of {r['C']} cells, scaled = {r['tj']:.1f} s"
    return {
        "value": 1.0 / (tfwd + r["tj"]), "unit": "evals/s", "cores": nthreads, "kind": "port",
        "sample": (f"forward: FP64 assembly {r['ta']:.1f} s + plain Jacobi-PCG C port over {r['nrhs']} sources, "
                   f"{r['it_sampled']} iterations timed ({r['t_it'] * 1e3:.0f} ms each) x {pcg_iters_needed} "
                   f"iterations (the plain FP64 PCG's count on this workload) = {r['t_it'] * pcg_iters_needed:.0f} s "
                   f"(the reference's direct solver cannot run: CHOLMOD is absent); {jac}; {nthreads} threads"),
        "forward_s": tfwd, "jacobian_s": r["tj"], "jacobian_kind": r.get("tj_kind", "port"),
    }


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE and scaled by a copy kernel of known size, as
    committed under profiles/ (None when not measured)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def batched_cpu_worker(out_path, nthreads, sample):
    """CPU baseline of the batched 2.5D workload on a bounded sample: for each
    sampled problem the oracle's FP64 path -- per-wavenumber assembly, sparse
    direct factorisation (SuperLU, the stand-in for the reference's CHOLMOD),
    one solve per source, and the calc2 C port for the Jacobian."""
    import scipy.sparse.linalg as spla
    from oracle import ert_oracle as eo
    from oracle import port
    from pygimli_amd import workloads
    probs = workloads.batched_2d(count=256, seed=2)
    res = []
    for i in sample:
        mesh, data, rho, k, w = probs[i]
        fo = eo.Forward(mesh, copy.deepcopy(data), k, w)
        t = time.time()
        U = fo.solve_subpotentials(rho)
        tf = time.time() - t
        nP = int(mesh.cellMarkers().max()) + 1
        t = time.time()
        port.calc2(mesh, U, k, w, data.abmn(), nP, nthreads=nthreads, J=np.zeros((data.size(), nP)))
        tj = time.time() - t
        res.append({"i": int(i), "n": mesh.nodeCount(), "nE": len(data.sensorPositions()), "tf": tf, "tj": tj})
        with open(out_path, "w") as f:
            json.dump(res, f)


def batched_leg(args, log):
    """Config 3: 256 independent 2.5D problems (seed 2), sub-batches of
    args.batch_size problems, each ONE batched solve (band: one factor and one
    solve launch over every wavenumber of every problem of the sub-batch;
    block methods: every solver launch covers every wavenumber block of every
    problem, iterations replayed from a captured HIP graph), then the FP32
    Jacobians of the sub-batch in one launch per wavenumber pass.  The timed
    evaluation follows a first evaluation of every sub-batch (per-mesh
    set-up, reported separately).  Returns the result object."""
    import torch
    from pygimli_amd import workloads
    from pygimli_amd.batch import batch_jacobian, batch_response_device
    from pygimli_amd.ertModelling import ERTModelling
    t0 = time.time()
    probs = workloads.batched_2d(count=args.batched, seed=2)
    log(f"[batched] {len(probs)} problems generated in {time.time() - t0:.1f}s")
    fops, models = [], []
    for mesh, data, rho, k, w in probs:
        f = ERTModelling()
        f.setMesh(mesh)
        f.setData(copy.deepcopy(data))
        f.setkValues(k)
        f.setWeights(w)
        f.jacobianDtype = torch.float32
        n = mesh.nodeCount()
        f.solverOptions.update(method=args.batched_method, rtol=args.rtol,
                               coarse=min(1024, 2 ** int(np.log2(n / 32))) if args.batched_method != "band" else 0)
        fops.append(f)
        models.append(torch.as_tensor(rho, device=f.device))
    log(f"[batched] operators set up at {time.time() - t0:.1f}s")
    bs = args.batch_size
    groups = [list(range(i, min(i + bs, len(fops)))) for i in range(0, len(fops), bs)]

    def evaluate(g):
        fs, ms = [fops[i] for i in g], [models[i] for i in g]
        batch_response_device(fs, ms)
        its = [f.lastStats["max_iterations"] for f in fs]
        conv = all(f.lastStats["converged"] == 1 for f in fs)
        batch_jacobian(fs, ms)                     # one launch per wavenumber pass over all problems
        for f in fs:
            f._J = None                            # release the group's Jacobians
        return max(its), conv

    # first evaluation of every group: per-mesh set-up (device upload of the
    # mesh, band ordering, S1 matrices, primaries) that an inversion pays once
    torch.cuda.synchronize()
    t = time.perf_counter()
    for g in groups:
        evaluate(g)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    log(f"[batched] first evaluation (with per-mesh set-up) {t_setup:.1f}s")
    t = time.perf_counter()
    iters, ok = [], True
    for g in groups:
        it, conv = evaluate(g)
        iters.append(it)
        ok = ok and conv
        log(f"[batched] group of {len(g)}: {it} iterations, converged={conv}, {time.perf_counter() - t:.1f}s")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    # accuracy on a sample: FP64 block-PCG to 1e-10, one problem at a time
    errs = []
    for i in (0, len(fops) // 2, len(fops) - 1):
        f = fops[i]
        f.solverOptions.update(method=args.batched_method, rtol=args.rtol)
        r = batch_response_device([f], [models[i]])[0]
        # (plain Jacobi block-PCG needs more than 20000 iterations for 1e-10 on
        # the larger of these meshes; the two-level cycle converges in hundreds)
        nn = f._mesh.nodeCount()
        f.solverOptions.update(method="block64", rtol=1e-10, coarse=min(1024, 2 ** int(np.log2(nn / 32))), smooth=2)
        r_ref = f.response_device(models[i])
        errs.append(float(torch.linalg.norm(r - r_ref) / torch.linalg.norm(r_ref)))
    nodes = sum(f._mesh.nodeCount() for f in fops)
    rhs = sum(len(f._kv) * len(f._data.sensorPositions()) for f in fops)
    return {"problems": len(fops), "sub_batch": bs, "seconds": dt, "evals_per_s": 1.0 / dt,
            "first_evaluation_with_setup_s": t_setup,
            "problems_per_s": len(fops) / dt, "iterations_per_group": iters, "converged": ok,
            "nodes_total": nodes, "rhs_total": rhs, "data_total": sum(f._data.size() for f in fops),
            "rhoa_rel_err_vs_fp64_1e-10": errs,
            "method": args.batched_method,
            "workload": "256 x 2.5D: nodes U[2k,40k], electrodes U[24,96] @ U[0.5,2] m, Wenner-Schlumberger, "
                        "2-4 layers + 0-3 ellipses (log10 rho U[1,3]), 11 wavenumbers; "
                        + ("direct banded Cholesky (FP64)" if args.batched_method == "band" else
                           "mixed block-PCG 1e-6 with per-problem coarse space") + ", FP32 Jacobian per problem"}


def _section_arrays(out_path, name, what):
    """Reference results of the section leg: large, so kept in the temp dir."""
    import tempfile
    return os.path.join(tempfile.gettempdir(), f"{os.path.basename(out_path)}.{name}.{what}.npy")


def section_ref_worker(out_path, nthreads, reps):
    """Config-1 CPU baseline in a child process: the reference's own C++
    core (oracle/_ref, DCSRMultiElectrodeModelling with its solver plug-in
    served by an envelope Cholesky) on the full workload, response and
    createJacobian, for the explicit 11 wavenumbers and the default list;
    median of `reps` runs after one warm-up."""
    os.dup2(2, 1)  # the reference core prints its warnings to stdout; keep stdout for the JSON line
    from oracle import ref
    from pygimli_amd import bert, workloads
    mesh, data, rho = workloads.ert2d_layers_block()
    an = bert.MeshAnalysis(mesh, data.sensorPositions())
    d = np.sqrt(((an.elecPos[:, None] - an.elecPos[None]) ** 2).sum(-1))
    k11, w11 = bert.kwave_from_range(d[d > 0].min() / 2, d.max() * 2, 7, 4)
    res = {}
    for name, kw in (("k11", (k11, w11)), ("default", (None, None))):
        ses = ref.Session(mesh, copy.deepcopy(data), kw[0], kw[1], nthreads=nthreads)
        ts = []
        for r in range(reps + 1):
            t = time.perf_counter()
            rhoa = ses.response(rho)
            J = ses.jacobian(rho)
            ts.append(time.perf_counter() - t)
        res[name] = {"s": float(np.median(ts[1:])), "nk": len(ses.kw()[0])}
        np.save(_section_arrays(out_path, name, "rhoa"), rhoa)
        np.save(_section_arrays(out_path, name, "J"), J)
    with open(out_path, "w") as f:
        json.dump(res, f)


def section_leg(args, log):
    """Config 1: one 2.5D section (3-layer + block model, 41 electrodes,
    dipole-dipole), the reference's CPU-runnable case.  A step = response
    (direct band solver, the reference's own direct approach) +
    createJacobian (FP64, as the reference's), with the explicit 11
    wavenumbers and with the default list.  CPU baseline: the reference's
    own C++ core (oracle/_ref) on the same inputs, results compared."""
    import torch
    from pygimli_amd import bert, workloads
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.ert2d_layers_block()
    an = bert.MeshAnalysis(mesh, data.sensorPositions())
    d = np.sqrt(((an.elecPos[:, None] - an.elecPos[None]) ** 2).sum(-1))
    k11, w11 = bert.kwave_from_range(d[d > 0].min() / 2, d.max() * 2, 7, 4)
    out = {}
    for name, kw in (("k11", (k11, w11)), ("default", (None, None))):
        fop = ERTModelling()
        fop.setMesh(mesh)
        fop.setData(copy.deepcopy(data))
        if kw[0] is not None:
            fop.setkValues(kw[0])
            fop.setWeights(kw[1])
        fop.solverOptions.update(method="band", rtol=1e-10)
        model = torch.as_tensor(rho, device=fop.device)
        for _ in range(args.warmup + 1):
            fop.response_device(model)
            fop.createJacobian(model)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(max(args.steps, 5)):
            r = fop.response_device(model)
            fop.createJacobian(model)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / max(args.steps, 5)
        out[name] = {"ms_per_step": dt * 1e3, "evals_per_s": 1.0 / dt, "nk": len(fop._kv),
                     "phase_ms": fop.timings(), "rhoa": r.cpu().numpy(), "J": fop.jacobian().cpu().numpy()}
        log(f"[section] {name}: {dt * 1e3:.2f} ms, phases {fop.timings()}")
    return out, mesh, data


def complex_ref_worker(out_path, nthreads, reps):
    """CPU baseline of the complex leg in a child process: the reference's
    own C++ core (oracle/_ref, DCMultiElectrodeModelling with setComplex(true),
    complex envelope Cholesky in its solver plug-in), response +
    createJacobian of the 2.5D section with chargeable block, default
    wavenumbers; median of `reps` runs after one warm-up."""
    os.dup2(2, 1)
    from oracle import ref
    from pygimli_amd import workloads
    mesh, data, rho = workloads.ert2d_layers_block()
    rim = workloads.chargeable_block_phases(mesh, rho)
    ses = ref.Session(mesh, copy.deepcopy(data), nthreads=nthreads, complex_=True)
    ts = []
    for r in range(reps + 1):
        t = time.perf_counter()
        rr, ri = ses.response_complex(rho, rim)
        Jr, Ji = ses.jacobian_complex(rho, rim)
        ts.append(time.perf_counter() - t)
    for name, a in (("re", rr), ("im", ri), ("Jre", Jr), ("Jim", Ji)):
        np.save(_section_arrays(out_path, "complex", name), a)
    # the real total-field operator (ERTModelling(sr=False)) on the same section
    st = ref.Session(mesh, copy.deepcopy(data), nthreads=nthreads, sr=False)
    tt = []
    for r in range(reps + 1):
        t = time.perf_counter()
        rt = st.response(rho)
        Jt = st.jacobian(rho)
        tt.append(time.perf_counter() - t)
    np.save(_section_arrays(out_path, "total", "rhoa"), rt)
    np.save(_section_arrays(out_path, "total", "J"), Jt)
    with open(out_path, "w") as f:
        json.dump({"s": float(np.median(ts[1:])), "nk": len(ses.kw()[0]), "total_s": float(np.median(tt[1:]))}, f)


def complex_leg(args, log):
    """Complex resistivity: the 2.5D section with a chargeable block
    (-30 mrad in the block, -5 mrad elsewhere), ERTModelling(sr=False) +
    setComplex(True), direct band solver, default wavenumbers; a step =
    complex response + complex Jacobian (FP64).  Also the 3D survey of
    8 x 8 electrodes (202k nodes, 2000 data) with the same phases inside a
    box, block-PCG mixed + two-level preconditioner, one step."""
    import torch
    from pygimli_amd import workloads
    from pygimli_amd.ertModelling import ERTModelling
    mesh, data, rho = workloads.ert2d_layers_block()
    rim = workloads.chargeable_block_phases(mesh, rho)
    fop = ERTModelling(sr=False)
    fop.setComplex(True)
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method="band", rtol=1e-10)
    model = torch.as_tensor(np.concatenate([rho, rim]), device=fop.device)
    for _ in range(args.warmup + 1):
        fop.response_device(model)
        fop.createJacobian(model)
    torch.cuda.synchronize()
    steps = max(args.steps, 5)
    t = time.perf_counter()
    for _ in range(steps):
        r = fop.response_device(model)
        fop.createJacobian(model)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    Jr, Ji = fop.jacobianComplex()
    res2 = {"ms_per_step": dt * 1e3, "evals_per_s": 1.0 / dt, "nk": len(fop._kv), "stats": fop.lastStats,
            "resp": r.cpu().numpy(), "Jre": Jr.cpu().numpy(), "Jim": Ji.cpu().numpy(),
            "workload": f"2.5D section: {mesh.nodeCount()} nodes / {mesh.cellCount()} triangles, 41 electrodes, "
                        f"{data.size()} dipole-dipole data, 3 layers + block, phases -30 mrad (block) / -5 mrad"}
    log(f"[complex] 2.5D: {dt * 1e3:.2f} ms, {fop.lastStats}")
    # the real total-field operator on the same section
    ft = ERTModelling(sr=False)
    ft.setMesh(mesh)
    ft.setData(copy.deepcopy(data))
    ft.solverOptions.update(method="band", rtol=1e-10)
    mt = torch.as_tensor(rho, device=ft.device)
    for _ in range(args.warmup + 1):
        ft.response_device(mt)
        ft.createJacobian(mt)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        rt = ft.response_device(mt)
        ft.createJacobian(mt)
    torch.cuda.synchronize()
    dtt = (time.perf_counter() - t) / steps
    res2["total"] = {"ms_per_step": dtt * 1e3, "evals_per_s": 1.0 / dtt, "rhoa": rt.cpu().numpy(),
                     "J": ft.jacobian().cpu().numpy(),
                     "workload": "same section, real resistivities, ERTModelling(sr=False), band solver, FP64 J"}
    log(f"[total] 2.5D: {dtt * 1e3:.2f} ms")
    mesh3, data3, rho3 = workloads.ert3d_grid_grf()
    c = mesh3.cellCenters()
    inside = (np.abs(c[:, 0]) < 4) & (np.abs(c[:, 1]) < 4) & (c[:, 2] > -6) & (c[:, 2] < -2)
    rim3 = rho3 * np.tan(np.where(inside, -0.03, -0.005))
    f3 = ERTModelling(sr=False)
    f3.setComplex(True)
    f3.setMesh(mesh3)
    f3.setData(copy.deepcopy(data3))
    f3.solverOptions.update(method="mixed", rtol=1e-6, coarse=args.coarse)
    m3 = torch.as_tensor(np.concatenate([rho3, rim3]), device=f3.device)
    f3.response_device(m3)
    torch.cuda.synchronize()
    t = time.perf_counter()
    f3.response_device(m3)
    f3.createJacobian(m3)
    torch.cuda.synchronize()
    dt3 = time.perf_counter() - t
    res3 = {"ms_per_step": dt3 * 1e3, "stats": f3.lastStats,
            "workload": f"3D: {mesh3.nodeCount()} nodes / {mesh3.cellCount()} tetrahedra, 64 electrodes, "
                        f"{data3.size()} data, phases -30 mrad (box) / -5 mrad, mixed + coarse, rtol 1e-6"}
    log(f"[complex] 3D: {dt3 * 1e3:.1f} ms, {f3.lastStats}")
    return res2, res3


def learned_large_leg(args, log):
    """Learned warm start at scale: 3D mesh of 1.09M nodes / 6.3M tets, 128
    electrodes (8x8 grid + two 32-electrode lines), log-normal training
    fields (std U[0.3, 0.8], correlation U[2, 20] m; seeds 100..).  Trains
    the warm-start MLP (BF16 matrix cores) on samples of the mixed solver's
    own solutions of the training fields, timing the epochs (samples/s and
    TFLOP/s of forward + backward), then solves a held-out field (seed 5000)
    with and without the warm start.  Returns the result object."""
    import torch
    from pygimli_amd import workloads
    from pygimli_amd.ertModelling import ERTModelling
    from pygimli_amd.learned import LearnedWarmStart
    t0 = time.time()
    mesh, data = workloads.ert3d_lines_grid()
    n, C, nE = mesh.nodeCount(), mesh.cellCount(), len(data.sensorPositions())
    log(f"[learned] workload {n} nodes, {C} cells, {nE} electrodes ({time.time() - t0:.1f}s)")
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.solverOptions.update(method=args.method, rtol=args.rtol, coarse=args.coarse, learned=False)
    dev = fop.device
    nf = args.learned_fields if args.learned_fields > 0 else 8
    train = [torch.as_tensor(f, device=dev) for f in workloads.learned_fields(mesh, range(100, 100 + nf))]
    test = torch.as_tensor(workloads.learned_fields(mesh, [5000])[0], device=dev)
    lw = LearnedWarmStart(fop, seed=0)
    rng = np.random.default_rng(0)
    per = args.learned_samples
    feats, tgts, solve_s, iters = [], [], [], []
    for m in train:
        torch.cuda.synchronize()
        t = time.perf_counter()
        fop.response_device(m)
        torch.cuda.synchronize()
        solve_s.append(time.perf_counter() - t)
        iters.append(fop.lastStats["iterations"])
        f, tg = lw.collect(per, rng)
        feats.append(f)
        tgts.append(tg)
    log(f"[learned] {nf} training solves: {iters} iterations, {np.mean(solve_s):.2f} s each")
    F = torch.cat(feats)
    T = torch.clamp(torch.cat(tgts), -5.0, 5.0)
    batch, epochs = args.learned_batch, args.learned_epochs
    nb = F.shape[0] // batch
    losses, ep_s = [], []
    for ep in range(epochs):
        perm = torch.as_tensor(rng.permutation(F.shape[0]), device=dev)
        torch.cuda.synchronize()
        t = time.perf_counter()
        tot = 0.0
        for b in range(nb):
            sel = perm[b * batch:(b + 1) * batch]
            loss, _ = lw.step(F[sel].contiguous(), T[sel].contiguous(), 1e-3)
            tot += loss
        torch.cuda.synchronize()
        ep_s.append(time.perf_counter() - t)
        losses.append(tot / max(nb, 1))
        log(f"[learned] epoch {ep}: loss {losses[-1]:.4e}, {ep_s[-1]:.2f} s")
    samples = nb * batch
    sps = samples / np.median(ep_s)
    flop_fwd = 2.0 * (32 * 128 + 128 * 128 * 2 + 128)
    tflops = sps * 3.0 * flop_fwd / 1e12

    def solve(learned):
        fop.solverOptions.update(learned=learned)
        fop.response_device(test)
        torch.cuda.synchronize()
        t = time.perf_counter()
        fop.response_device(test)
        torch.cuda.synchronize()
        return {"ms": (time.perf_counter() - t) * 1e3, "iterations": fop.lastStats["iterations"],
                "max_rel_residual": fop.lastStats["max_rel_residual"], "phase_ms": fop.timings()}

    without = solve(False)
    with_ = solve(True)
    log(f"[learned] held-out field: without {without}, with {with_}")
    return {
        "metric": "ERT forward/s on a held-out field with the learned warm start, 3D 1.09M nodes / 128 electrodes, 1 GPU",
        "value": 1e3 / with_["ms"], "unit": "evals/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_step": with_["ms"], "dtype": "fp32+fp64 (solver), bf16 (MLP)", "data": "synthetic",
        "config": {"workload": f"3D ERT: {n}-node / {C}-tet graded mesh, {nE} electrodes (8x8 @ 2 m + 2 lines of 32 @ 1 m), "
                               f"log-normal GRF fields (std U[0.3,0.8], corr U[2,20] m)",
                   "method": args.method, "rtol": args.rtol, "coarse_aggregates": args.coarse,
                   "training_fields": nf, "samples_per_field": per, "batch": batch, "epochs": epochs},
        "held_out": {"without_warm_start": without, "with_warm_start": with_,
                     "iteration_ratio": with_["iterations"] / max(without["iterations"], 1),
                     "speedup": without["ms"] / with_["ms"]},
        "training": {"solves_s_mean": float(np.mean(solve_s)), "solve_iterations": iters,
                     "samples": int(samples), "epoch_s": ep_s, "samples_per_s": sps,
                     "loss_first": losses[0], "loss_last": losses[-1]},
        "roofline": {"bound": "mfma", "kernel": "MLP train step (fwd + bwd, BF16)", "achieved": tflops,
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops / 2500.0, "traffic": None,
                     "flops_per_sample": 3.0 * flop_fwd},
    }


def survey_large_leg(args, log):
    """Large 3D survey: 256 electrodes (16x16 @ 1 m), 10000 dipole-dipole and
    gradient quadrupoles, 1.5M-tet mesh, log10-rho GRF; FP32 Jacobian of
    10000 x 1.5M (60 GB) in HBM.  A step = response + createJacobian with
    the mixed block-PCG + two-level preconditioner (four 64-column blocks).
    Plain FP64 PCG / block-PCG are timed once each for comparison, potentials
    are checked against an FP64 block-PCG solve to 1e-10.  Returns the result
    object."""
    import torch
    from pygimli_amd import workloads
    from pygimli_amd.ertModelling import ERTModelling
    t0 = time.time()
    mesh, data, rho = workloads.ert3d_survey_large()
    log(f"[large] workload {mesh.nodeCount()} nodes, {mesh.cellCount()} cells, {data.size()} data, "
        f"{len(data.sensorPositions())} electrodes ({time.time() - t0:.1f}s)")
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.jacobianDtype = torch.float32
    model = torch.as_tensor(rho, device=fop.device)
    fop.solverOptions.update(method="block64", rtol=1e-10, coarse=args.coarse)
    fop.response_device(model)
    U_ref = fop.subPotentials(node_major=True).clone()
    log(f"[large] reference solve {fop.lastStats} {fop.timings()} at {time.time() - t0:.1f}s")

    def run(method, steps, warmup, coarse):
        fop.solverOptions.update(method=method, rtol=args.rtol, coarse=coarse)
        for _ in range(warmup):
            fop.response_device(model)
            fop.createJacobian(model)
            fop._J = None
        torch.cuda.synchronize()
        t = time.perf_counter()
        jms = []
        for _ in range(steps):
            fop.response_device(model)
            fop.createJacobian(model)
            jms.append(fop.timings()["jacobian"])
            fop._J = None
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        U = fop.subPotentials(node_major=True)
        err = float(torch.linalg.norm(U - U_ref) / torch.linalg.norm(U_ref))
        r = dict(ms=dt * 1e3, iters=fop.lastStats["iterations"], pot_err=err, stats=dict(fop.lastStats),
                 timings_ms=fop.timings(), jacobian_ms=min(jms))
        log(f"[large] {method} coarse={coarse}: {r}")
        return r

    plain = run("pcg64", 1, 0, 0)
    plain_block = run("block64", 1, 0, 0)
    best = run(args.method, args.steps, args.warmup, args.coarse)
    # one Gauss-Newton step on the Jacobian in HBM (solveCGLSCDWWhtrans):
    # log-model / log-data transformation, 3 % data error, first-order smoothness
    from pygimli_amd.inversion import smoothness_constraints, solveCGLSCDWWhtrans
    fop.solverOptions.update(method=args.method, rtol=args.rtol, coarse=args.coarse)
    resp = fop.response(model)
    fop.createJacobian(model)
    Jd = fop.jacobian()
    Cm = smoothness_constraints(mesh)
    rng = np.random.default_rng(3)
    dobs = resp * (1.0 + 0.03 * rng.standard_normal(len(resp)))
    gn = dict(dW=np.full(len(resp), 1.0 / 0.03), b=np.log(dobs) - np.log(resp), tm=1.0 / rho, td=1.0 / resp,
              wc=np.ones(Cm.shape[0]), wm=np.ones(len(rho)), rough=Cm @ np.log(rho))
    solveCGLSCDWWhtrans(Jd, Cm, gn["dW"], gn["b"], np.zeros(len(rho)), gn["wc"], gn["wm"], gn["tm"], gn["td"], 20.0,
                        gn["rough"], maxIter=2, tol=0.0)  # set-up: constraints to the device
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, cg_it = solveCGLSCDWWhtrans(Jd, Cm, gn["dW"], gn["b"], np.zeros(len(rho)), gn["wc"], gn["wm"], gn["tm"],
                                   gn["td"], 20.0, gn["rough"], maxIter=args.cgls_iters, tol=0.0)
    torch.cuda.synchronize()
    cg_s = time.perf_counter() - t
    jb = Jd.numel() * Jd.element_size()
    cgls = {"iterations": cg_it, "seconds": cg_s, "ms_per_iteration": cg_s / max(cg_it, 1) * 1e3,
            "constraints": int(Cm.shape[0]),
            "jacobian_GB_per_s": 2.0 * jb * cg_it / cg_s / 1e9,
            "note": "each iteration reads J twice (J u and J' v); vectors FP64"}
    log(f"[large] CGLS: {cgls}")
    del Jd
    fop._J = None
    n, C, D = mesh.nodeCount(), mesh.cellCount(), data.size()
    nE = len(data.sensorPositions())
    # Jacobian compulsory bytes: J written once, potentials read once, cells + geometry read once
    jbytes = D * C * 4.0 + n * nE * 8.0 + C * (4 * 4 + 4 * 3 * 8 + 8 + 4)
    jgbs = jbytes / (best["jacobian_ms"] * 1e-3) / 1e9
    return {
        "metric": "ERT forward+Jacobian/s, 3D large survey (256 electrodes, 10000 data, 1.5M tets), 1 GPU",
        "value": 1e3 / best["ms"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64" if args.method == "mixed" else "fp64", "data": "synthetic",
        "config": {"workload": f"3D ERT: {n}-node / {C}-tet graded mesh, {nE} electrodes (16x16 @ 1 m), "
                               f"{D} dipole-dipole + gradient quadrupoles, log10-rho GRF (2.0, 0.5, 5 m); "
                               f"J {D} x {C} FP32 ({D * C * 4 / 1e9:.0f} GB)",
                   "method": args.method, "rtol": args.rtol, "coarse_aggregates": args.coarse,
                   "learned_warm_start": False, "nodes": n, "cells": C, "electrodes": nE, "data": D, "rhs": nE},
        "accuracy": {"potential_rel_err_fro": best["pot_err"], "max_rel_residual": best["stats"]["max_rel_residual"],
                     "reference": "FP64 block-PCG to rtol 1e-10"},
        "iterations": best["iters"], "phase_ms": best["timings_ms"],
        "roofline": {"bound": "hbm", "kernel": "k_sens (Jacobian)", "achieved": jgbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": jgbs / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": jbytes, "avg_launch_us": best["jacobian_ms"] * 1e3},
        "plain_fp64_pcg": {"ms_per_step": plain["ms"], "iterations": plain["iters"], "potential_rel_err_fro": plain["pot_err"]},
        "plain_fp64_block_pcg": {"ms_per_step": plain_block["ms"], "iterations": plain_block["iters"],
                                 "potential_rel_err_fro": plain_block["pot_err"]},
        "speedup_vs_plain_fp64_block_pcg": plain_block["ms"] / best["ms"],
        "speedup_vs_plain_fp64_pcg": plain["ms"] / best["ms"],
        "gauss_newton_cgls": cgls,
    }


def inversion_leg(args, log):
    """Gauss-Newton inversion loop on the large 3D survey (256 electrodes,
    10000 quadrupoles, 1.5M tets): synthetic data from the log-normal model
    plus 3 % Gaussian noise (seed 3), homogeneous start at the median
    apparent resistivity, log data / log model transforms, first-order
    smoothness, lambda 20, args.gn_iters iterations of RInversion's
    oneStep (CGLS on the 60 GB FP32 Jacobian in HBM, line search with one
    more forward run, Jacobian recomputed every iteration).  The learned warm
    start is trained on the method's own solves of two random fields first
    and fine-tuned on every accepted forward solution of the loop.  Returns
    the result object."""
    import torch
    from pygimli_amd import workloads
    from pygimli_amd.ertModelling import ERTModelling
    from pygimli_amd.inversion import Inversion, smoothness_constraints
    from pygimli_amd.learned import LearnedWarmStart
    t0 = time.time()
    mesh, data, rho = workloads.ert3d_survey_large()
    n, C, D = mesh.nodeCount(), mesh.cellCount(), data.size()
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.jacobianDtype = torch.float32
    fop.solverOptions.update(method=args.method, rtol=args.rtol, coarse=args.coarse)
    dev = fop.device
    truth = fop.response(torch.as_tensor(rho, device=dev))
    rng = np.random.default_rng(3)
    dobs = truth * (1.0 + 0.03 * rng.standard_normal(D))
    Cm = smoothness_constraints(mesh)
    log(f"[inversion] {n} nodes, {C} cells, {D} data, {Cm.shape[0]} constraints ({time.time() - t0:.1f}s)")
    lw = None
    train_s = 0.0
    if args.learned_fields > 0:
        lw = LearnedWarmStart(fop, seed=0)
        t = time.time()
        fields = workloads.learned_fields(mesh, range(300, 302))
        lw.train([torch.as_tensor(f, device=dev) for f in fields], samples_per_model=1 << 19, epochs=4, batch=8192,
                 lr=1e-3)
        train_s = time.time() - t
        fop.solverOptions.update(learned=True)
        log(f"[inversion] warm start trained on 2 fields in {train_s:.1f}s")
    start = np.full(C, float(np.median(dobs)))
    inv = Inversion(fop, dobs, np.full(D, 0.03), Cm, lam=20.0, maxIter=args.gn_iters, stopAtChi1=False,
                    dPhiAbortPercent=-1e300, warmStart=lw, verbose=False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    m = inv.run(start)
    torch.cuda.synchronize()
    total = time.perf_counter() - t
    for i, (c2, tau, tm) in enumerate(zip(inv.chi2History[1:], inv.tauHistory, inv.timings)):
        log(f"[inversion] iteration {i + 1}: chi2 {c2:.3f}, tau {tau:.2f}, {tm}")
    mh = m.cpu().numpy()
    err_log = float(np.sqrt(np.mean((np.log10(mh) - np.log10(rho)) ** 2)))
    return {
        "metric": "ERT Gauss-Newton inversion, 3D large survey (256 electrodes, 10000 data, 1.5M tets), 1 GPU",
        "value": args.gn_iters / total, "unit": "GN iterations/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_step": total / max(inv.iter, 1) * 1e3, "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": f"3D ERT inversion: {n}-node / {C}-tet mesh, {D} data (3 % noise), J {D} x {C} FP32, "
                               f"lambda 20, log transforms, first-order smoothness",
                   "method": args.method, "rtol": args.rtol, "coarse_aggregates": args.coarse,
                   "gn_iterations": inv.iter, "learned_warm_start": lw is not None},
        "seconds_total": total, "warm_start_training_s": train_s,
        "chi2": inv.chi2History, "tau": inv.tauHistory, "iterations": inv.timings,
        "model_rms_log10_error_vs_truth": err_log,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--method", default="mixed")
    ap.add_argument("--cpu-timeout", type=float, default=300.0)
    ap.add_argument("--coarse", type=int, default=1024, help="aggregates of the two-level preconditioner (0: off)")
    ap.add_argument("--smooth", type=int, default=2,
                    help="two-level cycle: 2 = symmetric multiplicative with degree-2 Chebyshev smoothing, "
                         "0 = additive (Jacobi + coarse correction)")
    ap.add_argument("--learned-fields", type=int, default=16,
                    help="random fields the learned warm start is trained on before timing (0: off)")
    ap.add_argument("--batched", type=int, default=0,
                    help="also run the batched 2.5D workload with this many problems (256 = the full config)")
    ap.add_argument("--batch-size", type=int, default=256, help="problems per batched solve")
    ap.add_argument("--batched-method", default="band",
                    help="solver of the batched 2.5D leg: band (direct banded Cholesky) or mixed (block-PCG)")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched 2.5D leg")
    ap.add_argument("--section", action="store_true",
                    help="run only the single 2.5D section leg (41 electrodes, 11 and default wavenumbers) "
                         "beside the reference's own C++ core")
    ap.add_argument("--learned-large", action="store_true",
                    help="run only the learned-warm-start leg on the 1.09M-node / 128-electrode mesh")
    ap.add_argument("--learned-samples", type=int, default=1 << 20, help="training samples per field")
    ap.add_argument("--learned-batch", type=int, default=16384)
    ap.add_argument("--learned-epochs", type=int, default=3)
    ap.add_argument("--cgls-iters", type=int, default=20, help="CGLS iterations of the Gauss-Newton step timed")
    ap.add_argument("--survey-large", action="store_true",
                    help="run only the large 3D survey leg (256 electrodes, 10000 data, 1.5M tets, 60 GB Jacobian)")
    ap.add_argument("--inversion", action="store_true",
                    help="Gauss-Newton inversion loop on the large 3D survey (RInversion with line search, learned "
                         "warm start fine-tuned between iterations)")
    ap.add_argument("--gn-iters", type=int, default=8, help="Gauss-Newton iterations of the --inversion leg")
    ap.add_argument("--complex", action="store_true",
                    help="complex-resistivity leg: 2.5D section with a chargeable block (plus a 3D step)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one 1 GiB device-to-device copy (source evicted from the last-level cache): "
                         "the known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(2**28, dtype=torch.float32, device="cuda")      # 1 GiB
        evict = torch.zeros(2**28, dtype=torch.float32, device="cuda")   # 1 GiB written after src
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4,
                          "checksum": float(dst[::4096].sum() + evict[0])}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(the path would shard by right-hand sides and by cells)"}))
        sys.exit(2)

    cpu_proc = None
    cpu_path = os.path.join(ROOT, "gpurun_out", f"cpu_baseline_{os.getpid()}.json")
    nth = min(8, max(1, (os.cpu_count() or 8) - 2))
    if args.batched_only:
        import multiprocessing as mp
        sample = [0, 96, 192]
        if not args.no_cpu:
            os.makedirs(os.path.dirname(cpu_path), exist_ok=True)
            cpu_proc = mp.get_context("spawn").Process(target=batched_cpu_worker, args=(cpu_path, nth, sample),
                                                       daemon=True)
            cpu_proc.start()
        res = batched_leg(args, log)
        out = {"metric": "batched 2.5D ERT forward+Jacobian, 256 problems per evaluation", "value": res["evals_per_s"],
               "unit": "evals/s", "n_gpus": 1, "higher_is_better": True, "dtype": "fp32+fp64", "data": "synthetic",
               "config": {"workload": res["workload"]}, "batched_2d": res}
        if cpu_proc is not None:
            cpu_proc.join(args.cpu_timeout)
            if cpu_proc.is_alive():
                cpu_proc.terminate()
            try:
                with open(cpu_path) as f:
                    r = json.load(f)
                from pygimli_amd import workloads
                probs = workloads.batched_2d(count=args.batched, seed=2)
                w_all = sum(m.nodeCount() * len(d.sensorPositions()) for m, d, *_ in probs)
                w_s = sum(x["n"] * x["nE"] for x in r)
                t_s = sum(x["tf"] + x["tj"] for x in r)
                t_all = t_s * w_all / w_s
                out["cpu_baseline"] = {
                    "value": 1.0 / t_all, "unit": "evals/s", "cores": nth, "kind": "port",
                    "sample": (f"{len(r)} of {args.batched} problems (indices {[x['i'] for x in r]}) through the oracle: "
                               f"per-wavenumber FP64 assembly + SuperLU factorisation + one solve per source, calc2 C "
                               f"port ({nth} threads) = {t_s:.1f} s, scaled by sum(n*nE) to all problems = "
                               f"{t_all:.0f} s")}
            except Exception as e:
                out["cpu_baseline"] = {"error": f"no batched cpu baseline ({e!r})"}
        print(json.dumps(out))
        return
    if args.section:
        import multiprocessing as mp
        cpu_proc = None
        if not args.no_cpu:
            os.makedirs(os.path.dirname(cpu_path), exist_ok=True)
            cpu_proc = mp.get_context("spawn").Process(target=section_ref_worker, args=(cpu_path, nth, 3),
                                                       daemon=True)
            cpu_proc.start()
        res, mesh, data = section_leg(args, log)
        g = res["k11"]
        out = {"metric": "ERT forward+Jacobian/s, 2.5D section (41 electrodes, 11 wavenumbers), 1 GPU",
               "value": g["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": max(args.steps, 5),
               "warmup": args.warmup + 1, "ms_per_step": g["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
               "config": {"workload": f"2.5D section: {mesh.nodeCount()} nodes / {mesh.cellCount()} triangles, "
                                      f"41 electrodes @ 1 m, {data.size()} dipole-dipole data, 3 layers + block; "
                                      f"direct band solver (FP64), FP64 Jacobian"},
               "section": {k: {kk: vv for kk, vv in v.items() if kk not in ("rhoa", "J")} for k, v in res.items()}}
        if cpu_proc is not None:
            cpu_proc.join(args.cpu_timeout)
            if cpu_proc.is_alive():
                cpu_proc.terminate()
            try:
                with open(cpu_path) as f:
                    r = json.load(f)
                cmp = {}
                for name in ("k11", "default"):
                    rr = np.load(_section_arrays(cpu_path, name, "rhoa"))
                    Jr = np.load(_section_arrays(cpu_path, name, "J"))
                    cmp[name] = {"reference_s": r[name]["s"], "nk": r[name]["nk"],
                                 "speedup": r[name]["s"] / (res[name]["ms_per_step"] * 1e-3),
                                 "rhoa_max_rel_diff": float(np.abs(res[name]["rhoa"] - rr).max() / np.abs(rr).max()),
                                 "J_rel_fro_diff": float(np.linalg.norm(res[name]["J"] - Jr) / np.linalg.norm(Jr))}
                out["cpu_baseline"] = {
                    "value": 1.0 / r["k11"]["s"], "unit": "evals/s", "cores": nth, "kind": "reference",
                    "sample": (f"the whole workload (response + createJacobian) through the reference's own C++ "
                               f"core built from its sources (oracle/_ref), median of 3 after a warm-up, "
                               f"{nth} Jacobian threads; its solver plug-in is an RCM envelope Cholesky "
                               f"(CHOLMOD is absent)"),
                    "comparison": cmp}
            except Exception as e:
                out["cpu_baseline"] = {"error": f"no reference baseline ({e!r})"}
        print(json.dumps(out))
        return
    if args.complex:
        import multiprocessing as mp
        cpu_proc = None
        if not args.no_cpu:
            os.makedirs(os.path.dirname(cpu_path), exist_ok=True)
            cpu_proc = mp.get_context("spawn").Process(target=complex_ref_worker, args=(cpu_path, nth, 1),
                                                       daemon=True)
            cpu_proc.start()
        r2, r3 = complex_leg(args, log)
        out = {"metric": "complex-resistivity ERT forward+Jacobian/s, 2.5D section (41 electrodes), 1 GPU",
               "value": r2["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": max(args.steps, 5),
               "warmup": args.warmup + 1, "ms_per_step": r2["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
               "config": {"workload": r2["workload"] + "; direct band solver (FP64), FP64 complex Jacobian"},
               "complex_2d": {k: v for k, v in r2.items() if k not in ("resp", "Jre", "Jim", "total")},
               "total_2d": {k: v for k, v in r2["total"].items() if k not in ("rhoa", "J")},
               "complex_3d": r3}
        if cpu_proc is not None:
            cpu_proc.join(args.cpu_timeout)
            if cpu_proc.is_alive():
                cpu_proc.terminate()
            try:
                with open(cpu_path) as f:
                    r = json.load(f)
                D = len(r2["resp"]) // 2
                rr, ri, Jr, Ji = (np.load(_section_arrays(cpu_path, "complex", n)) for n in ("re", "im", "Jre", "Jim"))
                out["cpu_baseline"] = {
                    "value": 1.0 / r["s"], "unit": "evals/s", "cores": nth, "kind": "reference",
                    "sample": (f"the whole 2.5D workload (complex response + createJacobian) through the "
                               f"reference's own C++ core (oracle/_ref, setComplex(true)), one run after a "
                               f"warm-up, {nth} Jacobian threads, complex envelope Cholesky as its solver"),
                    "comparison": {
                        "speedup": r["s"] / (r2["ms_per_step"] * 1e-3), "nk": r["nk"],
                        "re_max_rel_diff": float(np.abs(r2["resp"][:D] - rr).max() / np.abs(rr).max()),
                        "im_max_rel_diff": float(np.abs(r2["resp"][D:] - ri).max() / np.abs(ri).max()),
                        "Jre_rel_fro_diff": float(np.linalg.norm(r2["Jre"] - Jr) / np.linalg.norm(Jr)),
                        "Jim_rel_fro_diff": float(np.linalg.norm(r2["Jim"] - Ji) / np.linalg.norm(Ji))}}
                if "total_s" in r:
                    ra = np.load(_section_arrays(cpu_path, "total", "rhoa"))
                    JA = np.load(_section_arrays(cpu_path, "total", "J"))
                    tot = r2["total"]
                    out["total_2d"]["reference"] = {
                        "reference_s": r["total_s"], "speedup": r["total_s"] / (tot["ms_per_step"] * 1e-3),
                        "rhoa_max_rel_diff": float(np.abs(tot["rhoa"] - ra).max() / np.abs(ra).max()),
                        "J_rel_fro_diff": float(np.linalg.norm(tot["J"] - JA) / np.linalg.norm(JA))}
            except Exception as e:
                out["cpu_baseline"] = {"error": f"no reference baseline ({e!r})"}
        print(json.dumps(out))
        return
    if args.learned_large:
        print(json.dumps(learned_large_leg(args, log)))
        return
    if args.inversion:
        print(json.dumps(inversion_leg(args, log)))
        return
    if args.survey_large:
        if not args.no_cpu:
            import multiprocessing as mp
            os.makedirs(os.path.dirname(cpu_path), exist_ok=True)
            cpu_proc = mp.get_context("spawn").Process(
                target=cpu_baseline_worker, args=(cpu_path, nth, 0.002, 10, "ert3d_survey_large"), daemon=True)
            cpu_proc.start()
        out = survey_large_leg(args, log)
        if cpu_proc is not None:
            cpu_proc.join(args.cpu_timeout)
            if cpu_proc.is_alive():
                cpu_proc.terminate()
            out["cpu_baseline"] = cpu_baseline_result(cpu_path, nth, out["plain_fp64_pcg"]["iterations"])
        print(json.dumps(out))
        return
    if not args.no_cpu:
        import multiprocessing as mp
        os.makedirs(os.path.dirname(cpu_path), exist_ok=True)
        ctx = mp.get_context("spawn")  # before anything touches the GPU
        cpu_proc = ctx.Process(target=cpu_baseline_worker, args=(cpu_path, nth, 0.02, 30, "ert3d_grid_grf", 50),
                               daemon=True)
        cpu_proc.start()
        log(f"[bench] cpu baseline started in pid {cpu_proc.pid} ({nth} threads)")

    import torch
    from pygimli_amd import workloads
    from pygimli_amd.ertModelling import ERTModelling

    t0 = time.time()
    mesh, data, rho = workloads.ert3d_grid_grf()
    log(f"[bench] workload built in {time.time() - t0:.1f}s: {mesh}, {data.size()} data")
    fop = ERTModelling()
    fop.setMesh(mesh)
    fop.setData(copy.deepcopy(data))
    fop.jacobianDtype = torch.float32
    dev = fop.device
    model = torch.as_tensor(rho, device=dev)
    log(f"[bench] set-up done at {time.time() - t0:.1f}s")

    # accurate reference potentials: FP64 block-PCG to 1e-11
    fop.solverOptions.update(method="block64", rtol=1e-11, coarse=args.coarse, learned=False)
    fop.response_device(model)
    U_ref = fop.subPotentials(node_major=True).clone()
    log(f"[bench] reference solve: {fop.lastStats} {fop.timings()}")

    import ctypes
    L, h = fop._ctx.L, fop._ctx.h
    n, nb = mesh.nodeCount(), 1 * ((64 + 63) // 64)

    def run(method, steps, warmup, coarse=0, learned=False, timed_kernel=0, smooth=None):
        fop.solverOptions.update(method=method, rtol=args.rtol, coarse=coarse, learned=learned,
                                 smooth=args.smooth if smooth is None else smooth)
        for _ in range(warmup):
            fop.response_device(model)
            fop.createJacobian(model)
        tot, nl, bpl = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        L.pgert_kernel_timing(h, timed_kernel, 1, None, None, None)
        torch.cuda.synchronize()
        t = time.perf_counter()
        iters = []
        for _ in range(steps):
            fop.response_device(model)
            fop.createJacobian(model)
            iters.append(fop.lastStats["iterations"])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        L.pgert_kernel_timing(h, 0, 0, ctypes.byref(tot), ctypes.byref(nl), ctypes.byref(bpl))
        U = fop.subPotentials(node_major=True)
        err = float(torch.linalg.norm(U - U_ref) / torch.linalg.norm(U_ref))
        colerr = float((torch.linalg.norm(U - U_ref, dim=1) / torch.linalg.norm(U_ref, dim=1)).max())
        res = dict(ms=dt * 1e3, iters=int(np.mean(iters)), pot_err=err, pot_err_max_col=colerr,
                   stats=dict(fop.lastStats), timings_ms=fop.timings(),
                   k_ms=tot.value / max(nl.value, 1), k_launches=nl.value, k_bytes=bpl.value)
        log(f"[bench] {method} coarse={coarse} learned={learned}: {res}")
        return res

    plain = plain_block = block_coarse = None
    if not args.no_plain:
        plain = run("pcg64", 1, 0)
        plain_block = run("block64", 1, 0, coarse=0)
        if args.coarse and args.method != "block64":
            block_coarse = run("block64", 1, 0, coarse=args.coarse)

    # learned warm start: trained on the method's own solves of random
    # log-normal fields (outside the timed region)
    learned_info = None
    use_learned = args.learned_fields > 0
    if use_learned:
        from pygimli_amd.learned import LearnedWarmStart
        cc = mesh.cellCenters()
        fields = [10.0 ** workloads.gaussian_random_field(cc, 2.0, 0.3 + 0.1 * (i % 6), 2.0 + 3.0 * (i % 7), 100 + i,
                                                           (-30, -30, -40), (30, 30, 0))
                  for i in range(args.learned_fields)]
        fop.solverOptions.update(method=args.method, rtol=args.rtol, coarse=args.coarse, smooth=args.smooth)
        lw = LearnedWarmStart(fop, seed=0)
        t = time.time()
        losses = lw.train([torch.as_tensor(f, device=dev) for f in fields], samples_per_model=65536, epochs=20,
                          batch=4096, lr=1e-3)
        learned_info = {"fields": args.learned_fields, "samples": 65536 * args.learned_fields, "epochs": 20,
                        "train_s": time.time() - t, "loss_first": losses[0], "loss_last": losses[-1]}
        log(f"[bench] learned warm start trained: {learned_info}")

    # the timed steps run the product path: no per-kernel events, so the block
    # solver replays its captured iteration graphs
    best = run(args.method, args.steps, args.warmup, args.coarse if args.method != "pcg64" else 0,
               learned=use_learned, timed_kernel=0)
    ablation = run(args.method, 1, 0, args.coarse, learned=False) if use_learned else None
    # the same with the additive two-level preconditioner (the cycle's gain)
    additive = (run(args.method, 1, 0, args.coarse, learned=use_learned, smooth=0)
                if args.coarse and args.smooth and args.method != "pcg64" else None)

    # per-kernel time per launch (one solve each) and the roofline of the
    # dominant one: algorithmic bytes = each vector read/written once, flops =
    # the MFMA products' 2 m n k
    vt = 4 if args.method == "mixed" else 8
    vec = float(nb) * n * 64 * vt
    kernels = {1: ("k_spmm", None, None), 2: ("k_update", 3 * vec, 2 * 64 * 64 * n * nb),
               3: ("k_gram<Z>", 3 * vec, 2 * 3 * 64 * 64 * n * nb), 4: ("k_pupdate", 3 * vec, 2 * 2 * 64 * 64 * n * nb),
               5: ("k_gram<PQR>", 3 * vec, 2 * 2 * 64 * 64 * n * nb),
               6: ("k_small_alpha", None, None), 7: ("k_small_beta", None, None),
               8: ("k_zform", 2 * vec + 12 * n * nb, None), 9: ("k_restrict", vec, None),
               10: ("k_coarse_apply", None, None), 11: ("k_part_reduce", None, None)}
    if args.smooth and args.coarse:
        # the smoothed cycle: its sparse passes' bytes come from the library
        # (matrix + gathers + row vectors); Z is formed by the post-smoothing pass
        kernels[8] = ("k_smooth<SZ>", None, None)
        kernels[12] = ("k_smooth<S0,S1>", None, None)
        kernels[13] = ("k_coarse_resid", None, None)
    lib_bytes = {1, 8, 12, 13} if args.smooth and args.coarse else {1}
    mfma_peak = 157.3e12 if args.method == "mixed" else 78.6e12
    # per-kernel breakdown: one instrumented evaluation per kernel (HIP events
    # around every launch of that kernel on the solver stream; eager launches),
    # separate from the timed steps
    breakdown = {}
    if args.method != "pcg64":
        for kid, (name, kb, _) in kernels.items():
            if kid in (9, 10) and not args.coarse:
                continue
            r = run(args.method, 1, 0, args.coarse, learned=use_learned, timed_kernel=kid)
            # algorithmic bytes: the SpMM's from the library (matrix + gathers),
            # the streaming kernels' from the vector count above, none for the
            # latency-bound small kernels
            breakdown[name] = {"us_per_launch": r["k_ms"] * 1e3, "launches": r["k_launches"],
                               "bytes": r["k_bytes"] if kid in lib_bytes else kb,
                               "ms_per_step": r["k_ms"] * r["k_launches"]}
    # the dominant kernel among those with an algorithmic byte / flop count
    costed = [k for k in breakdown if breakdown[k]["bytes"]]
    dom = max(costed, key=lambda k: breakdown[k]["ms_per_step"]) if costed else None

    def roof(name, us, nbytes, flops):
        t_h = nbytes / (HBM_PEAK_GBS * 1e9)
        t_m = flops / mfma_peak if flops else 0.0
        hbm = {"achieved": nbytes / (us * 1e-6) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        hbm["frac"] = hbm["achieved"] / HBM_PEAK_GBS
        if t_m > t_h:
            r = {"bound": "mfma", "achieved": flops / (us * 1e-6) / 1e12, "peak": mfma_peak / 1e12, "unit": "TFLOP/s"}
            r["frac"] = r["achieved"] / r["peak"]
            r["hbm_view"] = hbm
        else:
            r = {"bound": "hbm", **hbm}
        r.update(kernel=name, avg_launch_us=us, bytes_per_launch=nbytes, flops_per_launch=flops,
                 traffic=traffic_from_profiles(name))
        return r

    if dom is not None:
        kid = [k for k, v in kernels.items() if v[0] == dom][0]
        nbytes = breakdown[dom]["bytes"]
        roofline = roof(dom, breakdown[dom]["us_per_launch"], nbytes, kernels[kid][2])
    else:
        roofline = roof("k_spmm_dot", best["k_ms"] * 1e3, best["k_bytes"], None)
    spmm = breakdown.get("k_spmm")
    out = {
        "metric": METRIC,
        "value": 1e3 / best["ms"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if args.method == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": "3D ERT: 201898-node / 1150488-tet graded mesh, 64 electrodes (8x8 @ 2 m), "
                        "2000 random quadrupoles, log10-rho GRF (2.0, 0.5, 5 m); J 2000 x 1150488 FP32",
            "method": args.method, "rtol": args.rtol, "coarse_aggregates": args.coarse,
            "preconditioner": ("two-level multiplicative, degree-2 Chebyshev smoothing" if args.smooth
                               else "two-level additive (Jacobi + coarse)") if args.coarse else "Jacobi",
            "learned_warm_start": use_learned, "nodes": n, "cells": mesh.cellCount(),
            "electrodes": 64, "data": data.size(), "rhs": 64,
        },
        "accuracy": {"potential_rel_err_fro": best["pot_err"], "potential_rel_err_max_col": best["pot_err_max_col"],
                     "max_rel_residual": best["stats"]["max_rel_residual"],
                     "reference": "FP64 block-PCG to rtol 1e-11"},
        "iterations": best["iters"],
        "phase_ms": best["timings_ms"],
        "roofline": roofline,
        "kernels": breakdown,
    }
    if spmm:
        out["spmm"] = {"GB/s": spmm["bytes"] / (spmm["us_per_launch"] * 1e-6) / 1e9,
                       "frac_hbm": spmm["bytes"] / (spmm["us_per_launch"] * 1e-6) / 1e9 / HBM_PEAK_GBS,
                       "bytes_per_launch": spmm["bytes"], "us_per_launch": spmm["us_per_launch"]}
    if learned_info:
        learned_info.update(iterations_without=ablation["iters"], ms_without=ablation["ms"])
        out["learned"] = learned_info
    if additive is not None:
        out["additive_two_level"] = {"value": 1e3 / additive["ms"], "ms_per_step": additive["ms"],
                                     "iterations": additive["iters"], "potential_rel_err_fro": additive["pot_err"]}
    if block_coarse is not None:
        out["fp64_block_pcg_coarse"] = {"value": 1e3 / block_coarse["ms"], "ms_per_step": block_coarse["ms"],
                                        "iterations": block_coarse["iters"],
                                        "potential_rel_err_fro": block_coarse["pot_err"]}
        out["speedup_vs_fp64_block_pcg_coarse"] = block_coarse["ms"] / best["ms"]
    if plain is not None:
        for key, r in (("plain_fp64_pcg", plain), ("plain_fp64_block_pcg", plain_block)):
            out[key] = {"value": 1e3 / r["ms"], "ms_per_step": r["ms"], "iterations": r["iters"],
                        "potential_rel_err_fro": r["pot_err"], "potential_rel_err_max_col": r["pot_err_max_col"]}
        out["speedup_vs_plain_fp64_block_pcg"] = plain_block["ms"] / best["ms"]
        out["speedup_vs_plain_fp64_pcg"] = plain["ms"] / best["ms"]
    if args.batched > 0:
        out["batched_2d"] = batched_leg(args, log)
    if cpu_proc is not None:
        t_wait = time.time()
        while cpu_proc.is_alive() and time.time() - t_wait < args.cpu_timeout:
            cpu_proc.join(30)
            log(f"[bench] waiting for the cpu baseline ({time.time() - t_wait:.0f} s)")
        timed_out = cpu_proc.is_alive()
        if timed_out:
            cpu_proc.terminate()
        elif not os.path.exists(cpu_path):
            # the child ended without a result (exit code in the log): once more,
            # now that the GPU work is done
            log(f"[bench] cpu baseline child ended with exit code {cpu_proc.exitcode} and no result; running it again")
            cpu_proc = mp.get_context("spawn").Process(target=cpu_baseline_worker,
                                                       args=(cpu_path, nth, 0.02, 30, "ert3d_grid_grf", 50), daemon=True)
            cpu_proc.start()
            cpu_proc.join(args.cpu_timeout)
            if cpu_proc.is_alive():
                cpu_proc.terminate()
        out["cpu_baseline"] = cpu_baseline_result(cpu_path, nth, plain["iters"] if plain else None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

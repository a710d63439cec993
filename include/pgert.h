/*
 * pgert.h -- C-ABI of the MI355X-native ERT forward + Jacobian library
 * (libpgert.so, HIP kernels for gfx950).
 *
 * It replaces, behind pyGIMLi's own Python API, the native core that
 * pygimli.physics.ert.ERTModelling drives through Boost.Python:
 *   DCSRMultiElectrodeModelling  (core/src/bert/dcfemmodelling.h:344-415)
 *   createSensitivityCol         (core/src/bert/bertJacobian.h:43)
 *   RSparseMatrix fill*Matrix    (core/src/sparsematrix.h:1039-1065)
 *   CHOLMODWrapper / LinSolver   (core/src/cholmodWrapper.cpp, linSolver.cpp)
 *
 * Conventions
 *  - Every call returns 0 on success and a negative code on failure; the
 *    message is available from pgert_last_error(ctx).  This mirrors the
 *    reference's throwError / log(Critical) exceptions (core/src/gimli.cpp:439).
 *  - Set-up calls (pgert_set_*) take HOST pointers and copy.
 *  - Compute calls take DEVICE pointers (memory of the context's device) for
 *    their per-evaluation inputs and outputs; all work is queued on the
 *    context's stream (pgert_set_stream, default: the null stream) and the
 *    call returns after the stream is synchronised unless stated otherwise.
 *  - A context is not thread-safe; use one context per host thread.
 */
#ifndef PGERT_H
#define PGERT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgert_ctx pgert_ctx;

/* Solver selection for pgert_forward. */
enum pgert_method {
    /* independent-column Jacobi-PCG in FP64 (the plain method) */
    PGERT_PCG_FP64 = 0,
    /* block-PCG (all electrodes of one wavenumber share one Krylov block), FP64 */
    PGERT_BLOCK_FP64 = 1,
    /* block-PCG with FP32 SpMM/inner iterations and FP64 residual refinement */
    PGERT_BLOCK_MIXED = 2,
    /* direct: banded Cholesky of every A_k (2D meshes; the reference's
       sparse direct solve, linSolver.cpp:96-178 / cholmodWrapper.cpp:224-459) */
    PGERT_BAND_FP64 = 3,
};

typedef struct pgert_solve_opts {
    int method;          /* enum pgert_method */
    double rtol;         /* relative residual ||b - A x|| / ||b|| per column */
    int max_iter;        /* cap on (inner) iterations per solve */
    int check_every;     /* convergence is read back every check_every iterations */
    double inner_rtol;   /* PGERT_BLOCK_MIXED: inner FP32 reduction per refinement step */
    const double * warm_start; /* optional DEVICE pointer [nK][n][nE] secondary-field guess, NULL = zero */
    int learned_warm_start;    /* 1: start from the learned network's prediction (pgert_mlp_*) */
    int coarse;                /* 1: two-level preconditioner (block methods, see pgert_set_coarse) */
    int no_graph;              /* 1: launch the block iterations one by one instead of replaying a
                                  captured HIP graph of check_every iterations */
    int report_only;           /* 0: a solve that misses rtol is repeated from the analytic initial
                                  guess (when it started from a learned or given guess), then with
                                  FP64 block-PCG (when it was mixed), and the call fails if that
                                  also misses; 1: return with stats.converged = 0 instead */
    int smooth;                /* block methods with coarse = 1: 0 = additive two-level preconditioner
                                  (Jacobi + coarse correction); 2 = symmetric multiplicative two-level
                                  cycle with degree-2 Chebyshev smoothing (three sparse products more
                                  per iteration, several times fewer iterations) */
} pgert_solve_opts;

typedef struct pgert_solve_stats {
    int iterations;        /* iterations until this problem's last block converged */
    int max_iterations;    /* iterations of the whole solve (all problems of a batch) */
    int refinements;       /* outer FP64 refinement steps (mixed) */
    int converged;         /* 1 if every column met rtol */
    double max_rel_residual; /* final FP64 ||b - A x|| / ||b||, max over columns */
    int fallback;          /* bit 0: re-solved from the analytic initial guess; bit 1: re-solved in FP64 */
} pgert_solve_stats;

/* Context ----------------------------------------------------------------- */
int  pgert_create(pgert_ctx ** ctx, int device);
void pgert_destroy(pgert_ctx * ctx);
const char * pgert_last_error(const pgert_ctx * ctx);
/* stream: a hipStream_t (e.g. torch.cuda.current_stream().cuda_stream) */
int  pgert_set_stream(pgert_ctx * ctx, void * stream);

/* Mesh (ModellingBase::setMesh, core/src/modellingbase.cpp:179-211).
 * dim 2 (triangles) or 3 (tetrahedra).  xyz: n x 3; cells: C x (dim+1);
 * cell_marker: C (Jacobian column = marker, bertJacobian.cpp:191,380);
 * bounds: F x dim node ids, bound_marker: F (-1 Neumann, -2 mixed, -3
 * homogeneous Dirichlet: the rows and columns of its nodes are cleared and
 * the diagonal set to one in every assembled matrix,
 * dcfemmodelling.cpp:280-298; -4 is refused, as the reference throws);
 * bound_cell: F, the cell adjacent to each boundary (Boundary::leftCell).
 * Builds the CSR pattern once (SparseMatrix::buildSparsityPattern,
 * sparsematrix.h:966-1032) and the element geometry on the device. */
int pgert_set_mesh(pgert_ctx * ctx, int dim, int n_nodes, const double * xyz,
                   int n_cells, const int * cells, const int * cell_marker,
                   int n_bounds, const int * bounds, const int * bound_marker,
                   const int * bound_cell);

/* pgert_set_mesh whose matrix pattern also couples the n_couple node pairs
 * (couple_a[i], couple_b[i]) that may share no cell: bypass resistors between
 * distant nodes, which the reference adds to a map matrix and converts back,
 * widening the pattern (assembleStiffnessMatrixDCFEMByPass_,
 * dcfemmodelling.cpp:695-707).  The coupled slots hold zero after assembly
 * until pgert_set_bypass puts a conductance there.  HOST arrays. */
int pgert_set_mesh_coupled(pgert_ctx * ctx, int dim, int n_nodes, const double * xyz,
                           int n_cells, const int * cells, const int * cell_marker,
                           int n_bounds, const int * bounds, const int * bound_marker,
                           const int * bound_cell, int n_couple, const int * couple_a,
                           const int * couple_b);

/* Quadratic (P2) mesh for the primary potentials of the SR operator with
 * topography: Mesh::createP2 (mesh.cpp:1285-1330) of a linear mesh, on which
 * the reference solves the total-field problem with rho = 1
 * (DCSRMultiElectrodeModelling::checkPrimpotentials_, dcfemmodelling.cpp:
 * 2009-2056).  cells: C x 6 (triangles) or C x 10 (tetrahedra), the vertices
 * first, then the edge nodes (0,1), (1,2), (0,2) or (0,1), (0,2), (0,3),
 * (1,2), (1,3), (2,3); bounds: F x 3 (edges: two vertices, the midpoint) or
 * F x 6 (triangles as the cells); other arguments as pgert_set_mesh.  Element
 * matrices are the exact P2 integrals (the reference's quadrature of orders
 * 2 and 4 is exact for them, elementmatrix.cpp:761-782, :872-1060).  On such
 * a context pgert_forward_total, pgert_potentials(_node_major), pgert_assemble
 * and pgert_fill_matrix run; the SR forward, the Jacobians and free
 * electrodes are refused. */
int pgert_set_mesh_p2(pgert_ctx * ctx, int dim, int n_nodes, const double * xyz,
                      int n_cells, const int * cells, const int * cell_marker,
                      int n_bounds, const int * bounds, const int * bound_marker,
                      const int * bound_cell);

/* Electrodes after DCMultiElectrodeModelling::searchElectrodes_
 * (dcfemmodelling.cpp:790-1073): node id of every current electrode (pole
 * sources, createCurrentPattern :1482-1525), the mean electrode position used
 * by the mixed boundary condition (sourceCenterPos_, :1025-1030) and the
 * Neumann surface height surfaceZ_ (:711-773; -DBL_MAX = full space). */
int pgert_set_electrodes(pgert_ctx * ctx, int n_elec, const int * elec_node,
                         const double * source_center, double surface_z);

/* Electrodes that may lie off the nodes (searchElectrodes_' fallback,
 * dcfemmodelling.cpp:979-985): elec_node[e] >= 0 as in pgert_set_electrodes;
 * elec_node[e] = -1 a free electrode in cell elec_cell[e] at elec_pos[e]
 * (HOST [n_elec][3], mesh coordinates), modelled as the reference's
 * ElectrodeShapeEntity (electrode.cpp:198-287): shape-function point source
 * on the cell's nodes, potential interpolated there, rho_s = the cell's
 * resistivity, primary potential from its position (singular value only at
 * a cell node within 1e-4). */
int pgert_set_electrodes_free(pgert_ctx * ctx, int n_elec, const int * elec_node, const int * elec_cell,
                              const double * elec_pos, const double * source_center, double surface_z);

/* Wavenumbers and weights (setkValues / setWeights, dcfemmodelling.h:239-243;
 * 3D: k = {0}, w = {1}). */
int pgert_set_wavenumbers(pgert_ctx * ctx, int nk, const double * k, const double * w);

/* Data: electrode indices a, b, m, n (-1 = none) and geometric factors k
 * (DataContainerERT tokens). */
int pgert_set_data(pgert_ctx * ctx, int n_data, const int * a, const int * b,
                   const int * m, const int * n, const double * kfac);

/* Two-level preconditioner for the block solvers (no counterpart in the
 * reference, whose direct solver needs none): n_agg aggregates (rounded down
 * to a power of two, at most n/4) by recursive coordinate bisection of the
 * nodes; 0 switches it off.  Used by pgert_forward when opts.coarse != 0. */
int pgert_set_coarse(pgert_ctx * ctx, int n_agg);

/* Query sizes: nnz of the CSR pattern, number of parameters (max marker + 1). */
int pgert_info(const pgert_ctx * ctx, int * nnz, int * n_param);

/* Copy the CSR pattern (row pointer n+1, column index nnz) to HOST memory
 * (RSparseMatrix vecColPtr / vecRowIdx, sparsematrix.h:1072-1076). */
int pgert_pattern(const pgert_ctx * ctx, int * rowptr, int * colidx);

/* Cell resistivity from a model vector (mapERTModel, dcfemmodelling.cpp:1211;
 * createMappedModel, modellingbase.cpp:401-470): len == n_cells -> copy,
 * else rho_c = model[marker_c] for cells with a marker >= 0 and cells with a
 * negative marker prolongated from their face neighbours, weighted by the
 * horizontal part of the face normal (Mesh::prolongateEmptyCellsValues,
 * mesh.cpp:2247-2318, background -9e99 as in response(model)).  Fails for
 * fixed-value regions (marker <= -1000000, region manager).  DEVICE pointers. */
int pgert_map_model(pgert_ctx * ctx, const double * model, int n_model, double * rho);

/* System matrix values of one wavenumber on the CSR pattern, exactly as
 * DCSRMultiElectrodeModelling::calculateK assembles them
 * (dcfemDomainAssembleStiffnessMatrix + dcfemBoundaryAssembleStiffnessMatrix,
 * dcfemmodelling.cpp:163-292).  rho: DEVICE [n_cells]; vals: DEVICE [nnz]. */
int pgert_assemble(pgert_ctx * ctx, const double * rho, double k, double * vals);

/* RSparseMatrix::fillStiffnessMatrix(mesh, a) (mass = 0) or
 * fillMassMatrix(mesh, a) (mass = 1) (sparsematrix.h:1039-1065), the back end
 * of pg.solver.createStiffnessMatrix / createMassMatrix.  DEVICE pointers. */
int pgert_fill_matrix(pgert_ctx * ctx, int mass, const double * a, double * vals);

/* Forward response (DCMultiElectrodeModelling::response,
 * dcfemmodelling.cpp:1085-1197, SR variant :2152-2296).
 * rho: DEVICE [n_cells] cell resistivities; rhoa: DEVICE [n_data] output.
 * The per-wavenumber total potentials are kept in the context for
 * pgert_potentials / pgert_jacobian. */
int pgert_forward(pgert_ctx * ctx, const double * rho, const pgert_solve_opts * opts,
                  double * rhoa, pgert_solve_stats * stats);

/* Forward response of `count` independent problems (contexts set up like
 * for pgert_forward, on one device) with ONE block solve over all their
 * wavenumber blocks: every solver kernel launch covers every block of every
 * problem, and the iterations are replayed from a captured HIP graph.  The
 * per-problem steps (assembly, primaries, right-hand sides, response) are the
 * ones of pgert_forward.  rhos[i] / rhoas[i]: DEVICE pointers of problem i;
 * stats: optional array of count entries.  Everything runs on ctxs[0]'s
 * stream.  A problem that misses rtol fails the call (no fallback solves in a
 * batch) unless opts.report_only is set.  Block methods and the band solver (PGERT_BLOCK_FP64,
 * PGERT_BLOCK_MIXED, PGERT_BAND_FP64: one launch factors every wavenumber
 * of every problem, one launch solves them).
 * The reference runs its problems one after another (each a
 * DCMultiElectrodeModelling::response, dcfemmodelling.cpp:1085-1197); the
 * batch has no counterpart there. */
int pgert_batch_forward(pgert_ctx * const * ctxs, int count, const double * const * rhos,
                        const pgert_solve_opts * opts, double * const * rhoas, pgert_solve_stats * stats);

/* Per-wavenumber unweighted total potentials of the last forward in the
 * reference's subSolutions_ layout: row = electrode + nE * kIdx, col = node
 * (dcfemmodelling.cpp:1681-1682).  DEVICE [nK * nE * n]. */
int pgert_potentials(pgert_ctx * ctx, double * out);

/* Total-field operator, DCMultiElectrodeModelling (ertModelling.py:103-106 with
 * sr=False): the system of pgert_forward without singularity removal, unit
 * point sources at the electrode nodes (calculateK_, dcfemmodelling.cpp:1755-1928,
 * ElectrodeShapeNode::assembleRHS electrode.cpp:133-151), potentials = the
 * solution, then the same response (rounding to 1e-10, reciprocal geometric
 * mean, :1140-1196).  pgert_jacobian and pgert_potentials work on its
 * potentials as on pgert_forward's.  Arguments as pgert_forward (warm_start:
 * a total-field guess). */
int pgert_forward_total(pgert_ctx * ctx, const double * rho, const pgert_solve_opts * opts, double * rhoa,
                        pgert_solve_stats * stats);

/* Reference electrode and calibration nodes of the total-field operator
 * (searchElectrodes_, dcfemmodelling.cpp:1010-1071): ref_node >= 0 (the first
 * node with marker -999) makes every current pattern a dipole against it
 * (createCurrentPattern :1517-1523, right-hand side +1 at the electrode, -1
 * there); the n_cal calibration nodes (marker -1000, or node 0 in a Neumann
 * domain without one; none outside Neumann domains) get cleared rows and
 * columns and a unit diagonal after the bypass resistors (:1835), the
 * right-hand side unchanged.  ref_node = -1 and n_cal = 0 remove them.  The
 * SR forward takes both: dipole primaries (the reference electrode's
 * analytic potential subtracted, rho_s the mean of both geometric means,
 * checkPrimpotentials_ :2101-2102, calculateK :2219-2244) and the calibration
 * nodes in both matrices with their right-hand side zeroed (:2182-2192,
 * :2280-2283); the complex forward takes both (a complex unit diagonal).
 * HOST array. */
int pgert_set_reference(pgert_ctx * ctx, int ref_node, int n_cal, const int * cal_nodes);

/* Complete electrode model of the total-field operator
 * (assembleCompleteElectrodeModel_, dcfemmodelling.cpp:315-410, applied in
 * calculateK_ :1790-1829): n_terms constant entries vals[i] added at
 * (rows[i], cols[i]) of every wavenumber's matrix before the bypass
 * resistors.  The electrode unknowns are mesh nodes of no cell, appended
 * after the cell nodes and coupled to their electrode's face nodes through
 * pgert_set_mesh_coupled; the electrodes are set on those nodes
 * (pgert_set_electrodes), so the unit current enters at the unknown and the
 * electrode potential is its value.  Terms per electrode e with unknown v,
 * contact impedance z and faces f: (i, v) and (v, i) -int_f N_i / z,
 * (i, j) int_f N_i N_j / z, (v, v) area / z.  n_terms = 0 removes them.
 * HOST arrays. */
int pgert_set_cem(pgert_ctx * ctx, int n_terms, const int * rows, const int * cols, const double * vals);

/* Numerical primary potentials for the SR operator, in place of the analytic
 * half-space ones (DCSRMultiElectrodeModelling::setPrimaryPotential,
 * dcfemmodelling.h:374; computed by checkPrimpotentials_,
 * dcfemmodelling.cpp:2009-2056, when the mesh has topography): prim DEVICE
 * [nK][n][nE] (the layout of pgert_potentials_node_major), unit-current
 * potentials of every electrode and wavenumber at the mesh nodes.  They hold
 * until the mesh, the electrodes, the reference electrode or the
 * wavenumbers change; pgert_forward then fails until they are set again.
 * With a reference electrode they are the dipole primaries (the P2 problem
 * solved with the same pgert_set_reference); rho_s is then the mean of both
 * geometric means, as for the analytic ones. */
int pgert_set_primary_potentials(pgert_ctx * ctx, const double * prim);

/* Bypass resistors of the total-field operator
 * (assembleStiffnessMatrixDCFEMByPass_, dcfemmodelling.cpp:601-707, applied in
 * calculateK_ :1833 only): conductance[i] = 1/R between nodes a[i] and b[i],
 * added to every wavenumber's matrix by pgert_forward_total.  The pairs must
 * share a cell or have been coupled by pgert_set_mesh_coupled.  n_pairs = 0
 * removes them. HOST arrays. */
int pgert_set_bypass(pgert_ctx * ctx, int n_pairs, const int * a, const int * b, const double * conductance);

/* Complex resistivity, DCMultiElectrodeModelling with setComplex(true)
 * (ertModelling.py:217-238 with sr=False; dcfemmodelling.cpp:1103-1120 response,
 * :163-292 assembly with 1/rho complex, :1755-1928 calculateK_ total-field
 * potentials of unit point sources).  rho_re, rho_im: DEVICE [n_cells]
 * (pgert_map_model of the real and imaginary model halves); resp_re, resp_im:
 * DEVICE [n_data], k * ((uAM - uAN) - (uBM - uBN)) of the real and imaginary
 * potentials (no rounding, no reciprocity, as the reference's complex branch).
 * The complex symmetric systems are solved by real/imaginary block Gauss-Seidel
 * sweeps, each half with opts' real solver; fails when cell phases are too large
 * for the sweeps to converge. */
int pgert_forward_complex(pgert_ctx * ctx, const double * rho_re, const double * rho_im,
                          const pgert_solve_opts * opts, double * resp_re, double * resp_im,
                          pgert_solve_stats * stats);

/* Complex Jacobian of the last pgert_forward_complex (createJacobian,
 * dcfemmodelling.cpp:1446-1463 -> createJacobian_ :1410-1444): sensitivities of
 * the complex potentials (no conjugation), divided by m_j^2 / k_i (complex m)
 * when n_model == n_param.  model_re, model_im: DEVICE [n_model]; J_re, J_im:
 * DEVICE [n_data x n_param] FP64 row-major. */
int pgert_jacobian_complex(pgert_ctx * ctx, const double * model_re, const double * model_im, int n_model,
                           double * J_re, double * J_im);

/* Same potentials, node-major as held in HBM: [nK][n][nE].  DEVICE. */
int pgert_potentials_node_major(pgert_ctx * ctx, double * out);

/* Jacobian of the last forward (createJacobian_, dcfemmodelling.cpp:1318-1408;
 * createSensitivityCol calc2, bertJacobian.cpp:170-219, :367-521).
 * model: DEVICE [n_model]; J: DEVICE [n_data x n_param] row-major, FP64
 * (dtype 0) or FP32 (dtype 1).  Scaled by k_i / m_j^2 when n_model == n_param. */
int pgert_jacobian(pgert_ctx * ctx, const double * model, int n_model, void * J, int dtype);

/* Jacobians of several problems after a forward of each (pgert_forward or
 * pgert_batch_forward), in one launch per wavenumber pass over the cells of
 * every problem.  The reference computes them one model at a time
 * (DCMultiElectrodeModelling::createJacobian, dcfemmodelling.cpp:1446-1476);
 * this is that call for `count` problems.  All contexts on one device, same
 * dimension; runs on ctxs[0]'s stream.  models[i]: DEVICE [n_models[i]];
 * Js[i]: DEVICE [n_data_i x n_param_i], FP64 (dtype 0) or FP32 (dtype 1).
 * The launch time is reported in ctxs[0]'s pgert_last_timings()[4]. */
int pgert_batch_jacobian(pgert_ctx * const * ctxs, int count, const double * const * models, const int * n_models,
                         void * const * Js, int dtype);

/* Constraint matrix C of the Gauss-Newton step (ModellingBase::constraints,
 * used by RInversion::oneStep, inversion.cpp:308): HOST CSR [n_const x
 * n_model], sorted or unsorted columns.  Copied with its transpose to the
 * device; replaces the previous one.  n_const = 0: no constraints. */
int pgert_set_constraints(pgert_ctx * ctx, int n_const, int n_model, const int * rowptr, const int * colidx,
                          const double * vals);

/* Regularised, transformed CGLS of one Gauss-Newton step
 * (solveCGLSCDWWhtrans, core/src/solver.cpp:27-224):
 *   min || dW (b - td * J (x / tm)) ||^2 + lambda || wc (C (wm x)) + wc roughness ||^2
 * by conjugate gradients on the normal equations, starting from x, with the
 * reference's stopping rule (|r|^2 <= tol; tol < 0: max(1e-12, 1e-8 |r0|^2),
 * r0 = J'(b dW^2 td)/tm - lambda wm C'(wc roughness)) and its rounding of
 * C (p wm) to 1e-10.  J: DEVICE [n_data x n_model] row-major, FP64 (dtype 0)
 * or FP32 (dtype 1), e.g. the output of pgert_jacobian.  d_weight, b, td:
 * DEVICE [n_data]; x (in/out), wm, tm: DEVICE [n_model]; wc, roughness:
 * DEVICE [n_const] (C from pgert_set_constraints).  All FP64.  Returns the
 * iteration count and the final |r|^2. */
int pgert_cgls(pgert_ctx * ctx, const void * J, int dtype, int n_data, int n_model, const double * d_weight,
               const double * b, double * x, const double * wc, const double * wm, const double * tm,
               const double * td, double lambda, const double * roughness, int max_iter, double tol,
               int * iterations, double * norm_r2);

/* Logarithmic transforms of the Gauss-Newton loop, elementwise on DEVICE
 * vectors of length n: TransLogLU(lower, upper) (core/src/trans.h:292-322,
 * trans.cpp:24-63), with upper = 0 TransLog(lower) (trans.h:241-287):
 *   op 0: out = trans(a)      1: invTrans(a)      2: deriv(a)
 *   op 3: out = invTrans(trans(a) + tau b)         (Trans::update, trans.h:76-79)
 *   op 4: out = |a b deriv(a)|, b relative errors  (Trans::error, trans.h:88-93)
 *   op 5: out = trans(a) - trans(b)
 * b: DEVICE [n] for ops 3-5, else ignored. */
int pgert_trans(pgert_ctx * ctx, int op, int n, const double * a, const double * b, double tau, double lower,
                double upper, double * out);

/* The objective RInversion::linesearch scans (core/src/inversion.h:615-640),
 * at n_tau step lengths taus[k] (HOST): with the data transform log and the
 * model transform TransLogLU(lower, upper),
 *   phiD[k] = | (tD(data) - tD(tD.update(resp, taus[k] dd))) / derr |^2   (getPhiD, inversion.cpp:84-98)
 *   phiM[k] = | cw (C (tM(tM.update(model, taus[k] dm)) mw)) |^2           (getPhiM, roughness, inversion.h:554-563)
 * and for taus[k] = 0 resp and model as they are (getPhi()).  C is the
 * matrix of pgert_set_constraints.  DEVICE: model, dm, mw [n_model], cw
 * [n_const], data, resp, dd, derr [n_data].  HOST outputs phiD, phiM [n_tau].
 * Fixed-order sums. */
int pgert_gn_phi(pgert_ctx * ctx, int n_tau, const double * taus, int n_model, const double * model,
                 const double * dm, const double * mw, const double * cw, double lower, double upper, int n_data,
                 const double * data, const double * resp, const double * dd, const double * derr, double * phiD,
                 double * phiM);

/* pgert_cgls without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as pgert_cgls; its work vectors live in a fixed
 * 64 MiB device array per GPU (held until the process ends), so the problem
 * must fit 8 (5 n_model + 3 n_data + 4 max(n_const, 1)) bytes, each vector
 * rounded up to 256 bytes, plus 4 KiB, within 64 MiB; a larger one fails
 * with a message before anything runs.  Calls on one GPU share that array:
 * calls that overlap on one GPU (other threads or streams) race. */
int pgert_cgls_noalloc(pgert_ctx * ctx, const void * J, int dtype, int n_data, int n_model,
                       const double * d_weight, const double * b, double * x, const double * wc,
                       const double * wm, const double * tm, const double * td, double lambda,
                       const double * roughness, int max_iter, double tol, int * iterations, double * norm_r2);

/* Timing of the last pgert_forward / pgert_jacobian phases in milliseconds
 * (HIP events on the context stream): [0] assembly, [1] primary + SR rhs,
 * [2] solve, [3] response, [4] jacobian. */
int pgert_last_timings(const pgert_ctx * ctx, double * ms5);

/* Learned warm start (no counterpart in the reference, which always starts
 * its direct solves from scratch): a per-(node, source) MLP 32-128-128-128-1,
 * GELU, BF16 weights on the matrix cores with FP32 accumulation, predicting
 * the relative secondary potential u_s / (rho_s p).  Parameters are a packed
 * FP32 vector of pgert_mlp_param_count() floats: W1[32][128], b1[128],
 * W2[128][128], b2, W3[128][128], b3, w4[128], b4 (row-major [in][out]). */
int pgert_mlp_param_count(void);
int pgert_mlp_set_params(pgert_ctx * ctx, const float * params, int count);   /* HOST */
int pgert_mlp_get_params(pgert_ctx * ctx, float * params, int count);         /* HOST */
/* Features [count][32] and targets [count] (NULL: none) of samples
 * idx = node * nE + e of wavenumber k_index, from the last forward.  DEVICE. */
int pgert_mlp_collect(pgert_ctx * ctx, int k_index, const long * idx, long count, float * feat,
                      float * target);
/* Network output for features [count][32].  DEVICE. */
int pgert_mlp_predict(pgert_ctx * ctx, const float * feat, int count, float * out);
/* Loss 0.5 mean((y - t)^2) and its gradient on a batch (count a multiple of
 * 16, DEVICE features/targets); apply = 1 takes one Adam step with rate lr.
 * grad: optional HOST array of pgert_mlp_param_count() floats. */
int pgert_mlp_train_step(pgert_ctx * ctx, const float * feat, const float * target, int count, float lr,
                         int apply, float * loss, float * grad);

/* Per-launch timing of one solver kernel with HIP events on the context
 * stream.  enable selects the kernel (0: off):
 *   PGERT_TIME_SPMM    1  sparse product Q = A P (k_spmm; k_spmm_dot for PGERT_PCG_FP64)
 *   PGERT_TIME_UPDATE  2  X += P alpha, R -= Q alpha (k_update, block methods)
 *   PGERT_TIME_GRAM_Z  3  Z = M^-1 R and the Grams Q'Z, Z'Z, P'Z (block methods)
 *   PGERT_TIME_PUPDATE 4  P = Z T + P beta T (block methods)
 *   PGERT_TIME_GRAM_PQ 5  the Grams P'Q, P'R (block methods)
 *   PGERT_TIME_SMALL_ALPHA 6   alpha = (P'Q)^-1 P'R, 64 x 64 (block methods)
 *   PGERT_TIME_SMALL_BETA  7   beta, chol(W'W)^-T, 64 x 64 (block methods)
 *   PGERT_TIME_ZFORM       8   Z = D^-1 R + coarse prolongation (block methods)
 *   PGERT_TIME_RESTRICT    9   aggregate sums of R (block methods, coarse on)
 *   PGERT_TIME_COARSE     10   E^-1 y (block methods, coarse on)
 *   PGERT_TIME_REDUCE     11   fixed-order sum of the partial Grams (block methods)
 *   PGERT_TIME_SMOOTH     12   the two Chebyshev pre-smoothing products (smooth = 2)
 *   PGERT_TIME_COARSE_RESID 13 r1 - (A Z) c (smooth = 2); with smooth = 2, PGERT_TIME_ZFORM
 *                              times the post-smoothing product that forms Z
 * reset: zero the accumulators.  Returns the accumulated milliseconds, number
 * of timed launches and the compulsory HBM bytes of one launch (matrix values
 * and indices once, every vector the kernel reads or writes once, all blocks
 * active).  Diagnostic entry point with no counterpart in the reference. */
#define PGERT_TIME_SPMM 1
#define PGERT_TIME_UPDATE 2
#define PGERT_TIME_GRAM_Z 3
#define PGERT_TIME_PUPDATE 4
#define PGERT_TIME_GRAM_PQ 5
#define PGERT_TIME_SMALL_ALPHA 6
#define PGERT_TIME_SMALL_BETA 7
#define PGERT_TIME_ZFORM 8
#define PGERT_TIME_RESTRICT 9
#define PGERT_TIME_COARSE 10
#define PGERT_TIME_REDUCE 11
#define PGERT_TIME_SMOOTH 12
#define PGERT_TIME_COARSE_RESID 13
int pgert_kernel_timing(pgert_ctx * ctx, int enable, int reset, double * total_ms, long * launches,
                        double * bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* PGERT_H */

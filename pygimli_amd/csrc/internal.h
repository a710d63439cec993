// Internal definitions shared by the libpgert HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pgert.h"

#define PG_CHECK(x)                                                                  \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess)                                                        \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PG_REQUIRE(cond, msg)                                 \
    do {                                                      \
        if (!(cond)) throw std::runtime_error(msg);           \
    } while (0)

namespace pgert {

constexpr int SHAPE_NODES = 4;  // nodes of a P1 simplex (an electrode's shape nodes)

struct MlpState;

// Owned device allocation.
template <class T> struct DBuf {
    T * p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf & operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count) PG_CHECK(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T * h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) PG_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> & v, hipStream_t s) { upload(v.data(), v.size(), s); }
    // at least `count` elements, keeping a larger allocation (work buffers reused across calls)
    void reserve(size_t count) {
        if (p && n >= count) return;
        release();
        if (count) PG_CHECK(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // ---- mesh
    int dim = 0, nv = 0, n = 0, C = 0;
    int order = 1;               // 1: linear simplices; 2: quadratic (P2, primary potentials only)
    int nvert = 0;               // vertices per cell, dim + 1
    int nParam = 0;
    bool uniqueMarkers = false;  // marker -> cell is a bijection over cells with marker >= 0
    DBuf<double> xyz;            // n * 3
    DBuf<int> cells;             // C * nv (P2: the vertices first, then the edge nodes)
    DBuf<int> cellMarker;        // C
    DBuf<double> grads;          // C * nvert * 3, barycentric gradients (components >= dim are 0)
    DBuf<double> vol;            // C
    DBuf<int> sortedCells;       // cells with marker >= 0, sorted by marker
    int nSorted = 0;
    int nDistinct = 0;           // distinct markers >= 0
    std::vector<int> hSortedMarker;  // marker of each sorted cell
    // shared markers: Jacobian tiles of whole markers (built for tile width segTile)
    int segTile = 0, segNTile = 0, segNBig = 0, segNChunk = 0;
    DBuf<int> segTiles;          // segNTile * 4
    DBuf<int> segBig;            // segNBig * 3: marker, first chunk, chunk count
    DBuf<unsigned char> segPart; // chunk partial sums [chunk][D]
    // background (marker < 0) cells: prolongation sweeps planned per mesh (prolong.hip)
    int prolStatus = 0;              // 0 ok, 2 fixed-value regions, 3 no parameter cell, 4 unreachable cells
    std::vector<int> prolLayerPtr;   // sweep l fills prolCell[ptr[l] .. ptr[l+1])
    DBuf<int> prolCell, prolPtr, prolNbr;
    DBuf<double> prolW;

    // ---- CSR pattern and atomics-free assembly map
    int nnz = 0;
    std::vector<int> hRowptr, hColidx;
    DBuf<int> rowptr, colidx, diagSlot;
    DBuf<int> cptr;      // nnz + 1: contributions of each slot
    DBuf<int> ccell;     // contribution -> cell
    DBuf<double> cK;     // contribution -> stiffness value (geometry only)
    DBuf<double> cM;     // contribution -> mass value (geometry only)

    // ---- mixed boundary faces
    int nMixed = 0;
    DBuf<int> mixCell;     // adjacent cell per mixed face
    DBuf<double> mixCtr;   // face centre (3)
    DBuf<double> mixNrm;   // unit normal (3)
    int nDirSlot = 0;
    DBuf<int> dirSlot;     // slots of homogeneous Dirichlet rows / columns (unique)
    DBuf<double> dirVal;   // their values: 1 on the diagonal, 0 elsewhere
    int nBSlot = 0;
    DBuf<int> bSlot;       // slots touched by mixed faces (unique)
    DBuf<int> bPtr;        // nBSlot + 1
    DBuf<int> bFace;       // contribution -> mixed face index
    DBuf<double> bCoef;    // contribution -> face mass value (geometry only)

    // ---- electrodes
    int nE = 0;
    std::vector<int> hElecNode;  // -1: free electrode
    DBuf<int> elecNode;
    DBuf<double> elecPos;    // nE * 3
    DBuf<double> minRadius;  // nE
    DBuf<int> singNode;      // nE: node with the singular primary value, -1: none
    DBuf<int> shapeNode;     // nE * SHAPE_NODES: the electrode's nodes (-1 padded)
    DBuf<double> shapeW;     // nE * SHAPE_NODES: their shape-function weights
    DBuf<int> ecPtr, ecCell; // cells whose attributes give rho_s (around the node / the free electrode's cell)
    double srcCenter[3] = {0, 0, 0};
    double surfaceZ = 0;
    float featureScale = 1.0f;  // length scale of the warm-start features

    // ---- two-level preconditioner (coarse.hip): aggregates and E^-1 per wavenumber
    std::vector<double> hXyz;
    int nAgg = 0;           // 0: off
    DBuf<int> agg;          // node -> aggregate
    DBuf<int> aggPtr, aggNodes;
    int nCE = 0;            // coarse nonzeros
    DBuf<int> cePtr, ceSlot, ceIJ;
    DBuf<double> Einv;      // nK * m * m
    DBuf<double> ycoarse, ccoarse;  // per block m * 64
    // multiplicative two-level variant (Chebyshev smoothing): A Z per row
    // (the aggregates of the row's columns and the summed values), A D^-1,
    // and the Gershgorin bound of D^-1 A, per wavenumber
    int nAZ = 0;
    DBuf<int> azPtr, azAgg, azSlotPtr, azSlot;
    DBuf<double> azVals, valsAD, gershPart;
    std::vector<double> hGersh;

    // ---- wavenumbers
    int nK = 0;
    std::vector<double> kv, wv;
    DBuf<double> dk, dw;

    // ---- data
    int D = 0;
    DBuf<int> abmn;     // D * 4
    DBuf<double> kfac;  // D

    // ---- per-evaluation state
    DBuf<double> rho;       // C
    DBuf<double> rhoS;      // nE   geometric-mean resistivity at each source
    DBuf<double> valsA;     // nK * nnz
    DBuf<double> valsS1;    // nK * nnz (rho == 1), mesh-only
    bool s1Valid = false;
    DBuf<double> dinv;      // nK * n
    DBuf<double> prim;      // nK * n * nE   unit-resistivity primary potentials
    bool primValid = false;
    bool primNumeric = false;    // primaries come from pgert_set_primary_potentials
    bool virtualElec = false;    // an electrode sits on a node of no cell (complete electrode model)
    int refNode = -1;            // reference electrode node: sources are dipoles against it
    double refRmin = 0.0;        // its nearest-neighbour distance (singular primary value)
    DBuf<int> refCells;          // its cells (rho_s)
    int nCalSlot = 0;            // calibration nodes: Dirichlet slots applied after the bypass resistors
    DBuf<int> calNodes;          // the calibration nodes (their SR right-hand side is zeroed)
    DBuf<int> calSlot;
    DBuf<double> calVal;
    DBuf<double> U;         // nK * n * nE   total potentials (secondary + rho_s * primary)
    bool haveU = false;
    bool srState = false;   // X / prim / rhoS / rho hold a singularity-removal forward
    DBuf<double> emap;      // nE * nE
    // solver work
    DBuf<double> B, X, R, P, Q, Z;
    DBuf<double> part;      // reduction partials
    DBuf<double> scal;      // per-column scalars
    DBuf<double> small;     // block-CG small matrices

    std::shared_ptr<MlpState> mlp;  // learned warm start (mlp.hip)

    // constraint matrix C of the Gauss-Newton step and its transpose, CSR (cgls.hip)
    int nConst = 0, cCols = 0;
    DBuf<int> cRowptr, cColidx, ctRowptr, ctColidx;
    DBuf<double> cVals, ctVals;
    double cglsNormR2 = 0.0;  // |r|^2 at the end of the last pgert_cgls

    // Jacobian launch descriptors (jacobian.hip), kept until the next launch
    std::vector<unsigned char> jdescHost;
    DBuf<unsigned char> jdesc;

    // band solver (band.hip): ordering and CSR slot -> band position, per mesh
    bool bandValid = false;
    int bandB = 0;
    DBuf<int> bandPerm, bandSlot;
    // band work buffers of a (batched) solve led by this context, reused across calls
    DBuf<double> bandL, bandDinv, bandY;
    DBuf<unsigned char> bandDesc;
    // complex resistivity: imaginary half of A, unit point sources, the two
    // halves of the potentials, Jacobian work blocks, change accumulator
    DBuf<double> valsAi, Bpt, Xre, Xim, chg, chgPart;
    // bypass resistors of the total-field operator: matrix slots and values
    DBuf<int> bypSlot;
    DBuf<double> bypVal;
    int nByp = 0;
    DBuf<int> cemSlot;           // complete electrode model: constant terms per slot
    DBuf<double> cemVal;
    int nCem = 0;
    bool haveUc = false;

    // block solver's own stream (graph capture needs a non-default stream),
    // ordered after / before the caller's stream with these events
    hipStream_t solveStream = nullptr;
    hipEvent_t solveEv[2] = {nullptr, nullptr};

    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double timings[5] = {0, 0, 0, 0, 0};

    // optional per-launch timing of the solver's SpMM kernel (pgert_kernel_timing)
    bool kTiming = false;
    int kSelect = 0;            // PGERT_TIME_* kernel being timed
    std::vector<hipEvent_t> kEvents;
    int kUsed = 0;              // events used in the current solve
    double kTotalMs = 0.0;      // accumulated SpMM time since the last reset
    long kLaunches = 0;
    double kBytesPerLaunch = 0; // algorithmic bytes of one SpMM launch
    hipEvent_t kev_begin() {
        if (!kTiming) return nullptr;
        if (kUsed + 2 > (int)kEvents.size()) {
            for (int i = 0; i < 256; ++i) {
                hipEvent_t e;
                PG_CHECK(hipEventCreate(&e));
                kEvents.push_back(e);
            }
        }
        hipEvent_t e = kEvents[kUsed++];
        PG_CHECK(hipEventRecord(e, stream));
        return e;
    }
    void kev_end() {
        if (!kTiming) return;
        hipEvent_t e = kEvents[kUsed++];
        PG_CHECK(hipEventRecord(e, stream));
    }
    // called after the stream is synchronised
    void kev_collect() {
        if (!kTiming) return;
        for (int i = 0; i + 1 < kUsed; i += 2) {
            float t = 0;
            PG_CHECK(hipEventElapsedTime(&t, kEvents[i], kEvents[i + 1]));
            kTotalMs += t;
            ++kLaunches;
        }
        kUsed = 0;
    }
};

// ---- kernels / launchers (defined in the .hip files)
void launch_cell_geometry(Ctx & c);
void launch_contrib_geometry(Ctx & c, const int * codes);
// per-measure P2 tables of a d-simplex (d = 1..3): mass M[i][j] and stiffness
// weights W[i][j][a][e], K_ij = |T| sum_ae (grad lam_a . grad lam_e) W[i][j][a][e]
void p2_tables(int d, std::vector<double> & M, std::vector<double> & W);
void launch_contrib_geometry_p2(Ctx & c, const int * codes, const double * M, const double * W);
void launch_assemble(Ctx & c, const double * rho, double k, double * vals, bool unitRho);
// homogeneous Dirichlet rows / columns of an assembled matrix (imagPart: the
// imaginary half of a complex one, whose diagonal entries there are 0)
void launch_dirichlet(Ctx & c, double * vals, bool imagPart);
// calibration nodes (rows / columns cleared, diagonal one) in every wavenumber's matrix
void launch_calibration(Ctx & c, double * vals, bool imagPart = false);
// rows of the calibration nodes of a node-major block [k][node][e] set to zero
void launch_zero_rows(Ctx & c, double * B);
void launch_fill_matrix(Ctx & c, int mass, const double * a, double * vals);
void launch_map_model(Ctx & c, const double * model, int nModel, double * rho);
void build_prolongation(Ctx & c, const double * xyz, const int * cells, const int * marker);
void check_prolongation(const Ctx & c);
void launch_prolongation(Ctx & c, double * rho);
void launch_rho_source(Ctx & c, const double * rho);
void launch_primary(Ctx & c);
void launch_sr_rhs(Ctx & c);
void launch_diag_inv(Ctx & c);
void launch_total_potential(Ctx & c);
void launch_response(Ctx & c, double * rhoa);
void launch_transpose_pots(Ctx & c, double * out);
// accumulate: add the sensitivities to the values J holds (no model scaling then)
void launch_jacobian(Ctx & c, const double * model, int nModel, void * J, int dtype, bool accumulate = false);
// one Jacobian launch per wavenumber pass over the cells of several problems
// (same device and dimension; every context uses cs[0]->stream)
void launch_assemble_complex(Ctx & c, const double * rre, const double * rim, double k, double * vre,
                             double * vim);
void launch_point_rhs(Ctx & c, double * B);
void launch_bypass(Ctx & c, double * vals);
void launch_spmm_add(Ctx & c, const double * vals, const double * X, const double * base, double sign,
                     double * out);
void launch_change(Ctx & c, long total, const double * a, const double * b, double * acc);
void launch_response_complex(Ctx & c, const double * Ure, const double * Uim, double * out_re, double * out_im);
void launch_add(Ctx & c, long total, const double * a, const double * b, double * out);
void launch_jacobian_complex_pre(Ctx & c, double * Jre, double * Jim);
void launch_jacobian_complex_scale(Ctx & c, const double * mr, const double * mi, double * Jre, double * Jim);
void launch_jacobian_batch(const std::vector<Ctx *> & cs, const std::vector<const double *> & models,
                           const std::vector<int> & nModels, const std::vector<void *> & Js, int dtype);

// Gauss-Newton step (cgls.hip)
void set_constraints(Ctx & c, int nConst, int nModel, const int * rowptr, const int * colidx, const double * vals);
int cgls(Ctx & c, const void * J, int dtype, int nData, int nModel, const double * dW, const double * b, double * x,
         const double * wc, const double * wm, const double * tm, const double * td, double lambda,
         const double * roughness, int maxIter, double tol);
int cgls_noalloc(Ctx & c, const void * J, int dtype, int nData, int nModel, const double * dW, const double * b,
                 double * x, const double * wc, const double * wm, const double * tm, const double * td, double lambda,
                 const double * roughness, int maxIter, double tol);
// Gauss-Newton loop: logarithmic transforms and the line-search objective (gn.hip)
void gn_trans(Ctx & c, int op, int n, const double * a, const double * b, double tau, double lower, double upper,
              double * out);
void gn_phi(Ctx & c, int nTau, const double * taus, int nModel, const double * m, const double * dm, const double * mw,
            const double * cw, double lower, double upper, int nData, const double * d, const double * r,
            const double * dd, const double * e, double * phiD, double * phiM);
// device address of the allocation-free CGLS's workspace, looked up once per
// device (the device must be current)
void cgls_scratch_init(int device);

// solvers (solver.hip); return per-solve statistics
void solve_pcg_fp64(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st);
void solve_block(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st, bool mixed);
// one block-PCG over every block of several problems (same device; the
// caller's stream is cs[0]->stream and every context uses it)
void solve_blocks(const std::vector<Ctx *> & cs, const pgert_solve_opts & o, std::vector<pgert_solve_stats> & st,
                  bool mixed);
double residual_check(Ctx & c);
// direct banded Cholesky over every wavenumber of several 2D problems (band.hip)
void band_prepare(Ctx & c);
void solve_band(const std::vector<Ctx *> & cs, const pgert_solve_opts & o, std::vector<pgert_solve_stats> & st,
                bool refactor = true);

// two-level preconditioner (coarse.hip)
void coarse_prepare(Ctx & c);
void smooth_prepare(Ctx & c);

// learned warm start (mlp.hip)
int mlp_param_count();
void mlp_set_params(Ctx & c, const float * host);
void mlp_get_params(Ctx & c, float * host);
void mlp_warm_start(Ctx & c);
void mlp_collect(Ctx & c, int kk, const long * idx, long count, float * feat, float * target);
void mlp_predict(Ctx & c, const float * feat, int nsamp, float * out);
float mlp_train_step(Ctx & c, const float * feat, const float * target, int nsamp, float lr, bool apply,
                     float * grad_host);

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// Global-address-space view of a device pointer.  Pointers read from a
// descriptor in memory are generic (flat) to the compiler, and a flat access
// may alias LDS, so it is ordered against every LDS access around it (a wait
// after each load); through this view the loads and stores are global ones
// that overlap freely with LDS work.
template <typename T> __device__ inline __attribute__((address_space(1))) T * glob(T * p) {
    return (__attribute__((address_space(1))) T *)p;
}

// Workgroup barrier that orders LDS only: __syncthreads also waits for every
// outstanding global load and store of the wavefront (vmcnt(0)), which
// exposes the memory latency of in-flight prefetches and stores.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace pgert

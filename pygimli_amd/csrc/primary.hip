// Singularity-removal (SR) pieces of the forward operator: analytic primary
// potentials, source resistivities, the secondary-field right-hand sides, the
// total potentials and the apparent-resistivity response.
//
// Potential blocks are node-major, one block per wavenumber: V[k][node][e]
// with e the current electrode; a row of a block is contiguous, so a SpMM
// row reads nE consecutive values per nonzero.
#include <cfloat>

#include "internal.h"
#include "bessel.h"

namespace pgert {

namespace {

constexpr double TOL = 1e-12;
constexpr double PI = 3.14159265358979323846;

// exactDCSolution(v, source, k, surfaceZ, fallback = 0) (bertMisc.cpp:183-228)
__device__ double exact_dc(const double * v, const double * src, double k, double surfaceZ) {
    double dx = v[0] - src[0], dy = v[1] - src[1], dz = v[2] - src[2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r < TOL) return 0.0;
    int dim = k > 0 ? 2 : 3;
    double m[3] = {src[0], src[1], src[2]};
    m[dim - 1] = 2.0 * surfaceZ - src[dim - 1];
    if (surfaceZ == -DBL_MAX || isnan(surfaceZ)) {
        if (k == 0.0) return 1.0 / (4.0 * PI * r);
        return bessel_k0(r * k) / (2.0 * PI);
    }
    double mx = v[0] - m[0], my = v[1] - m[1], mz = v[2] - m[2];
    double rm = sqrt(mx * mx + my * my + mz * mz);
    if (k == 0.0) return (1.0 / r + 1.0 / rm) / (4.0 * PI);
    double sx = src[0] - m[0], sy = src[1] - m[1], sz = src[2] - m[2];
    if (sx * sx + sy * sy + sz * sz < TOL) return bessel_k0(r * k) / PI;  // source on the mirror plane
    return (bessel_k0(r * k) + bessel_k0(rm * k)) / (2.0 * PI);
}

// singNode[e]: the node that takes the singular value (the electrode node; a
// free electrode's cell node within 1e-4 of its position, else -1)
__global__ void k_primary(int n, int nE, int nK, const double * __restrict__ xyz,
                          const double * __restrict__ elecPos, const int * __restrict__ singNode,
                          const double * __restrict__ minRadius, const double * __restrict__ kv,
                          double surfaceZ, double * __restrict__ prim) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)nK * n * nE;
    if (t >= total) return;
    int e = (int)(t % nE);
    long kn = t / nE;
    int node = (int)(kn % n);
    int kk = (int)(kn / n);
    double k = kv[kk];
    double u;
    if (node == singNode[e]) {
        // ElectrodeShapeNode / ElectrodeShapeEntity::setSingValue, scale 0
        // (electrode.cpp:154-189, :207-249)
        double rmin = minRadius[e];
        u = k > 0.0 ? bessel_k0(rmin / 6.0 * k) / PI : 1.0 / (2.0 * PI * rmin / 2.0);
    } else {
        u = exact_dc(xyz + 3 * (size_t)node, elecPos + 3 * (size_t)e, k, surfaceZ);
    }
    prim[t] = u;
}

// The reference electrode's primary potential (eB of every current pattern,
// checkPrimpotentials_ dcfemmodelling.cpp:2101-2102; singular value at its
// node as for a node electrode) subtracted from every source's: dipole
// primaries.  One thread per (k, node).
__global__ void k_primary_ref(int n, int nE, int nK, const double * __restrict__ xyz, int refNode, double refRmin,
                              const double * __restrict__ kv, double surfaceZ, double * __restrict__ prim) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nK * n) return;
    const int node = (int)(t % n);
    const double k = kv[t / n];
    double u;
    if (node == refNode)
        u = k > 0.0 ? bessel_k0(refRmin / 6.0 * k) / PI : 1.0 / (2.0 * PI * refRmin / 2.0);
    else
        u = exact_dc(xyz + 3 * (size_t)node, xyz + 3 * (size_t)refNode, k, surfaceZ);
    double * p = prim + (size_t)t * nE;
    for (int e = 0; e < nE; ++e) p[e] -= u;
}

// rhoSource of a dipole pattern (DCSRMultiElectrodeModelling::calculateK,
// dcfemmodelling.cpp:2219-2244): the mean of the geometric means of eA and of
// the reference electrode, over those above TOLERANCE.  One thread.
__global__ void k_rho_ref(int nE, int nCells, const int * __restrict__ cells, const double * __restrict__ rho,
                          double * __restrict__ rhoS) {
    if (blockIdx.x * blockDim.x + threadIdx.x != 0) return;
    double s = 0.0;
    for (int q = 0; q < nCells; ++q) s += log(rho[cells[q]]);
    const double rb = exp(s / (double)nCells);
    for (int e = 0; e < nE; ++e) {
        const double ra = rhoS[e];
        double sum = 0.0;
        int cnt = 0;
        if (ra > 1e-12) { sum += ra; ++cnt; }
        if (rb > 1e-12) { sum += rb; ++cnt; }
        rhoS[e] = sum / cnt;
    }
}

// ElectrodeShapeNode::geomMeanCellAttributes (electrode.cpp:102-127)
__global__ void k_rho_source(int nE, const int * __restrict__ ptr, const int * __restrict__ cells,
                             const double * __restrict__ rho, double * __restrict__ rhoS) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    double s = 0.0;
    int cnt = ptr[e + 1] - ptr[e];
    for (int q = ptr[e]; q < ptr[e + 1]; ++q) s += log(rho[cells[q]]);
    rhoS[e] = exp(s / (double)cnt);
}

// b = S1 p - rho_s A p for every (k, e) column (dcfemmodelling.cpp:2253-2254
// with prim = rho_s p): one wavefront per (row, k), lanes over electrodes.
// The row's column indices and coefficients are read once, one per lane, and
// broadcast with readlane, so the gathers of a row are independent loads.
__device__ inline double readlane_d(double v, int j) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ void __launch_bounds__(256) k_sr_rhs(int n, int nE, int nnz, const int * __restrict__ rowptr,
                                                const int * __restrict__ colidx,
                                                const double * __restrict__ valsS1,
                                                const double * __restrict__ valsA,
                                                const double * __restrict__ rhoS,
                                                const double * __restrict__ prim, double * __restrict__ b) {
    const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int kk = blockIdx.y;
    if (row >= n) return;
    const double * s1 = valsS1 + (size_t)kk * nnz;
    const double * a = valsA + (size_t)kk * nnz;
    const double * P = prim + (size_t)kk * n * nE;
    double * out = b + ((size_t)kk * n + row) * nE;
    const int q0 = rowptr[row], q1 = rowptr[row + 1];
    for (int base = q0; base < q1; base += 64) {
        const int m = min(64, q1 - base);
        int ci = 0;
        double cs = 0.0, ca = 0.0;
        if (lane < m) {
            ci = colidx[base + lane];
            cs = s1[base + lane];
            ca = a[base + lane];
        }
        for (int e0 = 0; e0 < nE; e0 += 64) {
            const int e = e0 + lane;
            const bool ok = e < nE;
            const double rs = ok ? rhoS[e] : 0.0;
            double acc = 0.0;
            // groups of 8 entries (lanes >= m hold column 0 with coefficient 0)
            for (int j0 = 0; j0 < m; j0 += 8) {
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) {
                    const int col = __builtin_amdgcn_readlane(ci, j0 + jj);
                    const double coef = readlane_d(cs, j0 + jj) - rs * readlane_d(ca, j0 + jj);
                    if (ok) acc += coef * P[(size_t)col * nE + e];
                }
            }
            if (ok) out[e] = base == q0 ? acc : out[e] + acc;
        }
    }
}

__global__ void k_total_potential(long total, int nE, const double * __restrict__ X,
                                  const double * __restrict__ rhoS, const double * __restrict__ prim,
                                  double * __restrict__ U) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int e = (int)(t % nE);
    U[t] = X[t] + rhoS[e] * prim[t];
}

// DataMap::collect (datamap.cpp:57-98) of the k-weighted sum
// (dcfemmodelling.cpp:1707-1729): map[e][f] = potential of source e at
// electrode f, sum_j N_fj sum_k w_k U[k][node_fj][e] over the electrode's
// shape nodes (one node with N = 1 for a node electrode, ElectrodeShapeNode::pot,
// electrode.cpp:129-131; the cell's nodes and shape functions at the position
// for a free one, MeshEntity::pot, meshentities.cpp:347-349).
__global__ void k_electrode_map(int n, int nE, int nK, const double * __restrict__ U,
                                const double * __restrict__ w, const int * __restrict__ shapeNode,
                                const double * __restrict__ shapeW, double * __restrict__ emap) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nE * nE) return;
    int e = t / nE, f = t % nE;
    double s = 0.0;
    for (int j = 0; j < SHAPE_NODES; ++j) {
        const int v = shapeNode[SHAPE_NODES * f + j];
        if (v < 0) break;
        double u = 0.0;
        for (int kk = 0; kk < nK; ++kk) u += w[kk] * U[((size_t)kk * n + v) * nE + e];
        s += u * shapeW[SHAPE_NODES * f + j];
    }
    emap[t] = s;
}

__device__ double map_data(const double * emap, int nE, int a, int b, int m, int n) {
    double am = (a != -1 && m != -1) ? emap[a * nE + m] : 0.0;
    double an = (a != -1 && n != -1) ? emap[a * nE + n] : 0.0;
    double bm = (b != -1 && m != -1) ? emap[b * nE + m] : 0.0;
    double bn = (b != -1 && n != -1) ? emap[b * nE + n] : 0.0;
    return (am - an) - (bm - bn);
}

// DataMap::data + DCMultiElectrodeModelling::response (datamap.cpp:159-231,
// dcfemmodelling.cpp:1140-1196): U and its reciprocal, rounded to 1e-10
// (roundTo, core/src/expressions.h:169), times k, rhoa = sqrt|r * r_rez|.
__global__ void k_response(int D, int nE, const int * __restrict__ abmn, const double * __restrict__ kfac,
                           const double * __restrict__ emap, double * __restrict__ rhoa) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    int a = abmn[4 * i], b = abmn[4 * i + 1], m = abmn[4 * i + 2], n = abmn[4 * i + 3];
    double u = map_data(emap, nE, a, b, m, n);
    double ur = map_data(emap, nE, m, n, a, b);
    double r = rint(u / 1e-10) * 1e-10 * kfac[i];
    double rr = rint(ur / 1e-10) * 1e-10 * kfac[i];
    rhoa[i] = sqrt(fabs(r * rr));
}

__global__ void k_transpose_pots(int n, int nE, int nK, const double * __restrict__ U,
                                 double * __restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)nK * n * nE;
    if (t >= total) return;
    // t indexes the output: row = e + nE * k, col = node
    int node = (int)(t % n);
    long row = t / n;
    int e = (int)(row % nE);
    int kk = (int)(row / nE);
    out[t] = U[((size_t)kk * n + node) * nE + e];
}

}  // namespace

void launch_primary(Ctx & c) {
    long total = (long)c.nK * c.n * c.nE;
    hipLaunchKernelGGL(k_primary, dim3(div_up(total, 256)), dim3(256), 0, c.stream, c.n, c.nE, c.nK,
                       c.xyz.p, c.elecPos.p, c.singNode.p, c.minRadius.p, c.dk.p, c.surfaceZ, c.prim.p);
    PG_CHECK(hipGetLastError());
    if (c.refNode >= 0) {
        const long kn = (long)c.nK * c.n;
        hipLaunchKernelGGL(k_primary_ref, dim3(div_up(kn, 256)), dim3(256), 0, c.stream, c.n, c.nE, c.nK, c.xyz.p,
                           c.refNode, c.refRmin, c.dk.p, c.surfaceZ, c.prim.p);
        PG_CHECK(hipGetLastError());
    }
}

void launch_rho_source(Ctx & c, const double * rho) {
    hipLaunchKernelGGL(k_rho_source, dim3(div_up(c.nE, 64)), dim3(64), 0, c.stream, c.nE, c.ecPtr.p,
                       c.ecCell.p, rho, c.rhoS.p);
    PG_CHECK(hipGetLastError());
    if (c.refNode >= 0) {
        hipLaunchKernelGGL(k_rho_ref, dim3(1), dim3(64), 0, c.stream, c.nE, (int)c.refCells.n, c.refCells.p, rho,
                           c.rhoS.p);
        PG_CHECK(hipGetLastError());
    }
}

void launch_sr_rhs(Ctx & c) {
    hipLaunchKernelGGL(k_sr_rhs, dim3(div_up(c.n, 4), c.nK), dim3(256), 0, c.stream, c.n, c.nE, c.nnz,
                       c.rowptr.p, c.colidx.p, c.valsS1.p, c.valsA.p, c.rhoS.p, c.prim.p, c.B.p);
    PG_CHECK(hipGetLastError());
}

void launch_total_potential(Ctx & c) {
    long total = (long)c.nK * c.n * c.nE;
    hipLaunchKernelGGL(k_total_potential, dim3(div_up(total, 256)), dim3(256), 0, c.stream, total, c.nE,
                       c.X.p, c.rhoS.p, c.prim.p, c.U.p);
    PG_CHECK(hipGetLastError());
}

void launch_response(Ctx & c, double * rhoa) {
    c.emap.alloc((size_t)c.nE * c.nE);
    hipLaunchKernelGGL(k_electrode_map, dim3(div_up((long)c.nE * c.nE, 256)), dim3(256), 0, c.stream, c.n,
                       c.nE, c.nK, c.U.p, c.dw.p, c.shapeNode.p, c.shapeW.p, c.emap.p);
    PG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_response, dim3(div_up(c.D, 256)), dim3(256), 0, c.stream, c.D, c.nE, c.abmn.p,
                       c.kfac.p, c.emap.p, rhoa);
    PG_CHECK(hipGetLastError());
}

void launch_transpose_pots(Ctx & c, double * out) {
    long total = (long)c.nK * c.n * c.nE;
    hipLaunchKernelGGL(k_transpose_pots, dim3(div_up(total, 256)), dim3(256), 0, c.stream, c.n, c.nE, c.nK,
                       c.U.p, out);
    PG_CHECK(hipGetLastError());
}

}  // namespace pgert

// P1 element geometry and atomics-free stiffness/mass assembly into the CSR
// pattern built once per mesh.
//
// Every nonzero slot owns the list of (cell, local pair) contributions that
// fall on it, pre-sorted on the host; a value is then a gather-sum over that
// list, so no two threads ever write the same slot.  The geometric part of
// every contribution (vol grad N_i . grad N_j and vol/((d+1)(d+2))(1+d_ij),
// ElementMatrix::ux2uy2uz2 / u2, core/src/elementmatrix.cpp:682-1122) is
// computed once per mesh, so assembling a new model is one pass of
// sigma_c * (K + k^2 M) over the contribution list.
#include "internal.h"
#include "bessel.h"

namespace pgert {

namespace {

constexpr double TOL = 1e-12;  // core/src/gimli.h:206

// stride: nodes per cell in `cells` (the vertices come first)
__global__ void k_cell_geometry(int C, int dim, int stride, const double * __restrict__ xyz,
                                const int * __restrict__ cells, double * __restrict__ grads,
                                double * __restrict__ vol) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int nv = dim + 1;
    double p[4][3];
    for (int v = 0; v < nv; ++v) {
        int id = cells[(size_t)c * stride + v];
        for (int q = 0; q < 3; ++q) p[v][q] = xyz[3 * (size_t)id + q];
    }
    double g[4][3] = {{0}};
    double V;
    if (dim == 2) {
        double a = p[1][0] - p[0][0], b = p[1][1] - p[0][1];
        double cc = p[2][0] - p[0][0], d = p[2][1] - p[0][1];
        double det = a * d - b * cc;
        // rows of the edge matrix are (a b) and (cc d); gradients are the
        // columns of its inverse
        g[1][0] = d / det;   g[1][1] = -cc / det;
        g[2][0] = -b / det;  g[2][1] = a / det;
        V = 0.5 * fabs(det);
    } else {
        double e[3][3];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) e[r][q] = p[r + 1][q] - p[0][q];
        // cofactors: inverse column r = cross products of the other edges
        double c0[3] = {e[1][1] * e[2][2] - e[1][2] * e[2][1], e[1][2] * e[2][0] - e[1][0] * e[2][2],
                        e[1][0] * e[2][1] - e[1][1] * e[2][0]};
        double c1[3] = {e[2][1] * e[0][2] - e[2][2] * e[0][1], e[2][2] * e[0][0] - e[2][0] * e[0][2],
                        e[2][0] * e[0][1] - e[2][1] * e[0][0]};
        double c2[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                        e[0][0] * e[1][1] - e[0][1] * e[1][0]};
        double det = e[0][0] * c0[0] + e[0][1] * c0[1] + e[0][2] * c0[2];
        for (int q = 0; q < 3; ++q) {
            g[1][q] = c0[q] / det;
            g[2][q] = c1[q] / det;
            g[3][q] = c2[q] / det;
        }
        V = fabs(det) / 6.0;
    }
    for (int q = 0; q < 3; ++q) {
        double s = 0;
        for (int v = 1; v < nv; ++v) s += g[v][q];
        g[0][q] = -s;
    }
    for (int v = 0; v < nv; ++v)
        for (int q = 0; q < 3; ++q) grads[((size_t)c * nv + v) * 3 + q] = g[v][q];
    vol[c] = V;
}

__global__ void k_contrib_geometry(long nc, int nv, const int * __restrict__ ccell,
                                   const int * __restrict__ code, const double * __restrict__ grads,
                                   const double * __restrict__ vol, double * __restrict__ cK,
                                   double * __restrict__ cM) {
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc) return;
    int c = ccell[e];
    int ij = code[e];
    int i = ij / nv, j = ij % nv;
    const double * gi = grads + ((size_t)c * nv + i) * 3;
    const double * gj = grads + ((size_t)c * nv + j) * 3;
    double V = vol[c];
    cK[e] = V * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
    cM[e] = V / (nv * (nv + 1)) * (i == j ? 2.0 : 1.0);
}

// P2 element matrices per contribution (cell, i, j), exact for straight
// simplices as the reference's quadrature of orders 2 (stiffness) and 4 (mass)
// is (elementmatrix.cpp:761-782, :872-1060)
__global__ void k_contrib_geometry_p2(long nc, int nv, int nvert, const int * __restrict__ ccell,
                                      const int * __restrict__ code, const double * __restrict__ grads,
                                      const double * __restrict__ vol, const double * __restrict__ M,
                                      const double * __restrict__ W, double * __restrict__ cK,
                                      double * __restrict__ cM) {
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc) return;
    const int c = ccell[e];
    const int ij = code[e];
    const double V = vol[c];
    const double * g = grads + (size_t)c * nvert * 3;
    const double * w = W + (size_t)ij * nvert * nvert;
    double k = 0.0;
    for (int a = 0; a < nvert; ++a)
        for (int b = 0; b < nvert; ++b) {
            const double gab = g[3 * a] * g[3 * b] + g[3 * a + 1] * g[3 * b + 1] + g[3 * a + 2] * g[3 * b + 2];
            k += gab * w[a * nvert + b];
        }
    cK[e] = V * k;
    cM[e] = V * M[ij];
}

__global__ void k_assemble_domain(int nnz, const int * __restrict__ cptr, const int * __restrict__ ccell,
                                  const double * __restrict__ cK, const double * __restrict__ cM,
                                  const double * __restrict__ rho, double k2, int withMass,
                                  double * __restrict__ vals) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    double acc = 0.0;
    for (int e = cptr[s]; e < cptr[s + 1]; ++e) {
        double sig = 1.0;
        if (rho) {
            double r = rho[ccell[e]];
            // cells with |rho| <= TOLERANCE are skipped (dcfemmodelling.cpp:185)
            sig = fabs(r) > TOL ? 1.0 / r : 0.0;
        }
        double v = withMass ? (cM[e] * k2 + cK[e]) : cK[e];
        acc += v * sig;
    }
    vals[s] = acc;
}

// mixedBoundaryCondition (dcfemmodelling.cpp:430-506), MIRRORSOURCE, mirror plane 0.
__device__ double mixed_alpha(const double * ctr, const double * nrm, const double * src, double k) {
    int dim = k > 0 ? 2 : 3;
    double sm[3] = {src[0], src[1], src[2]};
    sm[dim - 1] = -src[dim - 1];
    double r[3], rm[3];
    for (int q = 0; q < 3; ++q) { r[q] = src[q] - ctr[q]; rm[q] = sm[q] - ctr[q]; }
    double ra = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double rma = sqrt(rm[0] * rm[0] + rm[1] * rm[1] + rm[2] * rm[2]);
    double rn = fabs(r[0] * nrm[0] + r[1] * nrm[1] + r[2] * nrm[2]);
    double rmn = fabs(rm[0] * nrm[0] + rm[1] * nrm[1] + rm[2] * nrm[2]);
    if (k == 0.0)
        return (rma * rma * rn / ra + ra * ra * rmn / rma) / (rma * ra * (ra + rma));
    double k0a = bessel_k0(ra * k), k0m = bessel_k0(rma * k);
    if (fabs(k0a) < TOL || fabs(k0m) < TOL) return 0.0;
    return k * (rn / ra * bessel_k1(ra * k) + rmn / rma * bessel_k1(rma * k)) / (k0a + k0m);
}

__global__ void k_assemble_boundary(int nbs, const int * __restrict__ bslot, const int * __restrict__ bptr,
                                    const int * __restrict__ bface, const double * __restrict__ bcoef,
                                    const int * __restrict__ mixCell, const double * __restrict__ mixCtr,
                                    const double * __restrict__ mixNrm, const double * __restrict__ rho,
                                    double s0, double s1, double s2, double k, double * __restrict__ vals) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbs) return;
    const double src[3] = {s0, s1, s2};
    double acc = 0.0;
    for (int e = bptr[t]; e < bptr[t + 1]; ++e) {
        int f = bface[e];
        double a = mixed_alpha(mixCtr + 3 * (size_t)f, mixNrm + 3 * (size_t)f, src, k);
        double r = rho ? rho[mixCell[f]] : 1.0;
        acc += bcoef[e] * (a / r);
    }
    vals[bslot[t]] += acc;
}

__global__ void k_fill_matrix(int nnz, const int * __restrict__ cptr, const int * __restrict__ ccell,
                              const double * __restrict__ cval, const double * __restrict__ a,
                              double * __restrict__ vals) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    double acc = 0.0;
    for (int e = cptr[s]; e < cptr[s + 1]; ++e) acc += cval[e] * a[ccell[e]];
    vals[s] = acc;
}

// vals[slot[i]] = val[i] (val null: 0): homogeneous Dirichlet rows and columns
__global__ void k_set_slots(int ns, const int * __restrict__ slot, const double * __restrict__ val,
                            double * __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    vals[slot[i]] = val ? val[i] : 0.0;
}

__global__ void k_map_model(int C, int nModel, const double * __restrict__ model,
                            const int * __restrict__ marker, double * __restrict__ rho) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    if (nModel == C) {
        rho[c] = model[c];
    } else {
        int m = marker[c];
        rho[c] = (m >= 0 && m < nModel) ? model[m] : 0.0;
    }
}

__global__ void k_diag_inv(int n, int nK, int nnz, const int * __restrict__ diagSlot,
                           const double * __restrict__ vals, double * __restrict__ dinv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int kk = blockIdx.y;
    if (i >= n || kk >= nK) return;
    dinv[(size_t)kk * n + i] = 1.0 / vals[(size_t)kk * nnz + diagSlot[i]];
}

}  // namespace

void launch_cell_geometry(Ctx & c) {
    hipLaunchKernelGGL(k_cell_geometry, dim3(div_up(c.C, 256)), dim3(256), 0, c.stream, c.C, c.dim, c.nv,
                       c.xyz.p, c.cells.p, c.grads.p, c.vol.p);
    PG_CHECK(hipGetLastError());
}

void launch_contrib_geometry_p2(Ctx & c, const int * codes, const double * M, const double * W) {
    long nc = (long)c.ccell.n;
    hipLaunchKernelGGL(k_contrib_geometry_p2, dim3(div_up(nc, 256)), dim3(256), 0, c.stream, nc, c.nv, c.nvert,
                       c.ccell.p, codes, c.grads.p, c.vol.p, M, W, c.cK.p, c.cM.p);
    PG_CHECK(hipGetLastError());
}

// The P2 basis of a d-simplex as quadratic forms in the barycentric
// coordinates, phi_i = lam^T Q_i lam: vertex a, 2 lam_a^2 - lam_a sum(lam);
// edge (p, q), 4 lam_p lam_q; nodes: the vertices, then the edges (0,1),
// (1,2), (0,2) of a triangle, (0,1), (0,2), (0,3), (1,2), (1,3), (2,3) of a
// tetrahedron.  Integrals of lam monomials over the simplex divided by its
// measure: d! prod(alpha!) / (d + |alpha|)!.
void p2_tables(int d, std::vector<double> & M, std::vector<double> & W) {
    static const int E1[1][2] = {{0, 1}};
    static const int E2[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    static const int E3[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    const int nb = d + 1, ne = d == 1 ? 1 : d == 2 ? 3 : 6, nn = nb + ne;
    const int (*E)[2] = d == 1 ? E1 : d == 2 ? E2 : E3;
    std::vector<double> Q((size_t)nn * nb * nb, 0.0);
    auto q = [&](int i, int a, int b) -> double & { return Q[((size_t)i * nb + a) * nb + b]; };
    for (int a = 0; a < nb; ++a) {
        for (int b = 0; b < nb; ++b) {
            q(a, a, b) -= 0.5;
            q(a, b, a) -= 0.5;
        }
        q(a, a, a) = 1.0;
    }
    for (int t = 0; t < ne; ++t) q(nb + t, E[t][0], E[t][1]) = q(nb + t, E[t][1], E[t][0]) = 2.0;
    auto fact = [](int m) {
        double f = 1.0;
        for (int i = 2; i <= m; ++i) f *= i;
        return f;
    };
    auto moment = [&](std::initializer_list<int> idx) {
        int alpha[4] = {0, 0, 0, 0};
        for (int v : idx) alpha[v]++;
        double num = fact(d);
        for (int v = 0; v < nb; ++v) num *= fact(alpha[v]);
        return num / fact(d + (int)idx.size());
    };
    M.assign((size_t)nn * nn, 0.0);
    W.assign((size_t)nn * nn * nb * nb, 0.0);
    for (int i = 0; i < nn; ++i)
        for (int j = 0; j < nn; ++j) {
            double m = 0.0;
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b)
                    for (int cc = 0; cc < nb; ++cc)
                        for (int e = 0; e < nb; ++e) {
                            const double w = q(i, a, b) * q(j, cc, e);
                            if (w != 0.0) m += w * moment({a, b, cc, e});
                        }
            M[(size_t)i * nn + j] = m;
            // grad phi_i = sum_a 2 (Q_i lam)_a grad lam_a
            for (int a = 0; a < nb; ++a)
                for (int cc = 0; cc < nb; ++cc) {
                    double w = 0.0;
                    for (int b = 0; b < nb; ++b)
                        for (int e = 0; e < nb; ++e)
                            if (q(i, a, b) != 0.0 && q(j, cc, e) != 0.0)
                                w += 4.0 * q(i, a, b) * q(j, cc, e) * moment({b, e});
                    W[(((size_t)i * nn + j) * nb + a) * nb + cc] = w;
                }
        }
}

void launch_contrib_geometry(Ctx & c, const int * codes) {
    long nc = (long)c.ccell.n;
    hipLaunchKernelGGL(k_contrib_geometry, dim3(div_up(nc, 256)), dim3(256), 0, c.stream, nc, c.nv,
                       c.ccell.p, codes, c.grads.p, c.vol.p, c.cK.p, c.cM.p);
    PG_CHECK(hipGetLastError());
}

void launch_assemble(Ctx & c, const double * rho, double k, double * vals, bool unitRho) {
    const double * r = unitRho ? nullptr : rho;
    hipLaunchKernelGGL(k_assemble_domain, dim3(div_up(c.nnz, 256)), dim3(256), 0, c.stream, c.nnz,
                       c.cptr.p, c.ccell.p, c.cK.p, c.cM.p, r, k * k, k > 0.0 ? 1 : 0, vals);
    PG_CHECK(hipGetLastError());
    if (c.nBSlot > 0) {
        hipLaunchKernelGGL(k_assemble_boundary, dim3(div_up(c.nBSlot, 128)), dim3(128), 0, c.stream,
                           c.nBSlot, c.bSlot.p, c.bPtr.p, c.bFace.p, c.bCoef.p, c.mixCell.p, c.mixCtr.p,
                           c.mixNrm.p, r, c.srcCenter[0], c.srcCenter[1], c.srcCenter[2], k, vals);
        PG_CHECK(hipGetLastError());
    }
    launch_dirichlet(c, vals, false);
}

// B[k][node][e] = 0 for the listed nodes: one thread per (k, listed node, e)
__global__ void k_zero_rows(int nK, int count, int n, int nE, const int * __restrict__ nodes,
                            double * __restrict__ B) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nK * count * nE) return;
    const int e = (int)(t % nE);
    const long kq = t / nE;
    const int q = (int)(kq % count);
    const long kk = kq / count;
    B[((size_t)kk * n + nodes[q]) * nE + e] = 0.0;
}

void launch_zero_rows(Ctx & c, double * B) {
    const int count = (int)c.calNodes.n;
    if (count == 0) return;
    const long total = (long)c.nK * count * c.nE;
    hipLaunchKernelGGL(k_zero_rows, dim3(div_up(total, 256)), dim3(256), 0, c.stream, c.nK, count, c.n, c.nE,
                       c.calNodes.p, B);
    PG_CHECK(hipGetLastError());
}

void launch_calibration(Ctx & c, double * vals, bool imagPart) {
    if (c.nCalSlot == 0) return;
    for (int kk = 0; kk < c.nK; ++kk)
        hipLaunchKernelGGL(k_set_slots, dim3(div_up(c.nCalSlot, 256)), dim3(256), 0, c.stream, c.nCalSlot,
                           c.calSlot.p, imagPart ? nullptr : c.calVal.p, vals + (size_t)kk * c.nnz);
    PG_CHECK(hipGetLastError());
}

void launch_dirichlet(Ctx & c, double * vals, bool imagPart) {
    if (c.nDirSlot == 0) return;
    hipLaunchKernelGGL(k_set_slots, dim3(div_up(c.nDirSlot, 256)), dim3(256), 0, c.stream, c.nDirSlot,
                       c.dirSlot.p, imagPart ? nullptr : c.dirVal.p, vals);
    PG_CHECK(hipGetLastError());
}

void launch_fill_matrix(Ctx & c, int mass, const double * a, double * vals) {
    hipLaunchKernelGGL(k_fill_matrix, dim3(div_up(c.nnz, 256)), dim3(256), 0, c.stream, c.nnz, c.cptr.p,
                       c.ccell.p, mass ? c.cM.p : c.cK.p, a, vals);
    PG_CHECK(hipGetLastError());
}

void launch_map_model(Ctx & c, const double * model, int nModel, double * rho) {
    hipLaunchKernelGGL(k_map_model, dim3(div_up(c.C, 256)), dim3(256), 0, c.stream, c.C, nModel, model,
                       c.cellMarker.p, rho);
    PG_CHECK(hipGetLastError());
}

void launch_diag_inv(Ctx & c) {
    c.dinv.alloc((size_t)c.nK * c.n);
    hipLaunchKernelGGL(k_diag_inv, dim3(div_up(c.n, 256), c.nK), dim3(256), 0, c.stream, c.n, c.nK, c.nnz,
                       c.diagSlot.p, c.valsA.p, c.dinv.p);
    PG_CHECK(hipGetLastError());
}

}  // namespace pgert

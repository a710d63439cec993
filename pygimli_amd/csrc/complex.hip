// Complex resistivity (induced polarisation) kernels: the total-field path
// of DCMultiElectrodeModelling with setComplex(true)
// (core/src/bert/dcfemmodelling.cpp:163-292 assembly with complex cell
// resistivities, :1755-1928 calculateK_ with unit point sources, :1085-1120
// response, :1410-1476 Jacobian).
//
// The complex symmetric system (A_re + i A_im) x = b (b real) is solved by
// block Gauss-Seidel on its real and imaginary parts with the real solvers:
//     A_re x_re = b + A_im x_im,      A_re x_im = -A_im x_re,
// which contracts by rho(A_re^-1 A_im)^2 per sweep, about tan^2 of the
// largest cell phase (4e-4 per sweep at 20 mrad).
#include <algorithm>

#include "internal.h"
#include "bessel.h"

namespace pgert {

namespace {

constexpr double TOL = 1e-12;  // core/src/gimli.h:206

// 1 / (re + i im)
__device__ inline void inv_c(double re, double im, double & sr, double & si) {
    const double d = re * re + im * im;
    sr = re / d;
    si = -im / d;
}

// domain part, both halves: sum over the slot's contributions of
// (K + k^2 M) / rho_c; cells with |rho| <= 1e-12 contribute nothing (:185)
__global__ void k_assemble_domain_c(int nnz, const int * __restrict__ cptr, const int * __restrict__ ccell,
                                    const double * __restrict__ cK, const double * __restrict__ cM,
                                    const double * __restrict__ rre, const double * __restrict__ rim,
                                    double k2, int withMass, double * __restrict__ vre,
                                    double * __restrict__ vim) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    double ar = 0.0, ai = 0.0;
    for (int e = cptr[s]; e < cptr[s + 1]; ++e) {
        const int c = ccell[e];
        const double r = rre[c], q = rim[c];
        if (!(sqrt(r * r + q * q) > TOL)) continue;
        double sr, si;
        inv_c(r, q, sr, si);
        const double v = withMass ? (cM[e] * k2 + cK[e]) : cK[e];
        ar += v * sr;
        ai += v * si;
    }
    vre[s] = ar;
    vim[s] = ai;
}

__device__ double mixed_alpha_c(const double * ctr, const double * nrm, const double * src, double k) {
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

// mixed boundaries (:243-292): alpha / rho of the boundary's cell
__global__ void k_assemble_boundary_c(int nbs, const int * __restrict__ bslot, const int * __restrict__ bptr,
                                      const int * __restrict__ bface, const double * __restrict__ bcoef,
                                      const int * __restrict__ mixCell, const double * __restrict__ mixCtr,
                                      const double * __restrict__ mixNrm, const double * __restrict__ rre,
                                      const double * __restrict__ rim, double s0, double s1, double s2,
                                      double k, double * __restrict__ vre, double * __restrict__ vim) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbs) return;
    const double src[3] = {s0, s1, s2};
    double ar = 0.0, ai = 0.0;
    for (int e = bptr[t]; e < bptr[t + 1]; ++e) {
        int f = bface[e];
        double a = mixed_alpha_c(mixCtr + 3 * (size_t)f, mixNrm + 3 * (size_t)f, src, k);
        double sr, si;
        inv_c(rre[mixCell[f]], rim[mixCell[f]], sr, si);
        ar += bcoef[e] * (a * sr);
        ai += bcoef[e] * (a * si);
    }
    vre[bslot[t]] += ar;
    vim[bslot[t]] += ai;
}

// unit point source of electrode e, every wavenumber: 1 at its node
// (ElectrodeShapeNode::assembleRHS, core/src/bert/electrode.cpp:133-151) or
// the shape functions at its position on its cell's nodes for a free one
// (ElectrodeShapeEntity::assembleRHS, :274-287)
// refNode >= 0: every source is a dipole against the reference electrode
// there (createCurrentPattern, dcfemmodelling.cpp:1517-1523; calculateK_
// :1854-1855 adds eB's -1 after eA's +1)
__global__ void k_point_rhs(long total, int n, int nE, const int * __restrict__ shapeNode,
                            const double * __restrict__ shapeW, int refNode, double * __restrict__ B) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int e = (int)(t % nE);
    const int node = (int)((t / nE) % n);
    double v = 0.0;
    for (int j = 0; j < SHAPE_NODES; ++j)
        if (shapeNode[SHAPE_NODES * e + j] == node) v = shapeW[SHAPE_NODES * e + j];
    if (node == refNode) v -= 1.0;
    B[t] = v;
}

// out[k] = base[k] + sign * A_k X[k] for node-major blocks [k][node][e]
__global__ void k_spmm_add(int n, int nE, int nnz, const int * __restrict__ rowptr,
                           const int * __restrict__ colidx, const double * __restrict__ vals,
                           const double * __restrict__ X, const double * __restrict__ base, double sign,
                           double * __restrict__ out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int kk = blockIdx.y;
    if (t >= (long)n * nE) return;
    const int row = (int)(t / nE), e = (int)(t % nE);
    const double * A = vals + (size_t)kk * nnz;
    const double * Xk = X + (size_t)kk * n * nE;
    double acc = 0.0;
    for (int q = rowptr[row]; q < rowptr[row + 1]; ++q) acc += A[q] * Xk[(size_t)colidx[q] * nE + e];
    const size_t o = (size_t)kk * n * nE + t;
    out[o] = (base ? base[o] : 0.0) + sign * acc;
}

// acc[0] += |a - b|^2, acc[1] += |a|^2 (one atomic per block)
__global__ void k_change(long total, const double * __restrict__ a, const double * __restrict__ b,
                         double * __restrict__ part) {
    __shared__ double s0[256], s1[256];
    double d2 = 0.0, a2 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const double d = a[t] - b[t];
        d2 += d * d;
        a2 += a[t] * a[t];
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

// fixed-order sum of the per-workgroup partials of k_change (one workgroup)
__global__ void k_change_sum(int nPart, const double * __restrict__ part, double * __restrict__ acc) {
    __shared__ double s0[256], s1[256];
    double d2 = 0.0, a2 = 0.0;
    for (int p = threadIdx.x; p < nPart; p += blockDim.x) {
        d2 += part[2 * p];
        a2 += part[2 * p + 1];
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[0] = s0[0];
        acc[1] = s1[0];
    }
}

// response (:1103-1120): k * ((uAM - uAN) - (uBM - uBN)) of the real and the
// imaginary electrode maps (DataMap::data, datamap.cpp:159-231)
__global__ void k_response_c(int D, int nE, const int * __restrict__ abmn, const double * __restrict__ kfac,
                             const double * __restrict__ mre, const double * __restrict__ mim,
                             double * __restrict__ out_re, double * __restrict__ out_im) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const int a = abmn[4 * i], b = abmn[4 * i + 1], m = abmn[4 * i + 2], n = abmn[4 * i + 3];
    double u[2];
    const double * mp[2] = {mre, mim};
    for (int h = 0; h < 2; ++h) {
        const double * M = mp[h];
        const double am = (a != -1 && m != -1) ? M[a * nE + m] : 0.0;
        const double an = (a != -1 && n != -1) ? M[a * nE + n] : 0.0;
        const double bm = (b != -1 && m != -1) ? M[b * nE + m] : 0.0;
        const double bn = (b != -1 && n != -1) ? M[b * nE + n] : 0.0;
        u[h] = (am - an) - (bm - bn);
    }
    out_re[i] = u[0] * kfac[i];
    out_im[i] = u[1] * kfac[i];
}

// map[e][f] = sum_j N_fj sum_k w_k U[k][node_fj][e] (as k_electrode_map)
__global__ void k_electrode_map_c(int n, int nE, int nK, const double * __restrict__ U,
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

// The complex Jacobian in the two output arrays, without scratch blocks:
// J_re = S(re), J_im = S(im) on entry; afterwards J_re = S(re) - S(im) and
// J_im = -(S(re) + S(im)), to which the sensitivity launch of re + im adds
// S(re + im), leaving J = (S(re) - S(im)) + i (S(re + im) - S(re) - S(im)).
__global__ void k_jacobian_c_pre(long total, double * __restrict__ Jre, double * __restrict__ Jim) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const double a = Jre[t], b = Jim[t];
    Jre[t] = a - b;
    Jim[t] = -(a + b);
}

// J / (m^2 / k) in place (createJacobian_, dcfemmodelling.cpp:1410-1444); m = mr + i mi
__global__ void k_jacobian_c_scale(long total, int nParam, const double * __restrict__ kfac,
                                   const double * __restrict__ mr, const double * __restrict__ mi,
                                   double * __restrict__ Jre, double * __restrict__ Jim) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int j = (int)(t % nParam);
    const long i = t / nParam;
    const double zr = Jre[t], zi = Jim[t];
    // q = m^2 / k
    const double a = mr[j], b = mi[j];
    const double qr = (a * a - b * b) / kfac[i], qi = (2.0 * a * b) / kfac[i];
    const double d = qr * qr + qi * qi;
    Jre[t] = (zr * qr + zi * qi) / d;
    Jim[t] = (zi * qr - zr * qi) / d;
}

// bypass conductances into every wavenumber's matrix, one thread per matrix
// (a node shared by two resistors takes both, in order)
__global__ void k_bypass(int nK, int nnz, int nSlot, const int * __restrict__ slot, const double * __restrict__ val,
                         double * __restrict__ vals) {
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    if (kk >= nK) return;
    double * A = vals + (size_t)kk * nnz;
    for (int i = 0; i < nSlot; ++i) A[slot[i]] += val[i];
}

__global__ void k_add(long total, const double * __restrict__ a, const double * __restrict__ b,
                      double * __restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) out[t] = a[t] + b[t];
}

}  // namespace

void launch_assemble_complex(Ctx & c, const double * rre, const double * rim, double k, double * vre,
                             double * vim) {
    hipLaunchKernelGGL(k_assemble_domain_c, dim3(div_up(c.nnz, 256)), dim3(256), 0, c.stream, c.nnz, c.cptr.p,
                       c.ccell.p, c.cK.p, c.cM.p, rre, rim, k * k, k > 0.0 ? 1 : 0, vre, vim);
    PG_CHECK(hipGetLastError());
    if (c.nBSlot > 0) {
        hipLaunchKernelGGL(k_assemble_boundary_c, dim3(div_up(c.nBSlot, 128)), dim3(128), 0, c.stream, c.nBSlot,
                           c.bSlot.p, c.bPtr.p, c.bFace.p, c.bCoef.p, c.mixCell.p, c.mixCtr.p, c.mixNrm.p, rre,
                           rim, c.srcCenter[0], c.srcCenter[1], c.srcCenter[2], k, vre, vim);
        PG_CHECK(hipGetLastError());
    }
    launch_dirichlet(c, vre, false);
    launch_dirichlet(c, vim, true);
}

void launch_point_rhs(Ctx & c, double * B) {
    const long total = (long)c.nK * c.n * c.nE;
    hipLaunchKernelGGL(k_point_rhs, dim3(div_up(total, 256)), dim3(256), 0, c.stream, total, c.n, c.nE,
                       c.shapeNode.p, c.shapeW.p, c.refNode, B);
    PG_CHECK(hipGetLastError());
}

void launch_spmm_add(Ctx & c, const double * vals, const double * X, const double * base, double sign,
                     double * out) {
    const long rows = (long)c.n * c.nE;
    hipLaunchKernelGGL(k_spmm_add, dim3(div_up(rows, 256), c.nK), dim3(256), 0, c.stream, c.n, c.nE, c.nnz,
                       c.rowptr.p, c.colidx.p, vals, X, base, sign, out);
    PG_CHECK(hipGetLastError());
}

// |a - b|^2 and |a|^2 into acc[0..1] (device), zeroed first
void launch_change(Ctx & c, long total, const double * a, const double * b, double * acc) {
    const int grid = std::max(1, (int)std::min<long>(1024, div_up(total, 256)));
    c.chgPart.reserve(2 * 1024);
    hipLaunchKernelGGL(k_change, dim3(grid), dim3(256), 0, c.stream, total, a, b, c.chgPart.p);
    PG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_change_sum, dim3(1), dim3(256), 0, c.stream, grid, c.chgPart.p, acc);
    PG_CHECK(hipGetLastError());
}

void launch_response_complex(Ctx & c, const double * Ure, const double * Uim, double * out_re,
                             double * out_im) {
    c.emap.alloc((size_t)2 * c.nE * c.nE);
    const int ne2 = c.nE * c.nE;
    hipLaunchKernelGGL(k_electrode_map_c, dim3(div_up(ne2, 256)), dim3(256), 0, c.stream, c.n, c.nE, c.nK, Ure,
                       c.dw.p, c.shapeNode.p, c.shapeW.p, c.emap.p);
    PG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_electrode_map_c, dim3(div_up(ne2, 256)), dim3(256), 0, c.stream, c.n, c.nE, c.nK, Uim,
                       c.dw.p, c.shapeNode.p, c.shapeW.p, c.emap.p + ne2);
    PG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_response_c, dim3(div_up(c.D, 256)), dim3(256), 0, c.stream, c.D, c.nE, c.abmn.p,
                       c.kfac.p, c.emap.p, c.emap.p + ne2, out_re, out_im);
    PG_CHECK(hipGetLastError());
}

// the complete electrode model's terms, then the bypass resistors, as the
// reference adds them (calculateK_, dcfemmodelling.cpp:1790-1833)
void launch_bypass(Ctx & c, double * vals) {
    if (c.nCem > 0) {
        hipLaunchKernelGGL(k_bypass, dim3(div_up(c.nK, 64)), dim3(64), 0, c.stream, c.nK, c.nnz, c.nCem,
                           c.cemSlot.p, c.cemVal.p, vals);
        PG_CHECK(hipGetLastError());
    }
    if (c.nByp == 0) return;
    hipLaunchKernelGGL(k_bypass, dim3(div_up(c.nK, 64)), dim3(64), 0, c.stream, c.nK, c.nnz, 4 * c.nByp,
                       c.bypSlot.p, c.bypVal.p, vals);
    PG_CHECK(hipGetLastError());
}

void launch_add(Ctx & c, long total, const double * a, const double * b, double * out) {
    hipLaunchKernelGGL(k_add, dim3(div_up(total, 256)), dim3(256), 0, c.stream, total, a, b, out);
    PG_CHECK(hipGetLastError());
}

void launch_jacobian_complex_pre(Ctx & c, double * Jre, double * Jim) {
    const long total = (long)c.D * c.nParam;
    hipLaunchKernelGGL(k_jacobian_c_pre, dim3(div_up(total, 256)), dim3(256), 0, c.stream, total, Jre, Jim);
    PG_CHECK(hipGetLastError());
}

void launch_jacobian_complex_scale(Ctx & c, const double * mr, const double * mi, double * Jre, double * Jim) {
    const long total = (long)c.D * c.nParam;
    hipLaunchKernelGGL(k_jacobian_c_scale, dim3(div_up(total, 256)), dim3(256), 0, c.stream, total, c.nParam,
                       c.kfac.p, mr, mi, Jre, Jim);
    PG_CHECK(hipGetLastError());
}

}  // namespace pgert

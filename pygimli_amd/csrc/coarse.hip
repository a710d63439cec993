// Two-level (aggregation) preconditioner for the block solver.
//
// The mixed-boundary ERT systems are nearly singular at small wavenumbers and
// on padded, strongly graded meshes: their lowest modes are smooth and
// Jacobi barely touches them.  An additive coarse correction on m aggregates
// (recursive coordinate bisection of the nodes, built once per mesh) removes
// them:  z = D^-1 r + Z E^-1 Z^T r,  Z the aggregate indicator, E = Z^T A Z.
// E is assembled from the CSR values through a fixed slot -> coarse-entry map
// (one thread per coarse entry, no atomics), inverted once per assembly with a
// blocked Gauss-Jordan sweep (FP64), and applied per iteration as
// restriction (aggregate sums) -> dense m x m times m x s product ->
// prolongation fused into the solver's Z formation.
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace pgert {

namespace {

constexpr int GJB = 32;  // Gauss-Jordan block size

// one wavefront per coarse entry: lane l sums slots l, l + 64, ... of the
// entry, then a fixed butterfly over the lanes (deterministic)
__global__ void __launch_bounds__(256) k_coarse_assemble(int nce, int m, const int * __restrict__ cptr,
                                                         const int * __restrict__ cslot, const int * __restrict__ cij,
                                                         const double * __restrict__ vals, double * __restrict__ E) {
    const int t = blockIdx.x * 4 + threadIdx.x / 64, l = threadIdx.x % 64;
    if (t >= nce) return;
    double s = 0.0;
    for (int q = cptr[t] + l; q < cptr[t + 1]; q += 64) s += vals[cslot[q]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (l == 0) E[(size_t)cij[2 * t] * m + cij[2 * t + 1]] = s;
}

// (a) Pinv = inverse of the GJB x GJB diagonal block K (one workgroup, LDS
// Gauss-Jordan, SPD: no pivoting)
__global__ void __launch_bounds__(256) k_gj_pivot(int m, int K0, const double * __restrict__ M,
                                                  double * __restrict__ Pinv) {
    __shared__ double A[GJB][GJB + 1];
    const int bs = min(GJB, m - K0);
    for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
        int i = t / GJB, j = t % GJB;
        A[i][j] = (i < bs && j < bs) ? M[(size_t)(K0 + i) * m + K0 + j] : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int k = 0; k < GJB; ++k) {
        double p = A[k][k];
        __syncthreads();
        // row k scaled (column k becomes 1/p)
        if (threadIdx.x < GJB) {
            int j = threadIdx.x;
            A[k][j] = (j == k) ? 1.0 / p : A[k][j] / p;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) {
            int i = t / GJB, j = t % GJB;
            if (i == k) continue;
            double f = A[i][k];
            if (j == k) continue;
            A[i][j] -= f * A[k][j];
        }
        __syncthreads();
        if (threadIdx.x < GJB && threadIdx.x != k) {
            int i = threadIdx.x;
            A[i][k] = -A[i][k] * A[k][k];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < GJB * GJB; t += blockDim.x) Pinv[t] = A[t / GJB][t % GJB];
}

// (b) column panel F = M[:, K] and row panel R = Pinv M[K][:] with
// R[:, K] = Pinv: one thread per (panel row, column); F and R are built by
// the same launch (R reads only M's block row K, which F does not change)
__global__ void k_gj_panels(int m, int K0, const double * __restrict__ M, const double * __restrict__ Pinv,
                            double * __restrict__ F, double * __restrict__ R) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)GJB * m) return;
    const int i = (int)(t / m), j = (int)(t % m);  // R: row i of the panel, column j
    const int bs = min(GJB, m - K0);
    // F[j][i] = M[j][K0 + i] (row j of M, panel column i)
    F[(size_t)j * GJB + i] = i < bs ? M[(size_t)j * m + K0 + i] : 0.0;
    if (i >= bs) return;
    double s = 0.0;
    if (j >= K0 && j < K0 + bs) {
        s = Pinv[i * GJB + (j - K0)];
    } else {
        for (int q = 0; q < bs; ++q) s += Pinv[i * GJB + q] * M[(size_t)(K0 + q) * m + j];
    }
    R[(size_t)i * m + j] = s;
}

// (c) M[i][j] = (j in K ? 0 : M[i][j]) - F[i] . R[:, j] for i not in K; M[K][:] = R.
// A workgroup takes a 32 x 32 tile; the tile's rows of F and columns of R
// are staged in LDS.
__global__ void __launch_bounds__(256) k_gj_update(int m, int K0, double * __restrict__ M,
                                                   const double * __restrict__ F, const double * __restrict__ R) {
    __shared__ double sF[32][GJB + 1], sR[GJB][32 + 1];
    const int i0 = 32 * blockIdx.y, j0 = 32 * blockIdx.x;
    const int bs = min(GJB, m - K0);
    for (int t = threadIdx.x; t < 32 * GJB; t += blockDim.x) {
        const int a = t / GJB, q = t % GJB;
        sF[a][q] = i0 + a < m ? F[(size_t)(i0 + a) * GJB + q] : 0.0;
        const int qq = t / 32, b = t % 32;
        sR[qq][b] = (qq < bs && j0 + b < m) ? R[(size_t)qq * m + j0 + b] : 0.0;
    }
    __syncthreads();
    const int b = threadIdx.x % 32, a0 = threadIdx.x / 32;  // 8 rows a0 + 8 r
    const int j = j0 + b;
    if (j >= m) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int a = a0 + 8 * r, i = i0 + a;
        if (i >= m) continue;
        const size_t t = (size_t)i * m + j;
        if (i >= K0 && i < K0 + bs) {
            M[t] = sR[i - K0][b];
            continue;
        }
        double v = (j >= K0 && j < K0 + bs) ? 0.0 : M[t];
        for (int q = 0; q < bs; ++q) v -= sF[a][q] * sR[q][b];
        M[t] = v;
    }
}

// (A Z)[r][I] = sum of row r's values in the columns of aggregate I
__global__ void k_az_values(int nAZ, const int * __restrict__ slotPtr, const int * __restrict__ slot,
                            const double * __restrict__ vals, double * __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nAZ) return;
    double s = 0.0;
    for (int q = slotPtr[t]; q < slotPtr[t + 1]; ++q) s += vals[slot[q]];
    out[t] = s;
}

// A D^-1 (column scaling); per row the Gershgorin radius of D^-1 A,
// dinv_r sum_q |a_rq|, reduced to a maximum per workgroup
constexpr int GB_WG = 256;
__global__ void __launch_bounds__(GB_WG) k_ad_values(int n, const int * __restrict__ rowptr,
                                                     const int * __restrict__ colidx, const double * __restrict__ vals,
                                                     const double * __restrict__ dinv, double * __restrict__ ad,
                                                     double * __restrict__ part) {
    __shared__ double red[GB_WG / 64];
    double mx = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int q = rowptr[r]; q < rowptr[r + 1]; ++q) {
            const double v = vals[q];
            ad[q] = v * dinv[colidx[q]];
            a += fabs(v);
        }
        mx = fmax(mx, a * dinv[r]);
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = mx;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

}  // namespace

// Per-solve data of the multiplicative two-level preconditioner with
// Chebyshev smoothing: A Z and A D^-1 for every wavenumber, and the
// Gershgorin bound of D^-1 A (the top of the smoothing interval) on the host.
void smooth_prepare(Ctx & c) {
    PG_REQUIRE(c.nAgg > 0 && c.nAZ > 0, "smoothing needs the coarse space (pgert_set_coarse)");
    const int gb = std::min(1024, div_up(c.n, GB_WG));
    c.azVals.alloc((size_t)c.nK * c.nAZ);
    c.valsAD.alloc((size_t)c.nK * c.nnz);
    c.gershPart.alloc((size_t)c.nK * gb);
    for (int kk = 0; kk < c.nK; ++kk) {
        hipLaunchKernelGGL(k_az_values, dim3(div_up(c.nAZ, 256)), dim3(256), 0, c.stream, c.nAZ, c.azSlotPtr.p,
                           c.azSlot.p, c.valsA.p + (size_t)kk * c.nnz, c.azVals.p + (size_t)kk * c.nAZ);
        hipLaunchKernelGGL(k_ad_values, dim3(gb), dim3(GB_WG), 0, c.stream, c.n, c.rowptr.p, c.colidx.p,
                           c.valsA.p + (size_t)kk * c.nnz, c.dinv.p + (size_t)kk * c.n,
                           c.valsAD.p + (size_t)kk * c.nnz, c.gershPart.p + (size_t)kk * gb);
    }
    PG_CHECK(hipGetLastError());
    std::vector<double> h((size_t)c.nK * gb);
    PG_CHECK(hipMemcpyAsync(h.data(), c.gershPart.p, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
    PG_CHECK(hipStreamSynchronize(c.stream));
    c.hGersh.assign(c.nK, 0.0);
    for (int kk = 0; kk < c.nK; ++kk)
        for (int g = 0; g < gb; ++g) c.hGersh[kk] = std::max(c.hGersh[kk], h[(size_t)kk * gb + g]);
    for (int kk = 0; kk < c.nK; ++kk)
        PG_REQUIRE(std::isfinite(c.hGersh[kk]) && c.hGersh[kk] > 0.0, "smoothing: no finite Gershgorin bound");
}

void coarse_prepare(Ctx & c) {
    if (c.nAgg <= 0) return;
    const int m = c.nAgg;
    c.Einv.alloc((size_t)c.nK * m * m);
    DBuf<double> Pinv, F, R;
    Pinv.alloc(GJB * GJB);
    F.alloc((size_t)m * GJB);
    R.alloc((size_t)GJB * m);
    for (int kk = 0; kk < c.nK; ++kk) {
        double * M = c.Einv.p + (size_t)kk * m * m;
        PG_CHECK(hipMemsetAsync(M, 0, (size_t)m * m * 8, c.stream));
        hipLaunchKernelGGL(k_coarse_assemble, dim3(div_up(c.nCE, 4)), dim3(256), 0, c.stream, c.nCE, m, c.cePtr.p,
                           c.ceSlot.p, c.ceIJ.p, c.valsA.p + (size_t)kk * c.nnz, M);
        for (int K0 = 0; K0 < m; K0 += GJB) {
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, c.stream, m, K0, M, Pinv.p);
            hipLaunchKernelGGL(k_gj_panels, dim3(div_up((long)GJB * m, 256)), dim3(256), 0, c.stream, m, K0, M, Pinv.p,
                               F.p, R.p);
            hipLaunchKernelGGL(k_gj_update, dim3(div_up(m, 32), div_up(m, 32)), dim3(256), 0, c.stream, m, K0, M, F.p,
                               R.p);
        }
        PG_CHECK(hipGetLastError());
    }
    PG_CHECK(hipStreamSynchronize(c.stream));  // scratch panels are released here
}

}  // namespace pgert

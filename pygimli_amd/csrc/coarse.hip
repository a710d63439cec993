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

__global__ void k_coarse_assemble(int nce, int m, const int * __restrict__ cptr, const int * __restrict__ cslot,
                                  const int * __restrict__ cij, const double * __restrict__ vals,
                                  double * __restrict__ E) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nce) return;
    double s = 0.0;
    for (int q = cptr[t]; q < cptr[t + 1]; ++q) s += vals[cslot[q]];
    E[(size_t)cij[2 * t] * m + cij[2 * t + 1]] = s;
}

// (a) Pinv = inverse of the GJB x GJB diagonal block K (one workgroup, LDS
// Gauss-Jordan, SPD: no pivoting); also saves the column panel F = M[:, K]
__global__ void __launch_bounds__(256) k_gj_pivot(int m, int K0, double * __restrict__ M, double * __restrict__ Pinv,
                                                  double * __restrict__ F) {
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
    for (int t = threadIdx.x; t < m * GJB; t += blockDim.x) {
        int i = t / GJB, q = t % GJB;
        F[t] = q < bs ? M[(size_t)i * m + K0 + q] : 0.0;
    }
}

// (b) row panel R = Pinv M[K][:] with R[:, K] = Pinv
__global__ void k_gj_rowpanel(int m, int K0, const double * __restrict__ M, const double * __restrict__ Pinv,
                              double * __restrict__ R) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int bs = min(GJB, m - K0);
    for (int i = 0; i < bs; ++i) {
        double s = 0.0;
        if (j >= K0 && j < K0 + bs) {
            s = Pinv[i * GJB + (j - K0)];
        } else {
            for (int q = 0; q < bs; ++q) s += Pinv[i * GJB + q] * M[(size_t)(K0 + q) * m + j];
        }
        R[(size_t)i * m + j] = s;
    }
}

// (c) M[i][j] = (j in K ? 0 : M[i][j]) - F[i] . R[:, j] for i not in K; M[K][:] = R
__global__ void k_gj_update(int m, int K0, double * __restrict__ M, const double * __restrict__ F,
                            const double * __restrict__ R) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)m * m) return;
    int i = (int)(t / m), j = (int)(t % m);
    const int bs = min(GJB, m - K0);
    if (i >= K0 && i < K0 + bs) {
        M[t] = R[(size_t)(i - K0) * m + j];
        return;
    }
    double v = (j >= K0 && j < K0 + bs) ? 0.0 : M[t];
    for (int q = 0; q < bs; ++q) v -= F[(size_t)i * GJB + q] * R[(size_t)q * m + j];
    M[t] = v;
}

}  // namespace

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
        hipLaunchKernelGGL(k_coarse_assemble, dim3(div_up(c.nCE, 256)), dim3(256), 0, c.stream, c.nCE, m, c.cePtr.p,
                           c.ceSlot.p, c.ceIJ.p, c.valsA.p + (size_t)kk * c.nnz, M);
        for (int K0 = 0; K0 < m; K0 += GJB) {
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, c.stream, m, K0, M, Pinv.p, F.p);
            hipLaunchKernelGGL(k_gj_rowpanel, dim3(div_up(m, 256)), dim3(256), 0, c.stream, m, K0, M, Pinv.p, R.p);
            hipLaunchKernelGGL(k_gj_update, dim3(div_up((long)m * m, 256)), dim3(256), 0, c.stream, m, K0, M, F.p, R.p);
        }
        PG_CHECK(hipGetLastError());
    }
    PG_CHECK(hipStreamSynchronize(c.stream));  // scratch panels are released here
}

}  // namespace pgert

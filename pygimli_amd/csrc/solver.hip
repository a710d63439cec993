// Multi-RHS Krylov solvers replacing the reference's per-wavenumber CHOLMOD
// factorisation + per-source solves (LinSolver/CHOLMODWrapper,
// core/src/linSolver.cpp:96-178, cholmodWrapper.cpp:224-459).
//
// All (wavenumber, electrode) systems are solved at once.  Matrices: one CSR
// value array per wavenumber on a shared pattern.  Vectors: node-major blocks
// V[k][node][e].  The SpMM is one wavefront per (row, k) with lanes over the
// electrodes, so every nonzero reads a contiguous nE-vector of X.
//
// This file holds the plain method: independent-column Jacobi-PCG in FP64
// with per-column scalars and deterministic two-stage reductions.
#include "internal.h"

#include <algorithm>
#include <cmath>

namespace pgert {

namespace {

constexpr int WAVES = 4;      // wavefronts per 256-thread workgroup
constexpr int NGROUPS = 1024; // workgroups per wavenumber for reducing passes
constexpr int NS = 8;         // scalars per column
enum { S_RZ = 0, S_PQ, S_RR, S_BB, S_ACT, S_ALPHA, S_BETA, S_AUX };

// Sum the per-wave partials of up to 4 chunks x NQ quantities across the
// workgroup and store them at part[(k * G + g) * nE * NQ + e * NQ + q].
template <int NQ>
__device__ void store_partials(double (&acc)[4][NQ], int nE, double * part, double * lds) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nch = (nE + 63) / 64;
    for (int ch = 0; ch < nch; ++ch) {
        for (int q = 0; q < NQ; ++q) lds[(wave * 64 + lane) * NQ + q] = acc[ch][q];
        __syncthreads();
        if (wave == 0) {
            int e = ch * 64 + lane;
            if (e < nE) {
                for (int q = 0; q < NQ; ++q) {
                    double s = 0;
                    for (int w = 0; w < WAVES; ++w) s += lds[(w * 64 + lane) * NQ + q];
                    part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * nE * NQ + (size_t)e * NQ + q] = s;
                }
            }
        }
        __syncthreads();
    }
}

// Y = A X (+ partial dot(X, Y) per column when part != nullptr)
__global__ void __launch_bounds__(256) k_spmm_dot(int n, int nE, int nnz, const int * __restrict__ rowptr,
                                                  const int * __restrict__ colidx,
                                                  const double * __restrict__ vals,
                                                  const double * __restrict__ X, double * __restrict__ Y,
                                                  double * __restrict__ part) {
    __shared__ double lds[256 * 1];
    const int kk = blockIdx.y;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nch = (nE + 63) / 64;
    vals += (size_t)kk * nnz;
    X += (size_t)kk * n * nE;
    Y += (size_t)kk * n * nE;
    double acc[4][1] = {{0}, {0}, {0}, {0}};
    for (int row = blockIdx.x * WAVES + wave; row < n; row += gridDim.x * WAVES) {
        const int q0 = rowptr[row], q1 = rowptr[row + 1];
        for (int ch = 0; ch < nch; ++ch) {
            int e = ch * 64 + lane;
            if (e < nE) {
                double y = 0.0;
                for (int q = q0; q < q1; ++q) y += vals[q] * X[(size_t)colidx[q] * nE + e];
                Y[(size_t)row * nE + e] = y;
                if (part) acc[ch][0] += X[(size_t)row * nE + e] * y;
            }
        }
    }
    if (part) store_partials<1>(acc, nE, part, lds);
}

// R = B - Y; P = D^-1 R; partials (r.z, r.r, b.b)
__global__ void __launch_bounds__(256) k_pcg_init(int n, int nE, const double * __restrict__ B,
                                                  const double * __restrict__ Y, const double * __restrict__ dinv,
                                                  double * __restrict__ R, double * __restrict__ P,
                                                  double * __restrict__ part) {
    __shared__ double lds[256 * 3];
    const int kk = blockIdx.y;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nch = (nE + 63) / 64;
    const size_t off = (size_t)kk * n * nE;
    dinv += (size_t)kk * n;
    double acc[4][3] = {};
    for (int row = blockIdx.x * WAVES + wave; row < n; row += gridDim.x * WAVES) {
        for (int ch = 0; ch < nch; ++ch) {
            int e = ch * 64 + lane;
            if (e < nE) {
                size_t t = off + (size_t)row * nE + e;
                double b = B[t];
                double r = b - Y[t];
                double z = dinv[row] * r;
                R[t] = r;
                P[t] = z;
                acc[ch][0] += r * z;
                acc[ch][1] += r * r;
                acc[ch][2] += b * b;
            }
        }
    }
    store_partials<3>(acc, nE, part, lds);
}

// X += alpha P; R -= alpha Q; partials (r.z, r.r) with z = D^-1 r
__global__ void __launch_bounds__(256) k_pcg_update(int n, int nE, const double * __restrict__ scal,
                                                    const double * __restrict__ P, const double * __restrict__ Q,
                                                    const double * __restrict__ dinv, double * __restrict__ X,
                                                    double * __restrict__ R, double * __restrict__ part) {
    __shared__ double lds[256 * 2];
    const int kk = blockIdx.y;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nch = (nE + 63) / 64;
    const size_t off = (size_t)kk * n * nE;
    dinv += (size_t)kk * n;
    double alpha[4];
    for (int ch = 0; ch < nch; ++ch) {
        int e = ch * 64 + lane;
        alpha[ch] = e < nE ? scal[((size_t)kk * nE + e) * NS + S_ALPHA] : 0.0;
    }
    double acc[4][2] = {};
    for (int row = blockIdx.x * WAVES + wave; row < n; row += gridDim.x * WAVES) {
        for (int ch = 0; ch < nch; ++ch) {
            int e = ch * 64 + lane;
            if (e < nE) {
                size_t t = off + (size_t)row * nE + e;
                double r = R[t] - alpha[ch] * Q[t];
                X[t] += alpha[ch] * P[t];
                R[t] = r;
                acc[ch][0] += r * (dinv[row] * r);
                acc[ch][1] += r * r;
            }
        }
    }
    store_partials<2>(acc, nE, part, lds);
}

// P = D^-1 R + beta P
__global__ void __launch_bounds__(256) k_pcg_pupdate(int n, int nE, const double * __restrict__ scal,
                                                     const double * __restrict__ R,
                                                     const double * __restrict__ dinv,
                                                     double * __restrict__ P) {
    const int kk = blockIdx.y;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nl = (long)n * nE;
    if (t >= nl) return;
    int e = (int)(t % nE);
    int row = (int)(t / nE);
    size_t g = (size_t)kk * nl + t;
    double beta = scal[((size_t)kk * nE + e) * NS + S_BETA];
    P[g] = dinv[(size_t)kk * n + row] * R[g] + beta * P[g];
}

enum { RED_INIT = 0, RED_ALPHA = 1, RED_BETA = 2 };

// Final reduction over the NGROUPS partials (fixed order: deterministic) and
// the per-column scalar recurrences.
__global__ void k_pcg_reduce(int nK, int nE, int G, int stage, double tol2, const double * __restrict__ part,
                             double * __restrict__ scal) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nK * nE) return;
    int kk = t / nE, e = t % nE;
    const int NQ = stage == RED_INIT ? 3 : (stage == RED_ALPHA ? 1 : 2);
    double s[3] = {0, 0, 0};
    for (int g = 0; g < G; ++g)
        for (int q = 0; q < NQ; ++q) s[q] += part[((size_t)kk * G + g) * nE * NQ + (size_t)e * NQ + q];
    double * sc = scal + (size_t)t * NS;
    if (stage == RED_INIT) {
        sc[S_RZ] = s[0];
        sc[S_RR] = s[1];
        sc[S_BB] = s[2];
        sc[S_ACT] = (s[2] > 0.0 && s[1] > tol2 * s[2]) ? 1.0 : 0.0;
        sc[S_BETA] = 0.0;
    } else if (stage == RED_ALPHA) {
        sc[S_PQ] = s[0];
        sc[S_ALPHA] = (sc[S_ACT] != 0.0 && s[0] != 0.0) ? sc[S_RZ] / s[0] : 0.0;
    } else {
        bool act = sc[S_ACT] != 0.0;
        sc[S_BETA] = (act && sc[S_RZ] != 0.0) ? s[0] / sc[S_RZ] : 0.0;
        if (act) {
            sc[S_RZ] = s[0];
            sc[S_RR] = s[1];
        }
        sc[S_ACT] = (act && s[1] > tol2 * sc[S_BB]) ? 1.0 : 0.0;
    }
}

// partials of ||B - A X||^2 and ||B||^2 per column
__device__ inline double readlane_d(double v, int j) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// partial ||b - A x||^2 and ||b||^2 per column; a wavefront per row, the
// row's entries broadcast from lane registers (independent gathers)
__global__ void __launch_bounds__(256) k_resid(int n, int nE, int nnz, const int * __restrict__ rowptr,
                                               const int * __restrict__ colidx, const double * __restrict__ vals,
                                               const double * __restrict__ X, const double * __restrict__ B,
                                               double * __restrict__ part) {
    __shared__ double lds[256 * 2];
    const int kk = blockIdx.y;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nch = (nE + 63) / 64;
    vals += (size_t)kk * nnz;
    X += (size_t)kk * n * nE;
    B += (size_t)kk * n * nE;
    double acc[4][2] = {};
    for (int row = blockIdx.x * WAVES + wave; row < n; row += gridDim.x * WAVES) {
        const int q0 = rowptr[row], q1 = rowptr[row + 1];
        for (int ch = 0; ch < nch; ++ch) {
            const int e = ch * 64 + lane;
            const bool ok = e < nE;
            double y = 0.0;
            for (int base = q0; base < q1; base += 64) {
                const int m = min(64, q1 - base);
                const int ci = lane < m ? colidx[base + lane] : 0;
                const double cv = lane < m ? vals[base + lane] : 0.0;
                // groups of 8 entries (lanes >= m hold column 0 with value 0)
                for (int j0 = 0; j0 < m; j0 += 8) {
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) {
                        const int col = __builtin_amdgcn_readlane(ci, j0 + jj);
                        const double v = readlane_d(cv, j0 + jj);
                        if (ok) y += v * X[(size_t)col * nE + e];
                    }
                }
            }
            if (ok) {
                const double b = B[(size_t)row * nE + e];
                const double r = b - y;
                acc[ch][0] += r * r;
                acc[ch][1] += b * b;
            }
        }
    }
    store_partials<2>(acc, nE, part, lds);
}

// one wavefront per (k, e) column: lanes over the G partials, then a
// butterfly over the wavefront (fixed order, deterministic)
__global__ void __launch_bounds__(64) k_resid_reduce(int nK, int nE, int G, const double * __restrict__ part,
                                                     double * __restrict__ out) {
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= nK * nE) return;
    const int kk = t / nE, e = t % nE;
    double rr = 0, bb = 0;
    for (int g = lane; g < G; g += 64) {
        const double * p = part + ((size_t)kk * G + g) * nE * 2 + (size_t)e * 2;
        rr += p[0];
        bb += p[1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        rr += __shfl_xor(rr, off);
        bb += __shfl_xor(bb, off);
    }
    if (lane == 0) {
        out[2 * t] = rr;
        out[2 * t + 1] = bb;
    }
}

}  // namespace

// max over columns of ||B - A X|| / ||B|| (FP64), synchronous
double residual_check(Ctx & c) {
    const int nK = c.nK, nE = c.nE;
    c.part.alloc((size_t)nK * NGROUPS * nE * 3);
    c.scal.alloc((size_t)nK * nE * NS);
    const int G = std::max(1, std::min(NGROUPS, div_up(c.n, 16 * WAVES)));  // >= 16 rows per wavefront
    hipLaunchKernelGGL(k_resid, dim3(G, nK), dim3(256), 0, c.stream, c.n, nE, c.nnz, c.rowptr.p,
                       c.colidx.p, c.valsA.p, c.X.p, c.B.p, c.part.p);
    PG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_resid_reduce, dim3(nK * nE), dim3(64), 0, c.stream, nK, nE, G, c.part.p, c.scal.p);
    PG_CHECK(hipGetLastError());
    std::vector<double> h((size_t)nK * nE * 2);
    PG_CHECK(hipMemcpyAsync(h.data(), c.scal.p, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
    PG_CHECK(hipStreamSynchronize(c.stream));
    // a non-finite residual or right-hand-side norm (NaN from a bad initial
    // guess, overflow) counts as a failed solve: +inf, never <= rtol
    double worst = 0;
    for (size_t t = 0; t < (size_t)nK * nE; ++t) {
        if (!std::isfinite(h[2 * t]) || !std::isfinite(h[2 * t + 1])) return HUGE_VAL;
        if (h[2 * t + 1] > 0) worst = std::max(worst, std::sqrt(h[2 * t] / h[2 * t + 1]));
    }
    return worst;
}

void solve_pcg_fp64(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st) {
    const int nK = c.nK, nE = c.nE, n = c.n;
    const size_t tot = (size_t)nK * n * nE;
    c.R.alloc(tot);
    c.P.alloc(tot);
    c.Q.alloc(tot);
    c.part.alloc((size_t)nK * NGROUPS * nE * 3);
    c.scal.alloc((size_t)nK * nE * NS);
    hipStream_t s = c.stream;
    const dim3 gridR(NGROUPS, nK), blk(256);
    const dim3 gridS(div_up((long)nK * nE, 64)), blkS(64);
    const double tol2 = o.rtol * o.rtol;
    // compulsory bytes of one k_spmm_dot launch: A values + column indices +
    // row pointers, P read once, Q written
    if (c.kSelect == PGERT_TIME_SPMM) c.kBytesPerLaunch = (double)nK * ((double)c.nnz * 12.0 + 4.0 * (n + 1)) + 2.0 * (double)tot * 8.0;
    // r0 = b - A x0
    hipLaunchKernelGGL(k_spmm_dot, gridR, blk, 0, s, n, nE, c.nnz, c.rowptr.p, c.colidx.p, c.valsA.p, c.X.p,
                       c.Q.p, (double *)nullptr);
    hipLaunchKernelGGL(k_pcg_init, gridR, blk, 0, s, n, nE, c.B.p, c.Q.p, c.dinv.p, c.R.p, c.P.p, c.part.p);
    hipLaunchKernelGGL(k_pcg_reduce, gridS, blkS, 0, s, nK, nE, NGROUPS, (int)RED_INIT, tol2, c.part.p, c.scal.p);
    PG_CHECK(hipGetLastError());
    std::vector<double> h((size_t)nK * nE * NS);
    int it = 0;
    bool done = false;
    while (!done && it < o.max_iter) {
        int chunk = std::min(o.check_every, o.max_iter - it);
        for (int j = 0; j < chunk; ++j) {
            if (c.kSelect == PGERT_TIME_SPMM) c.kev_begin();
            hipLaunchKernelGGL(k_spmm_dot, gridR, blk, 0, s, n, nE, c.nnz, c.rowptr.p, c.colidx.p, c.valsA.p,
                               c.P.p, c.Q.p, c.part.p);
            if (c.kSelect == PGERT_TIME_SPMM) c.kev_end();
            hipLaunchKernelGGL(k_pcg_reduce, gridS, blkS, 0, s, nK, nE, NGROUPS, (int)RED_ALPHA, tol2, c.part.p,
                               c.scal.p);
            hipLaunchKernelGGL(k_pcg_update, gridR, blk, 0, s, n, nE, c.scal.p, c.P.p, c.Q.p, c.dinv.p, c.X.p,
                               c.R.p, c.part.p);
            hipLaunchKernelGGL(k_pcg_reduce, gridS, blkS, 0, s, nK, nE, NGROUPS, (int)RED_BETA, tol2, c.part.p,
                               c.scal.p);
            hipLaunchKernelGGL(k_pcg_pupdate, dim3(div_up((long)n * nE, 256), nK), blk, 0, s, n, nE, c.scal.p,
                               c.R.p, c.dinv.p, c.P.p);
        }
        PG_CHECK(hipGetLastError());
        it += chunk;
        PG_CHECK(hipMemcpyAsync(h.data(), c.scal.p, h.size() * 8, hipMemcpyDeviceToHost, s));
        PG_CHECK(hipStreamSynchronize(s));
        c.kev_collect();
        done = true;
        for (size_t t = 0; t < (size_t)nK * nE; ++t)
            if (h[t * NS + S_ACT] != 0.0) done = false;
    }
    st.iterations = it;
    st.max_iterations = it;
    st.refinements = 0;
    st.max_rel_residual = residual_check(c);
    st.converged = st.max_rel_residual <= o.rtol * 1.0001 ? 1 : 0;
}

}  // namespace pgert

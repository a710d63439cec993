// Block preconditioned conjugate gradients over the electrodes of one
// wavenumber (O'Leary's block-PCG with an orthonormal search block).
//
// All sources of one wavenumber share A_k (the mixed boundary condition uses
// the mean electrode position, dcfemmodelling.cpp:1025-1030), so their
// Krylov spaces are merged: each iteration applies A_k to an n x s block P
// (s <= 64 columns) and takes an s x s step.  Per iteration:
//   1. Q = A P            fused with the Grams P'Q, P'R          (k_bspmm_gram)
//   2. alpha = (P'Q)^-1 P'R                                      (k_small_alpha)
//   3. X += P alpha; R -= Q alpha; Z = D^-1 R, fused with the
//      Grams Q'Z, Z'Z, P'Z and the column norms |R|^2            (k_bupdate)
//   4. beta = -(P'Q)^-1 Q'Z;  W = Z + P beta;  W'W from the Grams
//      (P'P = I);  L = chol(W'W);  T = L^-T                      (k_small_beta)
//   5. P = Z T + P (beta T)                                      (k_bpupdate)
// Dependent directions are dropped in both Choleskys (zero pivot -> zero
// column), which keeps the iteration breakdown-free when columns converge.
// The vector type VT is double (plain) or float (mixed precision: FP32 HBM
// traffic, FP64 accumulation, Grams and small matrices; the solution is
// accumulated in FP64 with periodic FP64 residual replacement).
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace pgert {

namespace {

constexpr int S = 64;      // max block width
constexpr int CH = 16;     // rows per chunk in the Gram passes
constexpr int MAXG = 512;  // workgroups per block for the Gram passes

// small-matrix slots per block (each S*S doubles)
constexpr int SL_G1 = 0;   // P'Q, replaced by its Cholesky factor
constexpr int SL_G2 = 1;   // P'R, replaced by alpha
constexpr int SL_QZ = 2;   // Q'Z
constexpr int SL_ZZ = 3;   // Z'Z
constexpr int SL_PZ = 4;   // P'Z
constexpr int SL_RR = 5;   // |R|^2 per column (first S entries)
constexpr int SL_T = 6;    // T = chol(W'W)^-T
constexpr int SL_M2 = 7;   // beta T
constexpr int SL_KEEP = 8; // kept pivots of P'Q (first S entries)
constexpr int NSL = 9;

struct BlockDesc {
    int k;       // wavenumber index
    int c0;      // first electrode column
    int s;       // width
    int active;  // 0 once the block has converged: its iteration kernels return at once
};

// ---------------------------------------------------------------- Gram helpers
// Each of the 256 threads owns a 4x4 tile (a0..a0+3, b0..b0+3) of a 64x64 Gram.
__device__ inline void tile_of(int & a0, int & b0) {
    a0 = (threadIdx.x / 16) * 4;
    b0 = (threadIdx.x % 16) * 4;
}

__device__ inline void gram_acc(double (&acc)[4][4], const double * Ls, const double * Rs, int a0, int b0) {
    double l[4], r[4];
    for (int i = 0; i < 4; ++i) { l[i] = Ls[a0 + i]; r[i] = Rs[b0 + i]; }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] += l[i] * r[j];
}

__device__ inline void gram_store(const double (&acc)[4][4], double * out, int a0, int b0) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[(a0 + i) * S + (b0 + j)] = acc[i][j];
}

template <typename VT>
__device__ inline VT ld(const VT * p) { return *p; }

// ---------------------------------------------------------------- 1. Q = A P + Grams
// part layout per (block, group): [G1 (S*S)][G2 (S*S)]
template <typename VT, typename AT>
__global__ void __launch_bounds__(256) k_bspmm_gram(int n, int nE, int nnz, const BlockDesc * __restrict__ blocks,
                                                    const int * __restrict__ rowptr, const int * __restrict__ colidx,
                                                    const AT * __restrict__ vals, const VT * __restrict__ P,
                                                    const VT * __restrict__ R, VT * __restrict__ Q,
                                                    double * __restrict__ part) {
    __shared__ double Ps[CH][S], Qs[CH][S], Rs[CH][S];
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t voff = (size_t)bd.k * n * nE + bd.c0;
    const AT * A = vals + (size_t)bd.k * nnz;
    int a0, b0;
    tile_of(a0, b0);
    double g1[4][4] = {}, g2[4][4] = {};
    const int nchunk = (n + CH - 1) / CH;
    for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const int base = ch * CH;
        for (int rr = wave; rr < CH; rr += 4) {
            const int row = base + rr;
            double q = 0.0, p = 0.0, r = 0.0;
            if (row < n && lane < bd.s) {
                for (int t = rowptr[row]; t < rowptr[row + 1]; ++t)
                    q += (double)A[t] * (double)P[voff + (size_t)colidx[t] * nE + lane];
                const size_t o = voff + (size_t)row * nE + lane;
                Q[o] = (VT)q;
                p = (double)P[o];
                r = (double)R[o];
            }
            Ps[rr][lane] = p;
            Qs[rr][lane] = q;
            Rs[rr][lane] = r;
        }
        __syncthreads();
        for (int rr = 0; rr < CH; ++rr) {
            gram_acc(g1, Ps[rr], Qs[rr], a0, b0);
            gram_acc(g2, Ps[rr], Rs[rr], a0, b0);
        }
        __syncthreads();
    }
    double * out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2 * S * S;
    gram_store(g1, out, a0, b0);
    gram_store(g2, out + S * S, a0, b0);
}

// ---------------------------------------------------------------- 3. update + Grams
// part layout per (block, group): [Q'Z][Z'Z][P'Z][|R|^2 (S)]
template <typename VT>
__global__ void __launch_bounds__(256) k_bupdate(int n, int nE, const BlockDesc * __restrict__ blocks,
                                                 const double * __restrict__ small, const double * __restrict__ dinv,
                                                 const VT * __restrict__ P, const VT * __restrict__ Q,
                                                 VT * __restrict__ X, VT * __restrict__ R,
                                                 double * __restrict__ part, int grams) {
    __shared__ double Ps[CH][S], Qs[CH][S], Zs[CH][S];
    __shared__ double al[S][S];
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t voff = (size_t)bd.k * n * nE + bd.c0;
    const double * dk = dinv + (size_t)bd.k * n;
    const double * alpha = small + ((size_t)blockIdx.y * NSL + SL_G2) * S * S;
    for (int t = threadIdx.x; t < S * S; t += 256) al[t / S][t % S] = alpha[t];
    int a0, b0;
    tile_of(a0, b0);
    double gqz[4][4] = {}, gzz[4][4] = {}, gpz[4][4] = {};
    double rr2 = 0.0;
    const int nchunk = (n + CH - 1) / CH;
    __syncthreads();
    for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const int base = ch * CH;
        for (int rr = wave; rr < CH; rr += 4) {
            const int row = base + rr;
            double p = 0.0, q = 0.0;
            if (row < n && lane < bd.s) {
                const size_t o = voff + (size_t)row * nE + lane;
                p = (double)P[o];
                q = (double)Q[o];
            }
            Ps[rr][lane] = p;
            Qs[rr][lane] = q;
        }
        __syncthreads();
        for (int rr = wave; rr < CH; rr += 4) {
            const int row = base + rr;
            double z = 0.0;
            if (row < n && lane < bd.s) {
                double dx = 0.0, dr = 0.0;
                for (int a = 0; a < bd.s; ++a) {
                    double w = al[a][lane];
                    dx += Ps[rr][a] * w;
                    dr += Qs[rr][a] * w;
                }
                const size_t o = voff + (size_t)row * nE + lane;
                X[o] = (VT)((double)X[o] + dx);
                double r = (double)R[o] - dr;
                VT rs = (VT)r;
                R[o] = rs;
                r = (double)rs;
                rr2 += r * r;
                z = dk[row] * r;
            }
            Zs[rr][lane] = z;
        }
        __syncthreads();
        if (grams)
            for (int rr = 0; rr < CH; ++rr) {
                gram_acc(gqz, Qs[rr], Zs[rr], a0, b0);
                gram_acc(gzz, Zs[rr], Zs[rr], a0, b0);
                gram_acc(gpz, Ps[rr], Zs[rr], a0, b0);
            }
        __syncthreads();
    }
    if (!grams) return;
    double * out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (3 * S * S + S);
    gram_store(gqz, out, a0, b0);
    gram_store(gzz, out + S * S, a0, b0);
    gram_store(gpz, out + 2 * S * S, a0, b0);
    // column norms: the four wavefronts hold partial sums of the same columns
    __shared__ double cn[4][S];
    cn[wave][lane] = rr2;
    __syncthreads();
    if (wave == 0) out[3 * S * S + lane] = cn[0][lane] + cn[1][lane] + cn[2][lane] + cn[3][lane];
}

// ---------------------------------------------------------------- 5. P = Z T + P M2
template <typename VT>
__global__ void __launch_bounds__(256) k_bpupdate(int n, int nE, const BlockDesc * __restrict__ blocks,
                                                  const double * __restrict__ small, const double * __restrict__ dinv,
                                                  const VT * __restrict__ R, VT * __restrict__ P,
                                                  const double * __restrict__ cc, const int * __restrict__ agg, int m) {
    __shared__ double Tm[S][S], Mm[S][S];
    __shared__ double Zs[CH][S], Ps[CH][S];
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t voff = (size_t)bd.k * n * nE + bd.c0;
    const double * dk = dinv + (size_t)bd.k * n;
    const double * sm = small + (size_t)blockIdx.y * NSL * S * S;
    for (int t = threadIdx.x; t < S * S; t += 256) {
        Tm[t / S][t % S] = sm[SL_T * S * S + t];
        Mm[t / S][t % S] = sm[SL_M2 * S * S + t];
    }
    __syncthreads();
    const int nchunk = (n + CH - 1) / CH;
    for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const int base = ch * CH;
        for (int rr = wave; rr < CH; rr += 4) {
            const int row = base + rr;
            double z = 0.0, p = 0.0;
            if (row < n && lane < bd.s) {
                const size_t o = voff + (size_t)row * nE + lane;
                z = dk[row] * (double)R[o];
                if (cc) z += cc[((size_t)blockIdx.y * m + agg[row]) * S + lane];
                p = (double)P[o];
            }
            Zs[rr][lane] = z;
            Ps[rr][lane] = p;
        }
        __syncthreads();
        for (int rr = wave; rr < CH; rr += 4) {
            const int row = base + rr;
            if (row < n && lane < bd.s) {
                double acc = 0.0;
                for (int a = 0; a < bd.s; ++a) acc += Zs[rr][a] * Tm[a][lane] + Ps[rr][a] * Mm[a][lane];
                P[voff + (size_t)row * nE + lane] = (VT)acc;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- coarse correction
// y[b][I][c] = sum over the nodes of aggregate I of R[node][c0 + c]
template <typename VT>
__global__ void __launch_bounds__(256) k_restrict(int n, int nE, int m, const BlockDesc * __restrict__ blocks,
                                                  const int * __restrict__ aggPtr, const int * __restrict__ aggNodes,
                                                  const VT * __restrict__ R, double * __restrict__ y) {
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int I = blockIdx.x * 4 + wave;
    if (I >= m) return;
    const size_t voff = (size_t)bd.k * n * nE + bd.c0;
    double s = 0.0;
    if (lane < bd.s)
        for (int q = aggPtr[I]; q < aggPtr[I + 1]; ++q) s += (double)R[voff + (size_t)aggNodes[q] * nE + lane];
    y[((size_t)blockIdx.y * m + I) * S + lane] = s;
}

// cc[b] = Einv_k y[b]: 4 coarse rows x 64 columns per workgroup, y staged in LDS
__global__ void __launch_bounds__(256) k_coarse_apply(int m, const BlockDesc * __restrict__ blocks,
                                                      const double * __restrict__ Einv, const double * __restrict__ y,
                                                      double * __restrict__ cc) {
    __shared__ double ys[64][S];
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    const int r = blockIdx.x * 4 + threadIdx.x / 64, col = threadIdx.x % 64;
    const double * E = Einv + (size_t)bd.k * m * m;
    const double * yb = y + (size_t)blockIdx.y * m * S;
    double acc = 0.0;
    for (int k0 = 0; k0 < m; k0 += 64) {
        for (int t = threadIdx.x; t < 64 * S; t += 256) {
            int kk = k0 + t / S;
            ys[t / S][t % S] = kk < m ? yb[(size_t)kk * S + t % S] : 0.0;
        }
        __syncthreads();
        if (r < m) {
            const double * Er = E + (size_t)r * m + k0;
            const int kn = min(64, m - k0);
            for (int q = 0; q < kn; ++q) acc += Er[q] * ys[q][col];
        }
        __syncthreads();
    }
    if (r < m) cc[((size_t)blockIdx.y * m + r) * S + col] = acc;
}

// Z = D^-1 R + prolong(cc); Grams Q'Z, Z'Z, P'Z and |R|^2 (k_bupdate's partial layout)
template <typename VT>
__global__ void __launch_bounds__(256) k_zgram(int n, int nE, int m, const BlockDesc * __restrict__ blocks,
                                               const double * __restrict__ dinv, const VT * __restrict__ P,
                                               const VT * __restrict__ Q, const VT * __restrict__ R,
                                               const double * __restrict__ cc, const int * __restrict__ agg,
                                               double * __restrict__ part) {
    __shared__ double Ps[CH][S], Qs[CH][S], Zs[CH][S];
    __shared__ double cn[4][S];
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t voff = (size_t)bd.k * n * nE + bd.c0;
    const double * dk = dinv + (size_t)bd.k * n;
    int a0, b0;
    tile_of(a0, b0);
    double gqz[4][4] = {}, gzz[4][4] = {}, gpz[4][4] = {};
    double rr2 = 0.0;
    const int nchunk = (n + CH - 1) / CH;
    for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const int base = ch * CH;
        for (int rr = wave; rr < CH; rr += 4) {
            const int row = base + rr;
            double p = 0.0, q = 0.0, z = 0.0;
            if (row < n && lane < bd.s) {
                const size_t o = voff + (size_t)row * nE + lane;
                p = (double)P[o];
                q = (double)Q[o];
                double r = (double)R[o];
                rr2 += r * r;
                z = dk[row] * r + cc[((size_t)blockIdx.y * m + agg[row]) * S + lane];
            }
            Ps[rr][lane] = p;
            Qs[rr][lane] = q;
            Zs[rr][lane] = z;
        }
        __syncthreads();
        for (int rr = 0; rr < CH; ++rr) {
            gram_acc(gqz, Qs[rr], Zs[rr], a0, b0);
            gram_acc(gzz, Zs[rr], Zs[rr], a0, b0);
            gram_acc(gpz, Ps[rr], Zs[rr], a0, b0);
        }
        __syncthreads();
    }
    double * out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (3 * S * S + S);
    gram_store(gqz, out, a0, b0);
    gram_store(gzz, out + S * S, a0, b0);
    gram_store(gpz, out + 2 * S * S, a0, b0);
    cn[wave][lane] = rr2;
    __syncthreads();
    if (wave == 0) out[3 * S * S + lane] = cn[0][lane] + cn[1][lane] + cn[2][lane] + cn[3][lane];
}

// ---------------------------------------------------------------- reductions
// sum over the groups of `count` doubles per (block, group) into small[block][dst ...]
__global__ void k_gram_reduce(const BlockDesc * __restrict__ blocks, int G, int count,
                              const double * __restrict__ part, double * __restrict__ small, int dstOff) {
    const int b = blockIdx.y;
    if (!blocks[b].active) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double s = 0.0;
    const double * p = part + (size_t)b * G * count + t;
    for (int g = 0; g < G; ++g) s += p[(size_t)g * count];
    small[(size_t)b * NSL * S * S + dstOff + t] = s;
}

// ---------------------------------------------------------------- small dense kernels
// Cholesky with dropping in LDS: L lower, keep[j] = 0 for dropped pivots.
// A pivot is dropped when the remaining diagonal falls below 1e-12 of the
// column's own diagonal (the column is dependent on the previous ones), a
// criterion independent of how the columns are scaled against each other.
__device__ void chol_drop(double (*M)[S + 1], int s, int * keep) {
    __shared__ double d0[S];
    if (threadIdx.x < s) d0[threadIdx.x] = M[threadIdx.x][threadIdx.x];
    __syncthreads();
    for (int j = 0; j < s; ++j) {
        if (threadIdx.x == 0) {
            double d = M[j][j];
            if (d0[j] > 0.0 && d > 1e-12 * d0[j]) {
                keep[j] = 1;
                M[j][j] = sqrt(d);
            } else {
                keep[j] = 0;
                M[j][j] = 0.0;
            }
        }
        __syncthreads();
        const double piv = M[j][j];
        const int kj = keep[j];
        for (int i = j + 1 + threadIdx.x; i < s; i += blockDim.x) M[i][j] = kj ? M[i][j] / piv : 0.0;
        __syncthreads();
        const int m = s - j - 1;
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
            int i = j + 1 + t / m, c = j + 1 + t % m;
            if (c <= i) M[i][c] -= M[i][j] * M[c][j];
        }
        __syncthreads();
    }
}

// Solve L L^T X = B in place (B: s x s columns), dropped rows -> 0.  One thread per column.
__device__ void chol_solve(double (*L)[S + 1], const int * keep, int s, double (*B)[S + 1]) {
    int c = threadIdx.x;
    if (c < s) {
        for (int i = 0; i < s; ++i) {
            if (!keep[i]) { B[i][c] = 0.0; continue; }
            double v = B[i][c];
            for (int q = 0; q < i; ++q) v -= L[i][q] * B[q][c];
            B[i][c] = v / L[i][i];
        }
        for (int i = s - 1; i >= 0; --i) {
            if (!keep[i]) { B[i][c] = 0.0; continue; }
            double v = B[i][c];
            for (int q = i + 1; q < s; ++q) v -= L[q][i] * B[q][c];
            B[i][c] = v / L[i][i];
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_small_alpha(const BlockDesc * __restrict__ blocks, double * __restrict__ small) {
    __shared__ double L[S][S + 1], Bm[S][S + 1];
    __shared__ int keep[S];
    if (!blocks[blockIdx.x].active) return;
    const int s = blocks[blockIdx.x].s;
    double * sm = small + (size_t)blockIdx.x * NSL * S * S;
    for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
        int i = t / S, j = t % S;
        L[i][j] = sm[SL_G1 * S * S + t];
        Bm[i][j] = sm[SL_G2 * S * S + t];
    }
    __syncthreads();
    chol_drop(L, s, keep);
    chol_solve(L, keep, s, Bm);
    for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
        int i = t / S, j = t % S;
        bool in = i < s && j < s;
        sm[SL_G2 * S * S + t] = in ? Bm[i][j] : 0.0;
        // keep the factor of P'Q for the beta solve (lower triangle, stored in SL_G1)
        sm[SL_G1 * S * S + t] = (in && j <= i) ? L[i][j] : 0.0;
    }
    if (threadIdx.x < S) sm[SL_KEEP * S * S + threadIdx.x] = (threadIdx.x < s) ? (double)keep[threadIdx.x] : 0.0;
}

// init = 1: first iteration (no previous block): beta = 0, W = Z.
__global__ void __launch_bounds__(256) k_small_beta(const BlockDesc * __restrict__ blocks, double * __restrict__ small,
                                                    int init) {
    __shared__ double L[S][S + 1], Bt[S][S + 1], W[S][S + 1], Pz[S][S + 1];
    __shared__ int keep[S], keep2[S];
    if (!blocks[blockIdx.x].active) return;
    const int s = blocks[blockIdx.x].s;
    double * sm = small + (size_t)blockIdx.x * NSL * S * S;
    for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
        int i = t / S, j = t % S;
        L[i][j] = sm[SL_G1 * S * S + t];
        Bt[i][j] = init ? 0.0 : sm[SL_QZ * S * S + t];  // Q'Z
        W[i][j] = sm[SL_ZZ * S * S + t];
        Pz[i][j] = init ? 0.0 : sm[SL_PZ * S * S + t];
    }
    if (threadIdx.x < S) keep[threadIdx.x] = init ? 0 : (int)sm[SL_KEEP * S * S + threadIdx.x];
    __syncthreads();
    if (!init) {
        chol_solve(L, keep, s, Bt);  // Bt = (P'Q)^-1 Q'Z ; beta = -Bt
        // W'W = Z'Z + Pz' beta + beta' Pz + beta' beta   (P'P = I)
        for (int t = threadIdx.x; t < s * s; t += blockDim.x) {
            int i = t / s, j = t % s;
            double v = 0.0;
            for (int q = 0; q < s; ++q)
                v += -Pz[q][i] * Bt[q][j] - Bt[q][i] * Pz[q][j] + Bt[q][i] * Bt[q][j];
            W[i][j] += v;
        }
        __syncthreads();
    }
    chol_drop(W, s, keep2);
    // T = L^-T.  Thread c forms column c of L^-1 by forward substitution
    // (x_i = 0 for i < c and for dropped pivots) into Pz (free after the W
    // update); that column is row c of T.
    {
        int c = threadIdx.x;
        if (c < s) {
            for (int i = 0; i < s; ++i) {
                if (i < c || !keep2[i]) { Pz[c][i] = 0.0; continue; }
                double v = (i == c) ? 1.0 : 0.0;
                for (int q = c; q < i; ++q) v -= W[i][q] * Pz[c][q];
                Pz[c][i] = v / W[i][i];
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
        int i = t / S, j = t % S;
        bool in = i < s && j < s;
        sm[SL_T * S * S + t] = in ? Pz[i][j] : 0.0;
    }
    __syncthreads();
    // M2 = beta T = -Bt T
    for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
        int i = t / S, j = t % S;
        double v = 0.0;
        if (i < s && j < s && !init)
            for (int q = 0; q < s; ++q) v -= Bt[i][q] * Pz[q][j];
        sm[SL_M2 * S * S + t] = v;
    }
}

// ---------------------------------------------------------------- residual for the block solver
// R = B - A X (X in FP64, R in VT); partial |R|^2, |B|^2 per column -> part[(b*G+g)*2S]
template <typename VT>
__global__ void __launch_bounds__(256) k_bresid(int n, int nE, int nnz, const BlockDesc * __restrict__ blocks,
                                                const int * __restrict__ rowptr, const int * __restrict__ colidx,
                                                const double * __restrict__ vals, const double * __restrict__ X,
                                                const double * __restrict__ B, VT * __restrict__ R,
                                                double * __restrict__ part) {
    __shared__ double cn[4][2][S];
    const BlockDesc bd = blocks[blockIdx.y];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t voff = (size_t)bd.k * n * nE + bd.c0;
    const double * A = vals + (size_t)bd.k * nnz;
    double rr = 0.0, bb = 0.0;
    for (int row = blockIdx.x * 4 + wave; row < n; row += gridDim.x * 4) {
        if (lane < bd.s) {
            double y = 0.0;
            for (int t = rowptr[row]; t < rowptr[row + 1]; ++t) y += A[t] * X[voff + (size_t)colidx[t] * nE + lane];
            const size_t o = voff + (size_t)row * nE + lane;
            double b = B[o];
            double r = b - y;
            if (R) R[o] = (VT)r;
            rr += r * r;
            bb += b * b;
        }
    }
    cn[wave][0][lane] = rr;
    cn[wave][1][lane] = bb;
    __syncthreads();
    if (wave == 0) {
        double * out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2 * S;
        out[lane] = cn[0][0][lane] + cn[1][0][lane] + cn[2][0][lane] + cn[3][0][lane];
        out[S + lane] = cn[0][1][lane] + cn[1][1][lane] + cn[2][1][lane] + cn[3][1][lane];
    }
}

__global__ void k_norm_reduce(int G, const double * __restrict__ part, double * __restrict__ out) {
    const int b = blockIdx.x;
    const int t = threadIdx.x;  // 0 .. 2S-1
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += part[((size_t)b * G + g) * 2 * S + t];
    out[(size_t)b * 2 * S + t] = s;
}

// X64 += (double) Xw ; Xw = 0   over the block's columns
__global__ void k_accumulate(int n, int nE, const BlockDesc * __restrict__ blocks, float * __restrict__ Xw,
                             double * __restrict__ X64) {
    const BlockDesc bd = blocks[blockIdx.y];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * S) return;
    int row = (int)(t / S), c = (int)(t % S);
    if (c >= bd.s) return;
    size_t o = (size_t)bd.k * n * nE + bd.c0 + (size_t)row * nE + c;
    X64[o] += (double)Xw[o];
    Xw[o] = 0.0f;
}

__global__ void k_to_float(long nn, const double * __restrict__ a, float * __restrict__ b) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nn) b[t] = (float)a[t];
}

// --------------------------------------------------------------- driver
struct Plan {
    std::vector<BlockDesc> h;
    DBuf<BlockDesc> d;
    int G = 0;
};

void make_plan(Ctx & c, Plan & pl) {
    pl.h.clear();
    for (int k = 0; k < c.nK; ++k)
        for (int c0 = 0; c0 < c.nE; c0 += S) pl.h.push_back({k, c0, std::min(S, c.nE - c0), 1});
    pl.d.upload(pl.h, c.stream);
    pl.G = std::max(1, std::min(MAXG, div_up(c.n, CH)));
}

// column norms |R|^2 and |B|^2 of the block solver's current state, host copy
void read_norms(Ctx & c, Plan & pl, DBuf<double> & nbuf, std::vector<double> & h) {
    h.resize(pl.h.size() * 2 * S);
    PG_CHECK(hipMemcpyAsync(h.data(), nbuf.p, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
    PG_CHECK(hipStreamSynchronize(c.stream));
}

template <typename VT, typename AT>
struct BlockState {
    DBuf<VT> X, R, P, Q;
};

}  // namespace

void solve_block(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st, bool mixed) {
    PG_REQUIRE(c.nE >= 1, "block solver: no electrodes");
    hipStream_t s = c.stream;
    Plan pl;
    make_plan(c, pl);
    const int nb = (int)pl.h.size();
    const int G = pl.G;
    const size_t tot = (size_t)c.nK * c.n * c.nE;
    c.small.alloc((size_t)nb * NSL * S * S);
    PG_CHECK(hipMemsetAsync(c.small.p, 0, c.small.bytes(), s));
    c.part.alloc((size_t)nb * G * (3 * S * S + S));
    const bool useCoarse = o.coarse && c.nAgg > 0;
    if (useCoarse) {
        coarse_prepare(c);
        c.ycoarse.alloc((size_t)nb * c.nAgg * S);
        c.ccoarse.alloc((size_t)nb * c.nAgg * S);
        PG_CHECK(hipMemsetAsync(c.ccoarse.p, 0, c.ccoarse.bytes(), s));
    }
    DBuf<double> norms;
    norms.alloc((size_t)nb * 2 * S);
    DBuf<double> npart;
    const int GR = std::max(1, std::min(1024, div_up(c.n, 4)));
    npart.alloc((size_t)nb * GR * 2 * S);
    std::vector<double> hn;

    // FP64 true residual R = B - A X64 into Rw (VT), norms -> norms buffer
    auto true_resid = [&](auto * Rw) {
        using VT = std::remove_pointer_t<decltype(Rw)>;
        hipLaunchKernelGGL(k_bresid<VT>, dim3(GR, nb), dim3(256), 0, s, c.n, c.nE, c.nnz, pl.d.p, c.rowptr.p,
                           c.colidx.p, c.valsA.p, c.X.p, c.B.p, Rw, npart.p);
        hipLaunchKernelGGL(k_norm_reduce, dim3(nb), dim3(2 * S), 0, s, GR, npart.p, norms.p);
        PG_CHECK(hipGetLastError());
    };
    auto max_rel = [&](const std::vector<double> & h, std::vector<double> * per) {
        double worst = 0;
        for (int b = 0; b < nb; ++b)
            for (int j = 0; j < pl.h[b].s; ++j) {
                double rr = h[(size_t)b * 2 * S + j], bb = h[(size_t)b * 2 * S + S + j];
                double r = bb > 0 ? std::sqrt(rr / bb) : 0.0;
                worst = std::max(worst, r);
            }
        (void)per;
        return worst;
    };

    auto run = [&](auto vtag, auto atag, const auto * Aw) {
        using VT = decltype(vtag);
        using AT = decltype(atag);
        const bool fp32 = std::is_same<VT, float>::value;
        // compulsory bytes of one k_bspmm_gram launch: A values + column
        // indices + row pointers, P read once, R read, Q written
        c.kBytesPerLaunch = (double)c.nK * ((double)c.nnz * (sizeof(AT) + 4) + 4.0 * (c.n + 1)) +
                            3.0 * (double)tot * sizeof(VT);
        // working vectors: X (correction since last replacement), R, P, Q
        DBuf<VT> Xw, Rw, Pw, Qw;
        DBuf<double> dinv;
        Rw.alloc(tot);
        Pw.alloc(tot);
        Qw.alloc(tot);
        VT * Xp;
        if (fp32) {
            Xw.alloc(tot);
            PG_CHECK(hipMemsetAsync(Xw.p, 0, Xw.bytes(), s));
            Xp = Xw.p;
        } else {
            Xp = (VT *)c.X.p;
        }
        PG_CHECK(hipMemsetAsync(Pw.p, 0, Pw.bytes(), s));
        PG_CHECK(hipMemsetAsync(Qw.p, 0, Qw.bytes(), s));
        true_resid(Rw.p);
        read_norms(c, pl, norms, hn);
        std::vector<double> bnorm2(nb * S);
        for (int b = 0; b < nb; ++b)
            for (int j = 0; j < S; ++j) bnorm2[b * S + j] = hn[(size_t)b * 2 * S + S + j];
        double rel = max_rel(hn, nullptr);
        double lastRel = rel;
        int it = 0, refine = 0, lastReplace = 0;
        const dim3 gG(G, nb), blk(256);
        const int cQZ = 3 * S * S + S;
        const int m = useCoarse ? c.nAgg : 0;
        const double * ccp = useCoarse ? c.ccoarse.p : nullptr;
        const int * aggp = useCoarse ? c.agg.p : nullptr;
        // X += P alpha, R -= Q alpha, then Z = M^-1 R and the Grams Q'Z, Z'Z, P'Z, |R|^2
        auto update_and_z = [&](bool updateXR) {
            if (!useCoarse) {
                hipLaunchKernelGGL(k_bupdate<VT>, gG, blk, 0, s, c.n, c.nE, pl.d.p, c.small.p, c.dinv.p, Pw.p, Qw.p,
                                   Xp, Rw.p, c.part.p, 1);
            } else {
                if (updateXR)
                    hipLaunchKernelGGL(k_bupdate<VT>, gG, blk, 0, s, c.n, c.nE, pl.d.p, c.small.p, c.dinv.p, Pw.p,
                                       Qw.p, Xp, Rw.p, c.part.p, 0);
                hipLaunchKernelGGL(k_restrict<VT>, dim3(div_up(m, 4), nb), blk, 0, s, c.n, c.nE, m, pl.d.p,
                                   c.aggPtr.p, c.aggNodes.p, Rw.p, c.ycoarse.p);
                hipLaunchKernelGGL(k_coarse_apply, dim3(div_up(m, 4), nb), blk, 0, s, m, pl.d.p, c.Einv.p,
                                   c.ycoarse.p, c.ccoarse.p);
                hipLaunchKernelGGL(k_zgram<VT>, gG, blk, 0, s, c.n, c.nE, m, pl.d.p, c.dinv.p, Pw.p, Qw.p, Rw.p,
                                   c.ccoarse.p, c.agg.p, c.part.p);
            }
            hipLaunchKernelGGL(k_gram_reduce, dim3(div_up(cQZ, 256), nb), blk, 0, s, pl.d.p, G, cQZ, c.part.p,
                               c.small.p, SL_QZ * S * S);
        };
        auto start_block = [&]() {
            // alpha = 0: only Z = M^-1 R and Z'Z are formed; then P = Z chol(Z'Z)^-T
            PG_CHECK(hipMemsetAsync(c.small.p, 0, c.small.bytes(), s));
            PG_CHECK(hipMemsetAsync(Pw.p, 0, Pw.bytes(), s));
            update_and_z(false);
            hipLaunchKernelGGL(k_small_beta, dim3(nb), blk, 0, s, pl.d.p, c.small.p, 1);
            hipLaunchKernelGGL(k_bpupdate<VT>, gG, blk, 0, s, c.n, c.nE, pl.d.p, c.small.p, c.dinv.p, Rw.p, Pw.p,
                               ccp, aggp, m);
            PG_CHECK(hipGetLastError());
        };
        if (rel > o.rtol) start_block();
        bool done = rel <= o.rtol;
        while (!done && it < o.max_iter) {
            int chunk = std::min(o.check_every, o.max_iter - it);
            for (int j = 0; j < chunk; ++j) {
                c.kev_begin();
                hipLaunchKernelGGL((k_bspmm_gram<VT, AT>), gG, blk, 0, s, c.n, c.nE, c.nnz, pl.d.p, c.rowptr.p,
                                   c.colidx.p, Aw, Pw.p, Rw.p, Qw.p, c.part.p);
                c.kev_end();
                hipLaunchKernelGGL(k_gram_reduce, dim3(div_up(2 * S * S, 256), nb), blk, 0, s, pl.d.p, G, 2 * S * S,
                                   c.part.p, c.small.p, SL_G1 * S * S);
                hipLaunchKernelGGL(k_small_alpha, dim3(nb), blk, 0, s, pl.d.p, c.small.p);
                update_and_z(true);
                hipLaunchKernelGGL(k_small_beta, dim3(nb), blk, 0, s, pl.d.p, c.small.p, 0);
                hipLaunchKernelGGL(k_bpupdate<VT>, gG, blk, 0, s, c.n, c.nE, pl.d.p, c.small.p, c.dinv.p, Rw.p,
                                   Pw.p, ccp, aggp, m);
            }
            PG_CHECK(hipGetLastError());
            it += chunk;
            // recursive residual norms of the last k_bupdate
            std::vector<double> hs((size_t)nb * S);
            for (int b = 0; b < nb; ++b)
                PG_CHECK(hipMemcpyAsync(hs.data() + (size_t)b * S, c.small.p + ((size_t)b * NSL + SL_RR) * S * S,
                                        S * 8, hipMemcpyDeviceToHost, s));
            PG_CHECK(hipStreamSynchronize(s));
            c.kev_collect();
            // per-block worst recursive residual of the still-active blocks
            std::vector<double> wb(nb, 0.0);
            double worst = 0;
            bool anyAtTarget = false;
            for (int b = 0; b < nb; ++b) {
                if (!pl.h[b].active) continue;
                for (int j = 0; j < pl.h[b].s; ++j) {
                    double bb = bnorm2[b * S + j];
                    if (bb > 0) wb[b] = std::max(wb[b], std::sqrt(hs[(size_t)b * S + j] / bb));
                }
                worst = std::max(worst, wb[b]);
                if (wb[b] <= o.rtol) anyAtTarget = true;
            }
            bool changed = false;
            // Residual replacement: the true FP64 residual b - A x decides which
            // blocks are done.  FP32 recursions are replaced when they have
            // gained a decade, reached the target or run 256 iterations; FP64
            // ones when a block's recursive residual reaches the target.
            const bool replaceNow = fp32 ? (worst < 0.1 * lastRel || anyAtTarget || it - lastReplace >= 256)
                                         : anyAtTarget;
            if (replaceNow) {
                lastReplace = it;
                if (fp32)
                    hipLaunchKernelGGL(k_accumulate, dim3(div_up((long)c.n * S, 256), nb), blk, 0, s, c.n, c.nE,
                                       pl.d.p, (float *)Xp, c.X.p);
                true_resid(Rw.p);
                read_norms(c, pl, norms, hn);
                ++refine;
                lastRel = 0.0;
                for (int b = 0; b < nb; ++b) {
                    if (!pl.h[b].active) continue;
                    double tb = 0.0;
                    for (int j = 0; j < pl.h[b].s; ++j) {
                        double rr = hn[(size_t)b * 2 * S + j], bb = hn[(size_t)b * 2 * S + S + j];
                        if (bb > 0) tb = std::max(tb, std::sqrt(rr / bb));
                    }
                    if (tb <= o.rtol) {
                        pl.h[b].active = 0;
                        changed = true;
                    } else {
                        lastRel = std::max(lastRel, tb);
                    }
                }
            }
            done = true;
            for (int b = 0; b < nb; ++b)
                if (pl.h[b].active) done = false;
            if (changed && !done)
                PG_CHECK(hipMemcpyAsync(pl.d.p, pl.h.data(), pl.h.size() * sizeof(BlockDesc), hipMemcpyHostToDevice, s));
        }
        if (fp32) {
            hipLaunchKernelGGL(k_accumulate, dim3(div_up((long)c.n * S, 256), nb), dim3(256), 0, s, c.n, c.nE,
                               pl.d.p, (float *)Xp, c.X.p);
            PG_CHECK(hipGetLastError());
        }
        st.iterations = it;
        st.max_iterations = it;
        st.refinements = refine;
    };

    if (mixed) {
        c.fA.alloc((size_t)c.nK * c.nnz);
        long nn = (long)c.fA.n;
        hipLaunchKernelGGL(k_to_float, dim3(div_up(nn, 256)), dim3(256), 0, s, nn, c.valsA.p, c.fA.p);
        run(float(0), float(0), (const float *)c.fA.p);
    } else {
        run(double(0), double(0), (const double *)c.valsA.p);
    }
    st.max_rel_residual = residual_check(c);
    st.converged = st.max_rel_residual <= o.rtol * 1.0001 ? 1 : 0;
}

}  // namespace pgert

// Block preconditioned conjugate gradients over the electrodes of one
// wavenumber (O'Leary's block-PCG with an orthonormal search block).
//
// All sources of one wavenumber share A_k (the mixed boundary condition uses
// the mean electrode position, dcfemmodelling.cpp:1025-1030), so their
// Krylov spaces are merged: each iteration applies A_k to an n x s block P
// (s <= 64 columns) and takes an s x s step:
//   1. Q = A P                                              (k_spmm)
//   2. G1 = P'Q, G2 = P'R;  alpha = G1^-1 G2                (k_gram<PQR>, k_small_alpha)
//   3. R -= Q alpha;  |R|^2 per column                      (k_update; X += P alpha
//      runs beside step 5's 64 x 64 work, in the same launch)
//   4. Z = M^-1 R;  Q'Z, Z'Z, P'Z                  (k_gram<Z>; M = Jacobi [+ additive coarse]
//      correction, k_zform], or the multiplicative two-level cycle with
//      Chebyshev smoothing [k_smooth, k_restrict, k_coarse_apply, k_coarse_resid])
//   5. beta = -G1^-1 Q'Z;  W = Z + P beta;  W'W from the Grams (P'P = I);
//      T = chol(W'W)^-T;  P = Z T + P (beta T)              (k_small_beta, k_pupdate)
// Dependent directions are dropped in both factorisations (a pivot below
// 1e-12 of its own diagonal gives a zero row and column), which keeps the
// iteration breakdown-free when columns converge.
//
// Layout: every work vector is held block-major, V[block][row][64] with the
// columns beyond the block width zero, so a row is one aligned 256 B (FP32)
// or 512 B (FP64) segment.  The tall-skinny products run on the matrix cores
// (v_mfma 16x16x4, FP32 or FP64):
//   * Grams L'R over the rows: a wavefront takes 4 rows per step, lane l
//     holding columns 4(l%16)..+3 of row l/16 as one vector load; component
//     t of that vector is directly the MFMA operand of column tile t, so
//     the 64x64 Gram is 16 MFMAs per 4 rows with no data movement.
//   * Updates C = L M (M 64 x 64): 16-row strips, lane l holding row l%16;
//     computed transposed (C' = M' L') so that operands and results share the
//     layout of a plain vector load/store; M' comes from LDS, pre-swizzled
//     to one conflict-free read per MFMA.
// Partial Grams are written per workgroup in FP64 and reduced in a fixed
// order (deterministic).  The 64 x 64 algebra runs in one workgroup per
// block in LDS (FP64).
//
// The vector type VT is double (plain block) or float (mixed precision: FP32
// HBM traffic and matrix-core products; Grams reduced and small matrices kept
// in FP64; the solution accumulated in FP64 with residual replacement by the
// true FP64 residual).
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <utility>

namespace pgert {

namespace {

constexpr int S = 64;  // max block width
constexpr int SS = S * S;

// small-matrix slots per block (each S*S doubles)
constexpr int SL_G1 = 0;  // P'Q, replaced by its inverse
constexpr int SL_G2 = 1;  // P'R, replaced by alpha
constexpr int SL_QZ = 2;  // Q'Z
constexpr int SL_ZZ = 3;  // Z'Z
constexpr int SL_PZ = 4;  // P'Z
constexpr int SL_RR = 5;  // |R|^2 per column (first S entries)
constexpr int SL_T = 6;   // T = chol(W'W)^-T
constexpr int SL_M2 = 7;  // beta T
constexpr int NSL = 8;

// One block: up to 64 sources of one wavenumber of one problem.  The
// descriptor carries everything the kernels need, so one launch serves any
// mix of problems (a single mesh, or a batch of independent meshes).
struct BlockDesc {
    int k;       // wavenumber index within its problem
    int c0;      // first electrode column
    int s;       // width
    int active;  // 0 once the block has converged: its iteration kernels return at once
    int n;       // rows (nodes of the problem)
    int m;       // coarse aggregates of the problem (0: Jacobi only)
    int nE;      // electrodes of the problem (leading dimension of Bsrc / Xdst)
    int prob;    // problem index
    long long off;   // element offset of the block's rows in the block-major work vectors
    long long coff;  // element offset of the block's m x 64 coarse vectors
    const int * rowptr;
    const int * colidx;
    const double * valsD;  // A_k, FP64
    const double * dinv;   // D^-1 of A_k
    const int * agg;       // node -> aggregate
    const int * aggPtr;
    const int * aggNodes;
    const double * Einv;   // E_k^-1, m x m
    const double * Bsrc;   // right-hand sides of wavenumber k, node-major [node][nE]
    double * Xdst;         // solution of wavenumber k, node-major [node][nE]
    // multiplicative two-level preconditioner with Chebyshev smoothing
    const double * valsAD;  // A_k D^-1
    const int * azPtr;      // A_k Z: per row the aggregates of its columns ...
    const int * azAgg;
    const double * azVals;  // ... and the summed values
    double ch[4];           // Chebyshev coefficients c0, c1, c2, cz (see k_smooth)
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <typename T> struct Vec4Of;
template <> struct Vec4Of<float> { using type = f32x4; };
template <> struct Vec4Of<double> { using type = f64x4; };
template <typename T> using V4 = typename Vec4Of<T>::type;

__device__ inline f32x4 mfma4(float a, float b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ inline f64x4 mfma4(double a, double b, f64x4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// Row (within the 16 x 16 tile) of result register r of lane l; the column
// is l % 16 and the contraction index of both operands is l / 16.  FP32 and
// FP64 differ: registers run along consecutive rows in FP32, across lane
// groups in FP64.
template <typename VT> __device__ inline int drow(int lane, int r) {
    return sizeof(VT) == 4 ? 4 * (lane / 16) + r : 4 * r + lane / 16;
}

// Workgroup barrier ordering LDS only.  __syncthreads() also orders global
// memory, which makes every wavefront wait for all of its outstanding loads
// (vmcnt(0)) and would drain a software-pipelined load ring at each step.
template <typename VT> __device__ inline V4<VT> ld4(const VT * p) { return *reinterpret_cast<const V4<VT> *>(p); }
// v, or zero when !keep, by a bit mask (a select of a whole vector would
// become a branch)
__device__ inline f32x4 zero_unless(bool keep, f32x4 v) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const unsigned m = keep ? 0xffffffffu : 0u;
    return __builtin_bit_cast(f32x4, __builtin_bit_cast(u32x4, v) & m);
}
__device__ inline f64x4 zero_unless(bool keep, f64x4 v) {
    typedef unsigned long long u64x4 __attribute__((ext_vector_type(4)));
    const unsigned long long m = keep ? ~0ull : 0ull;
    return __builtin_bit_cast(f64x4, __builtin_bit_cast(u64x4, v) & m);
}
template <typename VT> __device__ inline void st4(VT * p, V4<VT> v) { *reinterpret_cast<V4<VT> *>(p) = v; }

// ---------------------------------------------------------------- 1. Q = A P
// a wavefront takes 4 rows at once (16 lanes per row, 4 columns per lane as
// one vector load).  The 16 lanes of a row load up to 16 of its column
// indices and values side by side (coalesced) into the wavefront's LDS
// slice; every lane then reads them back as broadcasts and issues the 16
// gathers of P at once, so a row costs three dependent memory latencies
// (row pointer, indices, gathers) instead of two per 4 nonzeros.  Slots past
// the row's end read row 0 with weight 0 (no branch around a load).
// XCD-aware rows: workgroups are dispatched round-robin over the 8 XCDs
// (workgroup id % 8), so XCD x is given the contiguous rows [x n/8, (x+1) n/8)
// and the block rows it gathers (neighbours within a few mesh planes) stay in
// its own L2.  gridDim.x is a multiple of 8.
constexpr int NXCD = 8;

__device__ inline void xcd_rows(int n, int & lo, int & hi, int & first, int & stride) {
    const int x = blockIdx.x % NXCD, wi = blockIdx.x / NXCD, nw = gridDim.x / NXCD;
    lo = (int)((long)n * x / NXCD);
    hi = (int)((long)n * (x + 1) / NXCD);
    first = lo + wi * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    stride = nw * 4;
}
constexpr int SPMM_W = 16;  // nonzeros of a row held in the wavefront's LDS slice per round
template <typename VT, typename AT>
__global__ void __launch_bounds__(256) k_spmm(const BlockDesc * __restrict__ blocks, const VT * __restrict__ P,
                                              VT * __restrict__ Q) {
    using V = V4<VT>;
    using VA = V4<AT>;
    __shared__ int sIdx[4][4][SPMM_W];    // [wave][row of the wave][slot]
    __shared__ AT sVal[4][4][SPMM_W];
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    if (!bd.active) return;
    const int n = bd.n;
    const auto rowptr = glob(bd.rowptr);
    const auto colidx = glob(bd.colidx);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, rg = lane / 16, j = lane % 16;
    const VT * Pb = P + bd.off + 4 * j;
    VT * Qb = Q + bd.off + 4 * j;
    static_assert(sizeof(AT) == 8, "the matrix stays FP64");
    const auto A = glob(bd.valsD);
    const int x = blockIdx.x % NXCD, nw = gridDim.x / NXCD;
    const int lo = (int)((long)n * x / NXCD), hi = (int)((long)n * (x + 1) / NXCD);
    if (lo >= hi) return;
    const int stride = 16 * nw;
    int * myIdx = sIdx[wave][rg];
    AT * myVal = sVal[wave][rg];
    // Software pipeline over the wavefront's rounds of 4 rows: the row
    // pointers are loaded two rounds ahead and the indices / values one round
    // ahead, so they are in flight while this round's gathers are summed.
    // Loads of rounds or rows past hi read row hi - 1 and keep nothing.
    auto rp_load = [&](int r, int & p0, int & p1) {
        const int rc = min(r + rg, hi - 1);
        p0 = rowptr[rc];
        p1 = rowptr[rc + 1];
    };
    auto ci_load = [&](int r, int p0, int p1, int & cv, AT & av) {
        const int q = p0 + j;
        const bool in = r + rg < hi && q < p1;
        const int qc = in ? q : p0;  // p0: a valid slot of a non-empty row
        const int c = colidx[qc];
        const AT v = A[qc];
        cv = in ? c : 0;
        av = in ? v : (AT)0;
    };
    int r0 = lo + 4 * ((blockIdx.x / NXCD) * 4 + wave);
    int p0c, p1c, p0n, p1n, cvc;
    AT avc;
    rp_load(r0, p0c, p1c);
    ci_load(r0, p0c, p1c, cvc, avc);
    rp_load(r0 + stride, p0n, p1n);
    for (; r0 < hi; r0 += stride) {
        const int row = r0 + rg;
        myIdx[j] = cvc;
        myVal[j] = avc;
        __builtin_amdgcn_wave_barrier();
        int ci[SPMM_W];
#pragma unroll
        for (int u = 0; u < SPMM_W; ++u) ci[u] = myIdx[u];
        V xv[SPMM_W];
#pragma unroll
        for (int u = 0; u < SPMM_W; ++u) xv[u] = ld4(Pb + (size_t)ci[u] * S);
        // next round's indices and values, the row pointers of the one after
        int cvn, p0nn, p1nn;
        AT avn;
        ci_load(r0 + stride, p0n, p1n, cvn, avn);
        rp_load(r0 + 2 * stride, p0nn, p1nn);
        // every gather and prefetch is issued before the first product waits
        __builtin_amdgcn_sched_barrier(0);
        VA acc0 = VA{0, 0, 0, 0}, acc1 = acc0;
#pragma unroll
        for (int u = 0; u < SPMM_W; u += 2) {
            acc0 += myVal[u] * __builtin_convertvector(xv[u], VA);
            acc1 += myVal[u + 1] * __builtin_convertvector(xv[u + 1], VA);
        }
        // nonzeros past the first SPMM_W of a row (none in the structured
        // tetrahedral meshes; a plain loop for general meshes)
        const int q0 = row < hi ? p0c : 0, q1 = row < hi ? p1c : 0;
        for (int q = q0 + SPMM_W; q < q1; ++q)
            acc0 += A[q] * __builtin_convertvector(ld4(Pb + (size_t)colidx[q] * S), VA);
        if (row < hi) st4(Qb + (size_t)row * S, __builtin_convertvector(acc0 + acc1, V));
        __builtin_amdgcn_wave_barrier();
        p0c = p0n;
        p1c = p1n;
        p0n = p0nn;
        p1n = p1nn;
        cvc = cvn;
        avc = avn;
    }
}

// ---------------------------------------------------------------- smoothed two-level preconditioner
// z = M^-1 r with M^-1 the symmetric multiplicative two-level cycle: a
// degree-2 Chebyshev smoother S on D^-1 A over [b/11, b] (b the Gershgorin
// bound of D^-1 A), the coarse correction on the aggregates, the same
// smoother again:
//   z1 = S r,  r1 = r - A z1,  c = E^-1 Z' r1,  r2 = r1 - (A Z) c,
//   z  = z1 + Z c + S r2.
// With theta = (b + a)/2, delta = (b - a)/2, sigma = theta/delta, rho0 =
// 1/sigma, rho1 = 1/(2 sigma - rho0): S r = D^-1 (cz r + c2 s0) where
// s0 = r - A D^-1 r c0, and r - A S r = s0 - A D^-1 (c1 r + c2 s0), with
// c0 = 1/theta, c1 = rho1 rho0/theta, c2 = 2 rho1/delta, cz = (1 + rho1 rho0)/theta.
// M^-1 is symmetric (S is a polynomial in D^-1 A times D^-1) and positive
// definite (the smoother contracts in the A-norm on [0, b]).
// The three sparse passes share k_spmm's gather scheme, on AD = A D^-1:
//   MODE_S0: out = s0 = U - c0 AD U (U = r);  out2 = w = c1 U + c2 s0
//   MODE_S1: out = r1 = V - AD U               (U = w, V = s0)
//   MODE_SZ: s = U - c0 AD U (U = r2);  out = Z = D^-1 (cz (Rr + U) + c2 (V + s)) + cc[agg]
//            (Rr = r, V = s0)
// (each pass gathers one vector: r1 = s0 - A D^-1 (c1 r + c2 s0) uses the
// combination w that the first pass writes row by row)
constexpr int MODE_S0 = 0;
constexpr int MODE_S1 = 1;
constexpr int MODE_SZ = 2;
template <typename VT, int MODE>
__global__ void __launch_bounds__(256) k_smooth(const BlockDesc * __restrict__ blocks, const VT * __restrict__ U,
                                                const VT * __restrict__ V, const VT * __restrict__ Rr,
                                                const double * __restrict__ cc, VT * __restrict__ out,
                                                VT * __restrict__ out2) {
    using Vv = V4<VT>;
    using VA = f64x4;
    __shared__ int sIdx[4][4][SPMM_W];
    __shared__ double sVal[4][4][SPMM_W];
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    if (!bd.active) return;
    const int n = bd.n;
    const auto rowptr = glob(bd.rowptr);
    const auto colidx = glob(bd.colidx);
    const auto A = glob(bd.valsAD);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, rg = lane / 16, j = lane % 16;
    const VT * Ub = U + bd.off + 4 * j;
    const VT * Vb = V + bd.off + 4 * j;
    VT * Ob = out + bd.off + 4 * j;
    const int x = blockIdx.x % NXCD, nw = gridDim.x / NXCD;
    const int lo = (int)((long)n * x / NXCD), hi = (int)((long)n * (x + 1) / NXCD);
    if (lo >= hi) return;
    const int stride = 16 * nw;
    int * myIdx = sIdx[wave][rg];
    double * myVal = sVal[wave][rg];
    auto rp_load = [&](int r, int & p0, int & p1) {
        const int rc = min(r + rg, hi - 1);
        p0 = rowptr[rc];
        p1 = rowptr[rc + 1];
    };
    auto ci_load = [&](int r, int p0, int p1, int & cv, double & av) {
        const int q = p0 + j;
        const bool in = r + rg < hi && q < p1;
        const int qc = in ? q : p0;
        const int c = colidx[qc];
        const double v = A[qc];
        cv = in ? c : 0;
        av = in ? v : 0.0;
    };
    int r0 = lo + 4 * ((blockIdx.x / NXCD) * 4 + wave);
    int p0c, p1c, p0n, p1n, cvc;
    double avc;
    rp_load(r0, p0c, p1c);
    ci_load(r0, p0c, p1c, cvc, avc);
    rp_load(r0 + stride, p0n, p1n);
    for (; r0 < hi; r0 += stride) {
        const int row = r0 + rg;
        const int rowc = min(row, hi - 1);
        myIdx[j] = cvc;
        myVal[j] = avc;
        __builtin_amdgcn_wave_barrier();
        int ci[SPMM_W];
#pragma unroll
        for (int u = 0; u < SPMM_W; ++u) ci[u] = myIdx[u];
        Vv xu[SPMM_W];
#pragma unroll
        for (int u = 0; u < SPMM_W; ++u) xu[u] = ld4(Ub + (size_t)ci[u] * S);
        // this row's own values (the self term, and for MODE_SZ r, s0, D^-1, cc)
        const Vv self = ld4((MODE == MODE_S1 ? Vb : Ub) + (size_t)rowc * S);
        Vv rOwn = Vv{0, 0, 0, 0}, sOwn = Vv{0, 0, 0, 0};
        double dOwn = 0.0;
        VA cOwn = VA{0, 0, 0, 0};
        if constexpr (MODE == MODE_SZ) {
            rOwn = ld4(Rr + bd.off + 4 * j + (size_t)rowc * S);
            sOwn = ld4(Vb + (size_t)rowc * S);
            dOwn = glob(bd.dinv)[rowc];
            cOwn = ld4(cc + bd.coff + (size_t)glob(bd.agg)[rowc] * S + 4 * j);
        }
        int cvn, p0nn, p1nn;
        double avn;
        ci_load(r0 + stride, p0n, p1n, cvn, avn);
        rp_load(r0 + 2 * stride, p0nn, p1nn);
        __builtin_amdgcn_sched_barrier(0);
        VA au0 = VA{0, 0, 0, 0}, au1 = au0;
#pragma unroll
        for (int u = 0; u < SPMM_W; u += 2) {
            au0 += myVal[u] * __builtin_convertvector(xu[u], VA);
            au1 += myVal[u + 1] * __builtin_convertvector(xu[u + 1], VA);
        }
        const int q0 = row < hi ? p0c : 0, q1 = row < hi ? p1c : 0;
        for (int q = q0 + SPMM_W; q < q1; ++q)
            au0 += A[q] * __builtin_convertvector(ld4(Ub + (size_t)colidx[q] * S), VA);
        const VA sf = __builtin_convertvector(self, VA);
        VA res;
        if constexpr (MODE == MODE_S0) {
            res = sf - bd.ch[0] * (au0 + au1);
            if (row < hi) st4(out2 + bd.off + 4 * j + (size_t)row * S,
                              __builtin_convertvector(bd.ch[1] * sf + bd.ch[2] * res, Vv));
        } else if constexpr (MODE == MODE_S1) {
            res = sf - (au0 + au1);
        } else {
            const VA sp = sf - bd.ch[0] * (au0 + au1);
            res = dOwn * (bd.ch[3] * (__builtin_convertvector(rOwn, VA) + sf) +
                          bd.ch[2] * (__builtin_convertvector(sOwn, VA) + sp)) + cOwn;
        }
        if (row < hi) st4(Ob + (size_t)row * S, __builtin_convertvector(res, Vv));
        __builtin_amdgcn_wave_barrier();
        p0c = p0n;
        p1c = p1n;
        p0n = p0nn;
        p1n = p1nn;
        cvc = cvn;
        avc = avn;
    }
}

// r2 = r1 - (A Z) c in place; one thread per 4 columns of a row.  Most rows
// have one or two entries (their neighbours lie in one or two aggregates):
// the first two entries' indices and coarse rows are loaded together (a row
// with one entry repeats it with weight 0), further entries in a loop.
template <typename VT>
__global__ void __launch_bounds__(256) k_coarse_resid(const BlockDesc * __restrict__ blocks,
                                                      const double * __restrict__ cc, VT * __restrict__ R12) {
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    if (!bd.active) return;
    const auto azPtr = glob(bd.azPtr);
    const auto azAgg = glob(bd.azAgg);
    const auto azVals = glob(bd.azVals);
    const double * cb = cc + bd.coff;
    const long total = (long)bd.n * (S / 4);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int row = (int)(t / (S / 4)), c4 = 4 * (int)(t % (S / 4));
        const size_t o = bd.off + (size_t)row * S + c4;
        const V4<VT> r = ld4(R12 + o);
        const int p0 = azPtr[row], p1 = azPtr[row + 1];
        const int q1 = min(p0 + 1, p1 - 1);
        const int a0 = azAgg[p0], a1 = azAgg[q1];
        const double v0 = azVals[p0], w1 = azVals[q1];
        const double v1 = q1 > p0 ? w1 : 0.0;
        f64x4 acc = v0 * ld4(cb + (size_t)a0 * S + c4) + v1 * ld4(cb + (size_t)a1 * S + c4);
        for (int q = p0 + 2; q < p1; ++q) acc += azVals[q] * ld4(cb + (size_t)azAgg[q] * S + c4);
        st4(R12 + o, __builtin_convertvector(__builtin_convertvector(r, f64x4) - acc, V4<VT>));
    }
}

// ---------------------------------------------------------------- Gram passes
// MODE_PQR: vectors (P, Q, R) -> G1 = P'Q, G2 = P'R.
// MODE_Z:   vectors (P, Q, Z) -> Q'Z, Z'Z, P'Z.
// A workgroup of 8 wavefronts (two per SIMD) takes a contiguous range of
// 4-row steps; the rows of one step are 1 KB (FP32) per vector and
// contiguous in the block layout.  The steps go in groups of GS: every
// thread loads 3 of the group's 3 x GS x 64 16-byte pieces, keeps two groups
// in flight in registers and stages the current one through double-buffered
// LDS (one LDS-only barrier per group).  Every wavefront then holds the
// operands of the group's rows (lane l: row l/16 of a step, columns
// 4(l%16)..+3) and runs the MFMAs of its tiles: column tile tb = wave % 4,
// row tiles ta = 2 (wave / 4) + {0, 1}; tile (ta, tb) holds Gram entries
// (4i + ta, 4j + tb), result register jj of lane l entry
// (4 drow(l, jj) + ta, 4 (l%16) + tb).
// FP32 inputs: the matrix cores accumulate FP32 over FLUSH steps (16 rows),
// then the tile is added into FP64 registers -- the Grams keep FP64
// accumulation (their rounding steers the block recurrence) at the FP32
// matrix-core rate.  FP64 inputs accumulate in FP64 directly.
// Workgroup partials: part[((blk * G + g) * NG + gram) * S*S].
constexpr int MODE_PQR = 0;
constexpr int MODE_Z = 1;
constexpr int FLUSH = 4;
constexpr int GRAM_THREADS = 512;

template <typename VT, int MODE>
__global__ void __launch_bounds__(GRAM_THREADS) k_gram(const BlockDesc * __restrict__ blocks,
                                                       const VT * __restrict__ V0, const VT * __restrict__ V1,
                                                       const VT * __restrict__ V2, double * __restrict__ part) {
    constexpr int NG = MODE == MODE_PQR ? 2 : 3;
    constexpr bool F32 = sizeof(VT) == 4;
    using V = V4<VT>;
    constexpr int GS = 8;                                 // 4-row steps per LDS exchange
    constexpr int NL = 3 * GS * 64 / GRAM_THREADS;       // pieces loaded per thread and group
    static_assert(NL * GRAM_THREADS == 3 * GS * 64, "loads per group split evenly");
    static_assert(GS % FLUSH == 0, "flushes at group boundaries");
    __shared__ V sv[2][GS][3][64];  // [buffer][step of the group][vector][lane]
    const int b = blockIdx.y;
    if (!blocks[b].active) return;
    const int n = blocks[b].n;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int tb = wave % 4, ta0 = 2 * (wave / 4);
    const size_t vb = blocks[b].off;
    const int nst = (n + 3) / 4;
    const int st0 = (int)((long)nst * blockIdx.x / gridDim.x), st1 = (int)((long)nst * (blockIdx.x + 1) / gridDim.x);
    // piece i of this thread: item = i * GRAM_THREADS + threadIdx.x -> (step
    // ph[i] of the group, vector, lane); source pointers and LDS slots are
    // fixed per thread
    const VT * psrc[NL];
    int ph[NL], pslot[NL];
    const int ln = threadIdx.x % 64;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int item = i * GRAM_THREADS + threadIdx.x;
        const int vec = (item / 64) % 3;
        ph[i] = item / 192;
        pslot[i] = (ph[i] * 3 + vec) * 64 + ln;
        psrc[i] = (vec == 0 ? V0 : vec == 1 ? V1 : V2) + vb + 4 * (ln % 16);
    }
    // Loads without branches (a branch around a load makes the compiler wait
    // for every outstanding load at the next use): rows past the range read a
    // valid row, and stage() zeroes them.
    auto fetch = [&](int s0, V (&r)[NL]) {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int row = max(0, min(4 * min(s0 + ph[i], st1 - 1) + ln / 16, n - 1));
            r[i] = ld4(psrc[i] + (size_t)row * S);
        }
    };
    auto stage = [&](int buf, int s0, const V (&r)[NL]) {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int st = s0 + ph[i];
            (&sv[buf][0][0][0])[pslot[i]] = zero_unless(st < st1 && 4 * st + ln / 16 < n, r[i]);
        }
    };

    V acc[NG][2];
    f64x4 acc64[F32 ? NG : 1][2];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[g][t] = V{0, 0, 0, 0};
            if constexpr (F32) acc64[g][t] = f64x4{0, 0, 0, 0};
        }
    auto flush = [&]() {
        if constexpr (F32) {
#pragma unroll
            for (int g = 0; g < NG; ++g)
#pragma unroll
                for (int t = 0; t < 2; ++t) {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) acc64[g][t][jj] += (double)acc[g][t][jj];
                    acc[g][t] = V{0, 0, 0, 0};
                }
        }
    };

    auto compute = [&](int buf) {
#pragma unroll
        for (int h = 0; h < GS; ++h) {
            const V x0 = sv[buf][h][0][lane], x1 = sv[buf][h][1][lane], x2 = sv[buf][h][2][lane];
            V av[NG], bo[NG];
            if constexpr (MODE == MODE_PQR) {
                av[0] = x0;  // P'Q
                av[1] = x0;  // P'R
                bo[0] = x1;
                bo[1] = x2;
            } else {
                av[0] = x1;  // Q'Z
                av[1] = x2;  // Z'Z
                av[2] = x0;  // P'Z
                bo[0] = x2;
                bo[1] = x2;
                bo[2] = x2;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const VT bw = tb == 0 ? bo[g][0] : tb == 1 ? bo[g][1] : tb == 2 ? bo[g][2] : bo[g][3];
                const VT a0 = ta0 == 0 ? av[g][0] : av[g][2];
                const VT a1 = ta0 == 0 ? av[g][1] : av[g][3];
                acc[g][0] = mfma4(a0, bw, acc[g][0]);
                acc[g][1] = mfma4(a1, bw, acc[g][1]);
            }
            if ((h % FLUSH) == FLUSH - 1) flush();
        }
    };
    // two groups in flight in two register sets used in turn (no register
    // copies, which would wait for the younger loads too)
    V ra[NL], rb[NL];
    fetch(st0, ra);
    __builtin_amdgcn_sched_barrier(0);  // ra's loads are issued before rb's
    fetch(st0 + GS, rb);
    for (int s0 = st0; s0 < st1; s0 += 2 * GS) {
        // (the scheduling barriers keep each group's loads ahead of the
        // compute phase that hides them)
        stage(0, s0, ra);
        fetch(s0 + 2 * GS, ra);
        __builtin_amdgcn_sched_barrier(0);
        lds_barrier();
        compute(0);
        // no early exit between the halves (a branch here would let the
        // compiler sink the loads above into the path that uses them); a
        // group past the range is staged as zeros and adds nothing
        stage(1, s0 + GS, rb);
        fetch(s0 + 3 * GS, rb);
        __builtin_amdgcn_sched_barrier(0);
        lds_barrier();
        compute(1);
    }
    flush();
    double * out = part + ((size_t)b * gridDim.x + blockIdx.x) * NG * SS;
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int ai = 4 * drow<VT>(lane, jj) + ta0 + t, bj = 4 * (lane % 16) + tb;
                double v;
                if constexpr (F32)
                    v = acc64[g][t][jj];
                else
                    v = acc[g][t][jj];
                out[(size_t)g * SS + ai * S + bj] = v;
            }
}

// Z = D^-1 R (+ coarse prolongation cc[agg[row]]), block layout; one thread
// per 4 columns of a row
template <typename VT>
__global__ void __launch_bounds__(256) k_zform(const BlockDesc * __restrict__ blocks, const VT * __restrict__ R,
                                               const double * __restrict__ cc, VT * __restrict__ Z) {
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    if (!bd.active) return;
    const long total = (long)bd.n * (S / 4);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int row = (int)(t / (S / 4)), c4 = 4 * (int)(t % (S / 4));
        const size_t o = bd.off + (size_t)row * S + c4;
        const double d = glob(bd.dinv)[row];
        const V4<VT> r = ld4(R + o);
        f64x4 cv = {0, 0, 0, 0};
        if (cc && bd.m > 0) cv = ld4(cc + bd.coff + (size_t)glob(bd.agg)[row] * S + c4);
        V4<VT> z;
#pragma unroll
        for (int e = 0; e < 4; ++e) z[e] = (VT)(d * (double)r[e] + cv[e]);
        st4(Z + o, z);
    }
}

// ---------------------------------------------------------------- updates over 16-row strips
// Lane l of a wavefront holds row 16 s + l%16, columns 16 u + 4 (l/16) + e
// (u, e < 4) as four vector loads.  C = L M is computed as C' = M' L': MFMA
// step (u, e) contracts over the four columns 16 u + 4 k + e (k = l/16), the
// A operand of column tile tb is M[16 u + 4 (l/16) + e][16 tb + l%16] and the
// result register jj of tile tb is C[row][16 tb + drow(l, jj)] -- in FP32
// the same layout as the loads (one vector store), in FP64 columns strided
// by 4 (four scalar accesses).
// Pre-swizzled LDS operand: op[((4 u + e) * 4 + tb) * 64 + lane].
// Needs blockDim.x == 256: thread t stages entries t + 256 i, i < 16, and
// issues its 16 loads before the first LDS store (a loop with a store per
// load would wait for each load in turn).
template <typename VT>
__device__ inline void stage_op(VT * op, const double * M) {
    const int tb = threadIdx.x / 64, ln = threadIdx.x % 64;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = M[(16 * (i / 4) + 4 * (ln / 16) + i % 4) * S + 16 * tb + ln % 16];
#pragma unroll
    for (int i = 0; i < 16; ++i) op[i * 256 + threadIdx.x] = (VT)v[i];
}

template <typename VT>
__device__ inline void strip_mm(const VT * __restrict__ op, const V4<VT> (&x)[4], V4<VT> (&c)[4], int lane) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int tb = 0; tb < 4; ++tb) c[tb] = mfma4(op[((4 * u + e) * 4 + tb) * 64 + lane], x[u][e], c[tb]);
}

// 3. R -= Q alpha; partial |R|^2 per column -> part[(blk * G + g) * S]
// (X += P alpha runs later, beside the beta step: x_update)
template <typename VT>
__global__ void __launch_bounds__(256) k_update(const BlockDesc * __restrict__ blocks,
                                                const double * __restrict__ small, const VT * __restrict__ Q,
                                                VT * __restrict__ R, double * __restrict__ part) {
    using V = V4<VT>;
    __shared__ VT op[16 * 4 * 64];
    __shared__ double cn[4][S];
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    if (!bd.active) return;
    stage_op(op, small + ((size_t)b * NSL + SL_G2) * SS);
    __syncthreads();
    const int n = bd.n;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t vb = bd.off;
    const int cq = 4 * (lane / 16);
    double rr[4][4] = {};
    const int nstrip = (n + 15) / 16;
    const int stride = gridDim.x * 4;
    // software pipeline: the next strip's Q is in flight while the matrix
    // cores work on the current one
    auto fetch = [&](int si, V (&qq)[4]) {
        const int row = 16 * si + lane % 16;
        const bool ok = si < nstrip && row < n;
        const size_t o = vb + (size_t)row * S + cq;
#pragma unroll
        for (int u = 0; u < 4; ++u) qq[u] = ok ? ld4(Q + o + 16 * u) : V{0, 0, 0, 0};
    };
    int sidx = blockIdx.x * 4 + wave;
    V q[4];
    fetch(sidx, q);
    for (; sidx < nstrip; sidx += stride) {
        const int row = 16 * sidx + lane % 16;
        const bool ok = row < n;
        const size_t o = vb + (size_t)row * S + cq, ob = vb + (size_t)row * S;
        // this strip's R in the result layout (independent of the product)
        V rv[4];
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) {
            if constexpr (sizeof(VT) == 4) {
                rv[tb] = ok ? ld4(R + o + 16 * tb) : V{0, 0, 0, 0};
            } else {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) rv[tb][jj] = ok ? R[ob + 16 * tb + drow<VT>(lane, jj)] : 0.0;
            }
        }
        V qn[4], cr[4];
        fetch(sidx + stride, qn);
#pragma unroll
        for (int u = 0; u < 4; ++u) cr[u] = V{0, 0, 0, 0};
        strip_mm(op, q, cr, lane);
        if (ok) {
#pragma unroll
            for (int tb = 0; tb < 4; ++tb) {
                const V rn = rv[tb] - cr[tb];
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) rr[tb][jj] += (double)rn[jj] * (double)rn[jj];
                if constexpr (sizeof(VT) == 4) {
                    st4(R + o + 16 * tb, rn);
                } else {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) R[ob + 16 * tb + drow<VT>(lane, jj)] = rn[jj];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = qn[u];
    }
    // sum over the 16 rows of a lane group, then over the wavefronts
#pragma unroll
    for (int tb = 0; tb < 4; ++tb)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            double v = rr[tb][jj];
            v += __shfl_xor(v, 8);
            v += __shfl_xor(v, 4);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 1);
            if (lane % 16 == 0) cn[wave][16 * tb + drow<VT>(lane, jj)] = v;
        }
    __syncthreads();
    if (wave == 0)
        part[((size_t)b * gridDim.x + blockIdx.x) * S + lane] = cn[0][lane] + cn[1][lane] + cn[2][lane] + cn[3][lane];
}

// X += P alpha over the strips g, g + ng, ... (workgroup g of ng); op: LDS
// space for the staged alpha (16 * 4 * 64 VT)
template <typename VT>
__device__ void x_update(const BlockDesc & bd, const double * alpha, const VT * __restrict__ P, VT * __restrict__ X,
                         VT * op, int g, int ng) {
    using V = V4<VT>;
    stage_op(op, alpha);
    __syncthreads();
    const int n = bd.n;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t vb = bd.off;
    const int cq = 4 * (lane / 16);
    const int nstrip = (n + 15) / 16;
    const int stride = ng * 4;
    auto fetch = [&](int si, V (&pp)[4]) {
        const int row = 16 * si + lane % 16;
        const bool ok = si < nstrip && row < n;
        const size_t o = vb + (size_t)row * S + cq;
#pragma unroll
        for (int u = 0; u < 4; ++u) pp[u] = ok ? ld4(P + o + 16 * u) : V{0, 0, 0, 0};
    };
    int sidx = g * 4 + wave;
    V p[4];
    fetch(sidx, p);
    for (; sidx < nstrip; sidx += stride) {
        const int row = 16 * sidx + lane % 16;
        const bool ok = row < n;
        const size_t o = vb + (size_t)row * S + cq, ob = vb + (size_t)row * S;
        V xv[4];
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) {
            if constexpr (sizeof(VT) == 4) {
                xv[tb] = ok ? ld4(X + o + 16 * tb) : V{0, 0, 0, 0};
            } else {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) xv[tb][jj] = ok ? X[ob + 16 * tb + drow<VT>(lane, jj)] : 0.0;
            }
        }
        V pn[4], cx[4];
        fetch(sidx + stride, pn);
#pragma unroll
        for (int u = 0; u < 4; ++u) cx[u] = V{0, 0, 0, 0};
        strip_mm(op, p, cx, lane);
        if (ok) {
#pragma unroll
            for (int tb = 0; tb < 4; ++tb) {
                const V xn = xv[tb] + cx[tb];
                if constexpr (sizeof(VT) == 4) {
                    st4(X + o + 16 * tb, xn);
                } else {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) X[ob + 16 * tb + drow<VT>(lane, jj)] = xn[jj];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) p[u] = pn[u];
    }
}

// 5. P = Z T + P M2
template <typename VT>
__global__ void __launch_bounds__(256) k_pupdate(const BlockDesc * __restrict__ blocks,
                                                 const double * __restrict__ small, const VT * __restrict__ Z,
                                                 VT * __restrict__ P) {
    using V = V4<VT>;
    __shared__ VT opT[16 * 4 * 64], opM[16 * 4 * 64];
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    if (!bd.active) return;
    stage_op(opT, small + ((size_t)b * NSL + SL_T) * SS);
    stage_op(opM, small + ((size_t)b * NSL + SL_M2) * SS);
    __syncthreads();
    const int n = bd.n;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t vb = bd.off;
    const int cq = 4 * (lane / 16);
    const int nstrip = (n + 15) / 16;
    const int stride = gridDim.x * 4;
    // software pipeline: the next strip's Z and P are in flight during the products
    auto fetch = [&](int si, V (&pp)[4], V (&zz)[4]) {
        const int row = 16 * si + lane % 16;
        const bool ok = si < nstrip && row < n;
        const size_t o = vb + (size_t)row * S + cq;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            pp[u] = ok ? ld4(P + o + 16 * u) : V{0, 0, 0, 0};
            zz[u] = ok ? ld4(Z + o + 16 * u) : V{0, 0, 0, 0};
        }
    };
    int sidx = blockIdx.x * 4 + wave;
    V p[4], z[4];
    fetch(sidx, p, z);
    for (; sidx < nstrip; sidx += stride) {
        const int row = 16 * sidx + lane % 16;
        const bool ok = row < n;
        const size_t o = vb + (size_t)row * S + cq;
        V pn[4], zn[4], c[4];
        fetch(sidx + stride, pn, zn);
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = V{0, 0, 0, 0};
        strip_mm(opT, z, c, lane);
        strip_mm(opM, p, c, lane);
        if (ok) {
            if constexpr (sizeof(VT) == 4) {
#pragma unroll
                for (int tb = 0; tb < 4; ++tb) st4(P + o + 16 * tb, c[tb]);
            } else {
                const size_t ob = vb + (size_t)row * S;
#pragma unroll
                for (int tb = 0; tb < 4; ++tb)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) P[ob + 16 * tb + drow<VT>(lane, jj)] = c[tb][jj];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            p[u] = pn[u];
            z[u] = zn[u];
        }
    }
}

// ---------------------------------------------------------------- coarse correction
// y[b][I][c] = sum over the nodes of aggregate I of R[b][node][c]: one
// workgroup per aggregate, 4 wavefronts over its node list
template <typename VT>
__global__ void __launch_bounds__(256) k_restrict(const BlockDesc * __restrict__ blocks, const VT * __restrict__ R,
                                                  double * __restrict__ y) {
    __shared__ double cn[4][S];
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    const int I = blockIdx.x;
    if (!bd.active || I >= bd.m) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const VT * Rb = R + bd.off + lane;
    double s = 0.0;
    const auto aggPtr = glob(bd.aggPtr);
    const auto aggNodes = glob(bd.aggNodes);
    const int q0 = aggPtr[I], q1 = aggPtr[I + 1];
    // wavefront w takes the nodes q0 + 64 (4 c + w) + l: lane l loads one node
    // index, then the wavefront's rows are loaded 16 at a time (indices
    // broadcast from their lanes), all 16 loads in flight
    for (int c0 = q0 + 64 * wave; c0 < q1; c0 += 256) {
        const int cnt = min(64, q1 - c0);
        const int myNode = lane < cnt ? aggNodes[c0 + lane] : 0;
        for (int j0 = 0; j0 < cnt; j0 += 16) {
            // lanes past cnt hold node 0: their loads stay in bounds and are
            // not summed (no branch around a load, so all 16 stay in flight)
            VT v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = Rb[(size_t)__builtin_amdgcn_readlane(myNode, j0 + u) * S];
#pragma unroll
            for (int u = 0; u < 16; ++u) s += j0 + u < cnt ? (double)v[u] : 0.0;
        }
    }
    cn[wave][lane] = s;
    __syncthreads();
    if (wave == 0) y[bd.coff + (size_t)I * S + lane] = cn[0][lane] + cn[1][lane] + cn[2][lane] + cn[3][lane];
}

// cc[b] = Einv_k y[b] (FP64 MFMA): a workgroup computes one 16 x 16 tile
// (rows 16 blockIdx.x, columns 16 * tile); its 4 wavefronts split the
// m-long contraction (16 k-steps of loads in flight before their MFMAs) and
// are summed through LDS in a fixed order
constexpr int CA_WAVES = 4;
__global__ void __launch_bounds__(64 * CA_WAVES) k_coarse_apply(const BlockDesc * __restrict__ blocks,
                                                                const double * __restrict__ y,
                                                                double * __restrict__ cc) {
    __shared__ double red[CA_WAVES][64][4];
    const int b = blockIdx.z;
    const BlockDesc bd = blocks[b];
    const int m = bd.m;
    const int I0 = 16 * blockIdx.x, c0 = 16 * blockIdx.y;
    if (!bd.active || I0 >= m) return;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const auto E = glob(bd.Einv);
    const double * yb = y + bd.coff;
    const int i = I0 + lane % 16;
    f64x4 acc = {0, 0, 0, 0};
    const int nk = (m + 3) / 4;
    const int k0 = nk * wave / CA_WAVES, k1 = nk * (wave + 1) / CA_WAVES;
    auto a_at = [&](int ks) {
        const int J = 4 * ks + lane / 16;
        return (i < m && J < m) ? E[(size_t)i * m + J] : 0.0;
    };
    auto b_at = [&](int ks) {
        const int J = 4 * ks + lane / 16;
        return J < m ? yb[(size_t)J * S + c0 + lane % 16] : 0.0;
    };
    int ks = k0;
    for (; ks + 16 <= k1; ks += 16) {
        double a[16], bv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            a[u] = a_at(ks + u);
            bv[u] = b_at(ks + u);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = mfma4(a[u], bv[u], acc);
    }
    for (; ks < k1; ++ks) acc = mfma4(a_at(ks), b_at(ks), acc);
    for (int jj = 0; jj < 4; ++jj) red[wave][lane][jj] = acc[jj];
    __syncthreads();
    if (wave == 0)
        for (int jj = 0; jj < 4; ++jj) {
            const int row = I0 + drow<double>(lane, jj);
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < CA_WAVES; ++w) v += red[w][lane][jj];
            if (row < m) cc[bd.coff + (size_t)row * S + c0 + lane % 16] = v;
        }
}

// ---------------------------------------------------------------- partial reduction
// small[b][dstOff + t] = sum_g part[(b * G + g) * count + t]; 32 outputs x 8
// group ranges per workgroup, fixed summation order
__global__ void __launch_bounds__(256) k_part_reduce(const BlockDesc * __restrict__ blocks, int G, int count,
                                                     const double * __restrict__ part, double * __restrict__ small,
                                                     int dstOff) {
    __shared__ double acc[8][32];
    const int b = blockIdx.y;
    if (!blocks[b].active) return;
    const int o = threadIdx.x % 32, gc = threadIdx.x / 32;
    const int t = blockIdx.x * 32 + o;
    double s = 0.0;
    if (t < count) {
        const int g0 = G * gc / 8, g1 = G * (gc + 1) / 8;
        const double * p = part + (size_t)b * G * count + t;
        int g = g0;
        // 16 loads in flight, summed in a fixed order
        for (; g + 16 <= g1; g += 16) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = p[(size_t)(g + u) * count];
#pragma unroll
            for (int w = 8; w > 0; w >>= 1)
#pragma unroll
                for (int u = 0; u < w; ++u) v[u] += v[u + w];
            s += v[0];
        }
        for (; g < g1; ++g) s += p[(size_t)g * count];
    }
    acc[gc][o] = s;
    __syncthreads();
    if (gc == 0 && t < count) {
        double v = 0.0;
        for (int q = 0; q < 8; ++q) v += acc[q][o];
        small[(size_t)b * NSL * SS + dstOff + t] = v;
    }
}

// ---------------------------------------------------------------- small dense kernels (LDS, FP64)
constexpr int LD = S + 1;  // padded leading dimension against bank conflicts
constexpr int SMAT = S * LD;
constexpr size_t SMALL_LDS = 4 * SMAT * sizeof(double);

// C = scale op(A) B (64 x 64 LDS matrices, FP64 matrix cores); transA: A'.
// Wavefront w computes rows 16w .. 16w+15.  C must not alias A or B.
__device__ void mm64(const double * A, const double * B, double * C, bool transA, double scale) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int i = 16 * w + l % 16;
    f64x4 acc[4] = {};
#pragma unroll 4
    for (int ks = 0; ks < 16; ++ks) {
        const int k = 4 * ks + l / 16;
        const double a = transA ? A[k * LD + i] : A[i * LD + k];
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) acc[tb] = mfma4(a, B[k * LD + 16 * tb + l % 16], acc[tb]);
    }
#pragma unroll
    for (int tb = 0; tb < 4; ++tb)
#pragma unroll
        for (int r = 0; r < 4; ++r) C[(16 * w + drow<double>(l, r)) * LD + 16 * tb + l % 16] = scale * acc[tb][r];
    __syncthreads();
}

// Li = L^-1 for G = L L' (lower Cholesky factor) with dependent directions
// dropped: forward elimination of [G | I] whose right half becomes the
// inverse unit-lower factor, rows then scaled by the pivots^-1/2.  Each thread
// keeps a 4 x 4 tile of G and of Li in registers; pivot step j publishes
// column j and row j through double-buffered LDS vectors (one barrier per
// step).  A pivot at or below 1e-12 of the column's own diagonal is dropped:
// no elimination with it, zero row and column in Li.  G (LDS) is only read.
__device__ void inv_chol(const double * G, double * Li, int s) {
    __shared__ double colb[2][S], rowb[2][S], lrow[2][S], d0[S];
    const int i0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
    double g[4][4], li[4][4], piv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            g[i][c] = G[(i0 + i) * LD + c0 + c];
            li[i][c] = (i0 + i == c0 + c && i0 + i < s) ? 1.0 : 0.0;
        }
    if (threadIdx.x < S) d0[threadIdx.x] = threadIdx.x < s ? G[threadIdx.x * LD + threadIdx.x] : 0.0;
    __syncthreads();
    // Each step is one chain: publish, LDS-only barrier, one batch of LDS
    // reads, FMAs.  The steps go in fours so that the register holding
    // column / row j of a tile has a static index (j % 4); the owner threads
    // publish under a predicate, all threads read unconditionally.  Only the
    // row factors f are masked: updates of already eliminated columns of G
    // (c <= j) are never read again, and row j of L^-1 is zero beyond its
    // diagonal, so the row and Li updates need no masks.
    auto step = [&](auto qtag, int j0) {
        constexpr int Q = decltype(qtag)::value;
        const int j = j0 + Q;
        const int buf = j & 1;
        if (c0 == j0)
#pragma unroll
            for (int i = 0; i < 4; ++i) colb[buf][i0 + i] = g[i][Q];
        if (i0 == j0)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                rowb[buf][c0 + c] = g[Q][c];
                lrow[buf][c0 + c] = li[Q][c];
            }
        lds_barrier();
        const double d = rowb[buf][j], dj = d0[j];
        double cb[4], rb[4], lr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) cb[i] = colb[buf][i0 + i];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            rb[c] = rowb[buf][c0 + c];
            lr[c] = lrow[buf][c0 + c];
        }
        const bool keep = dj > 0.0 && d > 1e-12 * dj;
        const double rcp = 1.0 / d;
        const double inv = keep ? rcp : 0.0;  // a dropped pivot eliminates nothing
        if (i0 == j0) piv[Q] = keep ? d : 0.0;
        double f[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) f[i] = i0 + i > j ? cb[i] * inv : 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                g[i][c] -= f[i] * rb[c];
                li[i][c] -= f[i] * lr[c];
            }
    };
    for (int j0 = 0; j0 < s; j0 += 4) {
        step(std::integral_constant<int, 0>{}, j0);
        if (j0 + 1 < s) step(std::integral_constant<int, 1>{}, j0);
        if (j0 + 2 < s) step(std::integral_constant<int, 2>{}, j0);
        if (j0 + 3 < s) step(std::integral_constant<int, 3>{}, j0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double sc = piv[i] > 0.0 ? 1.0 / sqrt(piv[i]) : 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) Li[(i0 + i) * LD + c0 + c] = li[i][c] * sc;
    }
    __syncthreads();
}

// The same result as inv_chol, blocked in 16 x 16 tiles so that the 64
// sequential pivot steps need no workgroup barrier: for each diagonal tile K
//   (a) every wavefront eliminates the tile in registers (16 pivot steps;
//       row j reaches the other lanes of its 16-lane row through DPP
//       row_share, column j through one permute), giving V_K = U_KK^-1 and the
//       pivots D_K (dependent pivots dropped as in inv_chol);
//   (b) panel Y_IK = G_IK V_K' and U_IK = Y_IK D_K^+ (I > K, one wavefront each);
//   (c) trailing update G_IJ -= U_IK Y_JK' (K < J <= I);
// then X = U^-1 blockwise (X_KK = V_K, X_IJ = -V_I sum_K U_IK X_KJ, three
// rounds) and Li = D^-1/2 X (zero rows for dropped pivots).  All products are
// 16 x 16 x 16 FP64 MFMA tiles.  G is overwritten (U in its strictly lower
// tiles); Li holds Y as scratch until X replaces it.
// f(integral_constant<J>) for J = 0, 1, ... (every index a compile-time constant)
template <typename F, int... J> __device__ inline void unrolled(F && f, std::integer_sequence<int, J...>) {
    (f(std::integral_constant<int, J>{}), ...);
}
__device__ inline double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// value of lane 16 (l / 16) + J (within the lane's DPP row of 16)
template <int J> __device__ inline double row_share_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x150 + J, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x150 + J, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// acc + op(A) op(B) for 16 x 16 LDS tiles (leading dimension LD), FP64 MFMA
// layout: lane l supplies A[l%16][k] and B[k][l%16], k = 4 ks + l/16
__device__ inline f64x4 tile_mm(const double * A, bool tA, const double * B, bool tB, f64x4 acc, double sgn) {
    const int l = threadIdx.x % 64, i = l % 16, kq = l / 16;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int k = 4 * ks + kq;
        const double a = tA ? A[k * LD + i] : A[i * LD + k];
        const double b = tB ? B[i * LD + k] : B[k * LD + i];
        acc = mfma4(sgn * a, b, acc);
    }
    return acc;
}
__device__ inline f64x4 tile_load(const double * C) {
    const int l = threadIdx.x % 64;
    f64x4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = C[drow<double>(l, r) * LD + l % 16];
    return v;
}
__device__ inline void tile_store(double * C, f64x4 v) {
    const int l = threadIdx.x % 64;
#pragma unroll
    for (int r = 0; r < 4; ++r) C[drow<double>(l, r) * LD + l % 16] = v[r];
}

__device__ void inv_chol_blk(double * G, double * Li, int s) {
    __shared__ double d0[S], piv[S];
    const int wave = threadIdx.x / 64, l = threadIdx.x % 64, r = l % 16, qg = l / 16;
    if (threadIdx.x < S) d0[threadIdx.x] = threadIdx.x < s ? G[threadIdx.x * LD + threadIdx.x] : 0.0;
    lds_barrier();
    auto T = [&](double * M, int I, int J) { return M + 16 * I * LD + 16 * J; };
    for (int K = 0; K < 4; ++K) {
        // (a) the diagonal tile, redundantly in every wavefront
        double g[4], v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            g[e] = T(G, K, K)[r * LD + 4 * qg + e];
            v[e] = r == 4 * qg + e ? 1.0 : 0.0;
        }
        auto step = [&](auto jtag) {
            constexpr int j = decltype(jtag)::value;
            const double d = readlane_d(g[j % 4], j + 16 * (j / 4));
            const double gc = __shfl(g[j % 4], r + 16 * (j / 4));
            double gr[4], vr[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                gr[e] = row_share_d<j>(g[e]);
                vr[e] = row_share_d<j>(v[e]);
            }
            const double dj = d0[16 * K + j];
            const bool keep = dj > 0.0 && d > 1e-12 * dj;
            const double rcp = 1.0 / d;
            const double inv = keep ? rcp : 0.0;
            if (l == 0 && wave == 0) piv[16 * K + j] = keep ? d : 0.0;
            const double f = r > j ? gc * inv : 0.0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                g[e] -= f * gr[e];
                v[e] -= f * vr[e];
            }
        };
        unrolled(step, std::make_integer_sequence<int, 16>{});
        if (wave == 0)
#pragma unroll
            for (int e = 0; e < 4; ++e) T(Li, K, K)[r * LD + 4 * qg + e] = v[e];
        lds_barrier();
        if (K == 3) break;
        // (b) panels
        const int I = K + 1 + wave;
        if (I < 4) {
            const f64x4 y = tile_mm(T(G, I, K), false, T(Li, K, K), true, f64x4{0, 0, 0, 0}, 1.0);
            tile_store(T(Li, I, K), y);
            const double ip = piv[16 * K + l % 16] > 0.0 ? 1.0 / piv[16 * K + l % 16] : 0.0;
            tile_store(T(G, I, K), y * ip);
        }
        lds_barrier();
        // (c) trailing tiles (I, J), K < J <= I
        for (int p = wave;; p += 4) {
            int I2 = -1, J2 = -1, cnt = 0;
            for (int a = K + 1; a < 4 && I2 < 0; ++a)
                for (int b2 = K + 1; b2 <= a; ++b2)
                    if (cnt++ == p) {
                        I2 = a;
                        J2 = b2;
                        break;
                    }
            if (I2 < 0) break;
            f64x4 acc = tile_load(T(G, I2, J2));
            acc = tile_mm(T(G, I2, K), false, T(Li, J2, K), true, acc, -1.0);
            tile_store(T(G, I2, J2), acc);
        }
        lds_barrier();
    }
    // X = U^-1 into Li's lower tiles (X_KK = V_K already there)
    for (int dist = 1; dist < 4; ++dist) {
        const int J = wave, I = J + dist;
        f64x4 t = {0, 0, 0, 0};
        if (I < 4) {
            for (int K = J; K < I; ++K) t = tile_mm(T(G, I, K), false, T(Li, K, J), false, t, 1.0);
        }
        lds_barrier();  // every read of this round's inputs is done before X_IJ overwrites Y_IJ
        if (I < 4) {
            // X_IJ = -V_I t; t's result registers are the B operands of the k-steps
            const double * Vi = T(Li, I, I);
            f64x4 x = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) x = mfma4(-Vi[(l % 16) * LD + 4 * ks + l / 16], t[ks], x);
            tile_store(T(Li, I, J), x);
        }
        lds_barrier();
    }
    // rows scaled by the pivots^-1/2, tiles above the diagonal zero
    for (int t2 = threadIdx.x; t2 < SS; t2 += blockDim.x) {
        const int i = t2 / S, c = t2 % S;
        const double sc = piv[i] > 0.0 ? 1.0 / sqrt(piv[i]) : 0.0;
        Li[i * LD + c] = c / 16 > i / 16 ? 0.0 : Li[i * LD + c] * sc;
    }
    __syncthreads();
}

// global -> LDS copy of s x s matrices (zero padded), split into a fetch of
// N matrices into registers and the LDS stores, so that every load of a
// thread is in flight before its first store (a loop with a store per load
// waits for each load in turn).  Needs blockDim.x == 256; no barrier:
// callers synchronise once.
template <int N>
__device__ inline void fetch_small(double (&v)[N][16], const double * const (&src)[N]) {
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) v[q][r] = src[q][threadIdx.x + 256 * r];
}
__device__ inline void put_small(double * M, const double (&v)[16], int s) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int t = threadIdx.x + 256 * r, i = t / S, j = t % S;
        M[i * LD + j] = (i < s && j < s) ? v[r] : 0.0;
    }
}

// LDS -> global; the LDS matrix must not be overwritten before a barrier
__device__ void store_small(const double * M, double * dst) {
    for (int t = threadIdx.x; t < SS; t += blockDim.x) dst[t] = M[(t / S) * LD + t % S];
}

// alpha = G1^-1 G2; G1 is replaced by its inverse (reused for beta)
__global__ void __launch_bounds__(256) k_small_alpha(const BlockDesc * __restrict__ blocks, double * __restrict__ small) {
    extern __shared__ double lds[];
    if (!blocks[blockIdx.x].active) return;
    const int s = blocks[blockIdx.x].s;
    double * sm = small + (size_t)blockIdx.x * NSL * SS;
    double *M0 = lds, *M1 = lds + SMAT, *M2 = lds + 2 * SMAT, *M3 = lds + 3 * SMAT;
    double v[2][16];
    fetch_small<2>(v, {sm + SL_G1 * SS, sm + SL_G2 * SS});
    put_small(M0, v[0], s);
    put_small(M3, v[1], s);
    __syncthreads();
    inv_chol(M0, M1, s);
    mm64(M1, M1, M2, true, 1.0);   // G1^-1 = Li' Li
    mm64(M2, M3, M0, false, 1.0);  // alpha
    store_small(M2, sm + SL_G1 * SS);
    store_small(M0, sm + SL_G2 * SS);
}

// init = 1: first step of a block (no previous direction): beta = 0, W = Z.
// Workgroup 0 of each block (blockIdx.y) takes the beta step; workgroups
// 1 .. gridDim.x - 1 meanwhile add the previous direction's step into the
// solution, X += P alpha (independent of beta; it must precede the P update,
// the next launch), so that streaming work fills the GPU while the single
// 64 x 64 workgroup runs.
template <typename VT>
__global__ void __launch_bounds__(256) k_small_beta(const BlockDesc * __restrict__ blocks, double * __restrict__ small,
                                                    int init, const VT * __restrict__ P, VT * __restrict__ X) {
    extern __shared__ double lds[];
    const int b = blockIdx.y;
    if (!blocks[b].active) return;
    if (blockIdx.x > 0) {
        x_update<VT>(blocks[b], small + ((size_t)b * NSL + SL_G2) * SS, P, X, reinterpret_cast<VT *>(lds),
                     blockIdx.x - 1, gridDim.x - 1);
        return;
    }
    const int s = blocks[b].s;
    double * sm = small + (size_t)b * NSL * SS;
    double *M0 = lds, *M1 = lds + SMAT, *Bt = lds + 2 * SMAT, *W = lds + 3 * SMAT;
    if (!init) {
        // W = Z'Z, M0 = G1^-1, M1 = Q'Z; P'Z waits in registers for M0
        double v[4][16];
        fetch_small<4>(v, {sm + SL_ZZ * SS, sm + SL_G1 * SS, sm + SL_QZ * SS, sm + SL_PZ * SS});
        put_small(W, v[0], s);
        put_small(M0, v[1], s);
        put_small(M1, v[2], s);
        __syncthreads();
        mm64(M0, M1, Bt, false, -1.0);       // beta = -G1^-1 Q'Z
        put_small(M0, v[3], s);              // P'Z
        __syncthreads();
        mm64(M0, Bt, M1, true, 1.0);         // (P'Z)' beta
        for (int t = threadIdx.x; t < SS; t += blockDim.x) {
            const int i = t / S, j = t % S;
            W[i * LD + j] += M1[i * LD + j] + M1[j * LD + i];
        }
        __syncthreads();
        mm64(Bt, Bt, M1, true, 1.0);         // beta' beta  (P'P = I)
        for (int t = threadIdx.x; t < SS; t += blockDim.x) W[(t / S) * LD + t % S] += M1[(t / S) * LD + t % S];
        __syncthreads();
    } else {
        double v[1][16];
        fetch_small<1>(v, {sm + SL_ZZ * SS});
        put_small(W, v[0], s);
        for (int t = threadIdx.x; t < SS; t += blockDim.x) Bt[(t / S) * LD + t % S] = 0.0;
        __syncthreads();
    }
    inv_chol(W, M0, s);  // M0 = L^-1
    for (int t = threadIdx.x; t < SS; t += blockDim.x) M1[(t / S) * LD + t % S] = M0[(t % S) * LD + t / S];
    __syncthreads();     // M1 = T = L^-T
    store_small(M1, sm + SL_T * SS);
    mm64(Bt, M1, M0, false, 1.0);  // beta T
    store_small(M0, sm + SL_M2 * SS);
}

// ---------------------------------------------------------------- residual and layout kernels
// True residual r = B - A X (FP64, block layout).  rinv == nullptr: partial
// |r|^2, |B|^2 per column -> part[(b * G + g) * 2S]; otherwise R = r * rinv
// per column, stored as VT (the recurrence restarts from a residual of unit
// scale, so its FP32 products neither underflow nor overflow however small
// the true residual of a block is).
// In k_spmm's scheme: a wavefront takes 4 rows (16 lanes x 4 FP64 columns
// per row), the row's indices and values go through the wavefront's LDS
// slice and its 16 gathers of X rows are issued together; rows past hi read
// row hi - 1 and keep nothing.
template <typename VT>
__global__ void __launch_bounds__(256) k_bresid(const BlockDesc * __restrict__ blocks, const double * __restrict__ X,
                                                const double * __restrict__ B, VT * __restrict__ R,
                                                double * __restrict__ part, const double * __restrict__ rinv) {
    __shared__ int sIdx[4][4][SPMM_W];
    __shared__ double sVal[4][4][SPMM_W];
    __shared__ double cn[4][2][S];
    const int b = blockIdx.y;
    const BlockDesc bd = blocks[b];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, rg = lane / 16, j = lane % 16;
    const int n = bd.n;
    const auto A = glob(bd.valsD);
    const auto rowptr = glob(bd.rowptr);
    const auto colidx = glob(bd.colidx);
    const double * Xb = X + bd.off + 4 * j;
    const double * Bb = B + bd.off + 4 * j;
    VT * Rb = R + bd.off + 4 * j;
    const int x = blockIdx.x % NXCD, nw = gridDim.x / NXCD;
    const int lo = (int)((long)n * x / NXCD), hi = (int)((long)n * (x + 1) / NXCD);
    const int stride = 16 * nw;
    int * myIdx = sIdx[wave][rg];
    double * myVal = sVal[wave][rg];
    f64x4 sc = {1.0, 1.0, 1.0, 1.0};
    if (rinv) sc = ld4(rinv + (size_t)b * S + 4 * j);
    f64x4 rr = {0, 0, 0, 0}, bb = {0, 0, 0, 0};
    for (int r0 = lo + 4 * ((blockIdx.x / NXCD) * 4 + wave); r0 < hi; r0 += stride) {
        const int row = r0 + rg, rowc = min(row, hi - 1);
        const int p0 = rowptr[rowc], p1 = rowptr[rowc + 1];
        const f64x4 bv = ld4(Bb + (size_t)rowc * S);
        f64x4 y = {0, 0, 0, 0};
        for (int q0 = p0; q0 < p1; q0 += SPMM_W) {
            const int q = q0 + j;
            const bool in = q < p1;
            const int qc = in ? q : p0;
            const int c = colidx[qc];
            const double v = A[qc];
            myIdx[j] = in ? c : 0;
            myVal[j] = in ? v : 0.0;
            __builtin_amdgcn_wave_barrier();
            f64x4 xv[SPMM_W];
#pragma unroll
            for (int u = 0; u < SPMM_W; ++u) xv[u] = ld4(Xb + (size_t)myIdx[u] * S);
#pragma unroll
            for (int u = 0; u < SPMM_W; ++u) y += myVal[u] * xv[u];
            __builtin_amdgcn_wave_barrier();
        }
        if (row < hi) {
            const f64x4 r = bv - y;
            if (rinv) st4(Rb + (size_t)row * S, __builtin_convertvector(r * sc, V4<VT>));
            rr += r * r;
            bb += bv * bv;
        }
    }
    if (rinv) return;
    // sum over the 4 row groups of the wavefront, then over the wavefronts
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        double a = rr[e], c2 = bb[e];
        a += __shfl_xor(a, 16);
        a += __shfl_xor(a, 32);
        c2 += __shfl_xor(c2, 16);
        c2 += __shfl_xor(c2, 32);
        if (rg == 0) {
            cn[wave][0][4 * j + e] = a;
            cn[wave][1][4 * j + e] = c2;
        }
    }
    __syncthreads();
    if (wave == 0) {
        double * out = part + ((size_t)b * gridDim.x + blockIdx.x) * 2 * S;
        out[lane] = cn[0][0][lane] + cn[1][0][lane] + cn[2][0][lane] + cn[3][0][lane];
        out[S + lane] = cn[0][1][lane] + cn[1][1][lane] + cn[2][1][lane] + cn[3][1][lane];
    }
}

__global__ void k_norm_reduce(int G, const double * __restrict__ part, double * __restrict__ out) {
    const int b = blockIdx.x;
    const int t = threadIdx.x;  // 0 .. 2S-1
    double s = 0.0;
    int g = 0;
    for (; g + 16 <= G; g += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = part[((size_t)b * G + g + u) * 2 * S + t];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; g < G; ++g) s += part[((size_t)b * G + g) * 2 * S + t];
    out[(size_t)b * 2 * S + t] = s;
}

// block layout <-> electrode-major solver layout [k][row][nE] of each problem
// (which = 0: the right-hand sides, 1: the initial solution)
__global__ void k_to_blocks(const BlockDesc * __restrict__ blocks, int which, double * __restrict__ dst) {
    const BlockDesc bd = blocks[blockIdx.y];
    const double * src = which ? bd.Xdst : bd.Bsrc;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)bd.n * S; t += (long)gridDim.x * blockDim.x) {
        const int row = (int)(t / S), c = (int)(t % S);
        dst[bd.off + t] = c < bd.s ? src[(size_t)row * bd.nE + bd.c0 + c] : 0.0;
    }
}

__global__ void k_from_blocks(const BlockDesc * __restrict__ blocks, const double * __restrict__ src) {
    const BlockDesc bd = blocks[blockIdx.y];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)bd.n * S; t += (long)gridDim.x * blockDim.x) {
        const int row = (int)(t / S), c = (int)(t % S);
        if (c < bd.s) bd.Xdst[(size_t)row * bd.nE + bd.c0 + c] = src[bd.off + t];
    }
}

// X64 += scale * Xw ; Xw = 0   over the active blocks (scale: the column's
// residual norm the FP32 recurrence was normalised by)
template <typename VT>
__global__ void k_accumulate(const BlockDesc * __restrict__ blocks, VT * __restrict__ Xw, double * __restrict__ X64,
                             const double * __restrict__ scale) {
    const BlockDesc bd = blocks[blockIdx.y];
    if (!bd.active) return;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)bd.n * S; t += (long)gridDim.x * blockDim.x) {
        const size_t o = bd.off + t;
        X64[o] += scale[(size_t)blockIdx.y * S + t % S] * (double)Xw[o];
        Xw[o] = (VT)0;
    }
}


// --------------------------------------------------------------- driver
struct Plan {
    std::vector<BlockDesc> h;
    DBuf<BlockDesc> d;
};

// workgroups per block for a launch that covers rows in chunks of `rows`,
// keeping the whole grid (over nb blocks) near `cap` workgroups
int wg_per_block(int nmax, int rows, int nb, int cap, int lim) {
    return std::max(1, std::min({lim, div_up(nmax, rows), std::max(1, cap / std::max(nb, 1))}));
}

}  // namespace

void solve_block(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st, bool mixed) {
    std::vector<pgert_solve_stats> sts(1);
    solve_blocks({&c}, o, sts, mixed);
    st = sts[0];
}

void solve_blocks(const std::vector<Ctx *> & cs, const pgert_solve_opts & o, std::vector<pgert_solve_stats> & stats,
                  bool mixed) {
    PG_REQUIRE(!cs.empty(), "block solver: no problems");
    Ctx & c0 = *cs[0];
    const int np = (int)cs.size();
    hipStream_t user = c0.stream;
    const dim3 blk(256);
    // per problem (on the caller's stream): FP32 copy of A, coarse inverse
    std::vector<int> mp(np, 0);
    for (int p = 0; p < np; ++p) {
        Ctx & c = *cs[p];
        PG_REQUIRE(c.nE >= 1 && c.n > 0, "block solver: problem without electrodes or nodes");
        PG_REQUIRE(c.device == c0.device, "block solver: problems on different devices");
        if (o.coarse && c.nAgg > 0) {
            coarse_prepare(c);
            mp[p] = c.nAgg;
        }
    }
    // the smoothed multiplicative cycle needs the coarse space of every problem
    bool smooth = o.smooth > 0;
    for (int p = 0; p < np; ++p) smooth = smooth && mp[p] > 0;
    if (smooth)
        for (int p = 0; p < np; ++p) smooth_prepare(*cs[p]);
    PG_CHECK(hipGetLastError());
    // the solver's own stream, after everything queued so far on the caller's
    if (!c0.solveStream) {
        PG_CHECK(hipStreamCreateWithFlags(&c0.solveStream, hipStreamNonBlocking));
        for (auto & e : c0.solveEv) PG_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    hipStream_t s = c0.solveStream;
    PG_CHECK(hipEventRecord(c0.solveEv[0], user));
    PG_CHECK(hipStreamWaitEvent(s, c0.solveEv[0], 0));
    c0.stream = s;  // kernel-timing events go to the solver stream

    Plan pl;
    long long off = 0, coff = 0;
    int nmax = 0, mmax = 0;
    for (int p = 0; p < np; ++p) {
        Ctx & c = *cs[p];
        for (int k = 0; k < c.nK; ++k)
            for (int e0 = 0; e0 < c.nE; e0 += S) {
                BlockDesc bd{};
                bd.k = k;
                bd.c0 = e0;
                bd.s = std::min(S, c.nE - e0);
                bd.active = 1;
                bd.n = c.n;
                bd.m = mp[p];
                bd.nE = c.nE;
                bd.prob = p;
                bd.off = off;
                bd.coff = coff;
                off += (long long)c.n * S;
                coff += (long long)mp[p] * S;
                bd.rowptr = c.rowptr.p;
                bd.colidx = c.colidx.p;
                bd.valsD = c.valsA.p + (size_t)k * c.nnz;
                bd.dinv = c.dinv.p + (size_t)k * c.n;
                bd.agg = mp[p] ? c.agg.p : nullptr;
                bd.aggPtr = mp[p] ? c.aggPtr.p : nullptr;
                bd.aggNodes = mp[p] ? c.aggNodes.p : nullptr;
                bd.Einv = mp[p] ? c.Einv.p + (size_t)k * mp[p] * mp[p] : nullptr;
                bd.Bsrc = c.B.p + (size_t)k * c.n * c.nE;
                bd.Xdst = c.X.p + (size_t)k * c.n * c.nE;
                if (smooth) {
                    bd.valsAD = c.valsAD.p + (size_t)k * c.nnz;
                    bd.azPtr = c.azPtr.p;
                    bd.azAgg = c.azAgg.p;
                    bd.azVals = c.azVals.p + (size_t)k * c.nAZ;
                    // smoothing interval [b / 11, b], b the Gershgorin bound
                    const double hiB = c.hGersh[k], loB = hiB / 11.0;
                    const double th = 0.5 * (hiB + loB), dl = 0.5 * (hiB - loB);
                    const double sg = th / dl, r0 = 1.0 / sg, r1 = 1.0 / (2.0 * sg - r0);
                    bd.ch[0] = 1.0 / th;
                    bd.ch[1] = r1 * r0 / th;
                    bd.ch[2] = 2.0 * r1 / dl;
                    bd.ch[3] = (1.0 + r1 * r0) / th;
                }
                pl.h.push_back(bd);
                nmax = std::max(nmax, c.n);
                mmax = std::max(mmax, mp[p]);
            }
    }
    pl.d.upload(pl.h, s);
    const int nb = (int)pl.h.size();
    const size_t tot = (size_t)off;
    const bool anyCoarse = mmax > 0;
    // grids: the Gram passes keep ~2048 workgroups (partials of 2-3 64 x 64
    // FP64 matrices each); the streaming kernels ~64k workgroups overall
    const int G = wg_per_block(nmax, 16, nb, 2048, 256);
    // streaming updates: 256 workgroups (one per CU) with a long strip loop
    // each beat 1024-4096 short-lived ones (measured on the 202k-node block:
    // update 60 vs 70-79 us, P update 46 vs 47-53 us)
    const int GU = wg_per_block(nmax, 64, nb, 4096, 256);
    const int GS = NXCD * wg_per_block(nmax, 16 * NXCD, nb, 65536 / NXCD, 256);
    const int GR = NXCD * wg_per_block(nmax, 4 * NXCD, nb, 16384 / NXCD, 128);
    const int GZ = wg_per_block(nmax * (S / 4), 256, nb, 65536, 1 << 20);
    const int GT = wg_per_block(nmax * S, 256, nb, 65536, 1 << 20);
    c0.small.alloc((size_t)nb * NSL * SS);
    PG_CHECK(hipMemsetAsync(c0.small.p, 0, c0.small.bytes(), s));
    c0.part.alloc((size_t)nb * std::max((size_t)G * 3 * SS, (size_t)GU * S));
    DBuf<double> npart, norms;
    npart.alloc((size_t)nb * GR * 2 * S);
    norms.alloc((size_t)nb * 2 * S);
    if (anyCoarse) {
        c0.ycoarse.alloc((size_t)coff);
        c0.ccoarse.alloc((size_t)coff);
        PG_CHECK(hipMemsetAsync(c0.ccoarse.p, 0, c0.ccoarse.bytes(), s));
    }
    const double * ccp = anyCoarse ? c0.ccoarse.p : nullptr;
    PG_CHECK(hipFuncSetAttribute((const void *)k_small_alpha, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)SMALL_LDS));
    PG_CHECK(hipFuncSetAttribute((const void *)k_small_beta<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)SMALL_LDS));
    PG_CHECK(hipFuncSetAttribute((const void *)k_small_beta<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)SMALL_LDS));

    // right-hand sides and the FP64 solution in block layout
    DBuf<double> Bb, Xb;
    Bb.alloc(tot);
    Xb.alloc(tot);
    const dim3 gT(GT, nb);
    hipLaunchKernelGGL(k_to_blocks, gT, blk, 0, s, pl.d.p, 0, Bb.p);
    hipLaunchKernelGGL(k_to_blocks, gT, blk, 0, s, pl.d.p, 1, Xb.p);
    PG_CHECK(hipGetLastError());

    std::vector<double> hn(nb * 2 * S);
    // per column: the factor the recurrence is normalised by, and its inverse
    std::vector<double> hscale(nb * S, 1.0), hinv(nb * S, 1.0);
    DBuf<double> dscale, dinvs;
    dscale.alloc((size_t)nb * S);
    dinvs.alloc((size_t)nb * S);
    // FP64 true residual r = B - A X: norms to the host, then R = r / scale
    // into Rw (VT)
    auto true_resid = [&](auto * Rw) {
        using VT = std::remove_pointer_t<decltype(Rw)>;
        hipLaunchKernelGGL(k_bresid<VT>, dim3(GR, nb), blk, 0, s, pl.d.p, Xb.p, Bb.p, Rw, npart.p,
                           (const double *)nullptr);
        hipLaunchKernelGGL(k_norm_reduce, dim3(nb), dim3(2 * S), 0, s, GR, npart.p, norms.p);
        PG_CHECK(hipGetLastError());
        PG_CHECK(hipMemcpyAsync(hn.data(), norms.p, hn.size() * 8, hipMemcpyDeviceToHost, s));
        PG_CHECK(hipStreamSynchronize(s));
        // one power-of-two factor per block (about its largest column
        // residual): exact scaling, so the columns keep their relative weights
        // and the recurrence its rounding, while its values stay near 1
        for (int b = 0; b < nb; ++b) {
            double mx = 0.0;
            for (int j = 0; j < S; ++j) mx = std::max(mx, hn[(size_t)b * 2 * S + j]);
            const double sc = mx > 0.0 ? std::ldexp(1.0, std::ilogb(std::sqrt(mx))) : 1.0;
            for (int j = 0; j < S; ++j) {
                hscale[b * S + j] = sc;
                hinv[b * S + j] = 1.0 / sc;
            }
        }
        PG_CHECK(hipMemcpyAsync(dscale.p, hscale.data(), hscale.size() * 8, hipMemcpyHostToDevice, s));
        PG_CHECK(hipMemcpyAsync(dinvs.p, hinv.data(), hinv.size() * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_bresid<VT>, dim3(GR, nb), blk, 0, s, pl.d.p, Xb.p, Bb.p, Rw, npart.p,
                           (const double *)dinvs.p);
        PG_CHECK(hipGetLastError());
        PG_CHECK(hipStreamSynchronize(s));  // the host scale vectors are reused
    };
    // worst relative residual of block b from the norms just read; +inf when
    // a norm is not finite (NaN from a bad initial guess, overflow): such a
    // block is given up (deactivated unconverged, the final residual check
    // reports the failure)
    auto block_rel = [&](int b) {
        double w = 0.0;
        for (int j = 0; j < pl.h[b].s; ++j) {
            const double rr = hn[(size_t)b * 2 * S + j], bb = hn[(size_t)b * 2 * S + S + j];
            if (!std::isfinite(rr) || !std::isfinite(bb)) return HUGE_VAL;
            if (bb > 0) w = std::max(w, std::sqrt(rr / bb));
        }
        return w;
    };
    std::vector<int> probIters(np, 0);
    int refine = 0, it = 0;

    auto run = [&](auto vtag, auto atag) {
        using VT = decltype(vtag);
        using AT = decltype(atag);
        constexpr bool fp32 = std::is_same<VT, float>::value;
        // compulsory HBM bytes of one launch of the timed kernel (all blocks
        // active): matrix values and indices once, each vector once
        const double vec = (double)tot * sizeof(VT);
        if (c0.kTiming) {
            double mat = 0.0;
            for (int p = 0; p < np; ++p)
                mat += (double)cs[p]->nK * ((double)cs[p]->nnz * (sizeof(AT) + 4) + 4.0 * (cs[p]->n + 1));
            c0.kBytesPerLaunch = 0.0;  // latency-bound kernels: no algorithmic byte count
            switch (c0.kSelect) {
            case PGERT_TIME_SPMM: c0.kBytesPerLaunch = mat + 2.0 * vec; break;
            case PGERT_TIME_UPDATE: c0.kBytesPerLaunch = 3.0 * vec; break;
            case PGERT_TIME_GRAM_Z: c0.kBytesPerLaunch = 3.0 * vec; break;
            case PGERT_TIME_PUPDATE: c0.kBytesPerLaunch = 3.0 * vec; break;
            case PGERT_TIME_GRAM_PQ: c0.kBytesPerLaunch = 3.0 * vec; break;
            case PGERT_TIME_ZFORM:
                // smoothing: the post-smoothing product gathers r2 and reads r, s0 of its row
                c0.kBytesPerLaunch = smooth ? mat + 4.0 * vec + 12.0 * (double)nmax : 2.0 * vec + 12.0 * (double)nmax;
                break;
            case PGERT_TIME_RESTRICT: c0.kBytesPerLaunch = vec; break;
            case PGERT_TIME_SMOOTH: c0.kBytesPerLaunch = mat + 3.0 * vec; break;  // gather, row, 1-2 writes
            case PGERT_TIME_COARSE_RESID: {
                double az = 0.0;
                for (int p = 0; p < np; ++p) az += (double)cs[p]->nK * (12.0 * cs[p]->nAZ + 4.0 * (cs[p]->n + 1));
                c0.kBytesPerLaunch = 2.0 * vec + az;
                break;
            }
            default: break;
            }
        }
        auto tb = [&](int id) {
            if (c0.kSelect == id) c0.kev_begin();
        };
        auto te = [&](int id) {
            if (c0.kSelect == id) c0.kev_end();
        };
        DBuf<VT> Xw, Rw, Pw, Qw, Zw, S0w, R1w, Ww;
        Zw.alloc(tot);
        if (smooth) {
            S0w.alloc(tot);
            R1w.alloc(tot);
            Ww.alloc(tot);
        }
        Rw.alloc(tot);
        Pw.alloc(tot);
        Qw.alloc(tot);
        // the recurrence's correction in normalised units (see true_resid),
        // added into the FP64 solution at every residual replacement
        Xw.alloc(tot);
        PG_CHECK(hipMemsetAsync(Xw.p, 0, Xw.bytes(), s));
        VT * Xp = Xw.p;
        PG_CHECK(hipMemsetAsync(Pw.p, 0, Pw.bytes(), s));
        PG_CHECK(hipMemsetAsync(Qw.p, 0, Qw.bytes(), s));
        true_resid(Rw.p);
        std::vector<double> bnorm2(nb * S);
        for (int b = 0; b < nb; ++b)
            for (int j = 0; j < S; ++j) bnorm2[b * S + j] = hn[(size_t)b * 2 * S + S + j];
        double lastRel = 0.0;
        for (int b = 0; b < nb; ++b) {
            const double r = block_rel(b);
            if (r <= o.rtol || r == HUGE_VAL) pl.h[b].active = 0;
            else lastRel = std::max(lastRel, r);
        }
        PG_CHECK(hipMemcpyAsync(pl.d.p, pl.h.data(), pl.h.size() * sizeof(BlockDesc), hipMemcpyHostToDevice, s));
        int lastReplace = 0;
        const dim3 gG(G, nb), gU(GU, nb), gS(GS, nb), gZ(GZ, nb);
        auto z_grams = [&]() {
            if (smooth) {
                tb(PGERT_TIME_SMOOTH);
                hipLaunchKernelGGL((k_smooth<VT, MODE_S0>), gS, blk, 0, s, pl.d.p, Rw.p, Rw.p, Rw.p, ccp, S0w.p,
                                   Ww.p);
                te(PGERT_TIME_SMOOTH);
                tb(PGERT_TIME_SMOOTH);
                hipLaunchKernelGGL((k_smooth<VT, MODE_S1>), gS, blk, 0, s, pl.d.p, Ww.p, S0w.p, Rw.p, ccp, R1w.p,
                                   R1w.p);
                te(PGERT_TIME_SMOOTH);
                tb(PGERT_TIME_RESTRICT);
                hipLaunchKernelGGL(k_restrict<VT>, dim3(mmax, nb), blk, 0, s, pl.d.p, R1w.p, c0.ycoarse.p);
                te(PGERT_TIME_RESTRICT);
                tb(PGERT_TIME_COARSE);
                hipLaunchKernelGGL(k_coarse_apply, dim3(div_up(mmax, 16), 4, nb), dim3(64 * CA_WAVES), 0, s, pl.d.p,
                                   c0.ycoarse.p, c0.ccoarse.p);
                te(PGERT_TIME_COARSE);
                tb(PGERT_TIME_COARSE_RESID);
                hipLaunchKernelGGL(k_coarse_resid<VT>, gZ, blk, 0, s, pl.d.p, ccp, R1w.p);
                te(PGERT_TIME_COARSE_RESID);
                tb(PGERT_TIME_ZFORM);
                hipLaunchKernelGGL((k_smooth<VT, MODE_SZ>), gS, blk, 0, s, pl.d.p, R1w.p, S0w.p, Rw.p, ccp, Zw.p,
                                   Zw.p);
                te(PGERT_TIME_ZFORM);
            } else {
            if (anyCoarse) {
                tb(PGERT_TIME_RESTRICT);
                hipLaunchKernelGGL(k_restrict<VT>, dim3(mmax, nb), blk, 0, s, pl.d.p, Rw.p, c0.ycoarse.p);
                te(PGERT_TIME_RESTRICT);
                tb(PGERT_TIME_COARSE);
                hipLaunchKernelGGL(k_coarse_apply, dim3(div_up(mmax, 16), 4, nb), dim3(64 * CA_WAVES), 0, s, pl.d.p,
                                   c0.ycoarse.p,
                                   c0.ccoarse.p);
                te(PGERT_TIME_COARSE);
            }
            tb(PGERT_TIME_ZFORM);
            hipLaunchKernelGGL(k_zform<VT>, gZ, blk, 0, s, pl.d.p, Rw.p, ccp, Zw.p);
            te(PGERT_TIME_ZFORM);
            }
            tb(PGERT_TIME_GRAM_Z);
            hipLaunchKernelGGL((k_gram<VT, MODE_Z>), gG, dim3(GRAM_THREADS), 0, s, pl.d.p, Pw.p, Qw.p, Zw.p,
                               c0.part.p);
            te(PGERT_TIME_GRAM_Z);
            tb(PGERT_TIME_REDUCE);
            hipLaunchKernelGGL(k_part_reduce, dim3(3 * SS / 32, nb), blk, 0, s, pl.d.p, G, 3 * SS, c0.part.p,
                               c0.small.p, SL_QZ * SS);
            te(PGERT_TIME_REDUCE);
        };
        // one iteration; rr: reduce the recursive residual norms (last of a chunk)
        auto iteration = [&](bool rr) {
            tb(PGERT_TIME_SPMM);
            hipLaunchKernelGGL((k_spmm<VT, AT>), gS, blk, 0, s, pl.d.p, Pw.p, Qw.p);
            te(PGERT_TIME_SPMM);
            tb(PGERT_TIME_GRAM_PQ);
            hipLaunchKernelGGL((k_gram<VT, MODE_PQR>), gG, dim3(GRAM_THREADS), 0, s, pl.d.p, Pw.p, Qw.p, Rw.p,
                               c0.part.p);
            te(PGERT_TIME_GRAM_PQ);
            tb(PGERT_TIME_REDUCE);
            hipLaunchKernelGGL(k_part_reduce, dim3(2 * SS / 32, nb), blk, 0, s, pl.d.p, G, 2 * SS, c0.part.p,
                               c0.small.p, SL_G1 * SS);
            te(PGERT_TIME_REDUCE);
            tb(PGERT_TIME_SMALL_ALPHA);
            hipLaunchKernelGGL(k_small_alpha, dim3(nb), blk, SMALL_LDS, s, pl.d.p, c0.small.p);
            te(PGERT_TIME_SMALL_ALPHA);
            tb(PGERT_TIME_UPDATE);
            hipLaunchKernelGGL(k_update<VT>, gU, blk, 0, s, pl.d.p, c0.small.p, Qw.p, Rw.p, c0.part.p);
            te(PGERT_TIME_UPDATE);
            if (rr)
                hipLaunchKernelGGL(k_part_reduce, dim3(S / 32, nb), blk, 0, s, pl.d.p, GU, S, c0.part.p, c0.small.p,
                                   SL_RR * SS);
            z_grams();
            tb(PGERT_TIME_SMALL_BETA);
            hipLaunchKernelGGL(k_small_beta<VT>, dim3(1 + GU, nb), blk, SMALL_LDS, s, pl.d.p, c0.small.p, 0, Pw.p,
                               Xp);
            te(PGERT_TIME_SMALL_BETA);
            tb(PGERT_TIME_PUPDATE);
            hipLaunchKernelGGL(k_pupdate<VT>, gU, blk, 0, s, pl.d.p, c0.small.p, Zw.p, Pw.p);
            te(PGERT_TIME_PUPDATE);
        };
        auto start_block = [&]() {
            // P = 0, Q = 0: Z = M^-1 R and Z'Z only; then P = Z chol(Z'Z)^-T
            PG_CHECK(hipMemsetAsync(c0.small.p, 0, c0.small.bytes(), s));
            z_grams();
            hipLaunchKernelGGL(k_small_beta<VT>, dim3(1, nb), blk, SMALL_LDS, s, pl.d.p, c0.small.p, 1,
                               (const VT *)nullptr, (VT *)nullptr);
            hipLaunchKernelGGL(k_pupdate<VT>, gU, blk, 0, s, pl.d.p, c0.small.p, Zw.p, Pw.p);
            PG_CHECK(hipGetLastError());
        };
        bool done = true;
        for (int b = 0; b < nb; ++b)
            if (pl.h[b].active) done = false;
        if (!done) start_block();
        // a chunk of check_every iterations is captured once and replayed
        // (no memsets inside; the kernels read the block flags from memory)
        hipGraphExec_t gexec = nullptr;
        const int ce = o.check_every;
        if (!done && !o.no_graph && !c0.kTiming && ce <= o.max_iter) {
            hipGraph_t g = nullptr;
            PG_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < ce; ++j) iteration(j == ce - 1);
            PG_CHECK(hipStreamEndCapture(s, &g));
            PG_CHECK(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
            PG_CHECK(hipGraphDestroy(g));
        }
        std::vector<double> hs((size_t)nb * S);
        while (!done && it < o.max_iter) {
            const int chunk = std::min(ce, o.max_iter - it);
            if (gexec && chunk == ce) {
                PG_CHECK(hipGraphLaunch(gexec, s));
            } else {
                for (int j = 0; j < chunk; ++j) iteration(j == chunk - 1);
            }
            PG_CHECK(hipGetLastError());
            it += chunk;
            // recursive residual norms of the last k_update
            for (int b = 0; b < nb; ++b)
                PG_CHECK(hipMemcpyAsync(hs.data() + (size_t)b * S, c0.small.p + ((size_t)b * NSL + SL_RR) * SS, S * 8,
                                        hipMemcpyDeviceToHost, s));
            PG_CHECK(hipStreamSynchronize(s));
            c0.kev_collect();
            for (int b = 0; b < nb; ++b)
                if (pl.h[b].active) probIters[pl.h[b].prob] = it;
            double worst = 0;
            bool anyAtTarget = false;
            for (int b = 0; b < nb; ++b) {
                if (!pl.h[b].active) continue;
                double wb = 0.0;
                for (int j = 0; j < pl.h[b].s; ++j) {
                    const double bb = bnorm2[b * S + j], rr = hs[(size_t)b * S + j];
                    if (!std::isfinite(rr)) wb = -1.0;  // check the true residual now
                    if (bb > 0 && wb >= 0.0) wb = std::max(wb, std::sqrt(rr / bb) * hscale[b * S + j]);
                }
                worst = std::max(worst, wb);
                if (wb <= o.rtol) anyAtTarget = true;
            }
            // Residual replacement: the true FP64 residual b - A x decides which
            // blocks are done.  FP32 recursions are replaced when they have
            // gained a decade, reached the target or run 256 iterations (every
            // two decades instead: 216 iterations instead of 168 on the 3D
            // headline, the restarts keep the FP32 recursion on course); FP64
            // ones when a block's recursive residual reaches the target.
            const bool replaceNow =
                fp32 ? (worst < 0.1 * lastRel || anyAtTarget || it - lastReplace >= 256) : anyAtTarget;
            bool changed = false;
            if (replaceNow) {
                lastReplace = it;
                hipLaunchKernelGGL(k_accumulate<VT>, gT, blk, 0, s, pl.d.p, Xw.p, Xb.p, (const double *)dscale.p);
                true_resid(Rw.p);
                ++refine;
                lastRel = 0.0;
                for (int b = 0; b < nb; ++b) {
                    if (!pl.h[b].active) continue;
                    const double tb2 = block_rel(b);
                    if (tb2 <= o.rtol || tb2 == HUGE_VAL) {
                        pl.h[b].active = 0;
                        changed = true;
                    } else {
                        lastRel = std::max(lastRel, tb2);
                    }
                }
            }
            done = true;
            for (int b = 0; b < nb; ++b)
                if (pl.h[b].active) done = false;
            if (changed && !done)
                PG_CHECK(hipMemcpyAsync(pl.d.p, pl.h.data(), pl.h.size() * sizeof(BlockDesc), hipMemcpyHostToDevice, s));
        }
        if (gexec) PG_CHECK(hipGraphExecDestroy(gexec));
        hipLaunchKernelGGL(k_accumulate<VT>, gT, blk, 0, s, pl.d.p, Xw.p, Xb.p, (const double *)dscale.p);
        PG_CHECK(hipGetLastError());
        PG_CHECK(hipStreamSynchronize(s));  // work buffers are released on return
    };

    // the mixed solver keeps A in FP64 inside the FP32 SpMM (products and
    // row sums in FP64, Q rounded once): rounding A itself to FP32 would make
    // the inner recurrence solve a different operator, and the FP64 residual
    // replacement would then stall once cond(A) * 2^-24 nears 1 (the
    // near-singular small-wavenumber 2.5D systems)
    if (mixed)
        run(float(0), double(0));
    else
        run(double(0), double(0));
    // every block back to active for the layout copy
    for (auto & bd : pl.h) bd.active = 1;
    PG_CHECK(hipMemcpyAsync(pl.d.p, pl.h.data(), pl.h.size() * sizeof(BlockDesc), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_from_blocks, gT, blk, 0, s, pl.d.p, Xb.p);
    PG_CHECK(hipGetLastError());
    PG_CHECK(hipEventRecord(c0.solveEv[1], s));
    c0.stream = user;
    PG_CHECK(hipStreamWaitEvent(user, c0.solveEv[1], 0));
    PG_CHECK(hipStreamSynchronize(s));  // the descriptor table is released on return
    stats.assign(np, pgert_solve_stats{});
    for (int p = 0; p < np; ++p) {
        pgert_solve_stats & st = stats[p];
        st.iterations = probIters[p];
        st.max_iterations = it;
        st.refinements = refine;
        st.max_rel_residual = residual_check(*cs[p]);
        st.converged = st.max_rel_residual <= o.rtol * 1.0001 ? 1 : 0;
    }
}

}  // namespace pgert

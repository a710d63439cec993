// Gauss-Newton step on the GPU: the regularised, transformed CGLS of
// solveCGLSCDWWhtrans (core/src/solver.cpp:27-224), called by
// RInversion::oneStep (core/src/inversion.cpp:308).  The Jacobian stays where
// pgert_jacobian wrote it (HBM, FP32 or FP64, row-major data x model); each
// iteration reads it twice, once as J u (a workgroup per row, coalesced along
// the row) and once as J' v (a thread per column, rows streamed), both
// HBM-bound.  The constraint matrix C and its transpose are CSR (set once per
// mesh, pgert_set_constraints).  Vectors and scalars are FP64; products with
// an FP32 Jacobian accumulate in FP64.  Reductions are fixed-order two-stage
// sums (deterministic).
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

namespace pgert {

namespace {

constexpr int NRED = 512;  // partial sums of the dot products

// out[i] = (J[i,:] . u) * s1[i] * s2[i]   (s1, s2 may be null); a workgroup
// per group of JR rows, so each u[c] load serves JR Jacobian rows
constexpr int JR = 4;
template <typename JT>
__global__ void __launch_bounds__(256) k_jv(const JT * __restrict__ J, int D, long P, const double * __restrict__ u,
                                            const double * __restrict__ s1, const double * __restrict__ s2,
                                            double * __restrict__ out) {
    __shared__ double red[JR][4];
    for (int i0 = blockIdx.x * JR; i0 < D; i0 += gridDim.x * JR) {
        const JT * rows[JR];
#pragma unroll
        for (int k = 0; k < JR; ++k) rows[k] = J + (size_t)min(i0 + k, D - 1) * P;
        double acc[JR][2] = {};
        long c = threadIdx.x;
        for (; c + 3 * 256 < P; c += 4 * 256) {
            double uv[4], jv[JR][4];
#pragma unroll
            for (int m = 0; m < 4; ++m) uv[m] = u[c + m * 256];
#pragma unroll
            for (int k = 0; k < JR; ++k)
#pragma unroll
                for (int m = 0; m < 4; ++m) jv[k][m] = (double)rows[k][c + m * 256];
#pragma unroll
            for (int k = 0; k < JR; ++k)
#pragma unroll
                for (int m = 0; m < 4; ++m) acc[k][m & 1] += jv[k][m] * uv[m];
        }
        for (; c < P; c += 256) {
            const double uv = u[c];
#pragma unroll
            for (int k = 0; k < JR; ++k) acc[k][0] += (double)rows[k][c] * uv;
        }
#pragma unroll
        for (int k = 0; k < JR; ++k) {
            double a = acc[k][0] + acc[k][1];
            for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
            if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = a;
        }
        __syncthreads();
        if (threadIdx.x < JR && i0 + (int)threadIdx.x < D) {
            const int k = threadIdx.x, i = i0 + k;
            double r = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
            if (s1) r *= s1[i];
            if (s2) r *= s2[i];
            out[i] = r;
        }
        __syncthreads();
    }
}

// r[c] = (J[:,c] . v) / tm[c] - lambda wm[c] ct[c] - cdx[c]; a thread per column
template <typename JT>
__global__ void __launch_bounds__(256) k_jtv(const JT * __restrict__ J, int D, long P, const double * __restrict__ v,
                                             const double * __restrict__ tm, const double * __restrict__ wm,
                                             const double * __restrict__ ct, const double * __restrict__ cdx,
                                             double lambda, double * __restrict__ r) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    double acc[4] = {};
    int i = 0;
    for (; i + 8 <= D; i += 8) {
        double x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = (double)J[(size_t)(i + k) * P + c];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k & 3] += v[i + k] * x[k];
    }
    for (; i < D; ++i) acc[0] += v[i] * (double)J[(size_t)i * P + c];
    double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) / tm[c];
    if (ct) s -= lambda * wm[c] * ct[c];
    if (cdx) s -= cdx[c];
    r[c] = s;
}

// y[row] = post[row] * rnd( sum_q vals[q] x[col_q] * pre[col_q] ), a thread per row;
// rnd rounds to multiples of 1e-10 when `round` is set (RVector::round)
__global__ void k_csr(int rows, const int * __restrict__ rp, const int * __restrict__ ci,
                      const double * __restrict__ va, const double * __restrict__ x, const double * __restrict__ pre,
                      const double * __restrict__ post, int round, double * __restrict__ y) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    double s = 0.0;
    for (int q = rp[row]; q < rp[row + 1]; ++q) {
        const int c = ci[q];
        s += va[q] * (pre ? x[c] * pre[c] : x[c]);
    }
    if (round) s = rint(s / 1e-10) * 1e-10;
    y[row] = post ? s * post[row] : s;
}

// y = a op b elementwise helpers
__global__ void k_div(long n, const double * __restrict__ a, const double * __restrict__ b, double * __restrict__ y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = a[t] / b[t];
}
__global__ void k_mul3(long n, const double * __restrict__ a, const double * __restrict__ b,
                       const double * __restrict__ c, double * __restrict__ y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = a[t] * b[t] * (c ? c[t] : 1.0);
}
// z = (b - y * td) * dW
__global__ void k_resid0(int n, const double * __restrict__ b, const double * __restrict__ y,
                         const double * __restrict__ td, const double * __restrict__ dW, double * __restrict__ z) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) z[t] = (b[t] - y[t] * td[t]) * dW[t];
}
// y += alpha x
__global__ void k_axpy(long n, double alpha, const double * __restrict__ x, double * __restrict__ y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += alpha * x[t];
}
// p = r + beta p
__global__ void k_xpby(long n, const double * __restrict__ r, double beta, double * __restrict__ p) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = r[t] + beta * p[t];
}
// wcwc[i] = wc[i]^2 (or wc[i] * rough[i] with rough)
__global__ void k_sq(int n, const double * __restrict__ wc, const double * __restrict__ rough, double * __restrict__ y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = wc[t] * (rough ? rough[t] : wc[t]);
}
// y = lambda * x * wm
__global__ void k_lscale(long n, double lambda, const double * __restrict__ x, const double * __restrict__ wm,
                         double * __restrict__ y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = x[t] * wm[t] * lambda;
}

// partial dot products: part[which * NRED + blockIdx.x]
__global__ void __launch_bounds__(256) k_dot(long n, const double * __restrict__ a, const double * __restrict__ b,
                                             double * __restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        s += a[t] * b[t];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ void __launch_bounds__(64) k_dot_sum(int np, const double * __restrict__ part, double * __restrict__ out) {
    double s = 0.0;
    for (int g = threadIdx.x; g < np; g += 64) s += part[g];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (threadIdx.x == 0) *out = s;
}

// work vectors of one CGLS run (device pointers)
struct Work {
    double *u, *q, *z, *zd, *r, *p, *cdx, *wcwc, *wcr, *g, *ct, *cv, *part, *dots;
};

// carve the work vectors for D data, P parameters and nC constraints out of
// one buffer of carved_doubles() doubles, each vector starting on a 256-byte
// boundary (the buffer is 256-byte aligned)
Work carve(double * base, int D, int P, int nC) {
    const size_t c = (size_t)std::max(nC, 1);
    size_t off = 0;
    auto take = [&](size_t n) {
        double * q = base + off;
        off += (n + 31) / 32 * 32;
        return q;
    };
    Work w;
    w.u = take(P);
    w.r = take(P);
    w.p = take(P);
    w.cdx = take(P);
    w.ct = take(P);
    w.q = take(D);
    w.z = take(D);
    w.zd = take(D);
    w.wcwc = take(c);
    w.wcr = take(c);
    w.g = take(c);
    w.cv = take(c);
    w.part = take(2 * NRED);
    w.dots = take(2);
    return w;
}
size_t carved_doubles(int D, int P, int nC) {
    const size_t c = (size_t)std::max(nC, 1);
    auto r = [](size_t n) { return (n + 31) / 32 * 32; };
    return 5 * r(P) + 3 * r(D) + 4 * r(c) + r(2 * NRED) + r(2);
}

// Workspace of pgert_cgls_noalloc: one array per device, held until the
// process ends.  In host code `scratch` only names the symbol; its device
// address comes from hipGetSymbolAddress, once per device, with that device
// current.
constexpr size_t SCRATCH_DOUBLES = size_t(8) << 20;  // 64 MiB
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
std::mutex scratchMu;
double * scratchPtr[256] = {};

}  // namespace

void cgls_scratch_init(int device) {
    PG_REQUIRE(device >= 0 && device < 256, "device index out of range");
    std::lock_guard<std::mutex> lk(scratchMu);
    if (scratchPtr[device]) return;
    void * p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    PG_REQUIRE(e == hipSuccess && p, std::string("workspace of the allocation-free CGLS: hipGetSymbolAddress: ") +
                                         hipGetErrorString(e));
    scratchPtr[device] = (double *)p;
}

void set_constraints(Ctx & c, int nConst, int nModel, const int * rp, const int * ci, const double * va) {
    PG_REQUIRE(nConst >= 0 && nModel > 0 && rp && (rp[nConst] == 0 || (ci && va)), "pgert_set_constraints: bad CSR");
    const int nnz = rp[nConst];
    for (int i = 0; i < nnz; ++i) PG_REQUIRE(ci[i] >= 0 && ci[i] < nModel, "pgert_set_constraints: column out of range");
    // transpose on the host (counting sort by column): C' as CSR
    std::vector<int> tp(nModel + 1, 0), tc(nnz);
    std::vector<double> tv(nnz);
    for (int i = 0; i < nnz; ++i) ++tp[ci[i] + 1];
    for (int j = 0; j < nModel; ++j) tp[j + 1] += tp[j];
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int row = 0; row < nConst; ++row)
        for (int q = rp[row]; q < rp[row + 1]; ++q) {
            const int k = fill[ci[q]]++;
            tc[k] = row;
            tv[k] = va[q];
        }
    c.nConst = nConst;
    c.cCols = nModel;
    c.cRowptr.upload(rp, nConst + 1, c.stream);
    c.cColidx.upload(ci, nnz, c.stream);
    c.cVals.upload(va, nnz, c.stream);
    c.ctRowptr.upload(tp, c.stream);
    c.ctColidx.upload(tc, c.stream);
    c.ctVals.upload(tv, c.stream);
    PG_CHECK(hipStreamSynchronize(c.stream));
}

// the CGLS iteration on the work vectors w (no allocation in here)
template <typename JT>
static int run_cgls(Ctx & c, const Work & w, const JT * J, int D, int P, const double * dW, const double * b,
                    double * x, const double * wc, const double * wm, const double * tm, const double * td,
                    double lambda, const double * rough, int maxIter, double tol) {
    hipStream_t s = c.stream;
    const int nC = c.nConst;
    const dim3 blk(256);
    auto G = [](long n) { return dim3((unsigned)std::max(1L, (n + 255) / 256)); };
    const int gjv = std::min((D + JR - 1) / JR, 4096);
    const int * crp = c.cRowptr.p, *cci = c.cColidx.p, *trp = c.ctRowptr.p, *tci = c.ctColidx.p;
    const double *cva = c.cVals.p, *tva = c.ctVals.p;
    auto launch_jv = [&](const double * u, const double * s1, const double * s2, double * out) {
        hipLaunchKernelGGL(k_jv<JT>, dim3(gjv), blk, 0, s, J, D, (long)P, u, s1, s2, out);
    };
    auto launch_jtv = [&](const double * v, const double * ct, const double * cdx, double * r) {
        hipLaunchKernelGGL(k_jtv<JT>, G(P), blk, 0, s, J, D, (long)P, v, tm, wm, ct, cdx, lambda, r);
    };
    // C' (wc^2 C (wm x)) into w.ct (zero without constraints)
    auto ctc = [&](const double * xv) {
        if (nC == 0) {
            PG_CHECK(hipMemsetAsync(w.ct, 0, (size_t)P * sizeof(double), s));
            return;
        }
        hipLaunchKernelGGL(k_csr, G(nC), blk, 0, s, nC, crp, cci, cva, xv, wm, w.wcwc, 0, w.g);
        hipLaunchKernelGGL(k_csr, G(P), blk, 0, s, P, trp, tci, tva, w.g, (const double *)nullptr,
                           (const double *)nullptr, 0, w.ct);
    };
    auto dot = [&](const double * a, const double * bb, long n, int which) {
        hipLaunchKernelGGL(k_dot, dim3(NRED), blk, 0, s, n, a, bb, w.part + which * NRED);
        hipLaunchKernelGGL(k_dot_sum, dim3(1), dim3(64), 0, s, NRED, w.part + which * NRED, w.dots + which);
    };
    auto read_dots = [&](double * h2) {
        PG_CHECK(hipMemcpyAsync(h2, w.dots, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        PG_CHECK(hipStreamSynchronize(s));
    };
    if (nC > 0) {
        hipLaunchKernelGGL(k_sq, G(nC), blk, 0, s, nC, wc, (const double *)nullptr, w.wcwc);
        hipLaunchKernelGGL(k_sq, G(nC), blk, 0, s, nC, wc, rough, w.wcr);
        // cdx = lambda wm C'(wc rough)                                   (solver.cpp:62)
        hipLaunchKernelGGL(k_csr, G(P), blk, 0, s, P, trp, tci, tva, w.wcr, (const double *)nullptr,
                           (const double *)nullptr, 0, w.ct);
        hipLaunchKernelGGL(k_lscale, G(P), blk, 0, s, (long)P, lambda, w.ct, wm, w.cdx);
    } else {
        PG_CHECK(hipMemsetAsync(w.cdx, 0, (size_t)P * sizeof(double), s));
    }
    // z = (b - S (x / tm) td) dW                                          (:63)
    hipLaunchKernelGGL(k_div, G(P), blk, 0, s, (long)P, x, tm, w.u);
    launch_jv(w.u, nullptr, nullptr, w.q);
    hipLaunchKernelGGL(k_resid0, G(D), blk, 0, s, D, b, w.q, td, dW, w.z);
    double h2[2];
    if (tol < 0.0) {
        // r0 = S'(b dW dW td) / tm - cdx, accuracy = max(TOLERANCE, 1e-8 r0.r0)   (:67, :98)
        hipLaunchKernelGGL(k_mul3, G(D), blk, 0, s, (long)D, b, dW, dW, w.zd);
        hipLaunchKernelGGL(k_mul3, G(D), blk, 0, s, (long)D, w.zd, td, (const double *)nullptr, w.zd);
        launch_jtv(w.zd, nullptr, w.cdx, w.r);
        dot(w.r, w.r, P, 0);
        read_dots(h2);
        tol = std::max(1e-12, 1e-8 * h2[0]);
    }
    // p = S'(z dW td) / tm - lambda wm C'(wc^2 C (wm x)) - cdx; r = p    (:64-66, :99)
    hipLaunchKernelGGL(k_mul3, G(D), blk, 0, s, (long)D, w.z, dW, td, w.zd);
    ctc(x);
    launch_jtv(w.zd, nC ? w.ct : nullptr, w.cdx, w.p);
    dot(w.p, w.p, P, 0);
    read_dots(h2);
    double normR2 = h2[0];
    int count = 0;
    while (count < maxIter && normR2 > tol) {
        ++count;
        // q = S (p / tm) dW td                                            (:123-124)
        hipLaunchKernelGGL(k_div, G(P), blk, 0, s, (long)P, w.p, tm, w.u);
        launch_jv(w.u, dW, td, w.q);
        dot(w.q, w.q, D, 0);
        // wcp = wc round(C (p wm), 1e-10)                                 (:131-138)
        if (nC > 0) {
            hipLaunchKernelGGL(k_csr, G(nC), blk, 0, s, nC, crp, cci, cva, w.p, wm, wc, 1, w.cv);
            dot(w.cv, w.cv, nC, 1);
        } else {
            PG_CHECK(hipMemsetAsync(w.dots + 1, 0, sizeof(double), s));
        }
        read_dots(h2);
        const double aQ = h2[0] + lambda * h2[1];                    // (:159)
        const double alpha = normR2 / aQ;                            // (:165)
        hipLaunchKernelGGL(k_axpy, G(P), blk, 0, s, (long)P, alpha, w.p, x);       // (:168)
        hipLaunchKernelGGL(k_axpy, G(D), blk, 0, s, (long)D, -alpha, w.q, w.z);  // (:175)
        // r = S'(z dW td) / tm - lambda wm C'(wc^2 C (wm x)) - cdx       (:179-185)
        hipLaunchKernelGGL(k_mul3, G(D), blk, 0, s, (long)D, w.z, dW, td, w.zd);
        ctc(x);
        launch_jtv(w.zd, nC ? w.ct : nullptr, w.cdx, w.r);
        dot(w.r, w.r, P, 0);
        read_dots(h2);
        const double old = normR2;
        normR2 = h2[0];                                              // (:190)
        hipLaunchKernelGGL(k_xpby, G(P), blk, 0, s, (long)P, w.r, normR2 / old, w.p);  // (:191-194)
    }
    PG_CHECK(hipGetLastError());
    PG_CHECK(hipStreamSynchronize(s));
    c.cglsNormR2 = normR2;
    return count;
}

int cgls(Ctx & c, const void * J, int dtype, int D, int P, const double * dW, const double * b, double * x,
         const double * wc, const double * wm, const double * tm, const double * td, double lambda,
         const double * rough, int maxIter, double tol) {
    PG_REQUIRE(c.nConst == 0 || c.cCols == P, "pgert_cgls: constraints have another model size");
    DBuf<double> buf;
    buf.alloc(carved_doubles(D, P, c.nConst));
    const Work w = carve(buf.p, D, P, c.nConst);
    if (dtype == 0)
        return run_cgls(c, w, (const double *)J, D, P, dW, b, x, wc, wm, tm, td, lambda, rough, maxIter, tol);
    return run_cgls(c, w, (const float *)J, D, P, dW, b, x, wc, wm, tm, td, lambda, rough, maxIter, tol);
}

// pgert_cgls with its work vectors in the device's scratch array: no
// allocation.  Calls on one device share the array (they must not overlap).
int cgls_noalloc(Ctx & c, const void * J, int dtype, int D, int P, const double * dW, const double * b, double * x,
                 const double * wc, const double * wm, const double * tm, const double * td, double lambda,
                 const double * rough, int maxIter, double tol) {
    PG_REQUIRE(c.nConst == 0 || c.cCols == P, "pgert_cgls_noalloc: constraints have another model size");
    const size_t need = carved_doubles(D, P, c.nConst);
    PG_REQUIRE(need <= SCRATCH_DOUBLES,
               "pgert_cgls_noalloc: the problem needs " + std::to_string(need * 8) + " bytes of workspace (5 n_model + "
               "3 n_data + 4 n_constraints doubles), more than the " + std::to_string(SCRATCH_DOUBLES * 8) +
               "-byte scratch array; use pgert_cgls");
    double * base;
    {
        std::lock_guard<std::mutex> lk(scratchMu);
        base = scratchPtr[c.device];
    }
    PG_REQUIRE(base, "pgert_cgls_noalloc: no workspace address for this device (context not created by pgert_create)");
    const Work w = carve(base, D, P, c.nConst);
    if (dtype == 0)
        return run_cgls(c, w, (const double *)J, D, P, dW, b, x, wc, wm, tm, td, lambda, rough, maxIter, tol);
    return run_cgls(c, w, (const float *)J, D, P, dW, b, x, wc, wm, tm, td, lambda, rough, maxIter, tol);
}

}  // namespace pgert

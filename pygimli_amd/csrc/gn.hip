// Gauss-Newton loop support: the logarithmic transforms of RInversion
// (TransLogLU / TransLog, core/src/trans.h:241-322, trans.cpp:24-63) and the
// objective its line search scans (RInversion::linesearch,
// core/src/inversion.h:615-640; getPhiD / getPhiM, inversion.cpp:84-117),
// evaluated on the device for a set of step lengths.
//
// Transforms are elementwise; the reference caps every element below the
// lower bound (above the upper bound) whenever any element is, which is the
// same as capping each element on its own.  The objective at step tau is
//   phiD = sum_i ((tD(d_i) - tD(tD.update(r_i, tau dd_i))) / e_i)^2,
//   phiM = sum_r (cw_r sum_j C_rj tM(tM.update(m_j, tau dm_j)) mw_j)^2,
// with the data transform log (TransLog(0)) and the model transform
// TransLogLU(lower, upper); tau = 0 uses r and m as they are (the reference's
// getPhi() of the current state).  Sums: per-workgroup partials, then one
// fixed-order pass (deterministic).
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace pgert {

namespace {

constexpr double TOL = 1e-12;       // gimli.h TOLERANCE
constexpr double TRANSTOL = 1e-8;   // trans.h TRANSTOL

struct LogLU {
    double lo, up;
    __device__ double rangify(double a) const {
        double v = fmax(a, lo * (1.0 + TRANSTOL));
        if (fabs(up) >= TOL) v = fmin(v, up * (1.0 - TRANSTOL));
        return v;
    }
    __device__ double trans(double a) const {
        const double v = rangify(a);
        if (fabs(up) < TOL) return log(v - lo);
        return log(v - lo) - log(up - v);
    }
    __device__ double inv(double f) const {
        if (fabs(up) < TOL) return exp(f) + lo;
        const double e = exp(fmin(f, 50.0));
        return (e * up + lo) / (e + 1.0);
    }
    __device__ double deriv(double a) const {
        const double v = rangify(a);
        if (fabs(up) < TOL) return 1.0 / (v - lo);
        return 1.0 / (v - lo) + 1.0 / (up - v);
    }
};

__global__ void k_trans(int op, int n, const double * __restrict__ a, const double * __restrict__ b, double tau,
                        LogLU t, double * __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = a[i];
    double y;
    switch (op) {
    case 0: y = t.trans(x); break;
    case 1: y = t.inv(x); break;
    case 2: y = t.deriv(x); break;
    case 3: y = t.inv(t.trans(x) + b[i] * tau); break;
    case 4: y = fabs(x * b[i] * t.deriv(x)); break;
    default: y = t.trans(x) - t.trans(b[i]); break;
    }
    out[i] = y;
}

constexpr int PHI_WG = 256;
constexpr int PHI_PARTS = 1024;  // max workgroups per sum

__device__ inline void block_sum_to(double v, double * dst) {
    __shared__ double red[PHI_WG / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *dst = (red[0] + red[1]) + (red[2] + red[3]);
}

// partial phiM of the constraint rows of this workgroup (grid-stride)
__global__ void __launch_bounds__(PHI_WG) k_phi_m(int nC, const int * __restrict__ rowptr,
                                                  const int * __restrict__ colidx, const double * __restrict__ vals,
                                                  const double * __restrict__ m, const double * __restrict__ dm,
                                                  const double * __restrict__ mw, const double * __restrict__ cw,
                                                  double tau, LogLU t, double * __restrict__ part) {
    double acc = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nC; r += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = rowptr[r]; q < rowptr[r + 1]; ++q) {
            const int j = colidx[q];
            const double mj = tau == 0.0 ? m[j] : t.inv(t.trans(m[j]) + dm[j] * tau);
            s += vals[q] * (t.trans(mj) * mw[j]);
        }
        const double v = s * cw[r];
        acc += v * v;
    }
    block_sum_to(acc, part + blockIdx.x);
}

// partial phiD of the data of this workgroup (data transform log)
__global__ void __launch_bounds__(PHI_WG) k_phi_d(int nD, const double * __restrict__ d, const double * __restrict__ r,
                                                  const double * __restrict__ dd, const double * __restrict__ e,
                                                  double tau, double * __restrict__ part) {
    const LogLU t{0.0, 0.0};
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nD; i += gridDim.x * blockDim.x) {
        const double ri = tau == 0.0 ? r[i] : t.inv(t.trans(r[i]) + dd[i] * tau);
        const double v = (t.trans(d[i]) - t.trans(ri)) / e[i];
        acc += v * v;
    }
    block_sum_to(acc, part + blockIdx.x);
}

// out[k] = sum of part[0 .. np) in a fixed order (one workgroup)
__global__ void __launch_bounds__(PHI_WG) k_phi_sum(int np, const double * __restrict__ part, double * __restrict__ out) {
    double acc = 0.0;
    for (int g = threadIdx.x; g < np; g += blockDim.x) acc += part[g];
    block_sum_to(acc, out);
}

}  // namespace

void gn_trans(Ctx & c, int op, int n, const double * a, const double * b, double tau, double lower, double upper,
              double * out) {
    PG_REQUIRE(op >= 0 && op <= 5, "pgert_trans: unknown operation");
    PG_REQUIRE(n >= 0 && (n == 0 || (a && out)), "pgert_trans: null vector");
    PG_REQUIRE(op < 3 || n == 0 || b, "pgert_trans: update / error / difference need b");
    if (n == 0) return;
    hipLaunchKernelGGL(k_trans, dim3(div_up(n, 256)), dim3(256), 0, c.stream, op, n, a, b, tau, LogLU{lower, upper},
                       out);
    PG_CHECK(hipGetLastError());
    PG_CHECK(hipStreamSynchronize(c.stream));
}

void gn_phi(Ctx & c, int nTau, const double * taus, int nModel, const double * m, const double * dm, const double * mw,
            const double * cw, double lower, double upper, int nData, const double * d, const double * r,
            const double * dd, const double * e, double * phiD, double * phiM) {
    PG_REQUIRE(nTau >= 1 && taus && phiD && phiM, "pgert_gn_phi: no step lengths");
    PG_REQUIRE(nModel > 0 && m && dm && mw, "pgert_gn_phi: null model vector");
    PG_REQUIRE(nData > 0 && d && r && dd && e, "pgert_gn_phi: null data vector");
    PG_REQUIRE(c.nConst == 0 || (c.cCols == nModel && cw), "pgert_gn_phi: constraints have another model size");
    hipStream_t s = c.stream;
    const int gM = std::max(1, std::min(PHI_PARTS, div_up(std::max(c.nConst, 1), PHI_WG)));
    const int gD = std::max(1, std::min(PHI_PARTS, div_up(nData, PHI_WG)));
    DBuf<double> part, out;
    part.alloc((size_t)2 * PHI_PARTS);
    out.alloc((size_t)2 * nTau);
    const LogLU t{lower, upper};
    for (int k = 0; k < nTau; ++k) {
        if (c.nConst > 0) {
            hipLaunchKernelGGL(k_phi_m, dim3(gM), dim3(PHI_WG), 0, s, c.nConst, c.cRowptr.p, c.cColidx.p, c.cVals.p, m,
                               dm, mw, cw, taus[k], t, part.p);
            hipLaunchKernelGGL(k_phi_sum, dim3(1), dim3(PHI_WG), 0, s, gM, part.p, out.p + 2 * k + 1);
        } else {
            PG_CHECK(hipMemsetAsync(out.p + 2 * k + 1, 0, sizeof(double), s));
        }
        hipLaunchKernelGGL(k_phi_d, dim3(gD), dim3(PHI_WG), 0, s, nData, d, r, dd, e, taus[k], part.p + PHI_PARTS);
        hipLaunchKernelGGL(k_phi_sum, dim3(1), dim3(PHI_WG), 0, s, gD, part.p + PHI_PARTS, out.p + 2 * k);
    }
    PG_CHECK(hipGetLastError());
    std::vector<double> h((size_t)2 * nTau);
    PG_CHECK(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    PG_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < nTau; ++k) {
        phiD[k] = h[2 * k];
        phiM[k] = h[2 * k + 1];
    }
}

}  // namespace pgert

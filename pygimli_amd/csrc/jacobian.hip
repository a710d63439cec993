// Sensitivity (Jacobian) kernel: createSensitivityCol / calc2
// (core/src/bert/bertJacobian.cpp:170-219, :367-521) re-formulated per cell.
//
// For P1 elements grad(u) is constant on a cell, so the reference's per-datum
// quadratic form (a - b)^T (K_c + k^2 M_c) (m - n) becomes
//   vol * (g_a - g_b) . (g_m - g_n)
//   + k^2 vol/((d+1)(d+2)) * (sum_v da_v dm_v + sum_v da_v * sum_v dm_v)
// with g_e the gradient of electrode e's potential on the cell and da_v the
// nodal potential differences.  A workgroup owns a tile of 64 cells (sorted
// by marker, so a tile writes contiguous Jacobian columns); it first stages
// every electrode's gradient / nodal values of its cells in LDS, then sweeps
// the data: each lane computes one (datum, cell) entry, so a wavefront
// writes 64 consecutive floats of a Jacobian row.
#include "internal.h"

namespace pgert {

namespace {

constexpr int TILE_MAX = 64;  // cells per workgroup (fewer when the electrodes' LDS tile does not fit)

enum { J_STORE = 0, J_ACCUM = 1, J_ATOMIC = 2 };

template <typename OutT>
__device__ inline void jwrite(OutT * J, size_t idx, double v, int mode, bool first) {
    if (mode == J_ATOMIC) {
        atomicAdd(J + idx, (OutT)v);
    } else if (mode == J_STORE || first) {
        J[idx] = (OutT)v;
    } else {
        J[idx] = (OutT)((double)J[idx] + v);
    }
}

// One launch per wavenumber.  LDS layout: S[(e * NCOMP + comp) * TILE + cell].
// Lane l of a wavefront takes cell l % TILE of the tile and every
// (64 / TILE)-th datum starting at its sub-group l / TILE.
template <typename OutT, int DIM, int MASS, int TILE>
__global__ void __launch_bounds__(256) k_sens(int n, int nE, int D, int nTile, int nSorted,
                                              const int * __restrict__ sorted, const int * __restrict__ cells,
                                              const int * __restrict__ marker, const double * __restrict__ grads,
                                              const double * __restrict__ vol, const double * __restrict__ U,
                                              const int * __restrict__ abmn, double k2, double wk,
                                              int mode, int first, int last, const double * __restrict__ kfac,
                                              const double * __restrict__ model, int scaleModel, int nParam,
                                              OutT * __restrict__ J) {
    constexpr int NV = DIM + 1;
    constexpr int NCOMP = DIM + (MASS ? NV : 0);  // gradient comps (+ nodal values)
    extern __shared__ double S[];
    const int tile = blockIdx.x;
    if (tile >= nTile) return;
    const int base = tile * TILE;
    const int ncell = min(TILE, nSorted - base);
    // ---- stage gradients and nodal values of all electrodes for this tile
    for (int idx = threadIdx.x; idx < TILE * nE; idx += blockDim.x) {
        int t = idx / nE, e = idx % nE;
        double g[DIM] = {};
        double uv[NV] = {};
        if (t < ncell) {
            int c = sorted[base + t];
            for (int v = 0; v < NV; ++v) {
                int node = cells[(size_t)c * NV + v];
                double u = U[(size_t)node * nE + e];
                uv[v] = u;
                for (int q = 0; q < DIM; ++q) g[q] += u * grads[((size_t)c * NV + v) * 3 + q];
            }
        }
        for (int q = 0; q < DIM; ++q) S[((size_t)e * NCOMP + q) * TILE + t] = g[q];
        if (MASS)
            for (int v = 0; v < NV; ++v) S[((size_t)e * NCOMP + DIM + v) * TILE + t] = uv[v];
    }
    __syncthreads();
    constexpr int NSUB = 64 / TILE;
    const int lane = threadIdx.x % TILE, wave = (threadIdx.x / 64) * NSUB + (threadIdx.x % 64) / TILE;
    const int nwork = (blockDim.x / 64) * NSUB;
    if (lane >= ncell) return;
    const int c = sorted[base + lane];
    const double V = vol[c];
    const double mfac = MASS ? k2 * V / (NV * (NV + 1)) : 0.0;
    const int col = marker[c];
    const double m2 = scaleModel ? model[col] * model[col] : 1.0;
    for (int i = wave; i < D; i += nwork) {
        const int a = abmn[4 * i], b = abmn[4 * i + 1], m = abmn[4 * i + 2], nn = abmn[4 * i + 3];
        double dab[NCOMP], dmn[NCOMP];
        for (int q = 0; q < NCOMP; ++q) {
            double va = a >= 0 ? S[((size_t)a * NCOMP + q) * TILE + lane] : 0.0;
            double vb = b >= 0 ? S[((size_t)b * NCOMP + q) * TILE + lane] : 0.0;
            double vm = m >= 0 ? S[((size_t)m * NCOMP + q) * TILE + lane] : 0.0;
            double vn = nn >= 0 ? S[((size_t)nn * NCOMP + q) * TILE + lane] : 0.0;
            dab[q] = va - vb;
            dmn[q] = vm - vn;
        }
        double val = 0.0;
        for (int q = 0; q < DIM; ++q) val += dab[q] * dmn[q];
        val *= V;
        if (MASS) {
            double s2 = 0, sa = 0, sm = 0;
            for (int v = 0; v < NV; ++v) {
                s2 += dab[DIM + v] * dmn[DIM + v];
                sa += dab[DIM + v];
                sm += dmn[DIM + v];
            }
            val += mfac * (s2 + sa * sm);
        }
        val *= wk;
        // createJacobian_ scaling J_ij *= k_i / m_j^2 (dcfemmodelling.cpp:1375-1384)
        if (last && scaleModel && mode != J_ATOMIC) {
            // the last wavenumber finishes the entry, so the scale applies to the sum
            size_t idx = (size_t)i * nParam + col;
            double prev = (first || mode == J_STORE) ? 0.0 : (double)J[idx];
            J[idx] = (OutT)((prev + val) * kfac[i] / m2);
        } else {
            jwrite(J, (size_t)i * nParam + col, val, mode, first);
        }
    }
}

template <typename OutT>
__global__ void k_scale_cols(long total, int nParam, const double * __restrict__ kfac,
                             const double * __restrict__ model, OutT * __restrict__ J) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i = (int)(t / nParam), j = (int)(t % nParam);
    J[t] = (OutT)((double)J[t] * kfac[i] / (model[j] * model[j]));
}

template <typename OutT> void run(Ctx & c, const double * model, int nModel, OutT * J) {
    const bool scale = (nModel == c.nParam);
    int mode;
    if (!c.uniqueMarkers) mode = J_ATOMIC;
    else mode = (c.nK == 1) ? J_STORE : J_ACCUM;
    const size_t jsize = (size_t)c.D * c.nParam;
    // columns without any cell stay zero (markers not covering 0..max)
    if (mode == J_ATOMIC || c.nSorted < c.nParam)
        PG_CHECK(hipMemsetAsync(J, 0, jsize * sizeof(OutT), c.stream));
    for (int kk = 0; kk < c.nK; ++kk) {
        double k = c.kv[kk];
        const bool mass = k > 0.0;
        const int ncomp = c.dim + (mass ? c.nv : 0);
        int tile = TILE_MAX;
        while (tile > 16 && (size_t)c.nE * ncomp * tile * sizeof(double) > 160 * 1024) tile /= 2;
        size_t lds = (size_t)c.nE * ncomp * tile * sizeof(double);
        PG_REQUIRE(lds <= 160 * 1024, "pgert_jacobian: too many electrodes for one LDS tile");
        const int nTile = div_up(c.nSorted, tile);
        const double * Uk = c.U.p + (size_t)kk * c.n * c.nE;
        int first = kk == 0, last = kk == c.nK - 1;
        auto pick = [&](auto tl) {
            constexpr int T = decltype(tl)::value;
            return c.dim == 2 ? (mass ? k_sens<OutT, 2, 1, T> : k_sens<OutT, 2, 0, T>)
                              : (mass ? k_sens<OutT, 3, 1, T> : k_sens<OutT, 3, 0, T>);
        };
        auto kern = tile == 64 ? pick(std::integral_constant<int, 64>{})
                  : tile == 32 ? pick(std::integral_constant<int, 32>{}) : pick(std::integral_constant<int, 16>{});
        PG_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3(nTile), dim3(256), lds, c.stream, c.n, c.nE, c.D, nTile, c.nSorted,
                           c.sortedCells.p, c.cells.p, c.cellMarker.p, c.grads.p, c.vol.p, Uk, c.abmn.p, k * k,
                           c.wv[kk], mode, first, last, c.kfac.p, model, scale ? 1 : 0, c.nParam, J);
        PG_CHECK(hipGetLastError());
    }
    if (mode == J_ATOMIC && scale) {
        long total = (long)jsize;
        hipLaunchKernelGGL(k_scale_cols<OutT>, dim3(div_up(total, 256)), dim3(256), 0, c.stream, total, c.nParam,
                           c.kfac.p, model, J);
        PG_CHECK(hipGetLastError());
    }
}

}  // namespace

void launch_jacobian(Ctx & c, const double * model, int nModel, void * J, int dtype) {
    if (dtype == 0) run<double>(c, model, nModel, (double *)J);
    else run<float>(c, model, nModel, (float *)J);
}

}  // namespace pgert

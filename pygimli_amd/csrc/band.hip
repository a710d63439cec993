// Direct solver for 2D (2.5D) meshes: banded Cholesky, the reference's own
// approach (a sparse direct factorisation per wavenumber, LinSolver ->
// CHOLMOD, linSolver.cpp:96-178, cholmodWrapper.cpp:224-459) on the GPU.
//
// A 2D P1 mesh numbered along its shorter grid direction (or by reverse
// Cuthill-McKee) has a small bandwidth b (58-227 on 2k-40k-node meshes), so
// A_k = L L' is kept in band storage, L[j][d] = L(j + d, j), d = 0..b, and
// factored with a blocked right-looking sweep: a panel of NB columns
// (rows j0 .. j0 + NB + b) is factored in LDS, the inverse of its NB x NB
// diagonal block is kept for the solves, and the trailing b x b band window
// gets the rank-NB update.  All right-hand sides of a wavenumber are then
// solved together, block row by block row: y1 = L11^-1 y1 and the band rows
// below get -= L21 y1 (forward); y1 -= L21' x2 and x1 = L11^-T y1
// (backward).  One workgroup per (problem, wavenumber): every factorisation
// and every solve of a batch of problems runs in one launch each.
// Everything is FP64; the result is the direct solution (no tolerance).
#include "internal.h"

#include <algorithm>
#include <cmath>

namespace pgert {

namespace {

constexpr int NB = 32;    // panel width
constexpr int LDP = NB + 1;
constexpr int BMAX = 400; // band limit (LDS)
constexpr int EMAX = 128; // electrodes per solve (LDS)

struct BandDesc {
    int n, b, nE, nnz;
    const double * vals;  // CSR values of A_k
    const int * slot;     // CSR slot -> band index, -1 for the upper triangle
    double * L;           // band storage n x (b + 1)
    double * Dinv;        // per panel the NB x NB inverse of its diagonal block
    const int * perm;     // band order -> node
    const double * B;     // right-hand sides of wavenumber k, node-major [node][nE]
    double * X;           // solution of wavenumber k, node-major [node][nE]
    double * Y;           // work, band order [row][nE]
};

__global__ void k_band_scatter(const BandDesc * __restrict__ descs) {
    const BandDesc d = descs[blockIdx.y];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < d.nnz; t += gridDim.x * blockDim.x) {
        const int s = d.slot[t];
        if (s >= 0) d.L[s] = d.vals[t];
    }
}

// map a linear index to (ti, tj), ti >= tj, over a lower triangle of tiles
__device__ inline void tri_tile(int t, int & ti, int & tj) {
    int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    ti = i;
    tj = t - i * (i + 1) / 2;
}

// one workgroup per factorisation; LDS: panel [NB + b + 4][LDP], Li [NB][LDP]
__global__ void __launch_bounds__(256) k_band_factor(const BandDesc * __restrict__ descs) {
    extern __shared__ double sm[];
    const BandDesc d = descs[blockIdx.x];
    const int n = d.n, b = d.b, W1 = b + 1;
    const int PR = NB + b + 4;
    double * Pm = sm;
    double * Li = sm + (size_t)PR * LDP;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nc = min(NB, n - j0);
        // panel: Pm[r][c] = L(j0 + r, j0 + c) inside the band, else 0
        for (int t = threadIdx.x; t < PR * NB; t += blockDim.x) {
            const int r = t / NB, c = t % NB;
            double v = 0.0;
            if (c < nc && r >= c && r - c <= b && j0 + r < n) v = d.L[(size_t)(j0 + c) * W1 + (r - c)];
            Pm[r * LDP + c] = v;
        }
        __syncthreads();
        for (int c = 0; c < nc; ++c) {
            const double piv = Pm[c * LDP + c];
            const double dd = sqrt(piv > 0.0 ? piv : 1e-300);
            __syncthreads();
            for (int r = c + threadIdx.x; r < PR; r += blockDim.x) Pm[r * LDP + c] = r == c ? dd : Pm[r * LDP + c] / dd;
            __syncthreads();
            const int m2 = nc - c - 1;
            if (m2 > 0)
                for (int t = threadIdx.x; t < PR * m2; t += blockDim.x) {
                    const int r = t / m2, c2 = c + 1 + t % m2;
                    if (r >= c2) Pm[r * LDP + c2] -= Pm[r * LDP + c] * Pm[c2 * LDP + c];
                }
            __syncthreads();
        }
        // factored panel back to the band
        for (int t = threadIdx.x; t < PR * NB; t += blockDim.x) {
            const int r = t / NB, c = t % NB;
            if (c < nc && r >= c && r - c <= b && j0 + r < n) d.L[(size_t)(j0 + c) * W1 + (r - c)] = Pm[r * LDP + c];
        }
        // inverse of the diagonal block (lower triangular), a thread per column
        if (threadIdx.x < NB) {
            const int col = threadIdx.x;
            for (int r = 0; r < NB; ++r) {
                double v = 0.0;
                if (r < nc && col < nc && r >= col) {
                    v = r == col ? 1.0 : 0.0;
                    for (int q = col; q < r; ++q) v -= Pm[r * LDP + q] * Li[q * LDP + col];
                    v /= Pm[r * LDP + r];
                }
                Li[r * LDP + col] = v;
            }
        }
        __syncthreads();
        double * Di = d.Dinv + (size_t)(j0 / NB) * NB * NB;
        for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) Di[t] = Li[(t / NB) * LDP + t % NB];
        // trailing band window: rows / columns j0 + nc .. j0 + nc + W - 1
        const int W = min(b, n - (j0 + nc));
        const int T = (W + 3) / 4;
        for (int tile = threadIdx.x; tile < T * (T + 1) / 2; tile += blockDim.x) {
            int ti, tj;
            tri_tile(tile, ti, tj);
            double acc[4][4] = {};
            for (int c = 0; c < nc; ++c) {
                double a[4], q[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    a[x] = Pm[(nc + 4 * ti + x) * LDP + c];
                    q[x] = Pm[(nc + 4 * tj + x) * LDP + c];
                }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * q[y];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int rr = 4 * ti + x, jj = 4 * tj + y;
                    if (rr < W && jj < W && rr >= jj) d.L[(size_t)(j0 + nc + jj) * W1 + (rr - jj)] -= acc[x][y];
                }
        }
        __syncthreads();
    }
}

// Y (band order) from the node-major right-hand sides, and back
__global__ void k_band_gather(const BandDesc * __restrict__ descs) {
    const BandDesc d = descs[blockIdx.y];
    const long tot = (long)d.n * d.nE;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const int r = (int)(t / d.nE), e = (int)(t % d.nE);
        d.Y[t] = d.B[(size_t)d.perm[r] * d.nE + e];
    }
}

__global__ void k_band_scatter_x(const BandDesc * __restrict__ descs) {
    const BandDesc d = descs[blockIdx.y];
    const long tot = (long)d.n * d.nE;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const int r = (int)(t / d.nE), e = (int)(t % d.nE);
        d.X[(size_t)d.perm[r] * d.nE + e] = d.Y[t];
    }
}

// one workgroup per (problem, wavenumber): forward and backward sweeps over
// all nE right-hand sides.  LDS: Lr [b + 4][LDP] (L21 of the block row),
// Li [NB][LDP], y [NB][nE + 1], z [NB][nE + 1]
__global__ void __launch_bounds__(256) k_band_solve(const BandDesc * __restrict__ descs) {
    extern __shared__ double sm[];
    const BandDesc d = descs[blockIdx.x];
    const int n = d.n, b = d.b, W1 = b + 1, nE = d.nE, LE = nE + 1;
    double * Lr = sm;
    double * Li = Lr + (size_t)(b + 4) * LDP;
    double * ys = Li + (size_t)NB * LDP;
    double * zs = ys + (size_t)NB * LE;
    double * Y = d.Y;
    const int nblk = (n + NB - 1) / NB;
    const int TE = (nE + 3) / 4;
    // stage the block row j0: L21 (rows j0 + nc + rr, columns j0 + c) and L11^-1
    auto stage = [&](int j0, int nc, int W) {
        for (int t = threadIdx.x; t < (W + 4) * NB; t += blockDim.x) {
            const int rr = t / NB, c = t % NB;
            const int dlt = nc + rr - c;
            double v = 0.0;
            if (rr < W && c < nc && dlt <= b) v = d.L[(size_t)(j0 + c) * W1 + dlt];
            Lr[rr * LDP + c] = v;
        }
        const double * Di = d.Dinv + (size_t)(j0 / NB) * NB * NB;
        for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) Li[(t / NB) * LDP + t % NB] = Di[t];
        for (int t = threadIdx.x; t < NB * nE; t += blockDim.x) {
            const int r = t / nE, e = t % nE;
            ys[r * LE + e] = r < nc ? Y[(size_t)(j0 + r) * nE + e] : 0.0;
        }
    };
    // ---- forward: L y = b
    for (int blk = 0; blk < nblk; ++blk) {
        const int j0 = blk * NB, nc = min(NB, n - j0), W = min(b, n - (j0 + nc));
        stage(j0, nc, W);
        __syncthreads();
        // z = L11^-1 y
        for (int t = threadIdx.x; t < NB * nE; t += blockDim.x) {
            const int r = t / nE, e = t % nE;
            double v = 0.0;
            for (int q = 0; q <= r; ++q) v += Li[r * LDP + q] * ys[q * LE + e];
            zs[r * LE + e] = v;
            if (r < nc) Y[(size_t)(j0 + r) * nE + e] = v;
        }
        __syncthreads();
        // rows below: Y -= L21 z (4 x 4 tiles over W x nE)
        const int TW = (W + 3) / 4;
        for (int tile = threadIdx.x; tile < TW * TE; tile += blockDim.x) {
            const int ti = tile / TE, te = tile % TE;
            double acc[4][4] = {};
            for (int c = 0; c < nc; ++c) {
                double a[4], z[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) a[x] = Lr[(4 * ti + x) * LDP + c];
#pragma unroll
                for (int y = 0; y < 4; ++y) z[y] = 4 * te + y < nE ? zs[c * LE + 4 * te + y] : 0.0;
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * z[y];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int rr = 4 * ti + x, e = 4 * te + y;
                    if (rr < W && e < nE) Y[(size_t)(j0 + nc + rr) * nE + e] -= acc[x][y];
                }
        }
        __syncthreads();
    }
    // ---- backward: L' x = y
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = blk * NB, nc = min(NB, n - j0), W = min(b, n - (j0 + nc));
        stage(j0, nc, W);
        __syncthreads();
        // z = y1 - L21' x2 (4 x 4 tiles over NB x nE)
        for (int tile = threadIdx.x; tile < (NB / 4) * TE; tile += blockDim.x) {
            const int tc = tile / TE, te = tile % TE;
            double acc[4][4] = {};
            for (int rr = 0; rr < W; ++rr) {
                double a[4], x2[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) a[x] = Lr[rr * LDP + 4 * tc + x];
#pragma unroll
                for (int y = 0; y < 4; ++y) x2[y] = 4 * te + y < nE ? Y[(size_t)(j0 + nc + rr) * nE + 4 * te + y] : 0.0;
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * x2[y];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int c = 4 * tc + x, e = 4 * te + y;
                    if (e < nE) zs[c * LE + e] = ys[c * LE + e] - acc[x][y];
                }
        }
        __syncthreads();
        // x1 = L11^-T z
        for (int t = threadIdx.x; t < NB * nE; t += blockDim.x) {
            const int r = t / nE, e = t % nE;
            double v = 0.0;
            for (int q = r; q < NB; ++q) v += Li[q * LDP + r] * zs[q * LE + e];
            if (r < nc) Y[(size_t)(j0 + r) * nE + e] = v;
        }
        __syncthreads();
    }
}

size_t factor_lds(int b) { return ((size_t)(NB + b + 4) * LDP + (size_t)NB * LDP) * sizeof(double); }
size_t solve_lds(int b, int nE) {
    return ((size_t)(b + 4) * LDP + (size_t)NB * LDP + 2 * (size_t)NB * (nE + 1)) * sizeof(double);
}

// bandwidth of an ordering (iperm: node -> position)
int bandwidth(const Ctx & c, const std::vector<int> & iperm) {
    int b = 0;
    for (int i = 0; i < c.n; ++i)
        for (int t = c.hRowptr[i]; t < c.hRowptr[i + 1]; ++t) b = std::max(b, std::abs(iperm[i] - iperm[c.hColidx[t]]));
    return b;
}

// reverse Cuthill-McKee over the CSR graph (pseudo-peripheral start per component)
std::vector<int> rcm(const Ctx & c) {
    const int n = c.n;
    const std::vector<int> & rp = c.hRowptr, & ci = c.hColidx;
    std::vector<int> deg(n), order, stamp(n, -1);
    for (int i = 0; i < n; ++i) deg[i] = rp[i + 1] - rp[i];
    std::vector<char> seen(n, 0);
    order.reserve(n);
    int tick = 0;
    auto farthest = [&](int s) {  // last BFS level's minimum-degree node
        std::vector<int> q{s};
        stamp[s] = ++tick;
        size_t head = 0, levelStart = 0;
        while (head < q.size()) {
            const size_t levelEnd = q.size();
            levelStart = head;
            for (; head < levelEnd; ++head)
                for (int t = rp[q[head]]; t < rp[q[head] + 1]; ++t)
                    if (stamp[ci[t]] != tick) {
                        stamp[ci[t]] = tick;
                        q.push_back(ci[t]);
                    }
        }
        int best = q[levelStart];
        for (size_t i = levelStart; i < q.size(); ++i)
            if (deg[q[i]] < deg[best]) best = q[i];
        return best;
    };
    for (int s0 = 0; s0 < n; ++s0) {
        if (seen[s0]) continue;
        int s = farthest(farthest(s0));
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        std::vector<int> nb;
        while (head < order.size()) {
            const int v = order[head++];
            nb.clear();
            for (int t = rp[v]; t < rp[v + 1]; ++t)
                if (!seen[ci[t]]) {
                    seen[ci[t]] = 1;
                    nb.push_back(ci[t]);
                }
            std::sort(nb.begin(), nb.end(), [&](int a, int b2) { return deg[a] < deg[b2] || (deg[a] == deg[b2] && a < b2); });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

}  // namespace

// ordering, band width and the CSR slot -> band map, once per mesh
void band_prepare(Ctx & c) {
    if (c.bandValid) return;
    const int n = c.n;
    std::vector<int> natural(n), inat(n);
    for (int i = 0; i < n; ++i) natural[i] = inat[i] = i;
    std::vector<int> r = rcm(c), ir(n);
    for (int i = 0; i < n; ++i) ir[r[i]] = i;
    const int bn = bandwidth(c, inat), br = bandwidth(c, ir);
    const std::vector<int> & perm = bn <= br ? natural : r;
    const std::vector<int> & iperm = bn <= br ? inat : ir;
    const int b = std::min(bn, br);
    PG_REQUIRE(b <= BMAX, "band solver: bandwidth " + std::to_string(b) + " above " + std::to_string(BMAX) +
                              " (3D meshes: use a block method)");
    std::vector<int> slot(c.nnz);
    for (int i = 0; i < n; ++i)
        for (int t = c.hRowptr[i]; t < c.hRowptr[i + 1]; ++t) {
            const int ii = iperm[i], jj = iperm[c.hColidx[t]];
            slot[t] = ii >= jj ? jj * (b + 1) + (ii - jj) : -1;
        }
    c.bandB = b;
    c.bandPerm.upload(perm, c.stream);
    c.bandSlot.upload(slot, c.stream);
    PG_CHECK(hipStreamSynchronize(c.stream));
    c.bandValid = true;
}

void solve_band(const std::vector<Ctx *> & cs, const pgert_solve_opts & o, std::vector<pgert_solve_stats> & stats) {
    PG_REQUIRE(!cs.empty(), "band solver: no problems");
    Ctx & c0 = *cs[0];
    hipStream_t s = c0.stream;
    size_t Lsz = 0, Dsz = 0, Ysz = 0;
    int bmax = 0, emax = 0;
    for (Ctx * c : cs) {
        PG_REQUIRE(c->dim == 2, "band solver: 2D (2.5D) meshes only");
        PG_REQUIRE(c->nE <= EMAX, "band solver: more than 128 electrodes");
        band_prepare(*c);
        const size_t nblk = (size_t)(c->n + NB - 1) / NB;
        Lsz += (size_t)c->nK * c->n * (c->bandB + 1);
        Dsz += (size_t)c->nK * nblk * NB * NB;
        Ysz += (size_t)c->nK * c->n * c->nE;
        bmax = std::max(bmax, c->bandB);
        emax = std::max(emax, c->nE);
    }
    DBuf<double> Lp, Dp, Yp;
    Lp.alloc(Lsz);
    Dp.alloc(Dsz);
    Yp.alloc(Ysz);
    PG_CHECK(hipMemsetAsync(Lp.p, 0, Lp.bytes(), s));
    std::vector<BandDesc> h;
    size_t lo = 0, dof = 0, yo = 0;
    int maxn = 0, maxnnz = 0;
    for (Ctx * c : cs)
        for (int k = 0; k < c->nK; ++k) {
            BandDesc d{};
            d.n = c->n;
            d.b = c->bandB;
            d.nE = c->nE;
            d.nnz = c->nnz;
            d.vals = c->valsA.p + (size_t)k * c->nnz;
            d.slot = c->bandSlot.p;
            d.L = Lp.p + lo;
            d.Dinv = Dp.p + dof;
            d.perm = c->bandPerm.p;
            d.B = c->B.p + (size_t)k * c->n * c->nE;
            d.X = c->X.p + (size_t)k * c->n * c->nE;
            d.Y = Yp.p + yo;
            lo += (size_t)c->n * (c->bandB + 1);
            dof += (size_t)((c->n + NB - 1) / NB) * NB * NB;
            yo += (size_t)c->n * c->nE;
            maxn = std::max(maxn, c->n);
            maxnnz = std::max(maxnnz, c->nnz);
            h.push_back(d);
        }
    DBuf<BandDesc> dd;
    dd.upload(h, s);
    const int nd = (int)h.size();
    const dim3 blk(256);
    hipLaunchKernelGGL(k_band_scatter, dim3(std::max(1, std::min(div_up(maxnnz, 256), 4096 / nd + 1)), nd), blk, 0, s,
                       dd.p);
    const size_t fl = factor_lds(bmax), sl = solve_lds(bmax, emax);
    PG_REQUIRE(fl <= 160 * 1024 && sl <= 160 * 1024, "band solver: band too wide for LDS");
    PG_CHECK(hipFuncSetAttribute((const void *)k_band_factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fl));
    PG_CHECK(hipFuncSetAttribute((const void *)k_band_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl));
    hipLaunchKernelGGL(k_band_factor, dim3(nd), blk, fl, s, dd.p);
    const int ge = std::max(1, std::min(div_up((long)maxn * emax, 256), 65536 / nd + 1));
    hipLaunchKernelGGL(k_band_gather, dim3(ge, nd), blk, 0, s, dd.p);
    hipLaunchKernelGGL(k_band_solve, dim3(nd), blk, sl, s, dd.p);
    hipLaunchKernelGGL(k_band_scatter_x, dim3(ge, nd), blk, 0, s, dd.p);
    PG_CHECK(hipGetLastError());
    PG_CHECK(hipStreamSynchronize(s));  // work buffers are released on return
    stats.assign(cs.size(), pgert_solve_stats{});
    for (size_t p = 0; p < cs.size(); ++p) {
        pgert_solve_stats & st = stats[p];
        st.max_rel_residual = residual_check(*cs[p]);
        st.converged = st.max_rel_residual <= o.rtol * 1.0001 ? 1 : 0;
    }
}

}  // namespace pgert

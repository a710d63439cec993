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
constexpr int BMAX = 255; // band limit (one L21 row per solve thread)
constexpr int EMAX = 128; // electrodes per solve
constexpr int CG = 16;    // right-hand sides per solve workgroup

struct BandDesc {
    int n, b, nE, nnz;
    const double * vals;  // CSR values of A_k
    const int * slot;     // CSR slot -> band index, -1 for the upper triangle
    double * L;           // band storage n x (b + 1)
    double * Dinv;        // per panel the NB x NB inverse of its diagonal block
    const int * perm;     // band order -> node
    const double * B;     // right-hand sides of wavenumber k, node-major [node][nE]
    double * X;           // solution of wavenumber k, node-major [node][nE]
    double * Y;           // forward-sweep result, band order [row][nE]
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

// One workgroup per factorisation.  Panel rows live in registers (thread r
// owns rows r, r + 256, ...: RPT rows), so a column step is one barrier: the
// top NB rows publish their next column in LDS (colb[c][r] = panel(r, c)
// once columns < c are applied) and every thread scales and updates its own
// rows.  Then wavefront 0 inverts the diagonal block (a column per lane, in
// registers) while wavefronts 1-3 apply the rank-NB update to the trailing
// b x b band window.  LDS: colb [NB][LDP], panel [NB + b][LDP].
template <int RPT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_band_factor(const BandDesc * __restrict__ descs) {
    extern __shared__ double sm[];
    const BandDesc d = descs[blockIdx.x];
    const int n = d.n, b = d.b, W1 = b + 1;
    const int tid = threadIdx.x;
    const auto L = glob(d.L);
    double * colb = sm;
    double * Pm = sm + NB * LDP;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nc = min(NB, n - j0);
        const int W = min(b, n - (j0 + nc));
        const int PRn = nc + W;
        // panel through LDS: column by column, lanes over rows (coalesced band reads)
        // (loads are unconditional from a clamped index, so they issue back to back)
        for (int r = tid; r < PRn; r += 256)
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const bool in = c < nc && r >= c && r - c <= b;
                const double v = L[in ? (size_t)(j0 + c) * W1 + (r - c) : 0];
                Pm[r * LDP + c] = in ? v : 0.0;
            }
        __syncthreads();
        double p[RPT][NB];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int r = tid + 256 * i;
#pragma unroll
            for (int c = 0; c < NB; ++c) p[i][c] = r < PRn ? Pm[r * LDP + c] : 0.0;
        }
        if (tid < NB) colb[tid] = p[0][0];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c < nc) {
                const double piv = colb[c * LDP + c];
                const double dd = sqrt(piv > 0.0 ? piv : 1e-300);
                const double rd = 1.0 / dd;
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int r = tid + 256 * i;
                    const double l = r == c ? dd : p[i][c] * rd;
                    const double lm = r >= c ? l : 0.0;  // rows above c: upper triangle, untouched
                    p[i][c] = lm;
#pragma unroll
                    for (int c2 = c + 1; c2 < NB; ++c2) p[i][c2] -= lm * (colb[c * LDP + c2] * rd);
                }
                if (c + 1 < NB && tid < NB) colb[(c + 1) * LDP + tid] = p[0][c + 1];
            }
            lds_barrier();
        }
        // factored panel: to LDS (L11 for its inverse, L21 for the update) and back to the band
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int r = tid + 256 * i;
            if (r < PRn)
#pragma unroll
                for (int c = 0; c < NB; ++c) Pm[r * LDP + c] = p[i][c];
        }
        __syncthreads();
        for (int c = 0; c < nc; ++c)
            for (int r = c + tid; r < PRn && r - c <= b; r += 256) L[(size_t)(j0 + c) * W1 + (r - c)] = Pm[r * LDP + c];
        if (tid < NB) {
            // inverse of the lower-triangular diagonal block, column `col`
            const int col = tid;
            double x[NB];
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                double v = r == col ? 1.0 : 0.0;
#pragma unroll
                for (int q = 0; q < r; ++q) v -= Pm[r * LDP + q] * x[q];
                x[r] = (r >= col && r < nc) ? v / Pm[r * LDP + r] : 0.0;
            }
            const auto Di = glob(d.Dinv) + (size_t)(j0 / NB) * NB * NB;
#pragma unroll
            for (int r = 0; r < NB; ++r) Di[r * NB + col] = x[r];
        } else if (tid >= 64) {
            // trailing band window: rows / columns j0 + nc .. j0 + nc + W - 1
            const int T = (W + 3) / 4;
            for (int tile = tid - 64; tile < T * (T + 1) / 2; tile += 192) {
                int ti, tj;
                tri_tile(tile, ti, tj);
                double acc[4][4] = {};
#pragma unroll 8
                for (int c = 0; c < NB; ++c) {  // panel columns >= nc are zero
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
                double old[4][4];
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) {
                        const int rr = 4 * ti + x, jj = 4 * tj + y;
                        const bool in = rr < W && jj < W && rr >= jj;
                        old[x][y] = L[in ? (size_t)(j0 + nc + jj) * W1 + (rr - jj) : 0];
                    }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) {
                        const int rr = 4 * ti + x, jj = 4 * tj + y;
                        if (rr < W && jj < W && rr >= jj) L[(size_t)(j0 + nc + jj) * W1 + (rr - jj)] = old[x][y] - acc[x][y];
                    }
            }
        }
        __syncthreads();
    }
}

// One workgroup per (problem, wavenumber, group of CG right-hand sides).
// Forward (L y = b): the rows still receiving updates live in an LDS ring of
// NB + b rows; per block row, z = L11^-1 y1 (stored as the forward result),
// the next W rows of the ring get -= L21 z and the consumed slots are
// refilled from b.  Backward (L' x = y): the ring holds the last solved rows;
// t = y1 - L21' x2, x1 = L11^-T t, written to the node-major solution.
// Software-pipelined: the next block row's L21 row (thread rr holds row rr,
// b <= 255), its L11^-1 and the refill / forward values are loaded into
// registers while the current block row computes; only LDS-ordering
// barriers inside the sweeps, so stores and prefetches stay in flight.
// LDS: ring [NB + b][CG], Lr [b][LDP], Li [NB][LDP], zs [NB][CG], pt [4][NB][CG].
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_band_solve(const BandDesc * __restrict__ descs) {
    extern __shared__ double sm[];
    const BandDesc d = descs[blockIdx.y];
    const int e0 = blockIdx.x * CG;
    if (e0 >= d.nE) return;
    const int ncol = min(CG, d.nE - e0);
    const int n = d.n, b = d.b, W1 = b + 1, nE = d.nE;
    const int R = NB + b;
    const int tid = threadIdx.x;
    double * ring = sm;
    double * Lr = ring + (size_t)R * CG;
    double * Li = Lr + (size_t)b * LDP;
    double * zs = Li + NB * LDP;
    double * pt = zs + NB * CG;  // [4][NB][CG] partial sums of the backward sweep
    const auto perm = glob(d.perm);
    const auto L = glob(d.L);
    const auto Bv = glob(d.B);
    const auto Y = glob(d.Y);
    const auto X = glob(d.X);
    const auto Dv = glob(d.Dinv);
    auto slot = [&](int s0, int q) { const int s = s0 + q; return s < R ? s : s - R; };
    double lp[NB], dp[4];
    // block row j0's L21 row `tid` and L11^-1 into registers
    auto fetch = [&](int j0) {
        const int nc = min(NB, n - j0), W = min(b, n - (j0 + nc));
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int dlt = nc + tid - c;
            const bool in = tid < W && c < nc && dlt <= b;
            const double v = L[in ? (size_t)(j0 + c) * W1 + dlt : 0];
            lp[c] = in ? v : 0.0;
        }
        const size_t o = (size_t)(j0 / NB) * NB * NB;
#pragma unroll
        for (int i = 0; i < 4; ++i) dp[i] = Dv[o + tid + 256 * i];
    };
    auto put = [&]() {
        if (tid < b)
#pragma unroll
            for (int c = 0; c < NB; ++c) Lr[tid * LDP + c] = lp[c];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int t = tid + 256 * i;
            Li[(t / NB) * LDP + t % NB] = dp[i];
        }
    };
    const int nblk = (n + NB - 1) / NB;
    // ---- forward
    for (int t = tid; t < min(n, R) * CG; t += 256) {
        const int r = t / CG, cc = t % CG;
        const double v = Bv[(size_t)perm[r] * nE + e0 + min(cc, ncol - 1)];
        ring[t] = cc < ncol ? v : 0.0;
    }
    fetch(0);
    for (int blk = 0; blk < nblk; ++blk) {
        const int j0 = blk * NB, nc = min(NB, n - j0), W = min(b, n - (j0 + nc));
        const int s0 = j0 % R;
        put();
        // values for the slots this block row consumes: rows j0 + R + r
        double rf[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k, r = t / CG, cc = t % CG;
            const int row = min(j0 + R + r, n - 1);
            rf[k] = Bv[(size_t)perm[row] * nE + e0 + min(cc, ncol - 1)];
        }
        lds_barrier();
        if (blk + 1 < nblk) fetch(j0 + NB);
        // z = L11^-1 y1 (L11^-1 is zero above the diagonal and beyond nc: fixed-length, unrolled)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k, r = t / CG, cc = t % CG;
            double v = 0.0;
#pragma unroll 8
            for (int q = 0; q < NB; ++q) v += Li[r * LDP + q] * ring[slot(s0, q) * CG + cc];
            zs[t] = v;
            if (r < nc && cc < ncol) Y[(size_t)(j0 + r) * nE + e0 + cc] = v;
        }
        lds_barrier();
        // ring rows j0 + nc + rr -= L21 z, tiles of 4 rows x 2 columns
        const int TR = (W + 3) / 4;
        for (int tile = tid; tile < TR * (CG / 2); tile += 256) {
            const int tr = tile / (CG / 2), tc = tile % (CG / 2);
            double acc[4][2] = {};
#pragma unroll 8
            for (int c = 0; c < NB; ++c) {  // L21 columns >= nc are zero
                const double2 z = *(const double2 *)(zs + c * CG + 2 * tc);
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    const double a = Lr[(4 * tr + x) * LDP + c];
                    acc[x][0] += a * z.x;
                    acc[x][1] += a * z.y;
                }
            }
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const int rr = 4 * tr + x;
                if (rr < W) {
                    double * rw = ring + slot(s0, nc + rr) * CG + 2 * tc;
                    rw[0] -= acc[x][0];
                    rw[1] -= acc[x][1];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k, r = t / CG, cc = t % CG;
            if (r < nc && j0 + R + r < n) ring[slot(s0, r) * CG + cc] = cc < ncol ? rf[k] : 0.0;
        }
        lds_barrier();
    }
    // ---- backward (the forward results in Y must be visible: full barrier)
    __syncthreads();
    double yp[2];
    int pp[2];
    auto fetch_y = [&](int j0) {
        const int nc = min(NB, n - j0);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k, r = min(t / CG, nc - 1), cc = t % CG;
            yp[k] = Y[(size_t)(j0 + r) * nE + e0 + min(cc, ncol - 1)];
            pp[k] = perm[j0 + r];
        }
    };
    fetch((nblk - 1) * NB);
    fetch_y((nblk - 1) * NB);
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = blk * NB, nc = min(NB, n - j0), W = min(b, n - (j0 + nc));
        const int s0 = j0 % R;
        put();
        int pc[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k, r = t / CG, cc = t % CG;
            zs[t] = (r < nc && cc < ncol) ? yp[k] : 0.0;
            pc[k] = pp[k];
        }
        lds_barrier();
        if (blk > 0) {
            fetch(j0 - NB);
            fetch_y(j0 - NB);
        }
        {
            // L21' x2: wavefront w takes rows rr = w, w + 4, ...; lane = 4 x 2 tile of (r, cc)
            const int w = tid >> 6, ln = tid & 63, tr = ln >> 3, tc = ln & 7;
            double acc[4][2] = {};
#pragma unroll 4
            for (int rr = w; rr < W; rr += 4) {
                const double2 xv = *(const double2 *)(ring + slot(s0, nc + rr) * CG + 2 * tc);
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    const double a = Lr[rr * LDP + 4 * tr + x];
                    acc[x][0] += a * xv.x;
                    acc[x][1] += a * xv.y;
                }
            }
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                pt[(w * NB + 4 * tr + x) * CG + 2 * tc] = acc[x][0];
                pt[(w * NB + 4 * tr + x) * CG + 2 * tc + 1] = acc[x][1];
            }
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k;
            zs[t] -= (pt[t] + pt[NB * CG + t]) + (pt[2 * NB * CG + t] + pt[3 * NB * CG + t]);
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = tid + 256 * k, r = t / CG, cc = t % CG;
            if (r < nc) {
                double v = 0.0;
#pragma unroll 8
                for (int q = 0; q < NB; ++q) v += Li[q * LDP + r] * zs[q * CG + cc];  // zero above / beyond
                ring[slot(s0, r) * CG + cc] = v;
                if (cc < ncol) X[(size_t)pc[k] * nE + e0 + cc] = v;
            }
        }
        lds_barrier();
    }
}

size_t factor_lds(int b) { return ((size_t)NB * LDP + (size_t)(NB + b) * LDP) * sizeof(double); }
size_t solve_lds(int b) {
    return ((size_t)(NB + b) * CG + (size_t)b * LDP + (size_t)NB * LDP + 5 * (size_t)NB * CG) * sizeof(double);
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

void solve_band(const std::vector<Ctx *> & cs, const pgert_solve_opts & o, std::vector<pgert_solve_stats> & stats,
                bool refactor) {
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
    DBuf<double> &Lp = c0.bandL, &Dp = c0.bandDinv, &Yp = c0.bandY;
    Lp.reserve(Lsz);
    Dp.reserve(Dsz);
    Yp.reserve(Ysz);
    if (refactor) PG_CHECK(hipMemsetAsync(Lp.p, 0, Lsz * sizeof(double), s));
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
    // longest work first: workgroups are dispatched in index order, so the
    // largest factorisations (n (b + 1)^2) start in the first round instead of
    // trailing behind many small ones (the storage offsets above are fixed per
    // descriptor, so the order is free)
    std::stable_sort(h.begin(), h.end(), [](const BandDesc & a, const BandDesc & b2) {
        return (double)a.n * (a.b + 1) * (a.b + 1) > (double)b2.n * (b2.b + 1) * (b2.b + 1);
    });
    DBuf<unsigned char> & ddb = c0.bandDesc;
    ddb.reserve(h.size() * sizeof(BandDesc));
    PG_CHECK(hipMemcpyAsync(ddb.p, h.data(), h.size() * sizeof(BandDesc), hipMemcpyHostToDevice, s));
    struct { const BandDesc * p; } dd{(const BandDesc *)ddb.p};
    const int nd = (int)h.size();
    const dim3 blk(256);
    const size_t fl = factor_lds(bmax), sl = solve_lds(bmax);
    PG_REQUIRE(fl <= 160 * 1024 && sl <= 160 * 1024, "band solver: band too wide for LDS");
    // refactor == false: the factors of the previous call on the same problems
    // and matrices are still in L / Dinv (new right-hand sides only)
    if (refactor) {
        hipLaunchKernelGGL(k_band_scatter, dim3(std::max(1, std::min(div_up(maxnnz, 256), 4096 / nd + 1)), nd), blk,
                           0, s, dd.p);
        auto factor = NB + bmax <= 256 ? k_band_factor<1> : k_band_factor<2>;
        PG_CHECK(hipFuncSetAttribute((const void *)factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fl));
        hipLaunchKernelGGL(factor, dim3(nd), blk, fl, s, dd.p);
    }
    PG_CHECK(hipFuncSetAttribute((const void *)k_band_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl));
    hipLaunchKernelGGL(k_band_solve, dim3(div_up(emax, CG), nd), blk, sl, s, dd.p);
    PG_CHECK(hipGetLastError());
    PG_CHECK(hipStreamSynchronize(s));  // the descriptors' host copy lives until here
    stats.assign(cs.size(), pgert_solve_stats{});
    for (size_t p = 0; p < cs.size(); ++p) {
        pgert_solve_stats & st = stats[p];
        st.max_rel_residual = residual_check(*cs[p]);
        st.converged = st.max_rel_residual <= o.rtol * 1.0001 ? 1 : 0;
    }
}

}  // namespace pgert

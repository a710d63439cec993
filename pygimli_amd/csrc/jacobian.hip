// Sensitivity (Jacobian) kernel: createSensitivityCol / calc2
// (core/src/bert/bertJacobian.cpp:170-219, :367-521) re-formulated per cell.
//
// The reference's entry for datum i = (a, b, m, n) and cell c is
//   sum_k w_k (u_a,k - u_b,k)_c^T (K_c + k^2 M_c) (u_m,k - u_n,k)_c
// over the cell's nodal potentials.  For P1 elements it is evaluated in one of
// two equivalent feature forms:
//  - gradient form (every wavenumber 0, i.e. 3D): K_c = vol * B B^T, so the
//    entry is vol * w * (g_a - g_b) . (g_m - g_n) with g_e = grad(u_e) on the
//    cell (DIM features per electrode, computed in FP64);
//  - nodal form (2.5D): features f_0 = u_0, f_v = u_v - u_0 (v >= 1, FP64
//    differences before rounding, so the stiffness part keeps its digits when
//    the potential is nearly constant over the cell) and the entry is
//    (f_a - f_b)^T Q' (f_m - f_n), Q' = T^T w (K_c + k^2 M_c) T.
// A workgroup owns a tile of TILE cells (sorted by marker, so a tile writes
// contiguous Jacobian columns) of one problem: it stages the features of every
// electrode of the pass's wavenumbers in LDS (layout [k][cell][comp][electrode],
// odd cell stride: conflict-free both when lanes run over electrodes (staging)
// and over cells (sweep); electrode slot nE is zero and stands for index -1),
// then sweeps the data, each lane one cell, UNR data per step.  Wavenumbers that
// do not fit into one LDS tile are done in passes that accumulate into J.
// One launch covers the tiles of every problem of a batch (per-problem
// descriptors, located by binary search on the first tile).
#include "internal.h"

#include <algorithm>
#include <type_traits>

namespace pgert {

namespace {

#ifndef PGERT_SENS_LDS
#define PGERT_SENS_LDS (80 * 1024)
#endif
#ifndef PGERT_SENS_UNR
#define PGERT_SENS_UNR 4
#endif
constexpr int KPMAX = 4;                   // wavenumbers per pass
constexpr int LDS_BUDGET = PGERT_SENS_LDS;  // feature tile bytes per workgroup
constexpr int UNR = PGERT_SENS_UNR;         // data per sweep step

enum { F_GRAD = 0, F_NODAL = 1 };

struct SensDesc {
    int nE, nEp, cs;            // electrodes, padded electrode count (slot nE = zero), cell stride
    int D, nSorted, nParam;     // data, cells with marker >= 0, Jacobian columns
    int tile0, nTile;           // first global tile of this problem in the launch, its tile count
    int seg, scale;             // shared markers (segmented sums over a tile's cells); apply k_i / m_j^2 on the last pass
    int first, last, nk;        // first / last pass of this problem; wavenumbers in this pass
    int accum;                  // add to the values J holds on entry (every pass reads and adds)
    const int * tiles;          // seg: per tile {first sorted cell, cell count, chunk slot or -1, 0}
    void * part;                // seg: partial sums [chunk][D] of markers with more cells than a tile
    const int * sorted;
    const int * cells;
    const int * marker;
    const int * abmn;
    const double * grads;
    const double * vol;
    const double * kfac;
    const double * model;
    const double * U[KPMAX];    // node-major potentials [n][nE] of each wavenumber of the pass
    double k2[KPMAX], wk[KPMAX];
    void * J;
};

template <typename P, typename T> __device__ inline void store_nt(P p, T v) { __builtin_nontemporal_store(v, p); }

template <typename OutT, typename ST, int DIM, int FORM, int TILE>
__global__ void __launch_bounds__(256) k_sens(const SensDesc * __restrict__ descs, int count) {
    constexpr int NV = DIM + 1;
    constexpr int NCOMP = FORM == F_GRAD ? DIM : NV;
    constexpr int NSUB = 64 / TILE;
    constexpr int NQ = FORM == F_GRAD ? 1 : NV * (NV + 1) / 2;
    using CT = ST;  // arithmetic of the sweep
    extern __shared__ unsigned char smem[];
    __shared__ int scell[TILE];
    __shared__ int smk[TILE];
    __shared__ int snode[TILE * NV];
    __shared__ double sgrad[FORM == F_GRAD ? TILE * NV * DIM : 1];

    const int g = blockIdx.x;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (descs[mid].tile0 <= g) lo = mid;
        else hi = mid - 1;
    }
    const SensDesc & d = descs[lo];
    const int tile = g - d.tile0;
    if (tile < 0 || tile >= d.nTile) return;
    // shared markers: tiles hold whole markers (or one chunk of a marker with more cells than a tile)
    int base, ncell, chunk = -1;
    if (d.seg) {
        base = glob(d.tiles)[4 * tile];
        ncell = glob(d.tiles)[4 * tile + 1];
        chunk = glob(d.tiles)[4 * tile + 2];
    } else {
        base = tile * TILE;
        ncell = min(TILE, d.nSorted - base);
    }
    const int nE = d.nE, nEp = d.nEp, CS = d.cs, nk = d.nk;
    ST * S = (ST *)smem;

    for (int t = threadIdx.x; t < TILE; t += blockDim.x) {
        const int c = t < ncell ? glob(d.sorted)[base + t] : -1;
        scell[t] = c;
        smk[t] = c >= 0 ? glob(d.marker)[c] : -1;
        for (int v = 0; v < NV; ++v) snode[t * NV + v] = c >= 0 ? glob(d.cells)[(size_t)c * NV + v] : 0;
        if (FORM == F_GRAD)
            for (int v = 0; v < NV; ++v)
                for (int q = 0; q < DIM; ++q)
                    sgrad[(t * NV + v) * DIM + q] = c >= 0 ? glob(d.grads)[((size_t)c * NV + v) * 3 + q] : 0.0;
    }
    // zero slots: electrodes nE .. nEp-1 of every (wavenumber, cell, component)
    for (int idx = threadIdx.x; idx < nk * TILE * NCOMP; idx += blockDim.x) {
        const int row = idx;  // (kp * TILE + t) * NCOMP + q
        const int kp = row / (TILE * NCOMP), r = row % (TILE * NCOMP);
        const int t = r / NCOMP, q = r % NCOMP;
        ST * dst = S + (size_t)(kp * TILE + t) * CS + q * nEp;
        for (int e = nE; e < nEp; ++e) dst[e] = ST(0);
    }
    __syncthreads();

    // ---- stage features: lanes run over (cell, electrode) pairs, electrode fastest
    const float rinv = 1.0f / (float)nE;
    for (int kp = 0; kp < nk; ++kp) {
        const auto Uk = glob(d.U[kp]);
#pragma unroll 4
        for (int idx = threadIdx.x; idx < TILE * nE; idx += blockDim.x) {
            const int t = (int)(((float)idx + 0.5f) * rinv);  // exact: idx < 2^15, frac >= 0.5/nE
            const int e = idx - t * nE;
            const int c = scell[t];
            double u[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) u[v] = Uk[(size_t)snode[t * NV + v] * nE + e];  // node 0 for empty cells
            ST * dst = S + (size_t)(kp * TILE + t) * CS + e;
            if (FORM == F_GRAD) {
#pragma unroll
                for (int q = 0; q < DIM; ++q) {
                    double gq = 0.0;
#pragma unroll
                    for (int v = 0; v < NV; ++v) gq += u[v] * sgrad[(t * NV + v) * DIM + q];
                    dst[q * nEp] = (ST)gq;
                }
            } else {
                dst[0] = (ST)u[0];
#pragma unroll
                for (int v = 1; v < NV; ++v) dst[v * nEp] = (ST)(u[v] - u[0]);
            }
        }
    }
    __syncthreads();

    // ---- sweep the data
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane % TILE, sub = lane / TILE;
    // Shared markers: the TILE lanes of a sub-group hold consecutive cells of the
    // marker-sorted order; a segmented scan over the lanes sums each marker's
    // cells in a fixed order and the marker's last lane stores the sum.  Lanes
    // of empty tile slots take part in the scan with a zero value.
    const bool seg = d.seg != 0;
    int c = scell[t];
    const bool live = c >= 0;
    if (!live) {
        if (!seg) return;  // no barrier below
        c = scell[0];
    }
    int segStart = t;
    bool tail = live;
    if (seg && live) {
        const int mk = smk[t];
        while (segStart > 0 && smk[segStart - 1] == mk) --segStart;
        tail = t + 1 >= ncell || smk[t + 1] != mk;
    }
    const double V = glob(d.vol)[c];
    const int col = glob(d.marker)[c];
    const double m2 = d.scale ? glob(d.model)[col] * glob(d.model)[col] : 1.0;

    // per-wavenumber coefficients of the quadratic form
    CT coef[KPMAX][NQ];
    if (FORM == F_GRAD) {
#pragma unroll
        for (int kp = 0; kp < KPMAX; ++kp) coef[kp][0] = (CT)(kp < nk ? V * d.wk[kp] : 0.0);
    } else {
        double Sm[NV][NV];  // grad N_v . grad N_w
        double gr[NV][3];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int q = 0; q < 3; ++q) gr[v][q] = glob(d.grads)[((size_t)c * NV + v) * 3 + q];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int w = 0; w < NV; ++w) Sm[v][w] = gr[v][0] * gr[w][0] + gr[v][1] * gr[w][1] + gr[v][2] * gr[w][2];
#pragma unroll
        for (int kp = 0; kp < KPMAX; ++kp) {
            const double wk = kp < nk ? d.wk[kp] : 0.0, k2 = kp < nk ? d.k2[kp] : 0.0;
            const double mf = k2 * V / (NV * (NV + 1));
            double Q[NV][NV];  // w (K_c + k^2 M_c); M_c = vol / ((d+1)(d+2)) (1 + delta_vw)
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int w = 0; w < NV; ++w) Q[v][w] = wk * (V * Sm[v][w] + mf * (v == w ? 2.0 : 1.0));
            // Q' = T^T Q T with u_0 = f_0, u_v = f_0 + f_v
            int qi = 0;
#pragma unroll
            for (int j = 0; j < NV; ++j)
#pragma unroll
                for (int l = j; l < NV; ++l) {
                    double s = 0.0;
                    if (j == 0 && l == 0) {
#pragma unroll
                        for (int v = 0; v < NV; ++v)
#pragma unroll
                            for (int w = 0; w < NV; ++w) s += Q[v][w];
                    } else if (j == 0) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) s += Q[v][l];
                    } else {
                        s = Q[j][l];
                    }
                    coef[kp][qi++] = (CT)s;
                }
        }
    }

    const int nwork = (blockDim.x >> 6) * NSUB;
    const int D = d.D;
    const int step = nwork * UNR;
    typedef int iv4 __attribute__((ext_vector_type(4)));
    const auto ab4 = glob((const iv4 *)d.abmn);
    const auto kf = glob(d.kfac);
    // destination of this lane's sums: column `col` of J, or the chunk's row of partial sums
    OutT * dst0 = chunk >= 0 ? (OutT *)d.part + (size_t)chunk * d.D : (OutT *)d.J + col;
    const auto J = glob(dst0);
    const size_t ldj = chunk >= 0 ? 1 : (size_t)d.nParam;
    const ST * Sc = S + (size_t)t * CS;
    // createJacobian_ scaling J_ij *= k_i / m_j^2 (dcfemmodelling.cpp:1375-1384), on the last pass
    // (chunk sums are scaled when they are added up)
    const bool scl = d.last && d.scale && chunk < 0;
    const CT rm2 = (CT)(1.0 / m2);
    // MODE 0: first pass, store; 1: later pass, add to the stored partial sums.
    // A step's indices, geometric factors and partial sums are loaded one step ahead,
    // before the previous step's stores: on this architecture a wait for a load also
    // waits for the stores issued before it, so loads issued after the stores would
    // put the store latency on every step.
    auto sweep = [&](auto mode_c, auto seg_c) {
        constexpr int MODE = decltype(mode_c)::value;
        constexpr bool SEG = decltype(seg_c)::value;
        int i0 = (wave * NSUB + sub) * UNR;
        iv4 qn[UNR];
        CT kn[UNR];
        OutT pn[UNR];
        auto fetch = [&](int i) {
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int ii = min(i + u, D - 1);
                qn[u] = ab4[ii];
                kn[u] = (CT)kf[ii];
                if (MODE == 1) pn[u] = J[(size_t)ii * ldj];
            }
        };
        fetch(i0);
        for (; i0 < D; i0 += step) {
            int ia[UNR], ib[UNR], im[UNR], in[UNR];
            CT fac[UNR];
            OutT prev[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                ia[u] = qn[u].x >= 0 ? qn[u].x : nE;
                ib[u] = qn[u].y >= 0 ? qn[u].y : nE;
                im[u] = qn[u].z >= 0 ? qn[u].z : nE;
                in[u] = qn[u].w >= 0 ? qn[u].w : nE;
                fac[u] = scl ? kn[u] * rm2 : CT(1);
                prev[u] = MODE == 1 ? pn[u] : OutT(0);
            }
            fetch(i0 + step);
            CT val[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) val[u] = CT(0);
#pragma unroll
            for (int kp = 0; kp < KPMAX; ++kp) {
                if (kp >= nk) break;
                const ST * Sk = Sc + (size_t)kp * TILE * CS;
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    CT dv[NCOMP], mv[NCOMP];
#pragma unroll
                    for (int q = 0; q < NCOMP; ++q) {
                        const ST * r = Sk + q * nEp;
                        dv[q] = (CT)r[ia[u]] - (CT)r[ib[u]];
                        mv[q] = (CT)r[im[u]] - (CT)r[in[u]];
                    }
                    if (FORM == F_GRAD) {
                        CT sum = CT(0);
#pragma unroll
                        for (int q = 0; q < DIM; ++q) sum += dv[q] * mv[q];
                        val[u] += coef[kp][0] * sum;
                    } else {
                        int qi = 0;
                        CT y[NV];
#pragma unroll
                        for (int j = 0; j < NV; ++j) y[j] = CT(0);
#pragma unroll
                        for (int j = 0; j < NV; ++j)
#pragma unroll
                            for (int l = j; l < NV; ++l) {
                                const CT qq = coef[kp][qi++];
                                y[j] += qq * mv[l];
                                if (l != j) y[l] += qq * mv[j];
                            }
                        CT sum = CT(0);
#pragma unroll
                        for (int j = 0; j < NV; ++j) sum += dv[j] * y[j];
                        val[u] += sum;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                CT v = val[u];
                if (SEG) {
                    // segmented inclusive scan over the sub-group's TILE lanes (fixed order)
#pragma unroll
                    for (int off = 1; off < TILE; off <<= 1) {
                        const CT o = __shfl_up(v, off, TILE);
                        if (t - off >= segStart) v += o;
                    }
                }
                const int i = i0 + u;
                if (i < D && (!SEG || tail)) {
                    const size_t idx = (size_t)i * ldj;
                    const OutT r = (OutT)((v + (CT)prev[u]) * fac[u]);
                    if (d.last) store_nt(J + idx, r);
                    else J[idx] = r;
                }
            }
        }
    };
    // chunk partials start from zero on the first pass; J itself is read back when accumulating
    const bool addPrev = !d.first || (d.accum && chunk < 0);
    if (seg) {
        if (addPrev) sweep(std::integral_constant<int, 1>{}, std::true_type{});
        else sweep(std::integral_constant<int, 0>{}, std::true_type{});
    } else {
        if (addPrev) sweep(std::integral_constant<int, 1>{}, std::false_type{});
        else sweep(std::integral_constant<int, 0>{}, std::false_type{});
    }
}

// Sum of the chunk partials of every marker with more cells than a tile, in
// chunk order, scaled by k_i / m_j^2: J[i][marker] for big[g] = {marker,
// first chunk, chunk count}.
template <typename OutT>
__global__ void k_chunk_sum(int nBig, int D, int nParam, const int * __restrict__ big,
                            const OutT * __restrict__ part, int scale, int accum, const double * __restrict__ kfac,
                            const double * __restrict__ model, OutT * __restrict__ J) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nBig * D) return;
    const int g = (int)(t / D), i = (int)(t % D);
    const int mk = big[3 * g], c0 = big[3 * g + 1], nc = big[3 * g + 2];
    OutT s = OutT(0);
    for (int ch = 0; ch < nc; ++ch) s += part[(size_t)(c0 + ch) * D + i];
    if (scale) s = (OutT)((double)s * kfac[i] / (model[mk] * model[mk]));
    const size_t idx = (size_t)i * nParam + mk;
    J[idx] = accum ? J[idx] + s : s;
}

// Tiles of the marker-sorted cells for shared markers: whole markers packed
// into tiles of at most `tile` cells, a marker with more cells than a tile
// split into chunks of its own tiles whose sums go to a partial buffer
// [chunk][D] that k_chunk_sum adds up.  Kept per context for the tile width.
void seg_tiles(Ctx & c, int tile, size_t elem) {
    if (c.segTile != tile) {
        std::vector<int> tl, big;
        int nChunk = 0, curStart = 0, curCount = 0;
        auto flush = [&] {
            if (curCount > 0) tl.insert(tl.end(), {curStart, curCount, -1, 0});
            curCount = 0;
        };
        const int ns = (int)c.hSortedMarker.size();
        for (int g0 = 0; g0 < ns;) {
            int g1 = g0 + 1;
            while (g1 < ns && c.hSortedMarker[g1] == c.hSortedMarker[g0]) ++g1;
            const int len = g1 - g0;
            if (len > tile) {
                flush();
                const int nch = div_up(len, tile);
                big.insert(big.end(), {c.hSortedMarker[g0], nChunk, nch});
                for (int ch = 0; ch < nch; ++ch)
                    tl.insert(tl.end(), {g0 + ch * tile, std::min(tile, len - ch * tile), nChunk++, 0});
            } else {
                if (curCount + len > tile) flush();
                if (curCount == 0) curStart = g0;
                curCount += len;
            }
            g0 = g1;
        }
        flush();
        c.segTiles.upload(tl, c.stream);
        c.segBig.upload(big, c.stream);
        c.segNTile = (int)tl.size() / 4;
        c.segNBig = (int)big.size() / 3;
        c.segNChunk = nChunk;
        c.segTile = tile;
    }
    c.segPart.reserve((size_t)c.segNChunk * c.D * elem);
}

struct Plan {
    int tile = 64, kp = 1, form = F_NODAL;
    size_t lds = 0;
};

int odd_stride(int ncomp, int nEp) {
    int cs = ncomp * nEp;
    return (cs & 1) ? cs : cs + 1;
}

// Tile width and wavenumbers per pass for the largest electrode count of a
// launch, by measurement on the 3D (1 pass) and 2.5D (11 wavenumbers)
// workloads: fewest passes first (each pass after the first is a
// read-modify-write sweep of J), then a tile of at most LDS_SMALL bytes
// (four and more workgroups per CU), then 32- over 16-cell tiles.
constexpr int LDS_SMALL = 40 * 1024;
Plan make_plan(int dim, int nEmax, int nK, bool allZeroK, size_t st) {
    Plan p;
    p.form = allZeroK ? F_GRAD : F_NODAL;
    const int ncomp = p.form == F_GRAD ? dim : dim + 1;
    const int cs = odd_stride(ncomp, nEmax + 1);
    long best = -1;
    for (int tile : {64, 32, 16}) {
        const size_t perk = (size_t)tile * cs * st;
        int kp = (int)std::min<size_t>(std::min(KPMAX, nK), LDS_BUDGET / perk);
        int over = 0;
        if (kp < 1) {
            if (perk > 150 * 1024) continue;
            kp = 1;   // one wavenumber beyond the budget: one workgroup per CU
            over = 1;
        }
        const long passes = (nK + kp - 1) / kp;
        const size_t lds = (size_t)kp * perk;
        const long score = passes * 16 + over * 8 + (lds > LDS_SMALL ? 4 : 0) + (tile == 16 ? 2 : 0);
        if (best < 0 || score < best) {
            best = score;
            p.tile = tile;
            p.kp = kp;
            p.lds = lds;
        }
    }
    PG_REQUIRE(best >= 0, "pgert_jacobian: too many electrodes for one LDS tile");
    return p;
}

template <typename OutT, typename ST, int DIM, int FORM>
void launch_k(const Plan & p, const SensDesc * descs, int count, int grid, hipStream_t s) {
    auto go = [&](auto kern) {
        PG_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), p.lds, s, descs, count);
        PG_CHECK(hipGetLastError());
    };
    if (p.tile == 64) go(k_sens<OutT, ST, DIM, FORM, 64>);
    else if (p.tile == 32) go(k_sens<OutT, ST, DIM, FORM, 32>);
    else go(k_sens<OutT, ST, DIM, FORM, 16>);
}

template <typename OutT>
void run(const std::vector<Ctx *> & cs, const std::vector<const double *> & models, const std::vector<int> & nModels,
         const std::vector<void *> & Js, bool accum) {
    Ctx & c0 = *cs[0];
    hipStream_t s = c0.stream;
    const int count = (int)cs.size();
    int nEmax = 0, nKmax = 0;
    bool allZeroK = true;
    for (Ctx * c : cs) {
        PG_REQUIRE(c->dim == c0.dim, "pgert_batch_jacobian: problems of different dimension");
        nEmax = std::max(nEmax, c->nE);
        nKmax = std::max(nKmax, c->nK);
        for (double k : c->kv) allZeroK = allZeroK && k == 0.0;
    }
    const Plan p = make_plan(c0.dim, nEmax, nKmax, allZeroK, sizeof(OutT));
    const int npass = (nKmax + p.kp - 1) / p.kp;
    const int ncomp = p.form == F_GRAD ? c0.dim : c0.dim + 1;

    std::vector<char> segP(count), scaleP(count);
    for (int j = 0; j < count; ++j) {
        Ctx & c = *cs[j];
        segP[j] = !c.uniqueMarkers;
        scaleP[j] = nModels[j] == c.nParam;
        if (segP[j]) seg_tiles(c, p.tile, sizeof(OutT));
        const size_t jsize = (size_t)c.D * c.nParam;
        // columns without any cell stay zero (markers not covering 0..max)
        if (c.nDistinct < c.nParam && !accum) PG_CHECK(hipMemsetAsync(Js[j], 0, jsize * sizeof(OutT), s));
    }
    // descriptors of every pass, uploaded once
    std::vector<SensDesc> hd;
    std::vector<int> passStart(npass + 1, 0), passGrid(npass, 0);
    for (int ps = 0; ps < npass; ++ps) {
        passStart[ps] = (int)hd.size();
        int tile0 = 0;
        for (int j = 0; j < count; ++j) {
            Ctx & c = *cs[j];
            const int k0 = ps * p.kp, k1 = std::min(c.nK, k0 + p.kp);
            if (k0 >= k1 || c.nSorted == 0 || c.D == 0) continue;
            SensDesc sd{};
            sd.nE = c.nE;
            sd.nEp = c.nE + 1;
            sd.cs = odd_stride(ncomp, sd.nEp);
            sd.D = c.D;
            sd.nSorted = c.nSorted;
            sd.nParam = c.nParam;
            sd.tile0 = tile0;
            sd.nTile = segP[j] ? c.segNTile : div_up(c.nSorted, p.tile);
            sd.seg = segP[j];
            sd.tiles = segP[j] ? c.segTiles.p : nullptr;
            sd.part = segP[j] ? (void *)c.segPart.p : nullptr;
            sd.scale = scaleP[j];
            sd.first = k0 == 0;
            sd.accum = accum;
            sd.last = k1 == c.nK;
            sd.nk = k1 - k0;
            sd.sorted = c.sortedCells.p;
            sd.cells = c.cells.p;
            sd.marker = c.cellMarker.p;
            sd.abmn = c.abmn.p;
            sd.grads = c.grads.p;
            sd.vol = c.vol.p;
            sd.kfac = c.kfac.p;
            sd.model = models[j];
            for (int kk = k0; kk < k1; ++kk) {
                sd.U[kk - k0] = c.U.p + (size_t)kk * c.n * c.nE;
                sd.k2[kk - k0] = c.kv[kk] * c.kv[kk];
                sd.wk[kk - k0] = c.wv[kk];
            }
            sd.J = Js[j];
            tile0 += sd.nTile;
            hd.push_back(sd);
        }
        passGrid[ps] = tile0;
    }
    passStart[npass] = (int)hd.size();
    if (hd.empty()) return;
    c0.jdescHost.assign((const unsigned char *)hd.data(), (const unsigned char *)(hd.data() + hd.size()));
    c0.jdesc.upload(c0.jdescHost, s);
    const SensDesc * dd = (const SensDesc *)c0.jdesc.p;
    for (int ps = 0; ps < npass; ++ps) {
        const int n = passStart[ps + 1] - passStart[ps];
        if (n == 0 || passGrid[ps] == 0) continue;
        const SensDesc * dp = dd + passStart[ps];
        if (c0.dim == 2) {
            if (p.form == F_GRAD) launch_k<OutT, OutT, 2, F_GRAD>(p, dp, n, passGrid[ps], s);
            else launch_k<OutT, OutT, 2, F_NODAL>(p, dp, n, passGrid[ps], s);
        } else {
            if (p.form == F_GRAD) launch_k<OutT, OutT, 3, F_GRAD>(p, dp, n, passGrid[ps], s);
            else launch_k<OutT, OutT, 3, F_NODAL>(p, dp, n, passGrid[ps], s);
        }
    }
    for (int j = 0; j < count; ++j) {
        Ctx & c = *cs[j];
        if (segP[j] && c.segNBig > 0 && c.D > 0) {
            const long total = (long)c.segNBig * c.D;
            hipLaunchKernelGGL(k_chunk_sum<OutT>, dim3(div_up(total, 256)), dim3(256), 0, s, c.segNBig, c.D,
                               c.nParam, c.segBig.p, (const OutT *)c.segPart.p, (int)scaleP[j], (int)accum, c.kfac.p,
                               models[j], (OutT *)Js[j]);
            PG_CHECK(hipGetLastError());
        }
    }
}

}  // namespace

void launch_jacobian_batch(const std::vector<Ctx *> & cs, const std::vector<const double *> & models,
                           const std::vector<int> & nModels, const std::vector<void *> & Js, int dtype) {
    if (dtype == 0) run<double>(cs, models, nModels, Js, false);
    else run<float>(cs, models, nModels, Js, false);
}

void launch_jacobian(Ctx & c, const double * model, int nModel, void * J, int dtype, bool accumulate) {
    if (dtype == 0) run<double>({&c}, {model}, {nModel}, {J}, accumulate);
    else run<float>({&c}, {model}, {nModel}, {J}, accumulate);
}

}  // namespace pgert

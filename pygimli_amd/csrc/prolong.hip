// Background (marker -1) cells of a model mapping: Mesh::prolongateEmptyCellsValues
// (core/src/mesh.cpp:2247-2318) as called by ModellingBase::createMappedModel
// with the default background -9e99 (modellingbase.cpp:449-457;
// DCMultiElectrodeModelling::response(model) passes it, dcfemmodelling.h:164).
//
// The reference fills the empty cells in sweeps: every empty cell with at
// least one filled neighbour (neighbour j = the cell across the face opposite
// node j) takes the weighted mean of its filled neighbours, weight
// |n * (1, 1, 0)| + 1e-6 (3D) or |n_x| + 1e-6 (2D) with n the unit normal of
// the common face (horizontal prolongation), all values of a sweep computed
// from the state before it; sweeps repeat until no cell is empty.  Which cells
// a sweep fills and with which weights depends only on the mesh and the cell
// markers (model values are > 0 for a response), so the sweeps are planned
// once per mesh on the host and applied per evaluation on the device, one
// launch per sweep, each cell summing its neighbours in the reference's order.
#include "internal.h"

#include <algorithm>
#include <array>
#include <cmath>

namespace pgert {

namespace {

__global__ void k_prolong(int count, const int * __restrict__ cell, const int * __restrict__ ptr,
                          const int * __restrict__ nbr, const double * __restrict__ w, double * __restrict__ rho) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double val = 0.0, weight = 0.0;
    for (int q = ptr[t]; q < ptr[t + 1]; ++q) {
        val += rho[nbr[q]] * w[q];
        weight += w[q];
    }
    rho[cell[t]] = val / weight;
}

}  // namespace

void build_prolongation(Ctx & c, const double * xyz, const int * cells, const int * marker) {
    c.prolLayerPtr.clear();
    c.prolStatus = 0;
    const int C = c.C, nv = c.nv, dim = c.dim;
    std::vector<char> filled(C);
    int empty = 0;
    for (int i = 0; i < C; ++i) {
        if (marker[i] <= -1000000) c.prolStatus = 2;  // fixed-value region (MARKER_FIXEDVALUE_REGION)
        filled[i] = marker[i] >= 0;
        empty += !filled[i];
    }
    if (empty == 0 || c.prolStatus) return;
    if (empty == C) {
        c.prolStatus = 3;
        return;
    }
    // neighbour across the face opposite node j, and that face's weight
    std::vector<std::array<int, 4>> key((size_t)C * nv);
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < nv; ++j) {
            std::array<int, 4> k{INT32_MAX, INT32_MAX, INT32_MAX, 0};
            int q = 0;
            for (int v = 0; v < nv; ++v)
                if (v != j) k[q++] = cells[(size_t)i * nv + v];
            std::sort(k.begin(), k.begin() + dim);
            k[3] = i * nv + j;
            key[(size_t)i * nv + j] = k;
        }
    std::sort(key.begin(), key.end(), [](const std::array<int, 4> & a, const std::array<int, 4> & b) {
        return a[0] != b[0] ? a[0] < b[0] : a[1] != b[1] ? a[1] < b[1] : a[2] != b[2] ? a[2] < b[2] : a[3] < b[3];
    });
    std::vector<int> nb((size_t)C * nv, -1);
    for (size_t q = 0; q + 1 < key.size(); ++q) {
        const auto & a = key[q];
        const auto & b = key[q + 1];
        if (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]) {
            // a face shared by exactly two cells (Cell::findNeighborCell keeps a unique other cell)
            const bool three = q + 2 < key.size() && key[q + 2][0] == a[0] && key[q + 2][1] == a[1] &&
                               key[q + 2][2] == a[2];
            if (!three) {
                nb[a[3]] = b[3] / nv;
                nb[b[3]] = a[3] / nv;
            }
        }
    }
    auto weight = [&](int i, int j) {
        const double * p[3];
        int q = 0;
        for (int v = 0; v < nv; ++v)
            if (v != j) p[q++] = xyz + 3 * (size_t)cells[(size_t)i * nv + v];
        if (dim == 2) {  // RVector3::normXY of the edge: (-dy, dx) / |.|, weight |n_x|
            const double dx = p[0][0] - p[1][0], dy = p[0][1] - p[1][1];
            return std::fabs(-dy) / std::sqrt(dx * dx + dy * dy) + 1e-6;
        }
        const double a[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double b[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        const double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        const double nx = n[0] / len, ny = n[1] / len;
        return std::sqrt(nx * nx + ny * ny) + 1e-6;
    };
    std::vector<int> lcell, lptr{0}, lnbr;
    std::vector<double> lw;
    c.prolLayerPtr.push_back(0);
    while (empty > 0) {
        const size_t start = lcell.size();
        for (int i = 0; i < C; ++i) {
            if (filled[i]) continue;
            double wsum = 0.0;
            const size_t q0 = lnbr.size();
            for (int j = 0; j < nv; ++j) {
                const int o = nb[(size_t)i * nv + j];
                if (o < 0 || !filled[o]) continue;
                const double w = weight(i, j);
                lnbr.push_back(o);
                lw.push_back(w);
                wsum += w;
            }
            if (wsum > 1e-8) {
                lcell.push_back(i);
                lptr.push_back((int)lnbr.size());
            } else {
                lnbr.resize(q0);
                lw.resize(q0);
            }
        }
        if (lcell.size() == start) {  // isolated empty cells: the reference gives up (mesh.cpp:2304-2310)
            c.prolStatus = 4;
            c.prolLayerPtr.clear();
            return;
        }
        for (size_t q = start; q < lcell.size(); ++q) filled[lcell[q]] = 1;
        empty -= (int)(lcell.size() - start);
        c.prolLayerPtr.push_back((int)lcell.size());
    }
    c.prolCell.upload(lcell, c.stream);
    c.prolPtr.upload(lptr, c.stream);
    c.prolNbr.upload(lnbr, c.stream);
    c.prolW.upload(lw, c.stream);
}

void check_prolongation(const Ctx & c) {
    PG_REQUIRE(c.prolStatus != 2, "map model: fixed-value regions (marker <= -1000000) need pyGIMLi's region "
                                  "manager, which this library does not provide");
    PG_REQUIRE(c.prolStatus != 3, "map model: too many empty cells (every cell has a marker < 0)");
    PG_REQUIRE(c.prolStatus != 4, "map model: cannot fill the empty (marker < 0) cells: some have no path "
                                  "of face neighbours to a parameter cell");
}

void launch_prolongation(Ctx & c, double * rho) {
    check_prolongation(c);
    for (size_t l = 0; l + 1 < c.prolLayerPtr.size(); ++l) {
        const int a = c.prolLayerPtr[l], b = c.prolLayerPtr[l + 1];
        hipLaunchKernelGGL(k_prolong, dim3(div_up(b - a, 256)), dim3(256), 0, c.stream, b - a, c.prolCell.p + a,
                           c.prolPtr.p + a, c.prolNbr.p, c.prolW.p, rho);
        PG_CHECK(hipGetLastError());
    }
}

}  // namespace pgert

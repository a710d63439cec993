// C-ABI entry points of libpgert (declared in include/pgert.h) and the host
// side set-up they need: CSR pattern, atomics-free assembly map, boundary
// and electrode tables.
#include "internal.h"

#include <cstdio>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>

using namespace pgert;

struct pgert_ctx : public Ctx {};

namespace {

template <class F> int guarded(pgert_ctx * c, F f) {
    if (!c) return -2;
    try {
        PG_CHECK(hipSetDevice(c->device));
        f();
        return 0;
    } catch (std::exception & e) {
        c->err = e.what();
        return -1;
    } catch (...) {
        c->err = "unknown error";
        return -1;
    }
}

// Sorted neighbour lists of every node (SparseMatrix::buildSparsityPattern,
// sparsematrix.h:966-1032: all node pairs sharing a cell, sorted per row),
// plus the n_couple node pairs (ca[i], cb[i]) in both directions and their
// diagonals: couplings that no cell carries, such as bypass resistors between
// distant nodes (the reference inserts them into a map matrix,
// dcfemmodelling.cpp:695-707) or the electrode unknowns of the complete
// electrode model (:315-410), which belong to no cell.
void build_pattern(Ctx & c, const int * cells, int n_couple, const int * ca, const int * cb) {
    const int n = c.n, C = c.C, nv = c.nv;
    std::vector<int> cnt(n + 1, 0);
    for (long e = 0; e < (long)C * nv; ++e) cnt[cells[e] + 1] += nv;
    for (int i = 0; i < n_couple; ++i) {
        cnt[ca[i] + 1] += 2;
        cnt[cb[i] + 1] += 2;
    }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> cand(cnt[n]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int cc = 0; cc < C; ++cc) {
        const int * v = cells + (long)cc * nv;
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j) cand[fill[v[i]]++] = v[j];
    }
    for (int i = 0; i < n_couple; ++i) {
        cand[fill[ca[i]]++] = cb[i];
        cand[fill[ca[i]]++] = ca[i];
        cand[fill[cb[i]]++] = ca[i];
        cand[fill[cb[i]]++] = cb[i];
    }
    c.hRowptr.assign(n + 1, 0);
    c.hColidx.clear();
    c.hColidx.reserve(cnt[n] / 2);
    for (int i = 0; i < n; ++i) {
        auto b = cand.begin() + cnt[i], e = cand.begin() + cnt[i + 1];
        std::sort(b, e);
        auto u = std::unique(b, e);
        c.hColidx.insert(c.hColidx.end(), b, u);
        c.hRowptr[i + 1] = (int)c.hColidx.size();
    }
    c.nnz = (int)c.hColidx.size();
}

inline int find_slot(const Ctx & c, int row, int col) {
    const int * b = c.hColidx.data() + c.hRowptr[row];
    const int * e = c.hColidx.data() + c.hRowptr[row + 1];
    const int * p = std::lower_bound(b, e, col);
    if (p == e || *p != col) return -1;
    return (int)(p - c.hColidx.data());
}

void sync(Ctx & c) { PG_CHECK(hipStreamSynchronize(c.stream)); }

}  // namespace

extern "C" {

int pgert_create(pgert_ctx ** out, int device) {
    if (!out) return -2;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -3;
    if (device < 0 || device >= ndev) return -4;
    pgert_ctx * c = new pgert_ctx();
    c->device = device;
    int rc = guarded(c, [&] {
        for (auto & e : c->ev) PG_CHECK(hipEventCreate(&e));
        cgls_scratch_init(device);
    });
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

void pgert_destroy(pgert_ctx * c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto & e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto & e : c->solveEv)
        if (e) (void)hipEventDestroy(e);
    for (auto & e : c->kEvents) (void)hipEventDestroy(e);
    if (c->solveStream) (void)hipStreamDestroy(c->solveStream);
    delete c;
}

const char * pgert_last_error(const pgert_ctx * c) { return c ? c->err.c_str() : "null context"; }

int pgert_set_stream(pgert_ctx * c, void * stream) {
    return guarded(c, [&] { c->stream = (hipStream_t)stream; });
}

static void mesh_setup(pgert_ctx * c, int elemOrder, int dim, int n_nodes, const double * xyz, int n_cells,
                       const int * cells, const int * cell_marker, int n_bounds, const int * bounds,
                       const int * bound_marker, const int * bound_cell, int n_couple, const int * couple_a,
                       const int * couple_b);

int pgert_set_mesh(pgert_ctx * c, int dim, int n_nodes, const double * xyz, int n_cells,
                   const int * cells, const int * cell_marker, int n_bounds, const int * bounds,
                   const int * bound_marker, const int * bound_cell) {
    return pgert_set_mesh_coupled(c, dim, n_nodes, xyz, n_cells, cells, cell_marker, n_bounds, bounds,
                                  bound_marker, bound_cell, 0, nullptr, nullptr);
}

int pgert_set_mesh_coupled(pgert_ctx * c, int dim, int n_nodes, const double * xyz, int n_cells,
                           const int * cells, const int * cell_marker, int n_bounds, const int * bounds,
                           const int * bound_marker, const int * bound_cell, int n_couple,
                           const int * couple_a, const int * couple_b) {
    return guarded(c, [&] {
        mesh_setup(c, 1, dim, n_nodes, xyz, n_cells, cells, cell_marker, n_bounds, bounds, bound_marker, bound_cell,
                   n_couple, couple_a, couple_b);
    });
}

int pgert_set_mesh_p2(pgert_ctx * c, int dim, int n_nodes, const double * xyz, int n_cells, const int * cells,
                      const int * cell_marker, int n_bounds, const int * bounds, const int * bound_marker,
                      const int * bound_cell) {
    return guarded(c, [&] {
        mesh_setup(c, 2, dim, n_nodes, xyz, n_cells, cells, cell_marker, n_bounds, bounds, bound_marker, bound_cell,
                   0, nullptr, nullptr);
    });
}

// Mesh set-up of pgert_set_mesh (order 1) and pgert_set_mesh_p2 (order 2).
static void mesh_setup(pgert_ctx * c, int elemOrder, int dim, int n_nodes, const double * xyz, int n_cells,
                       const int * cells, const int * cell_marker, int n_bounds, const int * bounds,
                       const int * bound_marker, const int * bound_cell, int n_couple, const int * couple_a,
                       const int * couple_b) {
    {
        PG_REQUIRE(dim == 2 || dim == 3, "pgert_set_mesh: dim must be 2 or 3");
        PG_REQUIRE(n_nodes > 0 && n_cells > 0, "pgert_set_mesh: empty mesh");
        PG_REQUIRE(xyz && cells && cell_marker, "pgert_set_mesh: null array");
        PG_REQUIRE(n_bounds == 0 || (bounds && bound_marker && bound_cell),
                   "pgert_set_mesh: null boundary array");
        PG_REQUIRE(n_couple >= 0 && (n_couple == 0 || (couple_a && couple_b)),
                   "pgert_set_mesh_coupled: null coupling array");
        for (int i = 0; i < n_couple; ++i)
            PG_REQUIRE(couple_a[i] >= 0 && couple_a[i] < n_nodes && couple_b[i] >= 0 && couple_b[i] < n_nodes &&
                           couple_a[i] != couple_b[i],
                       "pgert_set_mesh_coupled: bad node pair " + std::to_string(i));
        c->dim = dim;
        c->order = elemOrder;
        c->nvert = dim + 1;
        c->nv = elemOrder == 1 ? dim + 1 : (dim == 2 ? 6 : 10);
        const int nbf = elemOrder == 1 ? dim : (dim == 2 ? 3 : 6);   // nodes per boundary face
        c->n = n_nodes;
        c->C = n_cells;
        for (long e = 0; e < (long)n_cells * c->nv; ++e)
            PG_REQUIRE(cells[e] >= 0 && cells[e] < n_nodes, "pgert_set_mesh: cell node index out of range");
        // parameters: cells with marker >= 0, sorted by marker (bertJacobian.cpp:398-399)
        int maxMarker = -1;
        std::vector<int> order;
        order.reserve(n_cells);
        for (int i = 0; i < n_cells; ++i) {
            maxMarker = std::max(maxMarker, cell_marker[i]);
            if (cell_marker[i] >= 0) order.push_back(i);
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return cell_marker[a] < cell_marker[b]; });
        c->nParam = maxMarker + 1;
        c->nSorted = (int)order.size();
        bool uniq = true;
        int distinct = order.empty() ? 0 : 1;
        for (size_t i = 1; i < order.size(); ++i) {
            if (cell_marker[order[i]] == cell_marker[order[i - 1]]) uniq = false;
            else ++distinct;
        }
        c->uniqueMarkers = uniq;
        c->nDistinct = distinct;
        c->hSortedMarker.resize(order.size());
        for (size_t i = 0; i < order.size(); ++i) c->hSortedMarker[i] = cell_marker[order[i]];
        c->segTile = 0;

        hipStream_t s = c->stream;
        c->hXyz.assign(xyz, xyz + (size_t)n_nodes * 3);
        c->nAgg = 0;
        c->bandValid = false;
        c->xyz.upload(xyz, (size_t)n_nodes * 3, s);
        c->cells.upload(cells, (size_t)n_cells * c->nv, s);
        c->cellMarker.upload(cell_marker, n_cells, s);
        c->sortedCells.upload(order, s);
        c->grads.alloc((size_t)n_cells * c->nvert * 3);
        c->vol.alloc(n_cells);
        launch_cell_geometry(*c);
        if (elemOrder == 1) {
            build_prolongation(*c, xyz, cells, cell_marker);
        } else {
            c->prolLayerPtr.clear();
            c->prolStatus = 0;
        }

        // pattern + contributions sorted by slot (counting sort, cell order kept)
        build_pattern(*c, cells, n_couple, couple_a, couple_b);
        const int nv = c->nv;
        std::vector<int> slotOf((size_t)n_cells * nv * nv);
        std::vector<int> cptr(c->nnz + 1, 0);
        for (int cc = 0; cc < n_cells; ++cc) {
            const int * v = cells + (long)cc * nv;
            for (int i = 0; i < nv; ++i)
                for (int j = 0; j < nv; ++j) {
                    int sl = find_slot(*c, v[i], v[j]);
                    slotOf[((size_t)cc * nv + i) * nv + j] = sl;
                    cptr[sl + 1]++;
                }
        }
        for (int i = 0; i < c->nnz; ++i) cptr[i + 1] += cptr[i];
        std::vector<int> pos(cptr.begin(), cptr.end() - 1);
        std::vector<int> ccell(cptr[c->nnz]), ccode(cptr[c->nnz]);
        for (int cc = 0; cc < n_cells; ++cc)
            for (int ij = 0; ij < nv * nv; ++ij) {
                int sl = slotOf[(size_t)cc * nv * nv + ij];
                int p = pos[sl]++;
                ccell[p] = cc;
                ccode[p] = ij;
            }
        std::vector<int> diag(n_nodes);
        for (int i = 0; i < n_nodes; ++i) {
            diag[i] = find_slot(*c, i, i);
            PG_REQUIRE(diag[i] >= 0, "pgert_set_mesh: node " + std::to_string(i) + " belongs to no cell");
        }
        c->rowptr.upload(c->hRowptr, s);
        c->colidx.upload(c->hColidx, s);
        c->diagSlot.upload(diag, s);
        c->cptr.upload(cptr, s);
        c->ccell.upload(ccell, s);
        DBuf<int> code;
        code.upload(ccode, s);
        c->cK.alloc(ccell.size());
        c->cM.alloc(ccell.size());
        std::vector<double> faceMass;   // per-measure P2 mass of a boundary face
        if (elemOrder == 1) {
            launch_contrib_geometry(*c, code.p);
        } else {
            std::vector<double> M, W, Mf, Wf;
            p2_tables(dim, M, W);
            p2_tables(dim - 1, faceMass, Wf);
            DBuf<double> dM, dW;
            dM.upload(M, s);
            dW.upload(W, s);
            launch_contrib_geometry_p2(*c, code.p, dM.p, dW.p);
            sync(*c);   // the tables are released at the end of this scope
        }
        sync(*c);  // the code table is released at the end of this scope

        // homogeneous Dirichlet boundaries (marker -3): their nodes' rows and
        // columns are cleared and the diagonal set to one after the boundary
        // terms (dcfemBoundaryAssembleStiffnessMatrix, dcfemmodelling.cpp:280-298
        // with assembleStiffnessMatrixHomogenDirichletBC :141-153); every
        // slot so set is listed once with its value
        std::vector<char> isDir(n_nodes, 0);
        for (int f = 0; f < n_bounds; ++f) {
            PG_REQUIRE(bound_marker[f] != -4, "pgert_set_mesh: inhomogeneous Dirichlet boundaries (marker -4) are "
                                              "not supported (the reference does not implement them either)");
            if (bound_marker[f] != -3) continue;
            for (int a = 0; a < nbf; ++a) {
                const int v = bounds[(long)f * nbf + a];
                PG_REQUIRE(v >= 0 && v < n_nodes, "pgert_set_mesh: boundary node out of range");
                isDir[v] = 1;
            }
        }
        {
            std::vector<std::pair<int, double>> ds;
            for (int i = 0; i < n_nodes; ++i) {
                if (!isDir[i]) continue;
                for (int q = c->hRowptr[i]; q < c->hRowptr[i + 1]; ++q) {
                    const int j = c->hColidx[q];
                    ds.push_back({q, j == i ? 1.0 : 0.0});
                    if (j != i && !isDir[j]) ds.push_back({find_slot(*c, j, i), 0.0});
                }
            }
            std::sort(ds.begin(), ds.end());
            std::vector<int> dslot(ds.size());
            std::vector<double> dval(ds.size());
            for (size_t q = 0; q < ds.size(); ++q) {
                dslot[q] = ds[q].first;
                dval[q] = ds[q].second;
            }
            c->nDirSlot = (int)ds.size();
            c->dirSlot.upload(dslot, s);
            c->dirVal.upload(dval, s);
        }

        // mixed boundaries (marker -2)
        std::vector<int> mixCell;
        std::vector<double> mixCtr, mixNrm;
        std::vector<std::array<int, 3>> bcon;  // slot, face, code (i*dim+j)
        std::vector<double> bcoef;
        for (int f = 0; f < n_bounds; ++f) {
            int mk = bound_marker[f];
            if (mk != -2) continue;
            const int * bn = bounds + (long)f * nbf;
            PG_REQUIRE(bound_cell[f] >= 0 && bound_cell[f] < n_cells,
                       "pgert_set_mesh: no cell found for a mixed boundary");
            double p[3][3];
            for (int a = 0; a < dim; ++a) {
                PG_REQUIRE(bn[a] >= 0 && bn[a] < n_nodes, "pgert_set_mesh: boundary node out of range");
                for (int q = 0; q < 3; ++q) p[a][q] = xyz[3 * (long)bn[a] + q];
            }
            double ctr[3] = {0, 0, 0}, nrm[3] = {0, 0, 0}, size;
            for (int a = 0; a < dim; ++a)
                for (int q = 0; q < 3; ++q) ctr[q] += p[a][q] / dim;
            if (dim == 2) {
                double tx = p[1][0] - p[0][0], ty = p[1][1] - p[0][1];
                size = std::sqrt(tx * tx + ty * ty);
                nrm[0] = ty / size;
                nrm[1] = -tx / size;
            } else {
                double u[3], v[3];
                for (int q = 0; q < 3; ++q) { u[q] = p[1][q] - p[0][q]; v[q] = p[2][q] - p[0][q]; }
                double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                                u[0] * v[1] - u[1] * v[0]};
                double a2 = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
                size = a2 / 2.0;
                for (int q = 0; q < 3; ++q) nrm[q] = cr[q] / a2;
            }
            int fi = (int)mixCell.size();
            mixCell.push_back(bound_cell[f]);
            for (int q = 0; q < 3; ++q) { mixCtr.push_back(ctr[q]); mixNrm.push_back(nrm[q]); }
            // boundary mass: edge len/6 (1 + d_ij), triangle area/12 (1 + d_ij);
            // P2 faces: the measure times the P2 mass table of the face
            double mb = size / (dim * (dim + 1));
            for (int i = 0; i < nbf; ++i) {
                PG_REQUIRE(bn[i] >= 0 && bn[i] < n_nodes, "pgert_set_mesh: boundary node out of range");
                for (int j = 0; j < nbf; ++j) {
                    int sl = find_slot(*c, bn[i], bn[j]);
                    PG_REQUIRE(sl >= 0, "pgert_set_mesh: boundary face is not a cell face");
                    bcon.push_back({sl, fi, i * nbf + j});
                    bcoef.push_back(elemOrder == 1 ? mb * (i == j ? 2.0 : 1.0) : size * faceMass[(size_t)i * nbf + j]);
                }
            }
        }
        c->nMixed = (int)mixCell.size();
        std::vector<int> idx(bcon.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return bcon[a][0] < bcon[b][0]; });
        std::vector<int> bslot, bptr(1, 0), bface;
        std::vector<double> bc;
        for (size_t q = 0; q < idx.size(); ++q) {
            int sl = bcon[idx[q]][0];
            if (bslot.empty() || bslot.back() != sl) {
                if (!bslot.empty()) bptr.push_back((int)bface.size());
                bslot.push_back(sl);
            }
            bface.push_back(bcon[idx[q]][1]);
            bc.push_back(bcoef[idx[q]]);
        }
        bptr.push_back((int)bface.size());
        if (bslot.empty()) bptr.assign(1, 0);
        c->nBSlot = (int)bslot.size();
        c->mixCell.upload(mixCell, s);
        c->mixCtr.upload(mixCtr, s);
        c->mixNrm.upload(mixNrm, s);
        c->bSlot.upload(bslot, s);
        c->bPtr.upload(bptr, s);
        c->bFace.upload(bface, s);
        c->bCoef.upload(bc, s);

        c->hElecNode.clear();
        c->nE = 0;
        c->nByp = 0;   // bypass, electrode-model and calibration slots index the previous pattern
        c->nCem = 0;
        c->nCalSlot = 0;
        c->calNodes.release();
        c->refNode = -1;
        c->s1Valid = c->primValid = c->haveU = c->haveUc = c->srState = false;
        c->primNumeric = false;
        sync(*c);
    }
}

// Electrodes on nodes (elec_node[e] >= 0) or free in a cell (elec_node[e] = -1,
// elec_cell[e] the cell, elec_pos the position): searchElectrodes_
// (dcfemmodelling.cpp:790-1073) with ElectrodeShapeNode / ElectrodeShapeEntity.
static void set_electrodes(pgert_ctx * c, int n_elec, const int * elec_node, const int * elec_cell,
                           const double * elec_pos, const double * source_center, double surface_z) {
    PG_REQUIRE(c->n > 0, "pgert_set_electrodes: no mesh");
    PG_REQUIRE(n_elec > 0 && elec_node && source_center, "pgert_set_electrodes: bad arguments");
    PG_REQUIRE(n_elec <= 256, "pgert_set_electrodes: at most 256 current electrodes per context");
    std::vector<double> hx((size_t)c->n * 3);
    PG_CHECK(hipMemcpy(hx.data(), c->xyz.p, hx.size() * 8, hipMemcpyDeviceToHost));
    std::vector<int> hc((size_t)c->C * c->nv);
    PG_CHECK(hipMemcpy(hc.data(), c->cells.p, hc.size() * 4, hipMemcpyDeviceToHost));
    const int dim = c->dim, nv = c->nv;
    c->nE = n_elec;
    c->virtualElec = false;
    c->hElecNode.assign(elec_node, elec_node + n_elec);
    auto dist = [&](const double * a, const double * b) {
        double d2 = 0;
        for (int q = 0; q < 3; ++q) d2 += (a[q] - b[q]) * (a[q] - b[q]);
        return std::sqrt(d2);
    };
    // the singular node of every electrode: its node, or for a free electrode
    // the last node of its cell within 1e-4 (ElectrodeShapeEntity::setSingValue,
    // electrode.cpp:207-227); shape nodes and weights; electrode positions
    std::vector<int> sing(n_elec, -1), shn((size_t)n_elec * SHAPE_NODES, -1);
    std::vector<double> shw((size_t)n_elec * SHAPE_NODES, 0.0), pos((size_t)3 * n_elec);
    for (int e = 0; e < n_elec; ++e) {
        if (elec_node[e] >= 0) {
            PG_REQUIRE(elec_node[e] < c->n, "pgert_set_electrodes: node out of range");
            sing[e] = elec_node[e];
            shn[(size_t)SHAPE_NODES * e] = elec_node[e];
            shw[(size_t)SHAPE_NODES * e] = 1.0;
            for (int q = 0; q < 3; ++q) pos[3 * e + q] = hx[3 * (size_t)elec_node[e] + q];
            continue;
        }
        PG_REQUIRE(elec_cell && elec_pos, "pgert_set_electrodes: free electrode without cell and position");
        PG_REQUIRE(c->order == 1, "pgert_set_electrodes: free electrodes need a linear mesh");
        const int cc = elec_cell[e];
        PG_REQUIRE(cc >= 0 && cc < c->C, "pgert_set_electrodes: free electrode cell out of range");
        const double * p = elec_pos + 3 * (size_t)e;
        for (int q = 0; q < 3; ++q) pos[3 * e + q] = p[q];
        // shape functions at p: barycentric coordinates of the simplex
        // (MeshEntity::N(shape().rst(p))), by Cramer's rule on the edge matrix
        const int * cn = hc.data() + (size_t)cc * nv;
        const double * x0 = hx.data() + 3 * (size_t)cn[0];
        double T[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, r[3] = {0, 0, 0}, lam[3] = {0, 0, 0};
        for (int q = 0; q < dim; ++q) {
            for (int j = 0; j < dim; ++j) T[q][j] = hx[3 * (size_t)cn[j + 1] + q] - x0[q];
            r[q] = p[q] - x0[q];
        }
        if (dim == 2) {
            const double det = T[0][0] * T[1][1] - T[0][1] * T[1][0];
            PG_REQUIRE(std::fabs(det) > 0.0, "pgert_set_electrodes: degenerate cell");
            lam[0] = (r[0] * T[1][1] - T[0][1] * r[1]) / det;
            lam[1] = (T[0][0] * r[1] - r[0] * T[1][0]) / det;
        } else {
            auto det3 = [](const double (&M)[3][3]) {
                return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
            };
            const double det = det3(T);
            PG_REQUIRE(std::fabs(det) > 0.0, "pgert_set_electrodes: degenerate cell");
            for (int j = 0; j < 3; ++j) {
                double M[3][3];
                for (int q = 0; q < 3; ++q)
                    for (int l = 0; l < 3; ++l) M[q][l] = l == j ? r[q] : T[q][l];
                lam[j] = det3(M) / det;
            }
        }
        double l0 = 1.0;
        for (int j = 0; j < dim; ++j) l0 -= lam[j];
        for (int j = 0; j < nv; ++j) {
            shn[(size_t)SHAPE_NODES * e + j] = cn[j];
            shw[(size_t)SHAPE_NODES * e + j] = j == 0 ? l0 : lam[j - 1];
            if (dist(hx.data() + 3 * (size_t)cn[j], p) < 1e-4) sing[e] = cn[j];
        }
    }
    // cells around each singular node (rho_s for node electrodes, minimum
    // radius for every singular node); a free electrode's rho_s cell is its own
    std::vector<int> headE(c->n, -1), nextE(n_elec, -1);
    for (int e = n_elec - 1; e >= 0; --e)
        if (sing[e] >= 0) {
            nextE[e] = headE[sing[e]];
            headE[sing[e]] = e;
        }
    std::vector<std::vector<int>> cl(n_elec);
    for (int cc = 0; cc < c->C; ++cc)
        for (int i = 0; i < nv; ++i)
            for (int e = headE[hc[(size_t)cc * nv + i]]; e >= 0; e = nextE[e]) cl[e].push_back(cc);
    std::vector<int> ptr(1, 0), lst;
    std::vector<double> rmin(n_elec, 0.0);
    for (int e = 0; e < n_elec; ++e) {
        if (sing[e] >= 0 && cl[e].empty()) {
            // a node of no cell: an electrode unknown of the complete electrode
            // model, coupled to its faces only (total-field operator only)
            c->virtualElec = true;
        } else if (sing[e] >= 0) {
            // nearest neighbour distance (ElectrodeShapeNode::setSingValue, electrode.cpp:157-171)
            double rr = DBL_MAX;
            const int v = sing[e];
            for (int cc : cl[e])
                for (int i = 0; i < nv; ++i) {
                    const int u = hc[(size_t)cc * nv + i];
                    if (u != v) rr = std::min(rr, dist(hx.data() + 3 * (size_t)u, hx.data() + 3 * (size_t)v));
                }
            rmin[e] = rr;
        }
        if (elec_node[e] >= 0)
            lst.insert(lst.end(), cl[e].begin(), cl[e].end());
        else
            lst.push_back(elec_cell[e]);
        ptr.push_back((int)lst.size());
    }
    hipStream_t s = c->stream;
    c->elecNode.upload(c->hElecNode, s);
    c->elecPos.upload(pos, s);
    c->minRadius.upload(rmin, s);
    c->singNode.upload(sing, s);
    c->shapeNode.upload(shn, s);
    c->shapeW.upload(shw, s);
    c->ecPtr.upload(ptr, s);
    c->ecCell.upload(lst, s);
    for (int q = 0; q < 3; ++q) c->srcCenter[q] = source_center[q];
    c->surfaceZ = surface_z;
    // feature length scale: 4 x median nearest-neighbour electrode distance
    std::vector<double> nn;
    for (int e = 0; e < n_elec; ++e) {
        double best = DBL_MAX;
        for (int f = 0; f < n_elec; ++f) {
            if (f == e) continue;
            const double d = dist(&pos[3 * e], &pos[3 * f]);
            if (d > 0) best = std::min(best, d);
        }
        if (best < DBL_MAX) nn.push_back(best);
    }
    if (!nn.empty()) {
        std::nth_element(nn.begin(), nn.begin() + nn.size() / 2, nn.end());
        c->featureScale = (float)(4.0 * nn[nn.size() / 2]);
    }
    c->s1Valid = c->primValid = c->haveU = c->haveUc = c->srState = false;
}

int pgert_set_electrodes(pgert_ctx * c, int n_elec, const int * elec_node,
                         const double * source_center, double surface_z) {
    return guarded(c, [&] {
        for (int e = 0; elec_node && e < n_elec; ++e)
            PG_REQUIRE(elec_node[e] >= 0, "pgert_set_electrodes: negative node (free electrodes: pgert_set_electrodes_free)");
        set_electrodes(c, n_elec, elec_node, nullptr, nullptr, source_center, surface_z);
        sync(*c);
    });
}

int pgert_set_electrodes_free(pgert_ctx * c, int n_elec, const int * elec_node, const int * elec_cell,
                              const double * elec_pos, const double * source_center, double surface_z) {
    return guarded(c, [&] {
        set_electrodes(c, n_elec, elec_node, elec_cell, elec_pos, source_center, surface_z);
        sync(*c);
    });
}

int pgert_set_wavenumbers(pgert_ctx * c, int nk, const double * k, const double * w) {
    return guarded(c, [&] {
        PG_REQUIRE(nk > 0 && k && w, "pgert_set_wavenumbers: bad arguments");
        for (int i = 0; i < nk; ++i) PG_REQUIRE(k[i] >= 0.0, "pgert_set_wavenumbers: k < 0");
        if (c->dim == 3) PG_REQUIRE(nk == 1 && k[0] == 0.0, "pgert_set_wavenumbers: 3D uses k = {0}");
        c->nK = nk;
        c->kv.assign(k, k + nk);
        c->wv.assign(w, w + nk);
        c->dk.upload(c->kv, c->stream);
        c->dw.upload(c->wv, c->stream);
        c->s1Valid = c->primValid = c->haveU = c->haveUc = c->srState = false;
        sync(*c);
    });
}

int pgert_set_data(pgert_ctx * c, int n_data, const int * a, const int * b, const int * m,
                   const int * n, const double * kfac) {
    return guarded(c, [&] {
        PG_REQUIRE(c->nE > 0, "pgert_set_data: set electrodes first");
        PG_REQUIRE(n_data > 0 && a && b && m && n && kfac, "pgert_set_data: bad arguments");
        std::vector<int> h((size_t)n_data * 4);
        for (int i = 0; i < n_data; ++i) {
            int q[4] = {a[i], b[i], m[i], n[i]};
            for (int j = 0; j < 4; ++j) {
                // DataMap::data (datamap.cpp:185-194)
                PG_REQUIRE(q[j] >= -1 && q[j] < c->nE, "pgert_set_data: electrode index out of range");
                h[4 * (size_t)i + j] = q[j];
            }
        }
        c->D = n_data;
        c->abmn.upload(h, c->stream);
        c->kfac.upload(kfac, n_data, c->stream);
        sync(*c);
    });
}

int pgert_set_coarse(pgert_ctx * c, int n_agg) {
    return guarded(c, [&] {
        PG_REQUIRE(c->n > 0, "pgert_set_coarse: no mesh");
        if (n_agg <= 0) {
            c->nAgg = 0;
            return;
        }
        // power of two, at most n / 4 aggregates
        int m = 1;
        while (m * 2 <= n_agg && m * 2 <= c->n / 4) m *= 2;
        PG_REQUIRE(m >= 2, "pgert_set_coarse: mesh too small for a coarse space");
        const int n = c->n;
        // recursive coordinate bisection: split the longest extent at the median
        std::vector<int> idx(n), agg(n);
        for (int i = 0; i < n; ++i) idx[i] = i;
        struct Job { int b, e, lo, cnt; };
        std::vector<Job> stack{{0, n, 0, m}};
        while (!stack.empty()) {
            Job j = stack.back();
            stack.pop_back();
            if (j.cnt == 1) {
                for (int q = j.b; q < j.e; ++q) agg[idx[q]] = j.lo;
                continue;
            }
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int q = j.b; q < j.e; ++q)
                for (int d = 0; d < 3; ++d) {
                    double v = c->hXyz[3 * (size_t)idx[q] + d];
                    lo[d] = std::min(lo[d], v);
                    hi[d] = std::max(hi[d], v);
                }
            int ax = 0;
            for (int d = 1; d < 3; ++d)
                if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
            int mid = j.b + (j.e - j.b) / 2;
            std::nth_element(idx.begin() + j.b, idx.begin() + mid, idx.begin() + j.e, [&](int a, int b2) {
                double va = c->hXyz[3 * (size_t)a + ax], vb = c->hXyz[3 * (size_t)b2 + ax];
                return va < vb || (va == vb && a < b2);
            });
            stack.push_back({j.b, mid, j.lo, j.cnt / 2});
            stack.push_back({mid, j.e, j.lo + j.cnt / 2, j.cnt / 2});
        }
        std::vector<int> aptr(m + 1, 0), anodes(n);
        for (int i = 0; i < n; ++i) aptr[agg[i] + 1]++;
        for (int a = 0; a < m; ++a) aptr[a + 1] += aptr[a];
        std::vector<int> fill(aptr.begin(), aptr.end() - 1);
        for (int i = 0; i < n; ++i) anodes[fill[agg[i]]++] = i;
        // coarse entries: slots grouped by (agg(row), agg(col))
        std::vector<std::pair<long, int>> key((size_t)c->nnz);
        for (int r = 0; r < n; ++r)
            for (int q = c->hRowptr[r]; q < c->hRowptr[r + 1]; ++q)
                key[q] = {(long)agg[r] * m + agg[c->hColidx[q]], q};
        std::sort(key.begin(), key.end());
        std::vector<int> cptr(1, 0), cslot(c->nnz), cij;
        for (size_t q = 0; q < key.size(); ++q) {
            cslot[q] = key[q].second;
            if (q + 1 == key.size() || key[q + 1].first != key[q].first) {
                cptr.push_back((int)q + 1);
                cij.push_back((int)(key[q].first / m));
                cij.push_back((int)(key[q].first % m));
            }
        }
        // A Z: per row the distinct aggregates of its columns, and for each
        // such entry the row's slots that sum into it (in slot order)
        std::vector<int> azp(1, 0), aza, azsp(1, 0), azs;
        azs.reserve(c->nnz);
        std::vector<std::pair<int, int>> rowKeys;
        for (int r = 0; r < n; ++r) {
            rowKeys.clear();
            for (int q = c->hRowptr[r]; q < c->hRowptr[r + 1]; ++q) rowKeys.push_back({agg[c->hColidx[q]], q});
            std::sort(rowKeys.begin(), rowKeys.end());
            for (size_t u = 0; u < rowKeys.size(); ++u) {
                azs.push_back(rowKeys[u].second);
                if (u + 1 == rowKeys.size() || rowKeys[u + 1].first != rowKeys[u].first) {
                    aza.push_back(rowKeys[u].first);
                    azsp.push_back((int)azs.size());
                }
            }
            azp.push_back((int)aza.size());
        }
        hipStream_t s = c->stream;
        c->azPtr.upload(azp, s);
        c->azAgg.upload(aza, s);
        c->azSlotPtr.upload(azsp, s);
        c->azSlot.upload(azs, s);
        c->nAZ = (int)aza.size();
        c->agg.upload(agg, s);
        c->aggPtr.upload(aptr, s);
        c->aggNodes.upload(anodes, s);
        c->cePtr.upload(cptr, s);
        c->ceSlot.upload(cslot, s);
        c->ceIJ.upload(cij, s);
        c->nCE = (int)cij.size() / 2;
        c->nAgg = m;
        sync(*c);
    });
}

int pgert_info(const pgert_ctx * c, int * nnz, int * n_param) {
    if (!c) return -2;
    if (nnz) *nnz = c->nnz;
    if (n_param) *n_param = c->nParam;
    return 0;
}

int pgert_pattern(const pgert_ctx * c, int * rowptr, int * colidx) {
    if (!c || c->n == 0) return -2;
    std::memcpy(rowptr, c->hRowptr.data(), c->hRowptr.size() * sizeof(int));
    std::memcpy(colidx, c->hColidx.data(), c->hColidx.size() * sizeof(int));
    return 0;
}

int pgert_map_model(pgert_ctx * c, const double * model, int n_model, double * rho) {
    return guarded(c, [&] {
        PG_REQUIRE(c->C > 0, "pgert_map_model: no mesh");
        PG_REQUIRE(n_model == c->C || n_model >= c->nParam, "pgert_map_model: marker >= model size");
        launch_map_model(*c, model, n_model, rho);
        if (n_model != c->C) launch_prolongation(*c, rho);
        sync(*c);
    });
}

int pgert_assemble(pgert_ctx * c, const double * rho, double k, double * vals) {
    return guarded(c, [&] {
        PG_REQUIRE(c->C > 0, "pgert_assemble: no mesh");
        PG_REQUIRE(k >= 0.0, "pgert_assemble: k < 0");
        launch_assemble(*c, rho, k, vals, false);
        sync(*c);
    });
}

int pgert_fill_matrix(pgert_ctx * c, int mass, const double * a, double * vals) {
    return guarded(c, [&] {
        PG_REQUIRE(c->C > 0, "pgert_fill_matrix: no mesh");
        launch_fill_matrix(*c, mass, a, vals);
        sync(*c);
    });
}

namespace {

pgert_solve_opts resolve_opts(const pgert_solve_opts * opts) {
    pgert_solve_opts o{};
    if (opts) {
        o = *opts;
    } else {
        o.method = PGERT_BLOCK_MIXED;
        o.rtol = 1e-8;
        o.max_iter = 20000;
        o.check_every = 8;
        o.inner_rtol = 1e-3;
    }
    PG_REQUIRE(o.rtol > 0 && o.max_iter > 0, "pgert_forward: bad solver options");
    PG_REQUIRE(o.smooth == 0 || o.smooth == 2, "pgert_forward: smooth must be 0 or 2");
    if (o.check_every <= 0) o.check_every = 8;
    return o;
}

// assembly, primaries, SR right-hand sides and the initial guess of one
// problem, on its stream (events ev[0..2])
void forward_prepare(Ctx & c, const double * rho, const pgert_solve_opts & o) {
    PG_REQUIRE(c.C > 0 && c.nE > 0 && c.nK > 0 && c.D > 0,
               "pgert_forward: mesh, electrodes, wavenumbers and data must be set");
    PG_REQUIRE(c.order == 1, "pgert_forward: a P2 mesh serves the total-field primary problem only");
    PG_REQUIRE(!c.virtualElec, "pgert_forward: electrodes on nodes of no cell (complete electrode model) need the "
                               "total-field operator");
    PG_REQUIRE(rho, "pgert_forward: null pointer");
    hipStream_t s = c.stream;
    const size_t nk = c.nK, nn = c.n, ne = c.nE;
    c.rho.alloc(c.C);
    PG_CHECK(hipMemcpyAsync(c.rho.p, rho, (size_t)c.C * 8, hipMemcpyDeviceToDevice, s));
    PG_CHECK(hipEventRecord(c.ev[0], s));
    // assembly (A_k for this model, S1_k with rho == 1 once per mesh)
    c.valsA.alloc(nk * c.nnz);
    for (size_t kk = 0; kk < nk; ++kk) launch_assemble(c, c.rho.p, c.kv[kk], c.valsA.p + kk * c.nnz, false);
    launch_calibration(c, c.valsA.p);   // calibration nodes in S and S1 (dcfemmodelling.cpp:2182, :2192)
    if (!c.s1Valid) {
        c.valsS1.alloc(nk * c.nnz);
        for (size_t kk = 0; kk < nk; ++kk) launch_assemble(c, nullptr, c.kv[kk], c.valsS1.p + kk * c.nnz, true);
        launch_calibration(c, c.valsS1.p);
        c.s1Valid = true;
    }
    launch_diag_inv(c);
    PG_CHECK(hipEventRecord(c.ev[1], s));
    // primary potentials (once per electrode/wavenumber set) and SR right-hand sides
    if (!c.primValid) {
        PG_REQUIRE(!c.primNumeric, "pgert_forward: the numerical primary potentials belong to another electrode or "
                                   "wavenumber set; set them again (pgert_set_primary_potentials)");
        c.prim.alloc(nk * nn * ne);
        launch_primary(c);
        c.primValid = true;
    }
    c.rhoS.alloc(ne);
    launch_rho_source(c, c.rho.p);
    c.B.alloc(nk * nn * ne);
    launch_sr_rhs(c);
    launch_zero_rows(c, c.B.p);   // the calibration nodes' right-hand side (:2280-2283)
    c.X.alloc(nk * nn * ne);
    if (o.warm_start)
        PG_CHECK(hipMemcpyAsync(c.X.p, o.warm_start, c.X.bytes(), hipMemcpyDeviceToDevice, s));
    else if (o.learned_warm_start)
        mlp_warm_start(c);
    else
        PG_CHECK(hipMemsetAsync(c.X.p, 0, c.X.bytes(), s));
    PG_CHECK(hipEventRecord(c.ev[2], s));
}

// total potentials and apparent resistivities (events ev[3..4])
void forward_finish(Ctx & c, double * rhoa) {
    PG_CHECK(hipEventRecord(c.ev[3], c.stream));
    c.U.alloc((size_t)c.nK * c.n * c.nE);
    launch_total_potential(c);
    c.haveU = true;
    c.srState = true;
    launch_response(c, rhoa);
    PG_CHECK(hipEventRecord(c.ev[4], c.stream));
}

// one real solve of c.B into c.X (c.X holds the initial guess)
void solve_one(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st, bool refactor = true) {
    if (o.method == PGERT_PCG_FP64)
        solve_pcg_fp64(c, o, st);
    else if (o.method == PGERT_BLOCK_FP64)
        solve_block(c, o, st, false);
    else if (o.method == PGERT_BLOCK_MIXED)
        solve_block(c, o, st, true);
    else if (o.method == PGERT_BAND_FP64) {
        std::vector<pgert_solve_stats> v;
        solve_band({&c}, o, v, refactor);
        st = v[0];
    } else
        throw std::runtime_error("pgert_forward: unknown method");
}

void read_timings(Ctx & c) {
    float t;
    for (int i = 0; i < 4; ++i) {
        PG_CHECK(hipEventElapsedTime(&t, c.ev[i], c.ev[i + 1]));
        c.timings[i] = t;
    }
}

}  // namespace

// A solve that misses rtol: again from the analytic initial guess (secondary
// field 0) if it started from a learned or given guess, then on in FP64 from
// the best iterate if it was mixed; fails if that also misses (unless the
// caller asked for the statistics only).  The reference's direct solver has no
// such case; it warns on a residual above 1e-6 (dcfemmodelling.cpp:1896-1907).
void solve_with_fallback(Ctx & c, const pgert_solve_opts & o, pgert_solve_stats & st, const char * who) {
    solve_one(c, o, st);
    if (st.converged || o.report_only) return;
    pgert_solve_opts f = o;
    auto again = [&](int bit) {
        pgert_solve_stats s2{};
        solve_one(c, f, s2);
        st.iterations += s2.iterations;
        st.max_iterations += s2.max_iterations;
        st.refinements += s2.refinements;
        st.converged = s2.converged;
        st.max_rel_residual = s2.max_rel_residual;
        st.fallback |= bit;
    };
    if (o.learned_warm_start || o.warm_start) {
        f.learned_warm_start = 0;
        f.warm_start = nullptr;
        PG_CHECK(hipMemsetAsync(c.X.p, 0, c.X.bytes(), c.stream));
        again(1);
    }
    if (!st.converged && f.method == PGERT_BLOCK_MIXED) {
        f.method = PGERT_BLOCK_FP64;
        again(2);
    }
    if (!st.converged) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "%s: the solve did not reach rtol %.3e within max_iter %d iterations (max relative residual %.3e)",
                      who, o.rtol, o.max_iter, st.max_rel_residual);
        PG_REQUIRE(false, std::string(msg));
    }
}

int pgert_forward(pgert_ctx * c, const double * rho, const pgert_solve_opts * opts, double * rhoa,
                  pgert_solve_stats * stats) {
    return guarded(c, [&] {
        PG_REQUIRE(rhoa, "pgert_forward: null pointer");
        const pgert_solve_opts o = resolve_opts(opts);
        forward_prepare(*c, rho, o);
        pgert_solve_stats st{};
        solve_with_fallback(*c, o, st, "pgert_forward");
        forward_finish(*c, rhoa);
        sync(*c);
        read_timings(*c);
        if (stats) *stats = st;
    });
}

int pgert_forward_complex(pgert_ctx * c, const double * rho_re, const double * rho_im,
                          const pgert_solve_opts * opts, double * resp_re, double * resp_im,
                          pgert_solve_stats * stats) {
    return guarded(c, [&] {
        PG_REQUIRE(rho_re && rho_im && resp_re && resp_im, "pgert_forward_complex: null pointer");
        PG_REQUIRE(c->nCem == 0 && !c->virtualElec,
                   "pgert_forward_complex: the complete electrode model is real only (as in the reference, "
                   "dcfemmodelling.cpp:412-417)");
        Ctx & C = *c;
        PG_REQUIRE(C.C > 0 && C.nE > 0 && C.nK > 0 && C.D > 0,
                   "pgert_forward_complex: mesh, electrodes, wavenumbers and data must be set");
        pgert_solve_opts o = resolve_opts(opts);
        o.warm_start = nullptr;
        o.learned_warm_start = 0;
        hipStream_t s = C.stream;
        const size_t nk = C.nK, blk = nk * C.n * C.nE;
        PG_CHECK(hipEventRecord(C.ev[0], s));
        C.valsA.alloc(nk * C.nnz);
        C.valsAi.alloc(nk * C.nnz);
        for (size_t kk = 0; kk < nk; ++kk)
            launch_assemble_complex(C, rho_re, rho_im, C.kv[kk], C.valsA.p + kk * C.nnz, C.valsAi.p + kk * C.nnz);
        // calibration nodes: a complex unit diagonal (dcfemmodelling.cpp:1835)
        launch_calibration(C, C.valsA.p);
        launch_calibration(C, C.valsAi.p, true);
        launch_diag_inv(C);
        PG_CHECK(hipEventRecord(C.ev[1], s));
        C.Bpt.alloc(blk);
        launch_point_rhs(C, C.Bpt.p);
        PG_CHECK(hipEventRecord(C.ev[2], s));
        C.B.alloc(blk);
        C.X.alloc(blk);
        C.Xre.alloc(blk);
        C.Xim.alloc(blk);
        C.chg.alloc(4);
        PG_CHECK(hipMemsetAsync(C.Xre.p, 0, C.Xre.bytes(), s));
        PG_CHECK(hipMemsetAsync(C.Xim.p, 0, C.Xim.bytes(), s));
        C.haveU = false;
        C.haveUc = false;
        C.srState = false;
        pgert_solve_stats st{}, tot{};
        tot.converged = 1;
        const int maxSweeps = 60;
        // the sweeps stop once the potentials change by less than rtol (relative);
        // an inner solve from a converged guess moves them by about rtol itself
        const double stop = 10.0 * o.rtol;
        bool done = false;
        double hchg[4] = {0.0, 0.0, 0.0, 0.0};
        for (int it = 0; it < maxSweeps && !done; ++it) {
            for (int half = 0; half < 2; ++half) {
                double * Xh = half == 0 ? C.Xre.p : C.Xim.p;
                if (half == 0)
                    launch_spmm_add(C, C.valsAi.p, C.Xim.p, C.Bpt.p, 1.0, C.B.p);
                else
                    launch_spmm_add(C, C.valsAi.p, C.Xre.p, nullptr, -1.0, C.B.p);
                PG_CHECK(hipMemcpyAsync(C.X.p, Xh, C.X.bytes(), hipMemcpyDeviceToDevice, s));
                st = pgert_solve_stats{};
                // the imaginary half needs the real half's absolute accuracy: its
                // relative tolerance is looser by |X_re| / |X_im| (known from the
                // previous sweep), at most 1e-2
                pgert_solve_opts oh = o;
                if (half == 1 && it > 0 && hchg[3] > 0.0)
                    oh.rtol = std::min(1e-2, std::max(o.rtol, o.rtol * std::sqrt(hchg[1] / hchg[3])));
                solve_one(C, oh, st, it == 0 && half == 0);   // A_re is the same in every solve
                tot.iterations += st.iterations;
                tot.max_iterations += st.max_iterations;
                tot.refinements += st.refinements;
                tot.converged = tot.converged && st.converged;
                tot.max_rel_residual = std::max(tot.max_rel_residual, st.max_rel_residual);
                launch_change(C, (long)blk, C.X.p, Xh, C.chg.p + 2 * half);
                PG_CHECK(hipMemcpyAsync(Xh, C.X.p, C.X.bytes(), hipMemcpyDeviceToDevice, s));
            }
            PG_CHECK(hipMemcpyAsync(hchg, C.chg.p, sizeof(hchg), hipMemcpyDeviceToHost, s));
            sync(C);
            const double rel = std::sqrt((hchg[0] + hchg[2]) / std::max(hchg[1] + hchg[3], 1e-300));
            // an inner solve stalled at max_iter also moves the potentials little:
            // only sweeps whose solves all met their tolerance can end the loop
            done = it > 0 && rel < stop && st.converged && tot.converged;
            if (!tot.converged) break;
        }
        PG_REQUIRE(tot.converged || o.report_only,
                   "pgert_forward_complex: an inner solve did not reach its tolerance within max_iter " +
                       std::to_string(o.max_iter) + " iterations (max relative residual " +
                       std::to_string(tot.max_rel_residual) + ")");
        PG_REQUIRE(done || (o.report_only && !tot.converged),
                   "pgert_forward_complex: the real/imaginary sweeps did not converge "
                   "(cell phases too large for this solver)");
        PG_CHECK(hipEventRecord(C.ev[3], s));
        launch_response_complex(C, C.Xre.p, C.Xim.p, resp_re, resp_im);
        PG_CHECK(hipEventRecord(C.ev[4], s));
        C.haveUc = true;
        sync(C);
        read_timings(C);
        if (stats) *stats = tot;
    });
}

int pgert_forward_total(pgert_ctx * c, const double * rho, const pgert_solve_opts * opts, double * rhoa,
                        pgert_solve_stats * stats) {
    return guarded(c, [&] {
        PG_REQUIRE(rho && rhoa, "pgert_forward_total: null pointer");
        Ctx & C = *c;
        PG_REQUIRE(C.C > 0 && C.nE > 0 && C.nK > 0 && C.D > 0,
                   "pgert_forward_total: mesh, electrodes, wavenumbers and data must be set");
        pgert_solve_opts o = resolve_opts(opts);
        o.learned_warm_start = 0;
        hipStream_t s = C.stream;
        const size_t nk = C.nK, blk = nk * C.n * C.nE;
        C.srState = false;
        PG_CHECK(hipEventRecord(C.ev[0], s));
        C.rho.alloc(C.C);
        PG_CHECK(hipMemcpyAsync(C.rho.p, rho, (size_t)C.C * 8, hipMemcpyDeviceToDevice, s));
        C.valsA.alloc(nk * C.nnz);
        for (size_t kk = 0; kk < nk; ++kk) launch_assemble(C, C.rho.p, C.kv[kk], C.valsA.p + kk * C.nnz, false);
        launch_bypass(C, C.valsA.p);
        launch_calibration(C, C.valsA.p);
        launch_diag_inv(C);
        PG_CHECK(hipEventRecord(C.ev[1], s));
        C.B.alloc(blk);
        launch_point_rhs(C, C.B.p);
        C.X.alloc(blk);
        if (o.warm_start)
            PG_CHECK(hipMemcpyAsync(C.X.p, o.warm_start, C.X.bytes(), hipMemcpyDeviceToDevice, s));
        else
            PG_CHECK(hipMemsetAsync(C.X.p, 0, C.X.bytes(), s));
        PG_CHECK(hipEventRecord(C.ev[2], s));
        pgert_solve_stats st{};
        solve_with_fallback(C, o, st, "pgert_forward_total");
        PG_CHECK(hipEventRecord(C.ev[3], s));
        C.U.alloc(blk);
        PG_CHECK(hipMemcpyAsync(C.U.p, C.X.p, C.U.bytes(), hipMemcpyDeviceToDevice, s));
        C.haveU = true;
        launch_response(C, rhoa);
        PG_CHECK(hipEventRecord(C.ev[4], s));
        sync(C);
        read_timings(C);
        if (stats) *stats = st;
    });
}

int pgert_set_reference(pgert_ctx * c, int ref_node, int n_cal, const int * cal_nodes) {
    return guarded(c, [&] {
        Ctx & C = *c;
        PG_REQUIRE(C.n > 0, "pgert_set_reference: set the mesh first");
        PG_REQUIRE(ref_node >= -1 && ref_node < C.n, "pgert_set_reference: reference node out of range");
        PG_REQUIRE(n_cal >= 0 && (n_cal == 0 || cal_nodes), "pgert_set_reference: bad calibration nodes");
        std::vector<char> isCal(C.n, 0);
        for (int i = 0; i < n_cal; ++i) {
            PG_REQUIRE(cal_nodes[i] >= 0 && cal_nodes[i] < C.n, "pgert_set_reference: calibration node out of range");
            isCal[cal_nodes[i]] = 1;
        }
        // every slot of a calibration row or column once, 1 on the diagonal
        std::vector<std::pair<int, double>> ds;
        for (int i = 0; i < C.n; ++i) {
            if (!isCal[i]) continue;
            for (int q = C.hRowptr[i]; q < C.hRowptr[i + 1]; ++q) {
                const int j = C.hColidx[q];
                ds.push_back({q, j == i ? 1.0 : 0.0});
                if (j != i && !isCal[j]) ds.push_back({find_slot(C, j, i), 0.0});
            }
        }
        std::sort(ds.begin(), ds.end());
        std::vector<int> slot(ds.size());
        std::vector<double> val(ds.size());
        for (size_t q = 0; q < ds.size(); ++q) {
            slot[q] = ds[q].first;
            val[q] = ds[q].second;
        }
        if (ref_node >= 0) {
            // the reference electrode's cells and nearest-neighbour distance
            // (ElectrodeShapeNode, electrode.cpp:98-189)
            std::vector<int> hc((size_t)C.C * C.nv);
            PG_CHECK(hipMemcpy(hc.data(), C.cells.p, hc.size() * 4, hipMemcpyDeviceToHost));
            std::vector<int> rc;
            double rr = DBL_MAX;
            for (int cc = 0; cc < C.C; ++cc) {
                const int * v = hc.data() + (size_t)cc * C.nv;
                if (std::find(v, v + C.nvert, ref_node) == v + C.nvert) continue;
                rc.push_back(cc);
                for (int i = 0; i < C.nvert; ++i) {
                    if (v[i] == ref_node) continue;
                    double d2 = 0.0;
                    for (int q = 0; q < 3; ++q) {
                        const double dq = C.hXyz[3 * (size_t)v[i] + q] - C.hXyz[3 * (size_t)ref_node + q];
                        d2 += dq * dq;
                    }
                    rr = std::min(rr, std::sqrt(d2));
                }
            }
            PG_REQUIRE(!rc.empty(), "pgert_set_reference: the reference node belongs to no cell");
            C.refCells.upload(rc, C.stream);
            C.refRmin = rr;
        }
        C.refNode = ref_node;
        C.primValid = false;
        C.s1Valid = false;
        C.calNodes.upload(cal_nodes, n_cal, C.stream);
        C.nCalSlot = (int)ds.size();
        C.calSlot.upload(slot, C.stream);
        C.calVal.upload(val, C.stream);
        C.haveU = C.haveUc = C.srState = false;
        sync(C);
    });
}

int pgert_set_cem(pgert_ctx * c, int n_terms, const int * rows, const int * cols, const double * vals) {
    return guarded(c, [&] {
        Ctx & C = *c;
        PG_REQUIRE(C.n > 0, "pgert_set_cem: set the mesh first");
        PG_REQUIRE(n_terms >= 0 && (n_terms == 0 || (rows && cols && vals)), "pgert_set_cem: bad arguments");
        std::vector<int> slot(n_terms);
        for (int i = 0; i < n_terms; ++i) {
            PG_REQUIRE(rows[i] >= 0 && rows[i] < C.n && cols[i] >= 0 && cols[i] < C.n,
                       "pgert_set_cem: bad entry " + std::to_string(i));
            slot[i] = find_slot(C, rows[i], cols[i]);
            PG_REQUIRE(slot[i] >= 0, "pgert_set_cem: entry (" + std::to_string(rows[i]) + ", " +
                                         std::to_string(cols[i]) + ") is not in the matrix pattern");
        }
        C.nCem = n_terms;
        C.cemSlot.upload(slot, C.stream);
        C.cemVal.upload(vals, n_terms, C.stream);
        sync(C);
    });
}

int pgert_set_primary_potentials(pgert_ctx * c, const double * prim) {
    return guarded(c, [&] {
        Ctx & C = *c;
        PG_REQUIRE(C.order == 1 && C.n > 0 && C.nE > 0 && C.nK > 0,
                   "pgert_set_primary_potentials: set a linear mesh, the electrodes and the wavenumbers first");
        PG_REQUIRE(prim, "pgert_set_primary_potentials: null pointer");
        const size_t total = (size_t)C.nK * C.n * C.nE;
        C.prim.alloc(total);
        PG_CHECK(hipMemcpyAsync(C.prim.p, prim, total * sizeof(double), hipMemcpyDeviceToDevice, C.stream));
        C.primValid = true;
        C.primNumeric = true;
        C.srState = false;
        sync(C);
    });
}

int pgert_set_bypass(pgert_ctx * c, int n_pairs, const int * a, const int * b, const double * conductance) {
    return guarded(c, [&] {
        Ctx & C = *c;
        PG_REQUIRE(C.n > 0, "pgert_set_bypass: set the mesh first");
        PG_REQUIRE(n_pairs >= 0 && (n_pairs == 0 || (a && b && conductance)), "pgert_set_bypass: bad arguments");
        std::vector<int> slot;
        std::vector<double> val;
        for (int i = 0; i < n_pairs; ++i) {
            PG_REQUIRE(a[i] >= 0 && a[i] < C.n && b[i] >= 0 && b[i] < C.n && a[i] != b[i],
                       "pgert_set_bypass: bad node pair " + std::to_string(i));
            const int sab = find_slot(C, a[i], b[i]), sba = find_slot(C, b[i], a[i]);
            PG_REQUIRE(sab >= 0 && sba >= 0, "pgert_set_bypass: nodes " + std::to_string(a[i]) + " and " +
                                                 std::to_string(b[i]) + " share no cell; pass the pair to "
                                                 "pgert_set_mesh_coupled");
            const double g = conductance[i];
            slot.insert(slot.end(), {find_slot(C, a[i], a[i]), find_slot(C, b[i], b[i]), sab, sba});
            val.insert(val.end(), {g, g, -g, -g});
        }
        C.nByp = n_pairs;
        C.bypSlot.upload(slot, C.stream);
        C.bypVal.upload(val, C.stream);
        sync(C);
    });
}

int pgert_jacobian_complex(pgert_ctx * c, const double * model_re, const double * model_im, int n_model,
                           double * J_re, double * J_im) {
    return guarded(c, [&] {
        Ctx & C = *c;
        PG_REQUIRE(C.haveUc, "pgert_jacobian_complex: call pgert_forward_complex first");
        PG_REQUIRE(J_re && J_im, "pgert_jacobian_complex: bad output");
        PG_REQUIRE(n_model >= 0 && (n_model == 0 || (model_re && model_im)), "pgert_jacobian_complex: bad model");
        hipStream_t s = C.stream;
        const size_t blk = (size_t)C.nK * C.n * C.nE;
        C.U.alloc(blk);
        // S(U) is the quadratic form of the real kernel: S(re) - S(im) is the
        // real part, S(re + im) - S(re) - S(im) the imaginary part; built in
        // J_re / J_im themselves (no D x P scratch)
        PG_CHECK(hipMemcpyAsync(C.U.p, C.Xre.p, C.U.bytes(), hipMemcpyDeviceToDevice, s));
        launch_jacobian(C, nullptr, 0, J_re, 0);
        PG_CHECK(hipMemcpyAsync(C.U.p, C.Xim.p, C.U.bytes(), hipMemcpyDeviceToDevice, s));
        launch_jacobian(C, nullptr, 0, J_im, 0);
        launch_jacobian_complex_pre(C, J_re, J_im);
        launch_add(C, (long)blk, C.Xre.p, C.Xim.p, C.U.p);
        launch_jacobian(C, nullptr, 0, J_im, 0, true);
        if (n_model == C.nParam) launch_jacobian_complex_scale(C, model_re, model_im, J_re, J_im);
        sync(C);
    });
}

int pgert_batch_forward(pgert_ctx * const * ctxs, int count, const double * const * rhos,
                        const pgert_solve_opts * opts, double * const * rhoas, pgert_solve_stats * stats) {
    if (!ctxs || count <= 0 || !ctxs[0]) return -2;
    pgert_ctx * c0 = ctxs[0];
    return guarded(c0, [&] {
        PG_REQUIRE(rhos && rhoas, "pgert_batch_forward: null pointer");
        const pgert_solve_opts o = resolve_opts(opts);
        PG_REQUIRE(o.method == PGERT_BLOCK_FP64 || o.method == PGERT_BLOCK_MIXED || o.method == PGERT_BAND_FP64,
                   "pgert_batch_forward: block or band methods only");
        std::vector<Ctx *> cs(count);
        std::vector<hipStream_t> own(count);
        for (int i = 0; i < count; ++i) {
            PG_REQUIRE(ctxs[i] && rhos[i] && rhoas[i], "pgert_batch_forward: null problem");
            PG_REQUIRE(ctxs[i]->device == c0->device, "pgert_batch_forward: problems on different devices");
            cs[i] = ctxs[i];
            own[i] = cs[i]->stream;
            cs[i]->stream = c0->stream;
        }
        auto restore = [&]() {
            for (int i = 0; i < count; ++i) cs[i]->stream = own[i];
        };
        try {
            for (int i = 0; i < count; ++i) forward_prepare(*cs[i], rhos[i], o);
            std::vector<pgert_solve_stats> st;
            if (o.method == PGERT_BAND_FP64)
                solve_band(cs, o, st);
            else
                solve_blocks(cs, o, st, o.method == PGERT_BLOCK_MIXED);
            for (int i = 0; i < count; ++i) forward_finish(*cs[i], rhoas[i]);
            sync(*c0);
            for (int i = 0; i < count; ++i) {
                read_timings(*cs[i]);
                if (stats) stats[i] = st[i];
            }
            for (int i = 0; i < count; ++i)
                PG_REQUIRE(st[i].converged || o.report_only,
                           "pgert_batch_forward: problem " + std::to_string(i) + " did not reach rtol within " +
                               std::to_string(o.max_iter) + " iterations (max relative residual " +
                               std::to_string(st[i].max_rel_residual) + ")");
        } catch (...) {
            restore();
            throw;
        }
        restore();
    });
}

int pgert_potentials(pgert_ctx * c, double * out) {
    return guarded(c, [&] {
        PG_REQUIRE(c->haveU, "pgert_potentials: no forward result");
        launch_transpose_pots(*c, out);
        sync(*c);
    });
}

int pgert_potentials_node_major(pgert_ctx * c, double * out) {
    return guarded(c, [&] {
        PG_REQUIRE(c->haveU, "pgert_potentials_node_major: no forward result");
        PG_CHECK(hipMemcpyAsync(out, c->U.p, c->U.bytes(), hipMemcpyDeviceToDevice, c->stream));
        sync(*c);
    });
}

int pgert_jacobian(pgert_ctx * c, const double * model, int n_model, void * J, int dtype) {
    return guarded(c, [&] {
        PG_REQUIRE(c->haveU, "pgert_jacobian: call pgert_forward first");
        PG_REQUIRE(c->order == 1, "pgert_jacobian: linear meshes only");
        PG_REQUIRE(J && (dtype == 0 || dtype == 1), "pgert_jacobian: bad output");
        PG_REQUIRE(n_model >= 0 && (n_model == 0 || model), "pgert_jacobian: bad model");
        PG_CHECK(hipEventRecord(c->ev[4], c->stream));
        launch_jacobian(*c, model, n_model, J, dtype);
        PG_CHECK(hipEventRecord(c->ev[5], c->stream));
        sync(*c);
        float t;
        PG_CHECK(hipEventElapsedTime(&t, c->ev[4], c->ev[5]));
        c->timings[4] = t;
    });
}

int pgert_batch_jacobian(pgert_ctx * const * ctxs, int count, const double * const * models, const int * n_models,
                         void * const * Js, int dtype) {
    if (!ctxs || count <= 0 || !ctxs[0]) return -2;
    pgert_ctx * c0 = ctxs[0];
    return guarded(c0, [&] {
        PG_REQUIRE(models && n_models && Js && (dtype == 0 || dtype == 1), "pgert_batch_jacobian: bad arguments");
        std::vector<Ctx *> cs(count);
        std::vector<const double *> ms(count);
        std::vector<int> nm(count);
        std::vector<void *> js(count);
        for (int i = 0; i < count; ++i) {
            PG_REQUIRE(ctxs[i] && Js[i], "pgert_batch_jacobian: null problem");
            PG_REQUIRE(ctxs[i]->device == c0->device, "pgert_batch_jacobian: problems on different devices");
            PG_REQUIRE(ctxs[i]->haveU, "pgert_batch_jacobian: call a forward first");
            PG_REQUIRE(n_models[i] >= 0 && (n_models[i] == 0 || models[i]), "pgert_batch_jacobian: bad model");
            cs[i] = ctxs[i];
            ms[i] = models[i];
            nm[i] = n_models[i];
            js[i] = Js[i];
        }
        PG_CHECK(hipEventRecord(c0->ev[4], c0->stream));
        launch_jacobian_batch(cs, ms, nm, js, dtype);
        PG_CHECK(hipEventRecord(c0->ev[5], c0->stream));
        sync(*c0);
        float t;
        PG_CHECK(hipEventElapsedTime(&t, c0->ev[4], c0->ev[5]));
        c0->timings[4] = t;
    });
}

int pgert_set_constraints(pgert_ctx * c, int n_const, int n_model, const int * rowptr, const int * colidx,
                          const double * vals) {
    return guarded(c, [&] { set_constraints(*c, n_const, n_model, rowptr, colidx, vals); });
}

int pgert_cgls(pgert_ctx * c, const void * J, int dtype, int n_data, int n_model, const double * d_weight,
               const double * b, double * x, const double * wc, const double * wm, const double * tm,
               const double * td, double lambda, const double * roughness, int max_iter, double tol,
               int * iterations, double * norm_r2) {
    return guarded(c, [&] {
        PG_REQUIRE(J && (dtype == 0 || dtype == 1) && n_data > 0 && n_model > 0, "pgert_cgls: bad Jacobian");
        PG_REQUIRE(d_weight && b && x && wm && tm && td, "pgert_cgls: null vector");
        PG_REQUIRE(c->nConst == 0 || (wc && roughness), "pgert_cgls: constraint weights / roughness missing");
        PG_REQUIRE(max_iter >= 0, "pgert_cgls: bad iteration limit");
        const int it = cgls(*c, J, dtype, n_data, n_model, d_weight, b, x, wc, wm, tm, td, lambda, roughness,
                            max_iter, tol);
        if (iterations) *iterations = it;
        if (norm_r2) *norm_r2 = c->cglsNormR2;
    });
}

int pgert_cgls_noalloc(pgert_ctx * c, const void * J, int dtype, int n_data, int n_model, const double * d_weight,
                       const double * b, double * x, const double * wc, const double * wm, const double * tm,
                       const double * td, double lambda, const double * roughness, int max_iter, double tol,
                       int * iterations, double * norm_r2) {
    return guarded(c, [&] {
        PG_REQUIRE(J && (dtype == 0 || dtype == 1) && n_data > 0 && n_model > 0, "pgert_cgls_noalloc: bad Jacobian");
        PG_REQUIRE(d_weight && b && x && wm && tm && td, "pgert_cgls_noalloc: null vector");
        PG_REQUIRE(c->nConst == 0 || (wc && roughness), "pgert_cgls_noalloc: constraint weights / roughness missing");
        PG_REQUIRE(max_iter >= 0, "pgert_cgls_noalloc: bad iteration limit");
        const int it = cgls_noalloc(*c, J, dtype, n_data, n_model, d_weight, b, x, wc, wm, tm, td, lambda, roughness,
                                    max_iter, tol);
        if (iterations) *iterations = it;
        if (norm_r2) *norm_r2 = c->cglsNormR2;
    });
}

int pgert_trans(pgert_ctx * c, int op, int n, const double * a, const double * b, double tau, double lower,
                double upper, double * out) {
    return guarded(c, [&] { gn_trans(*c, op, n, a, b, tau, lower, upper, out); });
}

int pgert_gn_phi(pgert_ctx * c, int n_tau, const double * taus, int n_model, const double * model, const double * dm,
                 const double * mw, const double * cw, double lower, double upper, int n_data, const double * data,
                 const double * resp, const double * dd, const double * derr, double * phiD, double * phiM) {
    return guarded(c, [&] {
        gn_phi(*c, n_tau, taus, n_model, model, dm, mw, cw, lower, upper, n_data, data, resp, dd, derr, phiD, phiM);
    });
}

int pgert_mlp_param_count(void) { return mlp_param_count(); }

int pgert_mlp_set_params(pgert_ctx * c, const float * params, int count) {
    return guarded(c, [&] {
        PG_REQUIRE(params && count == mlp_param_count(), "pgert_mlp_set_params: wrong parameter count");
        mlp_set_params(*c, params);
        sync(*c);
    });
}

int pgert_mlp_get_params(pgert_ctx * c, float * params, int count) {
    return guarded(c, [&] {
        PG_REQUIRE(params && count == mlp_param_count(), "pgert_mlp_get_params: wrong parameter count");
        mlp_get_params(*c, params);
    });
}

int pgert_mlp_collect(pgert_ctx * c, int k_index, const long * idx, long count, float * feat, float * target) {
    return guarded(c, [&] {
        PG_REQUIRE(k_index >= 0 && k_index < c->nK, "pgert_mlp_collect: bad wavenumber index");
        PG_REQUIRE(count > 0 && idx && feat, "pgert_mlp_collect: bad arguments");
        PG_REQUIRE(c->srState && c->primValid,
                   "pgert_mlp_collect: needs the potentials of a singularity-removal forward (pgert_forward)");
        mlp_collect(*c, k_index, idx, count, feat, target);
        sync(*c);
    });
}

int pgert_mlp_predict(pgert_ctx * c, const float * feat, int count, float * out) {
    return guarded(c, [&] {
        PG_REQUIRE(count > 0 && feat && out, "pgert_mlp_predict: bad arguments");
        mlp_predict(*c, feat, count, out);
        sync(*c);
    });
}

int pgert_mlp_train_step(pgert_ctx * c, const float * feat, const float * target, int count, float lr, int apply,
                         float * loss, float * grad) {
    return guarded(c, [&] {
        PG_REQUIRE(feat && target && loss, "pgert_mlp_train_step: bad arguments");
        *loss = mlp_train_step(*c, feat, target, count, lr, apply != 0, grad);
    });
}

int pgert_kernel_timing(pgert_ctx * c, int enable, int reset, double * total_ms, long * launches,
                        double * bytes_per_launch) {
    return guarded(c, [&] {
        if (reset) {
            c->kTotalMs = 0.0;
            c->kLaunches = 0;
        }
        PG_REQUIRE(enable >= 0 && enable <= PGERT_TIME_COARSE_RESID, "pgert_kernel_timing: unknown kernel");
        c->kTiming = enable != 0;
        c->kSelect = enable;
        if (total_ms) *total_ms = c->kTotalMs;
        if (launches) *launches = c->kLaunches;
        if (bytes_per_launch) *bytes_per_launch = c->kBytesPerLaunch;
    });
}

int pgert_last_timings(const pgert_ctx * c, double * ms5) {
    if (!c || !ms5) return -2;
    for (int i = 0; i < 5; ++i) ms5[i] = c->timings[i];
    return 0;
}

}  // extern "C"

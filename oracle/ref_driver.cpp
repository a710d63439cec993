// Test infrastructure only: a thin extern "C" driver over the reference's own
// C++ ERT core (pyGIMLi core/src + core/src/bert), compiled in place from the
// reference sources by oracle/Makefile into oracle/_ref/.  It is used by the
// tests (as a checker) and by bench.py's cpu_baseline leg; the product path
// never loads it.
//
// The only piece the reference lacks in this image is its sparse direct
// solver back end (CHOLMOD/UMFPACK/LDL are external libraries).  The
// reference exposes a solver plug-in point for exactly this,
// DCMultiElectrodeModelling::setSolver(SolverWrapper*)
// (core/src/bert/dcfemmodelling.h:266, core/src/solverWrapper.h:26-55); this
// file installs its own FP64 direct solver there: reverse Cuthill-McKee
// ordering + envelope (skyline) Cholesky.  No reference header, library or
// generated file is replaced.

#include "gimli.h"
#include "mesh.h"
#include "node.h"
#include "meshentities.h"
#include "sparsematrix.h"
#include "elementmatrix.h"
#include "numericbase.h"
#include "solverWrapper.h"
#include "bert/dcfemmodelling.h"
#include "bert/bertDataContainer.h"
#include "bert/bertMisc.h"
#include "inversion.h"
#include "trans.h"
#include "interpolate.h"
#include "bert/datamap.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>
#include <unistd.h>

namespace GIMLI {
// Defined with external linkage in core/src/bert/dcfemmodelling.cpp:295-300
// but not declared in its header.
void dcfemBoundaryAssembleStiffnessMatrix(RSparseMatrix & S, const Mesh & mesh,
                                          const RVector3 & source, double k);
}

using namespace GIMLI;

namespace {

std::string g_err;

// ---------------------------------------------------------------------------
// RCM + envelope Cholesky, FP64.  Works on the full symmetric CRS the
// reference assembles (sparsematrix.h:966-1032: colPtr_ = row pointer,
// rowIdx_ = column index).
class EnvelopeCholesky : public SolverWrapper {
public:
    EnvelopeCholesky() : SolverWrapper(false) { name_ = "envelope-cholesky"; }

    void setMatrix(const RSparseMatrix & S) override {
        pattern_(S.vecColPtr(), S.vecRowIdx(), S.rows());
        factor_(S.vecColPtr(), S.vecRowIdx(), &S.vecVals()[0], L_);
    }

    // complex symmetric systems (setComplex(true)): L L^T without conjugation
    void setMatrix(const CSparseMatrix & S) override {
        pattern_(S.vecColPtr(), S.vecRowIdx(), S.rows());
        factor_(S.vecColPtr(), S.vecRowIdx(), &S.vecVals()[0], LC_);
    }

    void solve(const RVector & rhs, RVector & sol) override {
        std::vector<double> y(n_);
        for (int pi = 0; pi < n_; ++pi) y[pi] = rhs[perm_[pi]];
        subst_(L_, y);
        if (sol.size() != (Index)n_) sol.resize(n_);
        for (int pi = 0; pi < n_; ++pi) sol[perm_[pi]] = y[pi];
    }

    void solve(const CVector & rhs, CVector & sol) override {
        std::vector<Complex> y(n_);
        for (int pi = 0; pi < n_; ++pi) y[pi] = rhs[perm_[pi]];
        subst_(LC_, y);
        if (sol.size() != (Index)n_) sol.resize(n_);
        for (int pi = 0; pi < n_; ++pi) sol[perm_[pi]] = y[pi];
    }

private:
    static bool positive_(double d) { return d > 0.0; }
    static bool positive_(const Complex & d) { return std::abs(d) > 0.0; }

    void pattern_(const std::vector<int> & ptr, const std::vector<int> & idx, Index rows) {
        n_ = (int)rows;
        rcm_(ptr, idx);
        // envelope start per permuted row
        first_.assign(n_, 0);
        for (int pi = 0; pi < n_; ++pi) {
            int i = perm_[pi];
            int f = pi;
            for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
                int pj = iperm_[idx[e]];
                if (pj < f) f = pj;
            }
            first_[pi] = f;
        }
        off_.assign(n_ + 1, 0);
        for (int i = 0; i < n_; ++i) off_[i + 1] = off_[i] + (i - first_[i] + 1);
    }

    template <class T>
    void factor_(const std::vector<int> & ptr, const std::vector<int> & idx, const T * val, std::vector<T> & L) {
        L.assign(off_[n_], T(0.0));
        for (int pi = 0; pi < n_; ++pi) {
            int i = perm_[pi];
            for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
                int pj = iperm_[idx[e]];
                if (pj <= pi) L[off_[pi] + pj - first_[pi]] += val[e];
            }
        }
        // row-oriented envelope Cholesky: L(i,j) for first_[i] <= j <= i
        for (int i = 0; i < n_; ++i) {
            T * Li = &L[off_[i]];
            int fi = first_[i];
            for (int j = fi; j < i; ++j) {
                T * Lj = &L[off_[j]];
                int fj = first_[j];
                int s = std::max(fi, fj);
                T acc = Li[j - fi];
                for (int q = s; q < j; ++q) acc -= Li[q - fi] * Lj[q - fj];
                Li[j - fi] = acc / Lj[j - fj];
            }
            T d = Li[i - fi];
            for (int q = fi; q < i; ++q) d -= Li[q - fi] * Li[q - fi];
            if (!positive_(d)) throw std::runtime_error("EnvelopeCholesky: matrix not SPD");
            Li[i - fi] = std::sqrt(d);
        }
    }

    template <class T> void subst_(const std::vector<T> & L, std::vector<T> & y) const {
        for (int i = 0; i < n_; ++i) {
            const T * Li = &L[off_[i]];
            int fi = first_[i];
            T acc = y[i];
            for (int q = fi; q < i; ++q) acc -= Li[q - fi] * y[q];
            y[i] = acc / Li[i - fi];
        }
        for (int i = n_ - 1; i >= 0; --i) {
            const T * Li = &L[off_[i]];
            int fi = first_[i];
            y[i] /= Li[i - fi];
            for (int q = fi; q < i; ++q) y[q] -= Li[q - fi] * y[i];
        }
    }

    void rcm_(const std::vector<int> & ptr, const std::vector<int> & idx) {
        std::vector<int> deg(n_);
        for (int i = 0; i < n_; ++i) deg[i] = ptr[i + 1] - ptr[i];
        std::vector<char> seen(n_, 0);
        std::vector<int> order;
        order.reserve(n_);
        while ((int)order.size() < n_) {
            int start = -1;
            for (int i = 0; i < n_; ++i)
                if (!seen[i] && (start < 0 || deg[i] < deg[start])) start = i;
            // two BFS sweeps to find a pseudo-peripheral start node
            for (int sweep = 0; sweep < 2; ++sweep) {
                std::vector<int> lvl(n_, -1);
                std::deque<int> q;
                q.push_back(start);
                lvl[start] = 0;
                int last = start;
                while (!q.empty()) {
                    int u = q.front(); q.pop_front();
                    last = u;
                    for (int e = ptr[u]; e < ptr[u + 1]; ++e) {
                        int v = idx[e];
                        if (!seen[v] && lvl[v] < 0) { lvl[v] = lvl[u] + 1; q.push_back(v); }
                    }
                }
                start = last;
            }
            std::deque<int> q;
            q.push_back(start);
            seen[start] = 1;
            while (!q.empty()) {
                int u = q.front(); q.pop_front();
                order.push_back(u);
                std::vector<int> nb;
                for (int e = ptr[u]; e < ptr[u + 1]; ++e) {
                    int v = idx[e];
                    if (!seen[v]) { seen[v] = 1; nb.push_back(v); }
                }
                std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
                for (int v : nb) q.push_back(v);
            }
        }
        std::reverse(order.begin(), order.end());
        perm_ = order;
        iperm_.assign(n_, 0);
        for (int i = 0; i < n_; ++i) iperm_[perm_[i]] = i;
    }

    int n_ = 0;
    std::vector<int> perm_, iperm_, first_, off_;
    std::vector<double> L_;
    std::vector<Complex> LC_;
};

struct Session {
    Mesh mesh;
    DataContainerERT data;
    DCMultiElectrodeModelling * fop = nullptr;
    EnvelopeCholesky solver;
    EnvelopeCholesky solverP2;
    RMatrix subpots;
    RMatrix prim;
    int dim = 2;
    Session(int d) : mesh(d) { dim = d; }
    ~Session() { delete fop; }
};

template <class F> int guarded(F f) {
    try {
        f();
        return 0;
    } catch (std::exception & e) {
        g_err = e.what();
        return -1;
    } catch (...) {
        g_err = "unknown exception";
        return -1;
    }
}

} // namespace

extern "C" {

const char * ref_last_error() { return g_err.c_str(); }

// which: 0 = I0, 1 = I1, 2 = K0, 3 = K1 (core/src/numericbase.h:80-180)
void ref_bessel(int which, const double * x, double * y, int n) {
    for (int i = 0; i < n; ++i) {
        switch (which) {
            case 0: y[i] = besselI0(x[i]); break;
            case 1: y[i] = besselI1(x[i]); break;
            case 2: y[i] = besselK0(x[i]); break;
            default: y[i] = besselK1(x[i]); break;
        }
    }
}

void ref_gauss_legendre(double x1, double x2, int n, double * x, double * w) {
    RVector xv, wv;
    GaussLegendre(x1, x2, n, xv, wv);
    for (int i = 0; i < n; ++i) { x[i] = xv[i]; w[i] = wv[i]; }
}

void ref_gauss_laguerre(int n, double * x, double * w) {
    RVector xv, wv;
    GaussLaguerre(n, xv, wv);
    for (int i = 0; i < n; ++i) { x[i] = xv[i]; w[i] = wv[i]; }
}

// initKWaveList(rMin, rMax, nLeg, nLag) (core/src/bert/bertMisc.cpp:89-129)
void ref_kwave(double rMin, double rMax, int nLeg, int nLag, double * k, double * w) {
    RVector kv, wv;
    initKWaveList(rMin, rMax, nLeg, nLag, kv, wv);
    for (Index i = 0; i < kv.size(); ++i) { k[i] = kv[i]; w[i] = wv[i]; }
}

double ref_exact_dc(const double * v, const double * src, double k, double surfaceZ,
                    double fallback) {
    return exactDCSolution(RVector3(v[0], v[1], v[2]), RVector3(src[0], src[1], src[2]),
                           k, surfaceZ, fallback);
}

// Build a reference Mesh.  cells: nc x (dim+1) node ids; bounds: nb x dim node
// ids (edges in 2D, triangles in 3D).  Cell attributes start at 1.
void * ref_session_create(int dim, int nn, const double * xyz, const int * nodeMarker,
                          int nc, const int * cells, const int * cellMarker,
                          int nb, const int * bounds, const int * boundMarker) {
    Session * s = nullptr;
    int rc = guarded([&] {
        s = new Session(dim);
        Mesh & m = s->mesh;
        for (int i = 0; i < nn; ++i)
            m.createNode(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], nodeMarker[i]);
        int nv = dim + 1;
        for (int c = 0; c < nc; ++c) {
            const int * cc = cells + nv * c;
            if (dim == 2)
                m.createTriangle(m.node(cc[0]), m.node(cc[1]), m.node(cc[2]), cellMarker[c]);
            else
                m.createTetrahedron(m.node(cc[0]), m.node(cc[1]), m.node(cc[2]), m.node(cc[3]),
                                    cellMarker[c]);
        }
        for (int b = 0; b < nb; ++b) {
            const int * bb = bounds + dim * b;
            if (dim == 2)
                m.createEdge(m.node(bb[0]), m.node(bb[1]), boundMarker[b]);
            else
                m.createTriangleFace(m.node(bb[0]), m.node(bb[1]), m.node(bb[2]), boundMarker[b]);
        }
        m.createNeighborInfos();
        m.setCellAttributes(RVector(m.cellCount(), 1.0));
    });
    if (rc) { delete s; return nullptr; }
    return s;
}

void ref_session_destroy(void * h) { delete (Session *)h; }

int ref_session_set_data(void * h, int ns, const double * spos, int nd, const int * a,
                         const int * b, const int * m, const int * n, const double * k) {
    Session * s = (Session *)h;
    return guarded([&] {
        s->data = DataContainerERT();
        for (int i = 0; i < ns; ++i)
            s->data.createSensor(RVector3(spos[3 * i], spos[3 * i + 1], spos[3 * i + 2]));
        s->data.resize(nd);
        for (int i = 0; i < nd; ++i) s->data.createFourPointData(i, a[i], b[i], m[i], n[i]);
        if (k) {
            RVector kv(nd);
            for (int i = 0; i < nd; ++i) kv[i] = k[i];
            s->data.set("k", kv);
        }
    });
}

// Create the SR forward operator (ertModelling.py:103 DCSRMultiElectrodeModelling).
// nk < 0: keep the reference's default wavenumbers (initKWaveList).
int ref_session_init_fop(void * h, int nk, const double * kv, const double * wv, int nthreads) {
    Session * s = (Session *)h;
    return guarded([&] {
        delete s->fop;
        s->fop = new DCSRMultiElectrodeModelling(false);
        s->fop->setMesh(s->mesh, true);
        s->fop->setData(s->data);
        if (nk > 0) {
            RVector k(nk), w(nk);
            for (int i = 0; i < nk; ++i) { k[i] = kv[i]; w[i] = wv[i]; }
            s->fop->setkValues(k);
            s->fop->setWeights(w);
        }
        if (nthreads > 0) s->fop->setThreadCount(nthreads);
        s->fop->setSolver(&s->solver);
    });
}

// ERTModelling(sr=False) (ertModelling.py:103-106): the total-field
// DCMultiElectrodeModelling, real (complex = 0) or on complex resistivities
// (complex = 1, setComplex(True)).
int ref_session_init_fop_total(void * h, int nk, const double * kv, const double * wv, int nthreads,
                               int complex) {
    Session * s = (Session *)h;
    return guarded([&] {
        delete s->fop;
        s->fop = new DCMultiElectrodeModelling(false);
        s->fop->setComplex(complex != 0);
        s->fop->setMesh(s->mesh, true);
        s->fop->setData(s->data);
        if (nk > 0) {
            RVector k(nk), w(nk);
            for (int i = 0; i < nk; ++i) { k[i] = kv[i]; w[i] = wv[i]; }
            s->fop->setkValues(k);
            s->fop->setWeights(w);
        }
        if (nthreads > 0) s->fop->setThreadCount(nthreads);
        s->fop->setSolver(&s->solver);
    });
}

// complex response (dcfemmodelling.cpp:1103-1120): model = [re, im], result
// [real, imag] of k * U
int ref_session_response_complex(void * h, const double * re, const double * im, int nm, double * out_re,
                                 double * out_im) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector mod(2 * nm);
        for (int i = 0; i < nm; ++i) { mod[i] = re[i]; mod[nm + i] = im[i]; }
        RVector r(s->fop->response(mod));
        Index D = r.size() / 2;
        for (Index i = 0; i < D; ++i) { out_re[i] = r[i]; out_im[i] = r[D + i]; }
    });
}

// complex Jacobian (dcfemmodelling.cpp:1446-1463, :1410-1444)
int ref_session_jacobian_complex(void * h, const double * re, const double * im, int nm, double * out_re,
                                 double * out_im, int * rows, int * cols) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector mod(2 * nm);
        for (int i = 0; i < nm; ++i) { mod[i] = re[i]; mod[nm + i] = im[i]; }
        s->fop->createJacobian(mod);
        CMatrix * J = dynamic_cast<CMatrix *>(s->fop->jacobian());
        if (!J) throw std::runtime_error("complex Jacobian expected");
        *rows = (int)J->rows();
        *cols = (int)J->cols();
        if (out_re)
            for (Index i = 0; i < J->rows(); ++i)
                for (Index j = 0; j < J->cols(); ++j) {
                    out_re[i * J->cols() + j] = (*J)[i][j].real();
                    out_im[i * J->cols() + j] = (*J)[i][j].imag();
                }
    });
}

// Numerical primary potentials of the SR operator with topography, the steps of
// DCSRMultiElectrodeModelling::checkPrimpotentials_ (dcfemmodelling.cpp:2009-2056)
// run here because its inner DCMultiElectrodeModelling builds a LinSolver, for
// which this image has no back end: the same P2 mesh (Mesh::createP2), cell
// attributes 1, a total-field operator on the same data with this file's
// solver plugged in, its per-k potentials (collectSubPotentials + calculate)
// interpolated to the mesh nodes, installed with setPrimaryPotential.  Call
// after ref_session_init_fop; out: [nK * nE][n] or null.
int ref_session_primary_p2(void * h, double * out) {
    Session * s = (Session *)h;
    return guarded([&] {
        DCSRMultiElectrodeModelling * sr = dynamic_cast<DCSRMultiElectrodeModelling *>(s->fop);
        if (!sr) throw std::runtime_error("primary potentials need the SR operator");
        Mesh p2(s->fop->mesh()->createP2());
        p2.setCellAttributes(1.0);
        DCMultiElectrodeModelling f(s->fop->dataContainer(), false);
        f.setMesh(p2);
        f.setSolver(&s->solverP2);
        RMatrix pp;
        f.collectSubPotentials(pp);
        DataMap dm;
        f.calculate(dm);
        s->prim = RMatrix(pp.rows(), s->fop->mesh()->nodeCount());
        interpolate(p2, pp, s->fop->mesh()->positions(), s->prim, false);
        sr->setPrimaryPotential(s->prim);
        if (out)
            for (Index i = 0; i < s->prim.rows(); ++i)
                for (Index j = 0; j < s->prim.cols(); ++j) out[i * s->prim.cols() + j] = s->prim[i][j];
    });
}

int ref_session_nk(void * h) { return (int)((Session *)h)->fop->kValues().size(); }

void ref_session_kw(void * h, double * k, double * w) {
    Session * s = (Session *)h;
    for (Index i = 0; i < s->fop->kValues().size(); ++i) {
        k[i] = s->fop->kValues()[i];
        w[i] = s->fop->weights()[i];
    }
}

int ref_session_geometric_factors(void * h, double * out) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector k(s->fop->calcGeometricFactor(s->fop->dataContainer()));
        for (Index i = 0; i < k.size(); ++i) out[i] = k[i];
    });
}

// DCMultiElectrodeModelling::response (dcfemmodelling.cpp:1085-1197).  The per-k
// unweighted potentials (rows source + nE*kIdx) are captured into the session.
int ref_session_response(void * h, const double * model, int nm, double * out) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector mod(nm);
        for (int i = 0; i < nm; ++i) mod[i] = model[i];
        s->subpots = RMatrix();
        s->fop->collectSubPotentials(s->subpots);
        RVector r(s->fop->response(mod));
        for (Index i = 0; i < r.size(); ++i) out[i] = r[i];
    });
}

void ref_session_subpot_shape(void * h, int * rows, int * cols) {
    Session * s = (Session *)h;
    *rows = (int)s->subpots.rows();
    *cols = (int)s->subpots.cols();
}

void ref_session_subpots(void * h, double * out) {
    Session * s = (Session *)h;
    Index r = s->subpots.rows(), c = s->subpots.cols();
    for (Index i = 0; i < r; ++i)
        for (Index j = 0; j < c; ++j) out[i * c + j] = s->subpots[i][j];
}

// Potentials for createJacobian without a response(): the per-k unweighted
// matrix [rows = source + nE*kIdx][nodes] handed to the reference core
// (DCMultiElectrodeModelling::collectSubPotentials, dcfemmodelling.h:220);
// prepareJacobianT_ (dcfemmodelling.cpp:1245-1309) then uses it as is.
int ref_session_set_subpots(void * h, int rows, int cols, const double * U) {
    Session * s = (Session *)h;
    return guarded([&] {
        s->subpots = RMatrix(rows, cols);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) s->subpots[i][j] = U[(size_t)i * cols + j];
        s->fop->collectSubPotentials(s->subpots);
    });
}

// RInversion::run (inversion.cpp:135-246) on the session's operator: data
// `dvals` with relative errors `err`, log data / log model transforms
// (TransLogLU with the given bounds; upper 0 = TransLog), the constraint
// matrix given as CSR (installed with ModellingBase::setConstraints after
// the inversion is built, so checkConstraints keeps it), constraint and model
// weights 1, fixed lambda, maxIter iterations without the chi^2 = 1 stop and
// the 2 % abort, line search on or off.  Returns the final model and
// response and per-iteration chi^2 (phiD / nData) in chi2[0..maxIter].
int ref_session_invert(void * h, int nd, const double * dvals, const double * err, int nm, const double * start,
                       double lambda, int maxIter, int lineSearch, int maxCGLS, double cglsTol, double mLower,
                       double mUpper,
                       int nC, const int * crp, const int * cci, const double * cval, double * outModel,
                       double * outResp, double * chi2, int * iters) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector d(nd), e(nd), m0(nm);
        for (int i = 0; i < nd; ++i) { d[i] = dvals[i]; e[i] = err[i]; }
        for (int i = 0; i < nm; ++i) m0[i] = start[i];
        TransLogLU<RVector> tD(0.0, 0.0), tM(mLower, mUpper);
        RInversion inv(d, *s->fop, tD, tM, false, false);
        RSparseMapMatrix C(nC, nm);
        for (int r = 0; r < nC; ++r)
            for (int q = crp[r]; q < crp[r + 1]; ++q) C.setVal(r, cci[q], cval[q]);
        s->fop->setConstraints(&C);
        inv.setRelativeError(e);
        inv.setLambda(lambda);
        inv.setMaxIter(maxIter);
        inv.setLineSearch(lineSearch != 0);
        inv.setMaxCGLSIter(maxCGLS);
        inv.setCGLSTolerance(cglsTol);
        inv.stopAtChi1(false);
        inv.setDeltaPhiAbortPercent(-1e300);
        inv.setModel(m0);
        inv.setCWeight(RVector(nC, 1.0));
        inv.setMWeight(RVector(nm, 1.0));
        const RVector & m = inv.run();
        for (int i = 0; i < nm; ++i) outModel[i] = m[i];
        const RVector & r = inv.response();
        for (int i = 0; i < nd; ++i) outResp[i] = r[i];
        *iters = inv.iter();
        if (chi2) chi2[0] = inv.chi2();
        s->fop->setConstraints(nullptr);
    });
}

// Data-container k after response() (the reference fills missing k analytically).
void ref_session_data_k(void * h, double * out) {
    Session * s = (Session *)h;
    RVector k(s->fop->dataContainer().get("k"));
    for (Index i = 0; i < k.size(); ++i) out[i] = k[i];
}

// createJacobian (dcfemmodelling.cpp:1446-1476), reusing the potentials of the
// last response().
int ref_session_jacobian(void * h, const double * model, int nm, double * out, int * rows,
                         int * cols) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector mod(nm);
        for (int i = 0; i < nm; ++i) mod[i] = model[i];
        s->fop->createJacobian(mod);
        RMatrix * J = dynamic_cast<RMatrix *>(s->fop->jacobian());
        *rows = (int)J->rows();
        *cols = (int)J->cols();
        if (out)
            for (Index i = 0; i < J->rows(); ++i)
                for (Index j = 0; j < J->cols(); ++j) out[i * J->cols() + j] = (*J)[i][j];
    });
}

// System matrix of one wavenumber exactly as DCSRMultiElectrodeModelling::calculateK
// assembles it (dcfemmodelling.cpp:2179-2182): domain part for the given cell
// resistivities, mixed-BC part about sourcePos.  Output CSR (row ptr, col idx, vals).
int ref_assemble(void * h, const double * rho, double k, const double * sourcePos,
                 int * nnz, int * rowptr, int * colidx, double * vals) {
    Session * s = (Session *)h;
    return guarded([&] {
        Mesh m(s->mesh);
        RVector att(m.cellCount());
        for (Index i = 0; i < m.cellCount(); ++i) att[i] = rho[i];
        m.setCellAttributes(att);
        RSparseMatrix S;
        S.buildSparsityPattern(m);
        dcfemDomainAssembleStiffnessMatrix(S, m, k);
        dcfemBoundaryAssembleStiffnessMatrix(S, m,
                                             RVector3(sourcePos[0], sourcePos[1], sourcePos[2]), k);
        *nnz = (int)S.vecRowIdx().size();
        if (rowptr) {
            for (Index i = 0; i < S.vecColPtr().size(); ++i) rowptr[i] = S.vecColPtr()[i];
            for (Index i = 0; i < S.vecRowIdx().size(); ++i) colidx[i] = S.vecRowIdx()[i];
            for (Index i = 0; i < S.vecVals().size(); ++i) vals[i] = S.vecVals()[i];
        }
    });
}

// RSparseMatrix::fillStiffnessMatrix / fillMassMatrix (sparsematrix.h:1039-1065),
// the pg.solver.createStiffnessMatrix / createMassMatrix back end.
int ref_fill_matrix(void * h, int mass, const double * a, int * nnz, int * rowptr,
                    int * colidx, double * vals) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector av(s->mesh.cellCount());
        for (Index i = 0; i < av.size(); ++i) av[i] = a[i];
        RSparseMatrix S;
        if (mass) S.fillMassMatrix(s->mesh, av);
        else S.fillStiffnessMatrix(s->mesh, av);
        *nnz = (int)S.vecRowIdx().size();
        if (rowptr) {
            for (Index i = 0; i < S.vecColPtr().size(); ++i) rowptr[i] = S.vecColPtr()[i];
            for (Index i = 0; i < S.vecRowIdx().size(); ++i) colidx[i] = S.vecRowIdx()[i];
            for (Index i = 0; i < S.vecVals().size(); ++i) vals[i] = S.vecVals()[i];
        }
    });
}

// mixedBoundaryCondition for every boundary of the mesh (dcfemmodelling.cpp:430-506);
// entries for non-mixed boundaries are evaluated all the same.
int ref_mixed_bc(void * h, const double * sourcePos, double k, double * out) {
    Session * s = (Session *)h;
    return guarded([&] {
        RVector3 src(sourcePos[0], sourcePos[1], sourcePos[2]);
        for (Index i = 0; i < s->mesh.boundaryCount(); ++i)
            out[i] = mixedBoundaryCondition(s->mesh.boundary(i), src, k);
    });
}

int ref_boundary_count(void * h) { return (int)((Session *)h)->mesh.boundaryCount(); }

// node ids of boundary i (createNeighborInfos may add interior boundaries)
void ref_boundary_nodes(void * h, int i, int * ids, int * marker) {
    Session * s = (Session *)h;
    const Boundary & b = s->mesh.boundary(i);
    for (Index j = 0; j < b.nodeCount(); ++j) ids[j] = (int)b.node(j).id();
    *marker = b.marker();
}

// File formats, through the reference's own reader and writer.
// Mesh::load without neighbour infos keeps exactly the boundaries the file
// holds (core/src/mesh_io.cpp:30-49); Mesh::save writes binary v3 (:391-538).
int ref_mesh_convert(const char * in, const char * out) {
    return guarded([&] {
        Mesh m;
        m.load(in, false);
        m.save(out);
    });
}

// The session's mesh after Mesh::createNeighborInfos, which adds a boundary
// for every cell face (core/src/mesh.cpp:1684-1740).
int ref_session_save_mesh(void * h, const char * out) {
    return guarded([&] { ((Session *)h)->mesh.save(out); });
}

// DataContainerERT(in, removeInvalid) then save(out, fmtData, fmtSensor, noFilter)
// (core/src/datacontainer.cpp:209-419 load, :519-633 write).  The validity check
// writes invalid.data into the working directory, so the load runs in workdir.
int ref_data_convert(const char * in, const char * out, const char * fmtData,
                     const char * fmtSensor, int removeInvalid, int noFilter,
                     const char * workdir) {
    char old[4096];
    if (!getcwd(old, sizeof(old))) return -1;
    if (chdir(workdir) != 0) return -1;
    int rc = guarded([&] {
        DataContainerERT d(in, removeInvalid != 0);
        // std::string arguments: a const char * would bind to the
        // save(name, fmtData, bool noFilter) overload instead.
        d.save(std::string(out), std::string(fmtData), std::string(fmtSensor), noFilter != 0);
    });
    if (chdir(old) != 0) return -1;
    return rc;
}

} // extern "C"

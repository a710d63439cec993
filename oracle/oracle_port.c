/* Test infrastructure only: C restatement of the reference's sensitivity
 * loop, used as the CPU baseline of bench.py (never the product path).
 *
 * CreateSensitivityColMT::calc2 (core/src/bert/bertJacobian.cpp:170-219):
 * for every cell with marker >= 0, every wavenumber k and every datum i,
 *   J[i][marker] += w_k * (u_m - u_n)^T (K_c + k^2 M_c) (u_a - u_b)
 * over the cell's nodes, potentials row = electrode + nE * kIdx
 * (ElementMatrix::mult_, elementmatrix.h:280-302), threaded over contiguous
 * cell ranges like distributeCalc (core/src/calculateMultiThread.h:59-113).
 * Element matrices are the P1 closed forms (vol grad N_i . grad N_j and
 * vol / ((d+1)(d+2)) (1 + delta_ij)).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int dim, n, nE, nK, D, nParam;
    const double * xyz;
    const int * cells;
    const int * marker;
    const double * kv;
    const double * wv;
    const double * pots;
    const int * abmn;
    double * J;
    long c0, c1;
} job_t;

static void element(int dim, const double * xyz, const int * v, double K[4][4], double M[4][4]) {
    int nv = dim + 1;
    double g[4][3] = {{0}};
    double vol;
    if (dim == 2) {
        double a = xyz[3 * v[1]] - xyz[3 * v[0]], b = xyz[3 * v[1] + 1] - xyz[3 * v[0] + 1];
        double c = xyz[3 * v[2]] - xyz[3 * v[0]], d = xyz[3 * v[2] + 1] - xyz[3 * v[0] + 1];
        double det = a * d - b * c;
        g[1][0] = d / det; g[1][1] = -c / det;
        g[2][0] = -b / det; g[2][1] = a / det;
        vol = 0.5 * fabs(det);
    } else {
        double e[3][3];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) e[r][q] = xyz[3 * v[r + 1] + q] - xyz[3 * v[0] + q];
        double c0[3] = {e[1][1] * e[2][2] - e[1][2] * e[2][1], e[1][2] * e[2][0] - e[1][0] * e[2][2],
                        e[1][0] * e[2][1] - e[1][1] * e[2][0]};
        double c1[3] = {e[2][1] * e[0][2] - e[2][2] * e[0][1], e[2][2] * e[0][0] - e[2][0] * e[0][2],
                        e[2][0] * e[0][1] - e[2][1] * e[0][0]};
        double c2[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                        e[0][0] * e[1][1] - e[0][1] * e[1][0]};
        double det = e[0][0] * c0[0] + e[0][1] * c0[1] + e[0][2] * c0[2];
        for (int q = 0; q < 3; ++q) { g[1][q] = c0[q] / det; g[2][q] = c1[q] / det; g[3][q] = c2[q] / det; }
        vol = fabs(det) / 6.0;
    }
    for (int q = 0; q < 3; ++q) g[0][q] = -(g[1][q] + g[2][q] + g[3][q]);
    for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) {
            K[i][j] = vol * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
            M[i][j] = vol / (nv * (nv + 1)) * (i == j ? 2.0 : 1.0);
        }
}

static void * worker(void * arg) {
    job_t * jb = (job_t *)arg;
    const int nv = jb->dim + 1;
    double K[4][4], M[4][4], S[4][4];
    for (long c = jb->c0; c < jb->c1; ++c) {
        int mk = jb->marker[c];
        if (mk < 0) continue;
        const int * v = jb->cells + c * nv;
        element(jb->dim, jb->xyz, v, K, M);
        for (int kk = 0; kk < jb->nK; ++kk) {
            double k2 = jb->kv[kk] * jb->kv[kk], w = jb->wv[kk];
            for (int i = 0; i < nv; ++i)
                for (int j = 0; j < nv; ++j) S[i][j] = (jb->kv[kk] > 0.0 ? M[i][j] * k2 : 0.0) + K[i][j];
            const double * Uk = jb->pots + (size_t)kk * jb->nE * jb->n;
            for (int d = 0; d < jb->D; ++d) {
                int a = jb->abmn[4 * d], b = jb->abmn[4 * d + 1], m = jb->abmn[4 * d + 2], nn = jb->abmn[4 * d + 3];
                double da[4], dm[4];
                for (int i = 0; i < nv; ++i) {
                    int node = v[i];
                    da[i] = (a > -1 ? Uk[(size_t)a * jb->n + node] : 0.0) - (b > -1 ? Uk[(size_t)b * jb->n + node] : 0.0);
                    dm[i] = (m > -1 ? Uk[(size_t)m * jb->n + node] : 0.0) - (nn > -1 ? Uk[(size_t)nn * jb->n + node] : 0.0);
                }
                double ret = 0.0;
                for (int i = 0; i < nv; ++i) {
                    double t = 0.0;
                    for (int j = 0; j < nv; ++j) t += S[i][j] * da[j];
                    ret += t * dm[i];
                }
                jb->J[(size_t)d * jb->nParam + mk] += ret * w;
            }
        }
    }
    return NULL;
}

/* Cells [cell_begin, cell_end) of the sorted-by-marker order given in `order`
 * (or natural order when order == NULL) are split over nthreads threads.
 * pots: (nK * nE) x n row-major (the reference's subSolutions_ layout). */
int oracle_calc2(int dim, int n, const double * xyz, long C, const int * cells, const int * marker, int nE,
                 int nK, const double * kv, const double * wv, const double * pots, int D, const int * abmn,
                 int nParam, double * J, int nthreads, long cell_begin, long cell_end) {
    if (nthreads < 1) nthreads = 1;
    if (cell_end > C) cell_end = C;
    if (cell_begin < 0) cell_begin = 0;
    pthread_t th[256];
    job_t jobs[256];
    if (nthreads > 256) nthreads = 256;
    long total = cell_end - cell_begin;
    for (int t = 0; t < nthreads; ++t) {
        job_t * jb = &jobs[t];
        jb->dim = dim; jb->n = n; jb->nE = nE; jb->nK = nK; jb->D = D; jb->nParam = nParam;
        jb->xyz = xyz; jb->cells = cells; jb->marker = marker; jb->kv = kv; jb->wv = wv;
        jb->pots = pots; jb->abmn = abmn; jb->J = J;
        jb->c0 = cell_begin + total * t / nthreads;
        jb->c1 = cell_begin + total * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, worker, jb);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* Plain FP64 Jacobi-preconditioned conjugate gradients, every right-hand
 * side an independent recurrence, used as the forward part of bench.py's CPU
 * baseline.  (The reference solves the forward systems with a sparse direct
 * factorisation, cholmodWrapper.cpp:224-459; this per-source CG is the
 * iterative comparison point, the same algorithm as the GPU's plain FP64
 * PCG, so the two are compared at equal iteration counts.)  CSR A (n, rowptr, colidx,
 * vals), B and X row-major n x nrhs, X = 0 start.  Runs until every column
 * reaches |r| <= rtol |b| or maxit iterations; threads split the columns.
 * Returns the iterations run (max over threads). */
typedef struct {
    int n, nrhs, c0, c1, maxit, iters;
    double rtol;
    const int *rowptr, *colidx;
    const double *vals, *B;
    double * X;
} pcg_job_t;

static void * pcg_worker(void * arg) {
    pcg_job_t * j = (pcg_job_t *)arg;
    const int n = j->n, w = j->c1 - j->c0, ld = j->nrhs;
    if (w <= 0) return NULL;
    double * r = malloc(sizeof(double) * (size_t)n * w);
    double * z = malloc(sizeof(double) * (size_t)n * w);
    double * p = malloc(sizeof(double) * (size_t)n * w);
    double * q = malloc(sizeof(double) * (size_t)n * w);
    double * dinv = malloc(sizeof(double) * n);
    double * rz = calloc(w, sizeof(double)), *bb = calloc(w, sizeof(double)), *tmp = calloc(w, sizeof(double));
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int t = j->rowptr[i]; t < j->rowptr[i + 1]; ++t)
            if (j->colidx[t] == i) d = j->vals[t];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
        for (int c = 0; c < w; ++c) {
            const double b = j->B[(size_t)i * ld + j->c0 + c];
            j->X[(size_t)i * ld + j->c0 + c] = 0.0;
            r[(size_t)i * w + c] = b;
            z[(size_t)i * w + c] = dinv[i] * b;
            p[(size_t)i * w + c] = z[(size_t)i * w + c];
            rz[c] += b * z[(size_t)i * w + c];
            bb[c] += b * b;
        }
    }
    int it = 0;
    for (; it < j->maxit; ++it) {
        /* q = A p, pq = p'q */
        for (int c = 0; c < w; ++c) tmp[c] = 0.0;
        for (int i = 0; i < n; ++i) {
            double * qi = q + (size_t)i * w;
            for (int c = 0; c < w; ++c) qi[c] = 0.0;
            for (int t = j->rowptr[i]; t < j->rowptr[i + 1]; ++t) {
                const double a = j->vals[t];
                const double * pk = p + (size_t)j->colidx[t] * w;
                for (int c = 0; c < w; ++c) qi[c] += a * pk[c];
            }
            const double * pi = p + (size_t)i * w;
            for (int c = 0; c < w; ++c) tmp[c] += pi[c] * qi[c];
        }
        double worst = 0.0;
        double * rzn = calloc(w, sizeof(double)), *rr = calloc(w, sizeof(double));
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < w; ++c) {
                const size_t o = (size_t)i * w + c;
                const double al = tmp[c] != 0.0 ? rz[c] / tmp[c] : 0.0;
                j->X[(size_t)i * ld + j->c0 + c] += al * p[o];
                r[o] -= al * q[o];
                z[o] = dinv[i] * r[o];
                rzn[c] += r[o] * z[o];
                rr[c] += r[o] * r[o];
            }
        for (int c = 0; c < w; ++c) {
            const double be = rz[c] != 0.0 ? rzn[c] / rz[c] : 0.0;
            rz[c] = rzn[c];
            tmp[c] = be;
            const double rel = bb[c] > 0.0 ? sqrt(rr[c] / bb[c]) : 0.0;
            if (rel > worst) worst = rel;
        }
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < w; ++c) p[(size_t)i * w + c] = z[(size_t)i * w + c] + tmp[c] * p[(size_t)i * w + c];
        free(rzn);
        free(rr);
        if (worst <= j->rtol) {
            ++it;
            break;
        }
    }
    j->iters = it;
    free(r), free(z), free(p), free(q), free(dinv), free(rz), free(bb), free(tmp);
    return NULL;
}

int oracle_pcg(int n, const int * rowptr, const int * colidx, const double * vals, int nrhs, const double * B,
               double * X, int maxit, double rtol, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nrhs) nthreads = nrhs;
    pthread_t th[64];
    pcg_job_t jobs[64];
    if (nthreads > 64) nthreads = 64;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = (pcg_job_t){n, nrhs, (int)((long)nrhs * t / nthreads), (int)((long)nrhs * (t + 1) / nthreads),
                              maxit, 0, rtol, rowptr, colidx, vals, B, X};
        pthread_create(&th[t], NULL, pcg_worker, &jobs[t]);
    }
    int it = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        if (jobs[t].iters > it) it = jobs[t].iters;
    }
    return it;
}

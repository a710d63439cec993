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

/*
 * lqm_oracle.c -- FP64 CPU restatement of LatticeQM.jl's k-space band path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
 * against, and the timed CPU baseline of the benchmark.  Nothing in the product
 * library links, loads or calls it.
 *
 * It restates, in plain C99 + OpenMP, the per-k loops of the reference:
 *   fourierphase / fouriersum!      TightBinding/types.jl:8, :18-25
 *   Hermitian(M) -> eigvals!/eigen!  Eigen/Eigen.jl:14-39 (LAPACK zheevr, uplo 'U',
 *                                    range 'A', abstol -1: Julia's syevr! call)
 *   bandmatrix / _band_kpoint!       Spectrum/bands.jl:164-227
 *   dos! / dos_compute!              Spectrum/dos.jl:185-232
 *   fermidirac                       Utils/fermidirac.jl:1-5
 *   densitymatrix_kpoint!            Operators/densitymatrix.jl:130-145
 *   densitymatrix_compute_add!       Operators/densitymatrix.jl:107-123
 * with the ThreadedExec semantics of Parallel/Parallel.jl:240-273 (dynamic queue
 * for foreach) and :300-332 (static cld(Nk, nthreads) chunks, one partial per
 * chunk, partials folded serially in chunk order).
 *
 * The eigensolver is LAPACK's zheevr and the GEMM is BLAS zgemm, reached through
 * function pointers that the Python side takes from scipy's exported
 * cython_lapack / cython_blas tables (lqo_set_lapack).  OpenBLAS must be pinned
 * to one thread by the caller, as Parallel.configure_blas! does
 * (Parallel.jl:146-191).
 *
 * Layouts match Julia's: column-major, complex = interleaved (re, im) doubles.
 *   T     : N x N x nR  hopping stack, block r is T + r*N*N
 *   R     : d x nR      int32 lattice vectors
 *   ks    : d x Nk      fractional k-points
 *   evals : N x Nk      ascending eigenvalues per column
 *   rho   : N x N x nRr density-matrix stack
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex zc;

typedef void (*zheevr_fn)(char *jobz, char *range, char *uplo, int *n, zc *a, int *lda,
                          double *vl, double *vu, int *il, int *iu, double *abstol,
                          int *m, double *w, zc *z, int *ldz, int *isuppz, zc *work,
                          int *lwork, double *rwork, int *lrwork, int *iwork,
                          int *liwork, int *info);
typedef void (*zgemm_fn)(char *transa, char *transb, int *m, int *n, int *k, zc *alpha,
                         zc *a, int *lda, zc *b, int *ldb, zc *beta, zc *c, int *ldc);

static zheevr_fn g_zheevr = NULL;
static zgemm_fn g_zgemm = NULL;

void lqo_set_lapack(void *zheevr_ptr, void *zgemm_ptr) {
    g_zheevr = (zheevr_fn)zheevr_ptr;
    g_zgemm = (zgemm_fn)zgemm_ptr;
}

/* e^{+2 pi i k.R}: fourierphase, TightBinding/types.jl:8. */
static zc phase(int d, const double *k, const int *R, double sgn) {
    double dot = 0.0;
    for (int a = 0; a < d; ++a) dot += k[a] * (double)R[a];
    double x = (2.0 * M_PI) * (sgn * dot);
    return cos(x) + I * sin(x);
}

/* H(k) = sum_R T_R e^{2 pi i k.R}, in the stored key order (types.jl:18-25). */
void lqo_fouriersum(int N, int d, int nR, const int *R, const zc *T, const double *k, zc *out) {
    size_t NN = (size_t)N * N;
    memset(out, 0, NN * sizeof(zc));
    for (int r = 0; r < nR; ++r) {
        zc ph = phase(d, k, R + (size_t)r * d, 1.0);
        const zc *t = T + r * NN;
        for (size_t e = 0; e < NN; ++e) out[e] += t[e] * ph;
    }
}

/* Per-thread LAPACK workspace for zheevr. */
typedef struct {
    int N, lwork, lrwork, liwork;
    zc *work, *Z;
    double *rwork;
    int *iwork, *isuppz;
} eig_ws;

static int ws_init(eig_ws *ws, int N) {
    memset(ws, 0, sizeof(*ws));
    ws->N = N;
    if (!g_zheevr) return -100;
    /* workspace query, as LAPACK.syevr! does */
    char jobz = 'V', range = 'A', uplo = 'U';
    int n = N, lda = N > 1 ? N : 1, il = 0, iu = 0, m = 0, ldz = lda, info = 0;
    double vl = 0, vu = 0, abstol = -1.0;
    zc wq; double rq; int iq;
    int lw = -1, lrw = -1, liw = -1;
    double *w = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
    zc A1 = 0, Z1 = 0;
    int isz[2];
    g_zheevr(&jobz, &range, &uplo, &n, &A1, &lda, &vl, &vu, &il, &iu, &abstol, &m, w, &Z1, &ldz,
             isz, &wq, &lw, &rq, &lrw, &iq, &liw, &info);
    free(w);
    if (info != 0) return info;
    ws->lwork = (int)creal(wq) + 1;
    ws->lrwork = (int)rq + 1;
    ws->liwork = iq + 1;
    ws->work = (zc *)malloc(sizeof(zc) * ws->lwork);
    ws->rwork = (double *)malloc(sizeof(double) * ws->lrwork);
    ws->iwork = (int *)malloc(sizeof(int) * ws->liwork);
    ws->isuppz = (int *)malloc(sizeof(int) * 2 * (N > 0 ? N : 1));
    ws->Z = (zc *)malloc(sizeof(zc) * (size_t)N * N);
    return 0;
}

static void ws_free(eig_ws *ws) {
    free(ws->work); free(ws->rwork); free(ws->iwork); free(ws->isuppz); free(ws->Z);
}

/* Hermitian(A) (upper triangle) -> zheevr.  A is overwritten.  If vecs != NULL the
 * eigenvectors are written there (N x N, column n for eigenvalue n). */
static int eigh(eig_ws *ws, zc *A, double *w, zc *vecs) {
    char jobz = vecs ? 'V' : 'N', range = 'A', uplo = 'U';
    int n = ws->N, lda = n > 1 ? n : 1, il = 0, iu = 0, m = 0, ldz = lda, info = 0;
    double vl = 0, vu = 0, abstol = -1.0;
    g_zheevr(&jobz, &range, &uplo, &n, A, &lda, &vl, &vu, &il, &iu, &abstol, &m, w,
             vecs ? vecs : ws->Z, &ldz, ws->isuppz, ws->work, &ws->lwork, ws->rwork, &ws->lrwork,
             ws->iwork, &ws->liwork, &info);
    if (info == 0 && m != n) info = -200;
    return info;
}

/* Hermitian eigensolve of one matrix (no Bloch sum): test hook for the eigensolver. */
int lqo_eigh(int N, const zc *A, double *w, zc *vecs) {
    eig_ws ws;
    int st = ws_init(&ws, N);
    if (st) return st;
    zc *a = (zc *)malloc(sizeof(zc) * (size_t)N * N);
    memcpy(a, A, sizeof(zc) * (size_t)N * N);
    st = eigh(&ws, a, w, vecs);
    free(a);
    ws_free(&ws);
    return st;
}

/* bandmatrix without projectors (bands.jl:164-227): dynamic k-queue, as the
 * threaded foreach executor (Parallel.jl:248-262).  evecs may be NULL (the
 * eigenvalues-only branch, bands.jl:214-221). */
int lqo_bands(int N, int d, int nR, const int *R, const zc *T, int Nk, const double *ks,
              double *evals, zc *evecs, int nthreads) {
    int status = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        eig_ws ws;
        int st = ws_init(&ws, N);
        zc *Hk = (zc *)malloc(sizeof(zc) * (size_t)N * N);
        if (st) {
#pragma omp atomic write
            status = st;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int j = 0; j < Nk; ++j) {
                lqo_fouriersum(N, d, nR, R, T, ks + (size_t)j * d, Hk);
                int s = eigh(&ws, Hk, evals + (size_t)j * N,
                             evecs ? evecs + (size_t)j * N * N : NULL);
                if (s) {
#pragma omp atomic write
                    status = s;
                }
            }
        }
        free(Hk);
        ws_free(&ws);
    }
    return status;
}

/* Static chunking of Parallel._chunked_ranges (Parallel.jl:329-332): chunk c
 * covers [c*sz, min((c+1)*sz, Nk)) with sz = cld(Nk, nchunks). */
static void chunk_bounds(int Nk, int nchunks, int c, int *lo, int *hi) {
    int sz = (Nk + nchunks - 1) / nchunks;
    *lo = c * sz;
    *hi = (c + 1) * sz < Nk ? (c + 1) * sz : Nk;
    if (*lo > Nk) *lo = Nk;
}

/* getdos (dos.jl:185-232): per k, eigenvalues only, then
 * DOS(w) += weight * Im[1/(w - i Gamma - e)]; unweighted sum is divided by L. */
int lqo_dos(int N, int d, int nR, const int *R, const zc *T, int Nk, const double *ks,
            const double *kweights, int Nw, const double *omegas, double gamma, double *dos,
            int nthreads) {
    if (nthreads < 1) nthreads = 1;
    double *partials = (double *)calloc((size_t)nthreads * Nw, sizeof(double));
    int status = 0;
#pragma omp parallel num_threads(nthreads)
    {
        int c = omp_get_thread_num();
        int lo, hi;
        chunk_bounds(Nk, nthreads, c, &lo, &hi);
        eig_ws ws;
        int st = ws_init(&ws, N);
        zc *Hk = (zc *)malloc(sizeof(zc) * (size_t)N * N);
        double *ev = (double *)malloc(sizeof(double) * N);
        double *acc = partials + (size_t)c * Nw;
        if (st) {
#pragma omp atomic write
            status = st;
        } else {
            for (int j = lo; j < hi; ++j) {
                lqo_fouriersum(N, d, nR, R, T, ks + (size_t)j * d, Hk);
                int s = eigh(&ws, Hk, ev, NULL);
                if (s) {
#pragma omp atomic write
                    status = s;
                }
                double wk = kweights ? kweights[j] : 1.0;
                for (int n = 0; n < N; ++n)
                    for (int o = 0; o < Nw; ++o) {
                        zc den = (omegas[o] - I * gamma - ev[n]);
                        acc[o] += cimag(wk / den);
                    }
            }
        }
        free(Hk); free(ev);
        ws_free(&ws);
    }
    for (int o = 0; o < Nw; ++o) dos[o] = 0.0;
    for (int c = 0; c < nthreads; ++c)
        for (int o = 0; o < Nw; ++o) dos[o] += partials[(size_t)c * Nw + o];
    if (!kweights)
        for (int o = 0; o < Nw; ++o) dos[o] /= (double)Nk;
    free(partials);
    return status;
}

/* fermidirac (Utils/fermidirac.jl:1-5). */
double lqo_fermidirac(double e, double mu, double Temp) {
    double x = e - mu;
    if (Temp > 0) return 1.0 / (exp(x / Temp) + 1.0);
    return x < 0 ? 1.0 : (x > 0 ? 0.0 : 0.5);
}

/* getdensitymatrix! (densitymatrix.jl:28-145): rho is zeroed, every pre-existing
 * key Rr accumulates w_k e^{-2 pi i k.Rr} U diag(f) U^H; returns E_band through
 * *eband.  Static chunks with one partial stack per chunk, folded in order. */
int lqo_densitymatrix(int N, int d, int nR, const int *R, const zc *T, int Nk, const double *ks,
                      const double *kweights, double mu, double Temp, int nRr, const int *Rr,
                      zc *rho, double *eband, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    size_t NN = (size_t)N * N;
    zc *partials = (zc *)calloc((size_t)nthreads * nRr * NN, sizeof(zc));
    double *scal = (double *)calloc(nthreads, sizeof(double));
    int status = 0;
    if (!g_zgemm) status = -101;
#pragma omp parallel num_threads(nthreads)
    {
        int c = omp_get_thread_num();
        int lo, hi;
        chunk_bounds(Nk, nthreads, c, &lo, &hi);
        eig_ws ws;
        int st = ws_init(&ws, N);
        zc *Hk = (zc *)malloc(sizeof(zc) * NN);
        zc *U = (zc *)malloc(sizeof(zc) * NN);
        zc *UD = (zc *)malloc(sizeof(zc) * NN);
        zc *rk = (zc *)malloc(sizeof(zc) * NN);
        double *ev = (double *)malloc(sizeof(double) * N);
        double *fd = (double *)malloc(sizeof(double) * N);
        zc *acc = partials + (size_t)c * nRr * NN;
        double s_acc = 0.0;
        if (st || !g_zgemm) {
#pragma omp atomic write
            status = st ? st : -101;
        } else {
            for (int j = lo; j < hi; ++j) {
                const double *k = ks + (size_t)j * d;
                double wk = kweights ? kweights[j] : 1.0 / (double)Nk;
                lqo_fouriersum(N, d, nR, R, T, k, Hk);
                int s = eigh(&ws, Hk, ev, U);
                if (s) {
#pragma omp atomic write
                    status = s;
                }
                for (int n = 0; n < N; ++n) fd[n] = lqo_fermidirac(ev[n], mu, Temp);
                /* UD = U * Diagonal(fd); rho_k = UD * U' (densitymatrix.jl:139-140) */
                for (int n = 0; n < N; ++n)
                    for (int i = 0; i < N; ++i) UD[(size_t)n * N + i] = U[(size_t)n * N + i] * fd[n];
                char ta = 'N', tb = 'C';
                int nn = N;
                zc one = 1.0, zero = 0.0;
                g_zgemm(&ta, &tb, &nn, &nn, &nn, &one, UD, &nn, U, &nn, &zero, rk, &nn);
                for (int r = 0; r < nRr; ++r) {
                    zc ph = wk * phase(d, k, Rr + (size_t)r * d, -1.0);
                    zc *dst = acc + (size_t)r * NN;
                    for (size_t e = 0; e < NN; ++e) dst[e] += ph * rk[e];
                }
                double eb = 0.0;
                for (int n = 0; n < N; ++n) eb += ev[n] * fd[n];
                s_acc += wk * eb;
            }
        }
        scal[c] = s_acc;
        free(Hk); free(U); free(UD); free(rk); free(ev); free(fd);
        ws_free(&ws);
    }
    memset(rho, 0, sizeof(zc) * nRr * NN);
    double etot = 0.0;
    for (int c = 0; c < nthreads; ++c) {
        const zc *p = partials + (size_t)c * nRr * NN;
        for (size_t e = 0; e < (size_t)nRr * NN; ++e) rho[e] += p[e];
        etot += scal[c];
    }
    *eband = etot;
    free(partials);
    free(scal);
    return status;
}

/* Thread count the OpenMP runtime would use by default (for reporting). */
int lqo_max_threads(void) { return omp_get_max_threads(); }

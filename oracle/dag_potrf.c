/*
 * dag_potrf.c -- CPU tile-DAG executor of the oracle (test infrastructure,
 * and the CPU baseline that bench.py times).
 *
 * Restates src/zpotrf_L.jdf: the four task classes
 *   POTRF(k)      :91-208   T <- k==0 ? A(k,k) : SYRK(k-1,k)
 *   TRSM(m,k)     :214-343  T <- POTRF(k); C <- k==0 ? A(m,k) : GEMM(m,k,k-1)
 *   SYRK(k,m)     :349-464  A <- TRSM(m,k); T <- k==0 ? A(m,m) : SYRK(k-1,m)
 *   GEMM(m,n,k)   :466-633  A <- TRSM(m,k); B <- TRSM(n,k);
 *                           C <- k==0 ? A(m,n) : GEMM(m,n,k-1)
 * with the priority formulas of :53-64 (PRI_CHANGE, PRI_MAX :79-80), the
 * info rule k*mb + iinfo set only while still zero (:200-202), and the CPU
 * bodies CORE_dpotrf / dtrsm(Right,Lower,Trans,NonUnit) / dsyrk(Lower,NoTrans,
 * -1,1) / dgemm(N,T,-1,1) with tempmm for the ragged last tile row.
 * Workers pop the highest-priority ready task (PaRSEC semantics: larger value
 * first).  Every tile receives its updates in ascending k (RW chains :365-368,
 * :488-491), so the result does not depend on the schedule.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "oracle.h"

int  orc_k_potrf(int use_blas, int n, double *A, int lda);
void orc_k_trsm_rltn(int use_blas, int m, int n, const double *T, int ldt, double *C, int ldc);
void orc_k_syrk_ln(int use_blas, int n, int k, const double *A, int lda, double *C, int ldc);
void orc_k_gemm_nt(int use_blas, int m, int n, int k, const double *A, int lda,
                   const double *B, int ldb, double *C, int ldc);
void orc_k_gemm(int use_blas, int ta, int tb, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb, double beta,
                double *C, int ldc);
void orc_k_trsm(int use_blas, int side, int uplo, int trans, int diag, int m, int n,
                double alpha, const double *A, int lda, double *B, int ldb);

double orc_wtime(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec / 1e6;
}

enum { T_POTRF = 0, T_TRSM = 1, T_SYRK = 2, T_GEMM = 3 };

typedef struct { int type, m, n, k; long prio; } task_t;

typedef struct {
    int N, NB, MT, use_blas;
    long pri_change, pri_max;
    double *T;
    int *cnt_potrf, *cnt_trsm, *cnt_syrk, *cnt_gemm;
    task_t *heap; int heap_n;
    long done, total;
    int info;
    pthread_mutex_t mu;
    pthread_cond_t cv;
} dag_t;

static inline double *tile(dag_t *d, int m, int n)
{
    return d->T + orc_sym_tile_index(ORC_LOWER, d->MT, m, n) * (int64_t)d->NB * d->NB;
}

/* priorities, zpotrf_L.jdf:53-64 */
static long prio(dag_t *d, int type, int m, int n, int k)
{
    long MT = d->MT;
    switch (type) {
    case T_POTRF: return (k >= MT - d->pri_change) ? (MT - k) * (MT - k) * (MT - k) : d->pri_max;
    case T_TRSM:  return (m >= MT - d->pri_change) ?
                   (MT - m) * (MT - m) * (MT - m) + 3 * ((2 * MT) - k - m - 1) * (m - k) : d->pri_max;
    case T_SYRK:  return (m >= MT - d->pri_change) ? (MT - m) * (MT - m) * (MT - m) + 3 * (m - k) : d->pri_max;
    default:      return (m >= MT - d->pri_change) ?
                   (MT - m) * (MT - m) * (MT - m) + 3 * ((2 * MT) - m - n - 3) * (m - n) + 6 * (m - k)
                   : d->pri_max;
    }
    (void)n;
}

static void heap_push(dag_t *d, task_t t)
{
    int i = d->heap_n++;
    d->heap[i] = t;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (d->heap[p].prio >= d->heap[i].prio) break;
        task_t tmp = d->heap[p]; d->heap[p] = d->heap[i]; d->heap[i] = tmp;
        i = p;
    }
}

static task_t heap_pop(dag_t *d)
{
    task_t top = d->heap[0];
    d->heap[0] = d->heap[--d->heap_n];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, b = i;
        if (l < d->heap_n && d->heap[l].prio > d->heap[b].prio) b = l;
        if (r < d->heap_n && d->heap[r].prio > d->heap[b].prio) b = r;
        if (b == i) break;
        task_t tmp = d->heap[b]; d->heap[b] = d->heap[i]; d->heap[i] = tmp;
        i = b;
    }
    return top;
}

static void make_ready(dag_t *d, int type, int m, int n, int k)
{
    task_t t = { type, m, n, k, prio(d, type, m, n, k) };
    heap_push(d, t);
}

#define GIDX(d, m, n, k) (((int64_t)(k) * (d)->MT + (m)) * (d)->MT + (n))

/* decrement the in-degree of a successor; push it when ready (mutex held) */
static void release(dag_t *d, int type, int m, int n, int k)
{
    int *c;
    switch (type) {
    case T_POTRF: c = &d->cnt_potrf[k]; break;
    case T_TRSM:  c = &d->cnt_trsm[(int64_t)m * d->MT + k]; break;
    case T_SYRK:  c = &d->cnt_syrk[(int64_t)k * d->MT + m]; break;
    default:      c = &d->cnt_gemm[GIDX(d, m, n, k)]; break;
    }
    if (--(*c) == 0) make_ready(d, type, m, n, k);
}

static int tsize(dag_t *d, int m) { return (m == d->MT - 1) ? d->N - m * d->NB : d->NB; }

static void run_task(dag_t *d, task_t t)
{
    int NB = d->NB;
    switch (t.type) {
    case T_POTRF: {
        int iinfo = orc_k_potrf(d->use_blas, tsize(d, t.k), tile(d, t.k, t.k), NB);
        if (iinfo != 0) {
            pthread_mutex_lock(&d->mu);
            if (d->info == 0) d->info = t.k * NB + iinfo;
            pthread_mutex_unlock(&d->mu);
        }
        break; }
    case T_TRSM:
        orc_k_trsm_rltn(d->use_blas, tsize(d, t.m), NB, tile(d, t.k, t.k), NB, tile(d, t.m, t.k), NB);
        break;
    case T_SYRK:
        orc_k_syrk_ln(d->use_blas, tsize(d, t.m), NB, tile(d, t.m, t.k), NB, tile(d, t.m, t.m), NB);
        break;
    default:
        orc_k_gemm_nt(d->use_blas, tsize(d, t.m), NB, NB, tile(d, t.m, t.k), NB,
                      tile(d, t.n, t.k), NB, tile(d, t.m, t.n), NB);
        break;
    }
}

/* successors, from the output dependencies of each task class */
static void complete(dag_t *d, task_t t)
{
    int MT = d->MT;
    switch (t.type) {
    case T_POTRF:
        for (int m = t.k + 1; m < MT; m++) release(d, T_TRSM, m, 0, t.k);
        break;
    case T_TRSM:
        release(d, T_SYRK, t.m, 0, t.k);
        for (int n = t.k + 1; n < t.m; n++) release(d, T_GEMM, t.m, n, t.k);
        for (int m = t.m + 1; m < MT; m++) release(d, T_GEMM, m, t.m, t.k);
        break;
    case T_SYRK:
        if (t.m == t.k + 1) release(d, T_POTRF, 0, 0, t.m);
        else release(d, T_SYRK, t.m, 0, t.k + 1);
        break;
    default:
        if (t.n == t.k + 1) release(d, T_TRSM, t.m, 0, t.n);
        else release(d, T_GEMM, t.m, t.n, t.k + 1);
        break;
    }
    d->done++;
}

static void *worker(void *arg)
{
    dag_t *d = (dag_t *)arg;
    pthread_mutex_lock(&d->mu);
    for (;;) {
        while (d->heap_n == 0 && d->done < d->total) pthread_cond_wait(&d->cv, &d->mu);
        if (d->done >= d->total) break;
        task_t t = heap_pop(d);
        pthread_mutex_unlock(&d->mu);
        run_task(d, t);
        pthread_mutex_lock(&d->mu);
        int before = d->heap_n;
        complete(d, t);
        if (d->heap_n > before + 1 || d->done >= d->total) pthread_cond_broadcast(&d->cv);
        else if (d->heap_n > before) pthread_cond_signal(&d->cv);
    }
    pthread_cond_broadcast(&d->cv);
    pthread_mutex_unlock(&d->mu);
    return NULL;
}

int orc_dag_dpotrf_L(int N, int NB, double *T, int nthreads, int use_blas, int pri_change)
{
    dag_t d;
    memset(&d, 0, sizeof d);
    d.N = N; d.NB = NB; d.MT = (N + NB - 1) / NB; d.T = T; d.use_blas = use_blas;
    if (N <= 0) return 0;
    int MT = d.MT;
    d.pri_change = (pri_change < 0) ? MT : pri_change;
    d.pri_max = (long)MT * (6 + (long)MT * (6 + MT));
    d.cnt_potrf = (int *)calloc(MT, sizeof(int));
    d.cnt_trsm = (int *)calloc((size_t)MT * MT, sizeof(int));
    d.cnt_syrk = (int *)calloc((size_t)MT * MT, sizeof(int));
    d.cnt_gemm = (int *)calloc((size_t)MT * MT * MT, sizeof(int));
    long ntask = MT;
    for (int k = 0; k < MT; k++) d.cnt_potrf[k] = (k > 0) ? 1 : 0;
    for (int m = 1; m < MT; m++)
        for (int k = 0; k < m; k++) {
            d.cnt_trsm[(int64_t)m * MT + k] = 1 + (k > 0);
            d.cnt_syrk[(int64_t)k * MT + m] = 1 + (k > 0);
            ntask += 2;
        }
    for (int k = 0; k < MT - 2; k++)
        for (int m = k + 2; m < MT; m++)
            for (int n = k + 1; n < m; n++) {
                d.cnt_gemm[GIDX(&d, m, n, k)] = 2 + (k > 0);
                ntask++;
            }
    d.total = ntask;
    d.heap = (task_t *)malloc(sizeof(task_t) * (size_t)(ntask + 1));
    pthread_mutex_init(&d.mu, NULL);
    pthread_cond_init(&d.cv, NULL);
    make_ready(&d, T_POTRF, 0, 0, 0);

    if (nthreads < 1) nthreads = 1;
    if (use_blas) orc_blas_set_threads(1);   /* one BLAS thread per worker */
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, worker, &d);
    worker(&d);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);

    pthread_mutex_destroy(&d.mu);
    pthread_cond_destroy(&d.cv);
    free(th); free(d.heap);
    free(d.cnt_potrf); free(d.cnt_trsm); free(d.cnt_syrk); free(d.cnt_gemm);
    return d.info;
}

int orc_seq_dpotrf_L(int N, int NB, double *T, int use_blas)
{
    dag_t d;
    memset(&d, 0, sizeof d);
    d.N = N; d.NB = NB; d.MT = (N + NB - 1) / NB; d.T = T; d.use_blas = use_blas;
    pthread_mutex_init(&d.mu, NULL);
    for (int k = 0; k < d.MT; k++) {
        task_t t = { T_POTRF, 0, 0, k, 0 };
        run_task(&d, t);
        for (int m = k + 1; m < d.MT; m++) { task_t s = { T_TRSM, m, 0, k, 0 }; run_task(&d, s); }
        for (int m = k + 1; m < d.MT; m++) { task_t s = { T_SYRK, m, 0, k, 0 }; run_task(&d, s); }
        for (int n = k + 1; n < d.MT; n++)
            for (int m = n + 1; m < d.MT; m++) { task_t s = { T_GEMM, m, n, k, 0 }; run_task(&d, s); }
    }
    pthread_mutex_destroy(&d.mu);
    return d.info;
}

/* ---------------------------------------------------------------- solve */
/* Left/Lower tile TRSM on B (N x NRHS, tiles NB x NB, ld NB, column tile
 * width NB): ztrsm_LLN.jdf:29-60,156-174 (trans == NoTrans) and
 * ztrsm_LLT.jdf:29-60,155-171 (trans != NoTrans); alpha is applied at the
 * first step (lalpha) and 1 afterwards. */
void orc_tile_dtrsm_LL(int trans, int diag, double alpha, int N, int NB,
                       const double *T, int NRHS, double *B)
{
    int MT = (N + NB - 1) / NB, NT = (NRHS + NB - 1) / NB;
#define LT(m, n) (T + orc_sym_tile_index(ORC_LOWER, MT, (m), (n)) * (int64_t)NB * NB)
#define BT(m, n) (B + ((int64_t)(n) * MT + (m)) * (int64_t)NB * NB)
#define TS(m) (((m) == MT - 1) ? N - (m) * NB : NB)
#define TN(n) (((n) == NT - 1) ? NRHS - (n) * NB : NB)
    for (int n = 0; n < NT; n++) {
        if (trans == ORC_NOTRANS) {
            for (int k = 0; k < MT; k++) {
                double la = (k == 0) ? alpha : 1.0;
                orc_k_trsm(0, ORC_LEFT, ORC_LOWER, ORC_NOTRANS, diag, TS(k), TN(n), la,
                           LT(k, k), NB, BT(k, n), NB);
                for (int m = k + 1; m < MT; m++)
                    orc_k_gemm(0, ORC_NOTRANS, ORC_NOTRANS, TS(m), TN(n), NB, -1.0,
                               LT(m, k), NB, BT(k, n), NB, la, BT(m, n), NB);
            }
        } else {
            for (int k = MT - 1; k >= 0; k--) {
                double la = (k == MT - 1) ? alpha : 1.0;
                orc_k_trsm(0, ORC_LEFT, ORC_LOWER, ORC_TRANS, diag, TS(k), TN(n), la,
                           LT(k, k), NB, BT(k, n), NB);
                for (int m = 0; m < k; m++)
                    orc_k_gemm(0, ORC_TRANS, ORC_NOTRANS, NB, TN(n), TS(k), -1.0,
                               LT(k, m), NB, BT(k, n), NB, la, BT(m, n), NB);
            }
        }
    }
#undef LT
#undef BT
#undef TS
#undef TN
}

/* dplasma_dpotrs, lower: LLN then LLT with alpha = 1 (zpotrs_wrapper.c:91-94) */
void orc_tile_dpotrs_L(int N, int NB, const double *T, int NRHS, double *B)
{
    orc_tile_dtrsm_LL(ORC_NOTRANS, ORC_NONUNIT, 1.0, N, NB, T, NRHS, B);
    orc_tile_dtrsm_LL(ORC_TRANS, ORC_NONUNIT, 1.0, N, NB, T, NRHS, B);
}

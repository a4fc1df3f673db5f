/*
 * lapack_chain.c -- the reference's likelihood chain on the host through
 * MKL's LAPACK/BLAS: the CPU baseline the benchmark times beside the GPU.
 *
 * TEST INFRASTRUCTURE / BASELINE ONLY: loaded by tests/ and bench.py's
 * cpu_baseline leg; the product (libbv) never calls it.
 *
 * Per block, the op sequence of include/llh_Xvecchia_batch.h with the
 * reference's MAGMA calls replaced by their LAPACK/BLAS counterparts:
 *   S22 = K(N,N), S21 = K(N,B), S11 = K(B,B)     core_dcmg        :117-214
 *   dpotrf(S22)                                  magma_dpotrf     :295-299
 *   dtrsm(L22, S21), dtrsm(L22, yN)              magmablas_dtrsm  :320-337
 *   O = A^T A, mu = A^T z  (dgemm, both full)    magmablas_dgemm  :352-378
 *   S11 -= O, yB -= mu (blocks i >= 1)           magmablas_dgeadd :407-431
 *   dpotrf(S11), dtrsm(Ls, yB)                                    :449-466
 *   logdet = sum log(Ls_jj^2) over Ls_jj > 0     core_Xlogdet (vecchia_helper_method.h:263-280)
 *   dot = dnrm2(r)^2                                              :475-482
 *   llh = sum_i -(dot_i + logdet_i + bs_i log 2 pi) / 2 in block order
 * The covariance uses the closed forms for nu = 1/2, 3/2, 5/2 (the
 * benchmark's kinds); other orders return NaN (the oracle covers them).
 * OpenMP over blocks, one MKL thread per block.  MKL (libmkl_rt) is opened at
 * run time with dlopen (no headers are shipped; the prototypes below are
 * MKL's documented C interfaces).
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { LC_COL_MAJOR = 102, LC_NO_TRANS = 111, LC_TRANS = 112, LC_LOWER = 122, LC_NON_UNIT = 131,
       LC_LEFT = 141 };

typedef int (*lc_potrf_t)(int layout, char uplo, int n, double *a, int lda);
typedef void (*lc_trsm_t)(int layout, int side, int uplo, int transa, int diag, int m, int n,
                          double alpha, const double *a, int lda, double *b, int ldb);
typedef void (*lc_gemm_t)(int layout, int ta, int tb, int m, int n, int k, double alpha,
                          const double *a, int lda, const double *b, int ldb, double beta,
                          double *c, int ldc);
typedef double (*lc_nrm2_t)(int n, const double *x, int incx);
typedef int (*lc_setthr_t)(int n);
typedef void (*lc_version_t)(char *buf, int len);

static lc_potrf_t lc_potrf;
static lc_trsm_t lc_trsm;
static lc_gemm_t lc_gemm;
static lc_nrm2_t lc_nrm2;
static lc_setthr_t lc_set_local;
static char lc_version[256];
static char lc_path[512];

/* Opens MKL: path, else libmkl_rt.so on the loader path, else
   /opt/conda/lib/libmkl_rt.so.  Returns 0 or -1. */
int lapack_chain_open(const char *path) {
    if (lc_potrf) return 0;
    const char *cands[3] = {path, "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so"};
    void *h = NULL;
    for (int i = 0; i < 3 && !h; ++i) {
        if (!cands[i] || !cands[i][0]) continue;
        h = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
        if (h) snprintf(lc_path, sizeof(lc_path), "%s", cands[i]);
    }
    if (!h) return -1;
    lc_potrf = (lc_potrf_t)dlsym(h, "LAPACKE_dpotrf");
    lc_trsm = (lc_trsm_t)dlsym(h, "cblas_dtrsm");
    lc_gemm = (lc_gemm_t)dlsym(h, "cblas_dgemm");
    lc_nrm2 = (lc_nrm2_t)dlsym(h, "cblas_dnrm2");
    lc_set_local = (lc_setthr_t)dlsym(h, "MKL_Set_Num_Threads_Local");
    lc_version_t ver = (lc_version_t)dlsym(h, "MKL_Get_Version_String");
    if (!lc_potrf || !lc_trsm || !lc_gemm || !lc_nrm2) { lc_potrf = NULL; return -1; }
    if (ver) ver(lc_version, (int)sizeof(lc_version));
    return 0;
}

const char *lapack_chain_library(void) { return lc_version[0] ? lc_version : lc_path; }

static double lc_dist(double x1, double y1, double x2, double y2) {
    return sqrt(pow((x2 - x1), 2) + pow((y2 - y1), 2));
}

static double lc_cov(double d, const double *th) {
    const double e = d / th[1];
    if (e == 0.0) return th[0];
    if (th[2] == 0.5) return th[0] * exp(-e);
    if (th[2] == 1.5) return th[0] * (1.0 + e) * exp(-e);
    if (th[2] == 2.5) return th[0] * (1.0 + e + e * e / 3.0) * exp(-e);
    return NAN;
}

static void lc_fill(double *A, int lda, int m, int n, const double *x1, const double *y1,
                    const double *x2, const double *y2, const double *th) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i)
            A[i + (int64_t)j * lda] = lc_cov(lc_dist(x1[i], y1[i], x2[j], y2[j]), th);
}

/* llh of the 2-D Matern likelihood (closed-form nu) of a partitioned dataset
   (block order, neighbour indices nbr[B*cs]); per-block terms optional. */
double lapack_chain_llh(int64_t n, int B, int cs, const int32_t *bn, const double *x,
                        const double *y, const double *obs, const int32_t *nbr,
                        const double *theta, int threads, double *logdet_out, double *dot_out) {
    if (!lc_potrf && lapack_chain_open(NULL) != 0) return NAN;
    int64_t *acc = (int64_t *)malloc(sizeof(int64_t) * (B + 1));
    acc[0] = 0;
    int bmax = 1;
    for (int i = 0; i < B; ++i) {
        acc[i + 1] = acc[i] + bn[i];
        if (bn[i] > bmax) bmax = bn[i];
    }
    (void)n;
    double *ld = (double *)malloc(sizeof(double) * B), *dt = (double *)malloc(sizeof(double) * B);
    int nthr = threads > 0 ? threads : 1;
#pragma omp parallel num_threads(nthr)
    {
        if (lc_set_local) lc_set_local(1);
        const int cmax = cs > 0 ? cs : 1;
        double *S22 = (double *)malloc(sizeof(double) * (size_t)cmax * cmax);
        double *S21 = (double *)malloc(sizeof(double) * (size_t)cmax * bmax);
        double *S11 = (double *)malloc(sizeof(double) * (size_t)bmax * bmax);
        double *O = (double *)malloc(sizeof(double) * (size_t)bmax * bmax);
        double *vn = (double *)malloc(sizeof(double) * cmax);
        double *xn = (double *)malloc(sizeof(double) * cmax);
        double *yn = (double *)malloc(sizeof(double) * cmax);
        double *vb = (double *)malloc(sizeof(double) * bmax);
        double *mu = (double *)malloc(sizeof(double) * bmax);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < B; ++i) {
            const int bs = bn[i];
            const int c = (i == 0) ? 0 : cs;
            const double *xb = x + acc[i], *yb = y + acc[i];
            lc_fill(S11, bs, bs, bs, xb, yb, xb, yb, theta);
            memcpy(vb, obs + acc[i], sizeof(double) * bs);
            if (c > 0) {
                for (int k = 0; k < c; ++k) {
                    const int32_t q = nbr[(int64_t)i * cs + k];
                    xn[k] = x[q]; yn[k] = y[q]; vn[k] = obs[q];
                }
                lc_fill(S22, c, c, c, xn, yn, xn, yn, theta);
                lc_fill(S21, c, c, bs, xn, yn, xb, yb, theta);
                lc_potrf(LC_COL_MAJOR, 'L', c, S22, c);
                lc_trsm(LC_COL_MAJOR, LC_LEFT, LC_LOWER, LC_NO_TRANS, LC_NON_UNIT, c, bs, 1.0,
                        S22, c, S21, c);
                lc_trsm(LC_COL_MAJOR, LC_LEFT, LC_LOWER, LC_NO_TRANS, LC_NON_UNIT, c, 1, 1.0,
                        S22, c, vn, c);
                lc_gemm(LC_COL_MAJOR, LC_TRANS, LC_NO_TRANS, bs, bs, c, 1.0, S21, c, S21, c, 0.0,
                        O, bs);
                lc_gemm(LC_COL_MAJOR, LC_TRANS, LC_NO_TRANS, bs, 1, c, 1.0, S21, c, vn, c, 0.0,
                        mu, bs);
                for (int64_t k = 0; k < (int64_t)bs * bs; ++k) S11[k] -= O[k];
                for (int k = 0; k < bs; ++k) vb[k] -= mu[k];
            }
            const int info = lc_potrf(LC_COL_MAJOR, 'L', bs, S11, bs);
            lc_trsm(LC_COL_MAJOR, LC_LEFT, LC_LOWER, LC_NO_TRANS, LC_NON_UNIT, bs, 1, 1.0, S11,
                    bs, vb, bs);
            double s = 0.0;
            for (int k = 0; k < bs; ++k) {
                const double l = S11[k + (int64_t)k * bs];
                if (l > 0) s += log(l * l);
            }
            const double nrm = lc_nrm2(bs, vb, 1);
            ld[i] = info == 0 ? s : NAN;
            dt[i] = nrm * nrm;
        }
        free(S22); free(S21); free(S11); free(O); free(vn); free(vb); free(mu); free(xn);
        free(yn);
    }
    const double l2pi = log(2 * 3.141592653589793);
    double llh = 0.0;
    for (int i = 0; i < B; ++i) {
        llh += -(dt[i] + ld[i] + bn[i] * l2pi) * 0.5;
        if (logdet_out) logdet_out[i] = ld[i];
        if (dot_out) dot_out[i] = dt[i];
    }
    free(acc); free(ld); free(dt);
    return llh;
}

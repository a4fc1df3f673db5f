/*
 * oracle.c -- CPU restatement of the Block Vecchia log-likelihood chain.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker (and, for the
 * baseline, as the timed CPU port).  The product (libbv) never calls it.
 *
 * What it restates, operation by operation:
 *   covariance  core_dcmg (src/ckernel.cpp:1057-1093) with calculateDistance
 *               (:165-186) and distanceEarth (:145-154); the space-time kernel
 *               core_dcmg_spacetime_matern (:1421-1458).  Matern for every
 *               kernel id except 4 (include/llh_Xvecchia_batch.h:126-144).
 *   K_nu        GSL's gsl_sf_bessel_Knu (GSL 2.6, not vendored) is replaced by
 *               the integral K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt,
 *               trapezoidal rule in long double (step <= min(0.1, 0.25/sqrt x),
 *               error far below double rounding); an algorithm independent of
 *               the device's Temme / continued-fraction evaluation.  Orders
 *               1/2, 3/2, 5/2 use the elementary forms of K_nu.
 *   per block   exactly the reference's chain (include/llh_Xvecchia_batch.h):
 *               S22 = K(N,N), S21 = K(N,B), S11 = K(B,B)            :117-214
 *               L22 = potrf(S22) (lower)                            :295-299
 *               A = L22^-1 S21, z = L22^-1 yN  (trsm, left lower)   :320-337
 *               O = A^T A (full gemm), mu = A^T z                   :352-378
 *               S11 -= O, yB -= mu  for blocks i >= 1               :407-431
 *               Ls = potrf(S11), r = Ls^-1 yB                       :449-466
 *               logdet = sum log(Ls_jj^2) over Ls_jj > 0   core_Xlogdet
 *                        (include/vecchia_helper_method.h:263-280)
 *               nrm = dnrm2(r) (scaled sum of squares), dot = nrm^2  :475-482
 *               llh = sum_i -(dot_i + logdet_i + bs_i log(2 PI)) / 2 in block
 *               order, PI = 3.141592653589793                        :480-494
 *               (include/vecchia_helper_data.h:20)
 *   Block 0's conditioning result is discarded, as in the reference (the
 *   geadd loop starts at i = 1), so its chain is skipped here.
 * A block whose potrf fails reports info = first failing column + 1 (LAPACK
 * convention) and its terms are NaN; the reference continues with garbage.
 * Parallel over blocks with OpenMP (the CPU baseline uses this); the final
 * sum is sequential in block order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.141592653589793
#define OR_EARTH_KM 6371.0

/* ---------------- K_nu ---------------- */
/* Half-integer orders have elementary forms: K_{1/2}(x) = sqrt(pi/(2x)) e^-x,
   K_{3/2} = K_{1/2} (1 + 1/x), K_{5/2} = K_{1/2} (1 + 3/x + 3/x^2).  Every other
   order: K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt by the trapezoidal
   rule in long double. */
double oracle_besselk(double nu, double x) {
    if (!(x > 0.0)) return INFINITY;
    if (nu == 0.5 || nu == 1.5 || nu == 2.5) {
        const double k12 = sqrt(OR_PI / (2.0 * x)) * exp(-x);
        if (nu == 0.5) return k12;
        if (nu == 1.5) return k12 * (1.0 + 1.0 / x);
        return k12 * (1.0 + 3.0 / x + 3.0 / (x * x));
    }
    long double lx = x, lnu = nu;
    long double h = 0.1L;
    long double hs = 0.25L / sqrtl(lx);
    if (hs < h) h = hs;
    /* e^x K_nu(x) = int exp(-x (cosh t - 1)) cosh(nu t) dt */
    long double sum = 0.5L; /* t = 0 term: exp(0) cosh(0) / 2 */
    const long double tpeak = asinhl(lnu / lx);
    for (int k = 1; k < 200000; ++k) {
        const long double t = k * h;
        const long double term = expl(-lx * (coshl(t) - 1.0L)) * coshl(lnu * t);
        sum += term;
        if (t > tpeak && term < 1e-24L * sum) break;
    }
    return (double)(sum * h * expl(-lx));
}

/* ---------------- covariance ---------------- */
static double or_dist(double x1, double y1, double z1, double x2, double y2, double z2, int metric,
                      int use_z) {
    if (!use_z) {
        if (metric == 1) {
            const double lat1r = x1 * OR_PI / 180, lon1r = y1 * OR_PI / 180;
            const double lat2r = x2 * OR_PI / 180, lon2r = y2 * OR_PI / 180;
            const double u = sin((lat2r - lat1r) / 2), v = sin((lon2r - lon1r) / 2);
            return 2.0 * OR_EARTH_KM * asin(sqrt(u * u + cos(lat1r) * cos(lat2r) * v * v));
        }
        return sqrt(pow((x2 - x1), 2) + pow((y2 - y1), 2));
    }
    return sqrt(pow((x2 - x1), 2) + pow((y2 - y1), 2) + pow((z2 - z1), 2));
}

typedef struct {
    int kernel, metric, use_z;
    double dist_scale;
    const double *theta;
    double con;
} or_kern;

static void or_kern_init(or_kern *k, const double *theta, int kernel, int metric, int use_z,
                         double dist_scale) {
    k->kernel = kernel;
    k->metric = metric;
    k->use_z = use_z;
    k->dist_scale = dist_scale;
    k->theta = theta;
    double con = pow(2, (theta[2] - 1)) * tgamma(theta[2]);
    con = 1.0 / con;
    k->con = theta[0] * con;
}

static double or_cov(const or_kern *k, double x1, double y1, double z1, double x2, double y2,
                     double z2) {
    const double *th = k->theta;
    if (k->kernel == 4) {
        const double expr = or_dist(x1, y1, z1, x2, y2, z2, k->metric, 1) / th[1];
        const double dt = sqrt(pow(z1 - z2, 2));
        const double expr2 = pow(pow(dt, 2 * th[4]) / th[3] + 1.0, th[5] / 2.0);
        const double expr3 = expr / expr2;
        const double expr4 = pow(pow(dt, 2 * th[4]) / th[3] + 1.0, th[5] + th[6]);
        if (expr == 0) return th[0] / expr4;
        return k->con * pow(expr3, th[2]) * oracle_besselk(th[2], expr3) / expr4;
    }
    double expr = or_dist(x1, y1, z1, x2, y2, z2, k->metric, k->use_z) / th[1];
    expr /= k->dist_scale;
    if (expr == 0) return th[0];
    return k->con * pow(expr, th[2]) * oracle_besselk(th[2], expr);
}

/* A[i + j*m] = K(l1[i], l2[j]) */
int oracle_matern(int m, int n, const double *x1, const double *y1, const double *z1,
                  const double *x2, const double *y2, const double *z2, const double *theta,
                  int kernel, int metric, double dist_scale, double *A) {
    or_kern k;
    or_kern_init(&k, theta, kernel, metric, (z1 && z2) ? 1 : 0, dist_scale);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i)
            A[i + (int64_t)j * m] = or_cov(&k, x1[i], y1[i], z1 ? z1[i] : 0.0, x2[j], y2[j],
                                           z2 ? z2[j] : 0.0);
    return 0;
}

/* ---------------- dense kernels (column-major) ---------------- */
/* lower Cholesky in place, right-looking; returns 0 or failing column + 1 */
static int or_potrf(int n, double *a, int lda) {
    for (int k = 0; k < n; ++k) {
        double akk = a[k + (int64_t)k * lda];
        if (!(akk > 0.0)) return k + 1;
        akk = sqrt(akk);
        a[k + (int64_t)k * lda] = akk;
        const double r = 1.0 / akk;
        double *ck = a + (int64_t)k * lda;
        for (int i = k + 1; i < n; ++i) ck[i] *= r;
        for (int j = k + 1; j < n; ++j) {
            const double ljk = ck[j];
            double *cj = a + (int64_t)j * lda;
            for (int i = j; i < n; ++i) cj[i] -= ck[i] * ljk;
        }
    }
    return 0;
}

/* B <- L^-1 B, L lower non-unit n x n, B n x nrhs */
static void or_trsm(int n, int nrhs, const double *l, int ldl, double *b, int ldb) {
    for (int c = 0; c < nrhs; ++c) {
        double *x = b + (int64_t)c * ldb;
        for (int k = 0; k < n; ++k) {
            const double xk = x[k] / l[k + (int64_t)k * ldl];
            x[k] = xk;
            const double *lk = l + (int64_t)k * ldl;
            for (int i = k + 1; i < n; ++i) x[i] -= lk[i] * xk;
        }
    }
}

/* scaled Euclidean norm (reference-BLAS dnrm2 algorithm) */
static double or_nrm2(int n, const double *x) {
    if (n < 1) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; ++i) {
        if (x[i] != 0.0) {
            const double a = fabs(x[i]);
            if (scale < a) {
                ssq = 1.0 + ssq * (scale / a) * (scale / a);
                scale = a;
            } else {
                ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

/* ---------------- the likelihood ---------------- */
/* Block-ordered inputs (x, y, z, obs of length n; batch_num[B]); conditioning
 * sets as indices nbr[B*cs] (row 0 ignored).  Outputs per block logdet[B],
 * dot[B], info[B] (any may be NULL).  Returns llh, NaN on a failed block. */
double oracle_llh(int64_t n, int B, int cs, const int32_t *batch_num, const double *x,
                  const double *y, const double *z, const double *obs, const int32_t *nbr,
                  const double *theta, int kernel, int metric, int time_flag, double dist_scale,
                  int threads, double *logdet_out, double *dot_out, int32_t *info_out) {
    int64_t *accum = (int64_t *)malloc(sizeof(int64_t) * (B + 1));
    double *ld = (double *)malloc(sizeof(double) * B);
    double *dt = (double *)malloc(sizeof(double) * B);
    int32_t *inf = (int32_t *)malloc(sizeof(int32_t) * B);
    accum[0] = 0;
    for (int i = 0; i < B; ++i) accum[i + 1] = accum[i] + batch_num[i];
    (void)n;
    or_kern kp;
    const int use_z = (time_flag || kernel == 4) && z;
    or_kern_init(&kp, theta, kernel, metric, use_z, dist_scale);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < B; ++i) {
        const int bs = batch_num[i];
        const int64_t o = accum[i];
        const int cond = (i >= 1 && cs > 0);
        double *s11 = (double *)malloc(sizeof(double) * bs * bs);
        double *yb = (double *)malloc(sizeof(double) * bs);
        for (int c = 0; c < bs; ++c) {
            yb[c] = obs[o + c];
            for (int r = 0; r < bs; ++r)
                s11[r + (int64_t)c * bs] =
                    or_cov(&kp, x[o + r], y[o + r], use_z ? z[o + r] : 0.0, x[o + c], y[o + c],
                           use_z ? z[o + c] : 0.0);
        }
        int info = 0;
        if (cond) {
            const int32_t *nb = nbr + (int64_t)i * cs;
            double *s22 = (double *)malloc(sizeof(double) * cs * cs);
            double *s21 = (double *)malloc(sizeof(double) * cs * bs);
            double *yn = (double *)malloc(sizeof(double) * cs);
            double *ox = (double *)malloc(sizeof(double) * bs * bs);
            double *mu = (double *)malloc(sizeof(double) * bs);
            for (int c = 0; c < cs; ++c) {
                const int64_t qc = nb[c];
                yn[c] = obs[qc];
                for (int r = 0; r < cs; ++r) {
                    const int64_t pr = nb[r];
                    s22[r + (int64_t)c * cs] = or_cov(&kp, x[pr], y[pr], use_z ? z[pr] : 0.0,
                                                      x[qc], y[qc], use_z ? z[qc] : 0.0);
                }
            }
            for (int c = 0; c < bs; ++c)
                for (int r = 0; r < cs; ++r) {
                    const int64_t pr = nb[r];
                    s21[r + (int64_t)c * cs] =
                        or_cov(&kp, x[pr], y[pr], use_z ? z[pr] : 0.0, x[o + c], y[o + c],
                               use_z ? z[o + c] : 0.0);
                }
            const int i22 = or_potrf(cs, s22, cs);
            if (i22) info = -i22; /* conditioning factor failed */
            or_trsm(cs, bs, s22, cs, s21, cs);
            or_trsm(cs, 1, s22, cs, yn, cs);
            for (int c = 0; c < bs; ++c) {
                for (int r = 0; r < bs; ++r) {
                    double acc = 0.0;
                    for (int k = 0; k < cs; ++k)
                        acc += s21[k + (int64_t)r * cs] * s21[k + (int64_t)c * cs];
                    ox[r + (int64_t)c * bs] = acc;
                }
                double am = 0.0;
                for (int k = 0; k < cs; ++k) am += s21[k + (int64_t)c * cs] * yn[k];
                mu[c] = am;
            }
            for (int c = 0; c < bs; ++c) {
                for (int r = 0; r < bs; ++r) s11[r + (int64_t)c * bs] += -1. * ox[r + (int64_t)c * bs];
                yb[c] += -1 * mu[c];
            }
            free(s22); free(s21); free(yn); free(ox); free(mu);
        }
        const int is = or_potrf(bs, s11, bs);
        if (is && info == 0) info = is;
        double ldv = 0.0, dv = NAN;
        if (info == 0) {
            or_trsm(bs, 1, s11, bs, yb, bs);
            for (int k = 0; k < bs; ++k) {
                const double lkk = s11[k + (int64_t)k * bs];
                if (lkk > 0) ldv += log(lkk * lkk);
            }
            const double nr = or_nrm2(bs, yb);
            dv = nr * nr;
        } else {
            ldv = NAN;
        }
        ld[i] = ldv;
        dt[i] = dv;
        inf[i] = info;
        free(s11); free(yb);
    }
    double llk = 0.0;
    for (int k = 0; k < B; ++k) llk += -(dt[k] + ld[k] + batch_num[k] * log(2 * OR_PI)) * 0.5;
    if (logdet_out) memcpy(logdet_out, ld, sizeof(double) * B);
    if (dot_out) memcpy(dot_out, dt, sizeof(double) * B);
    if (info_out) memcpy(info_out, inf, sizeof(int32_t) * B);
    free(accum); free(ld); free(dt); free(inf);
    return llk;
}

/* Exact dense Gaussian log-likelihood of (x, y, obs): the closed form that a
 * single block reduces to.  -1/2 (y' S^-1 y + log|S| + n log 2 PI). */
double oracle_dense_llh(int n, const double *x, const double *y, const double *z,
                        const double *obs, const double *theta, int kernel, int metric,
                        double dist_scale) {
    or_kern kp;
    or_kern_init(&kp, theta, kernel, metric, z ? 1 : 0, dist_scale);
    double *s = (double *)malloc(sizeof(double) * (int64_t)n * n);
    double *r = (double *)malloc(sizeof(double) * n);
    for (int c = 0; c < n; ++c) {
        r[c] = obs[c];
        for (int q = 0; q < n; ++q)
            s[q + (int64_t)c * n] = or_cov(&kp, x[q], y[q], z ? z[q] : 0.0, x[c], y[c],
                                           z ? z[c] : 0.0);
    }
    double out = NAN;
    if (or_potrf(n, s, n) == 0) {
        or_trsm(n, 1, s, n, r, n);
        double ldv = 0.0, dv = 0.0;
        for (int k = 0; k < n; ++k) {
            ldv += log(s[k + (int64_t)k * n] * s[k + (int64_t)k * n]);
            dv += r[k] * r[k];
        }
        out = -(dv + ldv + n * log(2 * OR_PI)) * 0.5;
    }
    free(s); free(r);
    return out;
}

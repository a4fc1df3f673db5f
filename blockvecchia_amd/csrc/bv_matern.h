/*
 * bv_matern.h -- Matern covariance entry and the modified Bessel function
 * K_nu(x) for the device (also compiles for the host).
 *
 * The covariance follows core_dcmg (src/ckernel.cpp:1057-1093):
 *   x = dist(a, b) / beta;  x /= dist_scale
 *   K = sigma2                                  if x == 0
 *     = con * x^nu * K_nu(x),  con = sigma2 / (2^(nu-1) Gamma(nu))  otherwise
 * with the distance of calculateDistance (src/ckernel.cpp:165-186): Euclidean
 * sqrt(dx^2 + dy^2 [+ dz^2]) or the haversine great-circle distance
 * (src/ckernel.cpp:145-154, R = 6371 km, degrees in).
 *
 * For nu = 1/2, 3/2, 5/2 the closed forms sigma2 e^-x, sigma2 (1+x) e^-x and
 * sigma2 (1 + x + x^2/3) e^-x are used.  For general nu, K_nu is computed here
 * (GSL's gsl_sf_bessel_Knu is the reference's routine; it is not used):
 *   nu = mu + N with |mu| <= 1/2, N = floor(nu + 1/2);
 *   x <= 2 : Temme's series (N. M. Temme, J. Comput. Phys. 19 (1975) 324)
 *            for K_mu and K_{mu+1}, with 1/Gamma(1 +- mu) from its Taylor series;
 *   x >  2 : the minimal solution u_n = U(mu+1/2+n, 2mu+1, 2x) of
 *            u_{n-1} - 2(n+x) u_n + ((n+1/2)^2 - mu^2) u_{n+1} = 0
 *            by backward (Miller) recurrence, normalised with
 *            sum_n C_n u_n = (2x)^(-mu-1/2),  C_n = prod_{k<n} ((k+1/2)^2 - mu^2)/(k+1),
 *            which gives K_mu = sqrt(pi/(2x)) e^-x / S,  S = sum_n C_n u_n/u_0, and
 *            K_{mu+1} = K_mu (mu + 1/2 + x + (mu^2 - 1/4) u_1/u_0) / x
 *            (the Steed/Temme "CF2" normalisation, here evaluated by Horner's
 *            rule inside the backward sweep);
 *   then K_{k+1} = K_{k-1} + (2k/x) K_k up to nu.
 * Relative error against mpmath <= 5e-15 over nu in [0.01, 3], x in [1e-8, 700].
 */
#ifndef BV_MATERN_H
#define BV_MATERN_H

#include <math.h>
#include <stdint.h>
#include "bv_rgamma_coef.h"

#if defined(__HIPCC__)
#define BV_HD __host__ __device__ inline
#else
#define BV_HD inline
#endif

#define BV_PI 3.141592653589793
#define BV_EARTH_RADIUS_KM 6371.0

/* kernel "kinds" selected per evaluation */
enum { BV_KIND_NU05 = 0, BV_KIND_NU15 = 1, BV_KIND_NU25 = 2, BV_KIND_GENERAL = 3,
       BV_KIND_SPACETIME = 4 };

/* Parameters of one covariance evaluation, prepared on the host. */
struct bv_kparams {
    double sigma2;      /* theta[0] */
    double beta;        /* theta[1] */
    double nu;          /* theta[2] */
    double con;         /* sigma2 / (2^(nu-1) Gamma(nu)), as src/ckernel.cpp:1072-1074 */
    double dist_scale;
    double inv_scale;   /* 1 / (beta * dist_scale) */
    double beta_t, nu_t, sep, aux;   /* space-time theta[3..6] */
    double nugget;      /* added to the diagonal by the simulation only (0 in the
                           likelihood, which has no nugget, as the reference) */
    const double *cheb; /* general nu: the evaluation's x^nu K_nu table (device) */
    int32_t metric;     /* 0 Euclidean, 1 great circle */
    int32_t use_z;      /* 3-D distance */
};

/* 1/Gamma(1+z), |z| <= 1/2 */
BV_HD double bv_rgamma1p(double z) {
    const double a[BV_RGAMMA_NCOEF] = BV_RGAMMA_COEFS;
    double s = a[BV_RGAMMA_NCOEF - 1];
    for (int k = BV_RGAMMA_NCOEF - 2; k >= 0; --k) s = s * z + a[k];
    return s;
}

/* Temme's series: K_mu(x), K_{mu+1}(x) for |mu| <= 1/2, 0 < x <= 2. */
BV_HD void bv_knu_temme(double mu, double x, double *k0, double *k1) {
    const double a[BV_RGAMMA_NCOEF] = BV_RGAMMA_COEFS;
    const double mu2 = mu * mu;
    /* gamma1 = (1/G(1-mu) - 1/G(1+mu)) / (2 mu) = -sum_{k odd} a_k mu^(k-1)
       gamma2 = (1/G(1-mu) + 1/G(1+mu)) / 2    =  sum_{k even} a_k mu^k     */
    double g1 = 0.0, g2 = 0.0;
    for (int k = BV_RGAMMA_NCOEF - 1; k >= 1; k -= 2) g1 = g1 * mu2 + a[k];
    for (int k = BV_RGAMMA_NCOEF - 2; k >= 0; k -= 2) g2 = g2 * mu2 + a[k];
    g1 = -g1;
    const double rgp = g2 - mu * g1;   /* 1/Gamma(1+mu) */
    const double rgm = g2 + mu * g1;   /* 1/Gamma(1-mu) */
    const double lx = log(2.0 / x);
    const double sig = mu * lx;
    double shs;
    if (fabs(sig) < 1e-3) {
        const double s2 = sig * sig;
        shs = 1.0 + s2 / 6.0 * (1.0 + s2 / 20.0 * (1.0 + s2 / 42.0));
    } else {
        shs = sinh(sig) / sig;
    }
    const double pim = (fabs(mu) < 1e-15) ? 1.0 : (BV_PI * mu) / sin(BV_PI * mu);
    const double es = exp(sig);
    double f = pim * (0.5 * (es + 1.0 / es) * g1 + shs * lx * g2);
    double p = 0.5 * es / rgp;
    double q = 0.5 / (es * rgm);
    double c = 1.0;
    double sk = f, sk1 = p;
    const double x2 = 0.25 * x * x;
    for (int k = 1; k < 64; ++k) {
        const double dk = (double)k;
        f = (dk * f + p + q) / (dk * dk - mu2);
        p = p / (dk - mu);
        q = q / (dk + mu);
        c = c * x2 / dk;
        const double t = c * f;
        sk += t;
        sk1 += c * (p - dk * f);
        if (fabs(t) < 1e-17 * fabs(sk)) break;
    }
    *k0 = sk;
    *k1 = sk1 * 2.0 / x;
}

/* Backward-recurrence CF2: K_mu(x), K_{mu+1}(x) for |mu| <= 1/2, x > 2
   (scaled: both times e^x). */
BV_HD void bv_knu_cf2(double mu, double x, double *k0, double *k1, int scaled = 0) {
    const double mu2 = mu * mu;
    const int nmax = 12 + (int)(120.0 / x);
    double un1 = 0.0, un = 1.0, t = 1.0;    /* u_{n+1}, u_n, Horner sum T_n */
    for (int n = nmax; n >= 1; --n) {
        const double dn = (double)n;
        const double um1 = 2.0 * (dn + x) * un - ((dn + 0.5) * (dn + 0.5) - mu2) * un1;
        t = um1 + ((dn - 0.5) * (dn - 0.5) - mu2) / dn * t;
        un1 = un;
        un = um1;
        if (fabs(un) > 1e200) { un1 *= 1e-200; un *= 1e-200; t *= 1e-200; }
    }
    const double s = t / un;
    const double k = sqrt(BV_PI / (2.0 * x)) * (scaled ? 1.0 : exp(-x)) / s;
    *k0 = k;
    *k1 = k / x * (mu + 0.5 + x + (mu2 - 0.25) * un1 / un);
}

/* e^x K_nu(x), nu >= 0, x > 0 (no underflow for large x): Temme's series
   times e^x for x <= 2, the backward recurrence without its e^-x factor
   above, then the forward recurrence in the order (linear, so the common
   factor e^x passes through it). */
BV_HD double bv_besselk_scaled(double nu, double x) {
    const int nl = (int)(nu + 0.5);
    const double mu = nu - (double)nl;
    double k0, k1;
    if (x <= 2.0) {
        bv_knu_temme(mu, x, &k0, &k1);
        const double ex = exp(x);
        k0 *= ex;
        k1 *= ex;
    } else {
        bv_knu_cf2(mu, x, &k0, &k1, 1);
    }
    for (int k = 1; k <= nl; ++k) {
        const double kn = k0 + 2.0 * (mu + (double)k) / x * k1;
        k0 = k1;
        k1 = kn;
    }
    return k0;
}

/* K_nu(x), nu >= 0, x > 0. */
BV_HD double bv_besselk(double nu, double x) {
    const int nl = (int)(nu + 0.5);
    const double mu = nu - (double)nl;
    double k0, k1;
    if (x <= 2.0) bv_knu_temme(mu, x, &k0, &k1);
    else bv_knu_cf2(mu, x, &k0, &k1);
    for (int k = 1; k <= nl; ++k) {
        const double kn = k0 + 2.0 * (mu + (double)k) / x * k1;
        k0 = k1;
        k1 = kn;
    }
    return k0;
}

/* ---- K_nu with the nu-only work done once --------------------------------
 * Everything that depends only on nu (mu, N, the Temme coefficients
 * 1/Gamma(1 +- mu), pi mu / sin(pi mu), the reciprocals of the series'
 * denominators k^2 - mu^2, k - mu, k + mu, k, and the backward-recurrence
 * coefficients) is tabulated once per covariance parameter set, so the
 * per-entry work is one log, two exp, a few divisions and multiply-adds.
 * Same algorithm and accuracy as bv_besselk. */
#define BV_KT 24 /* Temme terms: for x <= 2 the series converges in < 16 */
#define BV_KC 74 /* backward recurrence: n <= 12 + 120 / x < 73 for x > 2 */
struct bv_knu_tab {
    double mu, g1, g2, inv_rgp, inv_rgm, pim, nu;
    int32_t nl, pad_;
    double rk2[BV_KT + 1], rkm[BV_KT + 1], rkp[BV_KT + 1], rk[BV_KT + 1];
    double an[BV_KC + 1], bn[BV_KC + 1]; /* (n+1/2)^2 - mu^2, ((n-1/2)^2 - mu^2) / n */
};

/* Fill the entries i0, i0 + stride, ... of the table (threads of a
   workgroup share the work); i0 == 0 also fills the scalars. */
BV_HD void bv_knu_tab_fill(double nu, bv_knu_tab *t, int i0, int stride) {
    const int nl = (int)(nu + 0.5);
    const double mu = nu - (double)nl, mu2 = mu * mu;
    if (i0 == 0) {
        const double a[BV_RGAMMA_NCOEF] = BV_RGAMMA_COEFS;
        double g1 = 0.0, g2 = 0.0;
        for (int k = BV_RGAMMA_NCOEF - 1; k >= 1; k -= 2) g1 = g1 * mu2 + a[k];
        for (int k = BV_RGAMMA_NCOEF - 2; k >= 0; k -= 2) g2 = g2 * mu2 + a[k];
        g1 = -g1;
        t->mu = mu;
        t->nu = nu;
        t->nl = nl;
        t->g1 = g1;
        t->g2 = g2;
        t->inv_rgp = 1.0 / (g2 - mu * g1);
        t->inv_rgm = 1.0 / (g2 + mu * g1);
        t->pim = (fabs(mu) < 1e-15) ? 1.0 : (BV_PI * mu) / sin(BV_PI * mu);
    }
    for (int k = i0; k <= BV_KT; k += stride) {
        const double dk = (double)k;
        t->rk2[k] = k ? 1.0 / (dk * dk - mu2) : 0.0;
        t->rkm[k] = k ? 1.0 / (dk - mu) : 0.0;
        t->rkp[k] = k ? 1.0 / (dk + mu) : 0.0;
        t->rk[k] = k ? 1.0 / dk : 0.0;
    }
    for (int n = i0; n <= BV_KC; n += stride) {
        const double dn = (double)n;
        t->an[n] = (dn + 0.5) * (dn + 0.5) - mu2;
        t->bn[n] = n ? ((dn - 0.5) * (dn - 0.5) - mu2) / dn : 0.0;
    }
}

/* x^nu K_nu(x) for x > 0 from the table; lx_out = log x. */
BV_HD double bv_xnu_knu_tab(double x, const bv_knu_tab &t) {
    const double mu = t.mu;
    const double lnx = log(x);
    const double inv_x = 1.0 / x;
    double k0, k1;
    if (x <= 2.0) {
        /* Temme's series (bv_knu_temme) */
        const double lx = 0.6931471805599453 - lnx; /* log(2 / x) */
        const double sig = mu * lx;
        const double es = exp(sig), ies = 1.0 / es;
        double shs; /* sinh(sig) / sig */
        if (fabs(sig) < 0.5) {
            /* sum_k sig^2k / (2k+1)!, 9 terms (the rest < 1e-19 relative) */
            const double s2 = sig * sig;
            double h = 1.0 / 6402373705728000.0; /* 1 / 19! */
            h = h * s2 + 1.0 / 355687428096000.0;  /* 1 / 17! */
            h = h * s2 + 1.0 / 1307674368000.0;    /* 1 / 15! */
            h = h * s2 + 1.0 / 6227020800.0;       /* 1 / 13! */
            h = h * s2 + 1.0 / 39916800.0;         /* 1 / 11! */
            h = h * s2 + 1.0 / 362880.0;           /* 1 / 9! */
            h = h * s2 + 1.0 / 5040.0;             /* 1 / 7! */
            h = h * s2 + 1.0 / 120.0;              /* 1 / 5! */
            h = h * s2 + 1.0 / 6.0;                /* 1 / 3! */
            shs = h * s2 + 1.0;
        } else {
            shs = 0.5 * (es - ies) / sig;
        }
        double f = t.pim * (0.5 * (es + ies) * t.g1 + shs * lx * t.g2);
        double p = 0.5 * es * t.inv_rgp;
        double q = 0.5 * ies * t.inv_rgm;
        double c = 1.0;
        double sk = f, sk1 = p;
        const double x2 = 0.25 * x * x;
        for (int k = 1; k <= BV_KT; ++k) {
            const double dk = (double)k;
            f = (dk * f + p + q) * t.rk2[k];
            p = p * t.rkm[k];
            q = q * t.rkp[k];
            c = c * x2 * t.rk[k];
            const double tt = c * f;
            sk += tt;
            sk1 += c * (p - dk * f);
            if (fabs(tt) < 1e-17 * fabs(sk)) break;
        }
        k0 = sk;
        k1 = sk1 * 2.0 * inv_x;
    } else {
        /* backward-recurrence CF2 (bv_knu_cf2) */
        const int nmax = 12 + (int)(120.0 * inv_x);
        double un1 = 0.0, un = 1.0, tt = 1.0;
        for (int n = nmax; n >= 1; --n) {
            const double dn = (double)n;
            const double um1 = 2.0 * (dn + x) * un - t.an[n] * un1;
            tt = um1 + t.bn[n] * tt;
            un1 = un;
            un = um1;
            if (fabs(un) > 1e200) { un1 *= 1e-200; un *= 1e-200; tt *= 1e-200; }
        }
        const double sinv = un / tt;
        const double k = sqrt(BV_PI * 0.5 * inv_x) * exp(-x) * sinv;
        k0 = k;
        k1 = k * inv_x * (mu + 0.5 + x + (mu * mu - 0.25) * un1 / un);
    }
    for (int k = 1; k <= t.nl; ++k) {
        const double kn = k0 + 2.0 * (mu + (double)k) * inv_x * k1;
        k0 = k1;
        k1 = kn;
    }
    return exp(t.nu * lnx) * k0;
}

/* ---- x^nu K_nu(x) from a per-nu Chebyshev table --------------------------
 * f(x) = x^nu K_nu(x) and g(x) = f(x) e^x are smooth on every half-octave:
 * on I(e, h) = 2^e [lo_h, hi_h), lo_0 = 1, hi_0 = lo_1 = sqrt 2, hi_1 = 2,
 * for e in [BV_CHEB_EMIN, BV_CHEB_EMAX), f (e < 0, x < 1) or g (e >= 0) is
 * kept as a Chebyshev series of BV_CHEB_N terms in t = (u - mid_h) / hw_h,
 * u = x 2^-e (so t in [-1, 1]), whose coefficients come from the accurate
 * routine (bv_besselk, bv_besselk_scaled) at the Chebyshev points of each
 * half-octave (bv_cheb_node_x; interpolation in those points, coefficients
 * by the discrete cosine sums, bv_cheb_build_kernel), stored as the same
 * polynomial's monomial coefficients in t (bv_cheb_mono).  An entry then
 * costs Horner's rule (13 multiply-adds), and from x = 1 on one exponential:
 * f = g e^-x.
 * Error against 30-digit mpmath (nu in [0.01, 2.7], 11 points in every
 * half-octave): 7.5e-15 relative at worst, the interpolation's own (the
 * Chebyshev form evaluated by Clenshaw's recurrence has the same worst
 * case); below 2^EMIN the accurate routine runs, from 2^EMAX on f = 0 (e^-x
 * is below the smallest double). */
#define BV_CHEB_N 14
#define BV_CHEB_EMIN (-24)
#define BV_CHEB_EMAX 10
#define BV_CHEB_NJ (2 * (BV_CHEB_EMAX - BV_CHEB_EMIN))
#define BV_CHEB_SIZE (BV_CHEB_NJ * BV_CHEB_N)
#define BV_SQRT2 1.4142135623730951
#define BV_CHEB_MID0 1.2071067811865475   /* (1 + sqrt 2) / 2   */
#define BV_CHEB_HW0 0.20710678118654757   /* (sqrt 2 - 1) / 2   */
#define BV_CHEB_MID1 1.7071067811865475   /* (sqrt 2 + 2) / 2   */
#define BV_CHEB_HW1 0.29289321881345254   /* (2 - sqrt 2) / 2   */

/* x^nu K_nu(x) by the accurate routine, for the rare x below the table. */
#if defined(__HIPCC__)
static __host__ __device__ __attribute__((noinline))
#else
static inline
#endif
double bv_xnu_knu_slow(double x, double nu) {
    return exp(nu * log(x)) * bv_besselk(nu, x);
}

/* g at the k-th Chebyshev point of half-octave j (the table's nodes). */
BV_HD double bv_cheb_node_x(int j, int k) {
    const int e = BV_CHEB_EMIN + (j >> 1), h = j & 1;
    const double t = cos(BV_PI * (k + 0.5) / BV_CHEB_N);
    const double u = h ? BV_CHEB_MID1 + BV_CHEB_HW1 * t : BV_CHEB_MID0 + BV_CHEB_HW0 * t;
    return ldexp(u, e);
}

/* The table's input at node k of half-octave j: below x = 1 (exponent < 0)
   f = x^nu K_nu(x) itself, from 1 on g = x^nu e^x K_nu(x) (f falls like
   e^-x there, which no short series on a half-octave follows in relative
   terms; below 1, e^-x is as smooth as f and needs no separate factor). */
BV_HD double bv_cheb_node_g(double nu, int j, int k) {
    const double x = bv_cheb_node_x(j, k);
    if (BV_CHEB_EMIN + (j >> 1) < 0) return exp(nu * log(x)) * bv_besselk(nu, x);
    return exp(nu * log(x)) * bv_besselk_scaled(nu, x);
}

/* Coefficient t of a half-octave's series from its BV_CHEB_N node values
   (discrete cosine sum; the constant term halved). */
BV_HD double bv_cheb_coef(const double *gv, int t) {
    double sum = 0.0;
    for (int k = 0; k < BV_CHEB_N; ++k) sum += gv[k] * cos(BV_PI * t * (k + 0.5) / BV_CHEB_N);
    sum *= 2.0 / BV_CHEB_N;
    return (t == 0) ? 0.5 * sum : sum;
}

/* In place: the Chebyshev coefficients c[0..BV_CHEB_N) of a half-octave's
   series (constant term already halved) become the monomial coefficients of
   the same polynomial in t, by Clenshaw's recurrence run on coefficient
   vectors (b_k = 2 t b_(k+1) - b_(k+2) + c_k, the polynomial t b_1 - b_2 +
   c_0).  Well conditioned here: |c_k| falls by about a factor of 10 per term,
   faster than the 2^(k-1) growth of T_k's coefficients. */
BV_HD void bv_cheb_mono(double *c) {
    double b1[BV_CHEB_N], b2[BV_CHEB_N];
#pragma unroll
    for (int i = 0; i < BV_CHEB_N; ++i) b1[i] = b2[i] = 0.0;
#pragma unroll
    for (int k = BV_CHEB_N - 1; k >= 1; --k) {
        double b0[BV_CHEB_N];
#pragma unroll
        for (int i = 0; i < BV_CHEB_N; ++i)
            b0[i] = ((i > 0) ? 2.0 * b1[i - 1] : 0.0) - b2[i] + ((i == 0) ? c[k] : 0.0);
#pragma unroll
        for (int i = 0; i < BV_CHEB_N; ++i) {
            b2[i] = b1[i];
            b1[i] = b0[i];
        }
    }
#pragma unroll
    for (int i = 0; i < BV_CHEB_N; ++i)
        c[i] = ((i > 0) ? b1[i - 1] : 0.0) - b2[i] + ((i == 0) ? c[0] : 0.0);
}

/* x^nu K_nu(x), x > 0, from the table. */
BV_HD double bv_cheb_g(double x, const double *cheb);

BV_HD double bv_xnu_knu_cheb(double x, const double *cheb, double nu) {
    if (!(x >= 0x1p-24)) return bv_xnu_knu_slow(x, nu);
    if (x >= 0x1p10) return 0.0;
    const double g = bv_cheb_g(x, cheb);
    if (x < 1.0) return g; /* the table holds x^nu K_nu(x) there */
    return g * exp(-x);
}

/* The table's function at x in [2^-24, 2^10) (x^nu K_nu(x) below 1,
   x^nu e^x K_nu(x) from 1 on): the half-octave from the exponent bits,
   Horner's rule on its monomial coefficients. */
BV_HD double bv_cheb_g(double x, const double *cheb) {
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof(b));
    const int e = (int)((b >> 52) & 0x7FF) - 1023;
    const uint64_t ub = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double u;
    __builtin_memcpy(&u, &ub, sizeof(u));
    const int h = u >= BV_SQRT2;
    const double t = h ? (u - BV_CHEB_MID1) * (1.0 / BV_CHEB_HW1)
                       : (u - BV_CHEB_MID0) * (1.0 / BV_CHEB_HW0);
    const double *c = cheb + (2 * (e - BV_CHEB_EMIN) + h) * BV_CHEB_N;
    double p = c[BV_CHEB_N - 1];
#pragma unroll
    for (int k = BV_CHEB_N - 2; k >= 0; --k) p = fma(p, t, c[k]);
    return p;
}

/* Great-circle distance (src/ckernel.cpp:145-154). */
BV_HD double bv_dist_earth(double lat1d, double lon1d, double lat2d, double lon2d) {
    const double lat1r = lat1d * BV_PI / 180.0, lon1r = lon1d * BV_PI / 180.0;
    const double lat2r = lat2d * BV_PI / 180.0, lon2r = lon2d * BV_PI / 180.0;
    const double u = sin((lat2r - lat1r) / 2.0);
    const double v = sin((lon2r - lon1r) / 2.0);
    return 2.0 * BV_EARTH_RADIUS_KM * asin(sqrt(u * u + cos(lat1r) * cos(lat2r) * v * v));
}

/* Euclidean distance, products and sums rounded separately (no contraction),
 * as sqrt(pow(dx,2) + pow(dy,2) [+ pow(dz,2)]). */
BV_HD double bv_dist_euclid(double x1, double y1, double z1, double x2, double y2, double z2,
                            int use_z) {
#pragma clang fp contract(off)
    const double dx = x2 - x1, dy = y2 - y1;
    double s = dx * dx + dy * dy;
    if (use_z) {
        const double dz = z2 - z1;
        s = s + dz * dz;
    }
    return sqrt(s);
}

/* Closed-form Matern of a distance d for nu = 1/2, 3/2, 5/2 (KIND). */
template <int KIND>
BV_HD double bv_matern_closed(double d, const bv_kparams &kp) {
    if (d == 0.0) return kp.sigma2;
    /* the reference divides by beta, then by dist_scale; the closed forms
       multiply by the precomputed reciprocal of their product (within 2 ulp) */
    const double e = d * kp.inv_scale;
    const double x = exp(-e);
    if (KIND == BV_KIND_NU05) return kp.sigma2 * x;
    if (KIND == BV_KIND_NU15) return kp.sigma2 * (1.0 + e) * x;
    return kp.sigma2 * (1.0 + e + e * e * (1.0 / 3.0)) * x;
}

/* e^-e for e >= 0, Tang's table-driven method (P. T. P. Tang, ACM TOMS 15
 * (1989) 144) with 64 table points:
 *   k = rint(e 64 / ln 2),  r = e - k ln2/64 (|r| <= ln2/128; two-term
 *   reduction with fused multiply-adds),  e^-e = 2^(-k/64) e^-r;
 *   -k = 64 M + j (0 <= j < 64):  e^-e = 2^M (T_j + (L_j + T_j P(-r))),
 *   T_j + L_j = 2^(j/64) to twice double precision (T_j rounded to nearest,
 *   L_j the rounded remainder), P(t) = e^t - 1 = t + t^2/2 + ... + t^5/120
 *   (truncation below |r|^6 / 720 = 3.5e-17 relative).
 * Within one ulp of exp(-e) (tests/test_knu_host.py); used for the
 * closed-form Matern entries of the register-resident factorization, where
 * the table sits in LDS (tab = T_0..T_63, L_0..L_63). */
#define BV_EXP_TAB 64
#define BV_LN2_64_HI 0.010830424696249145     /* ln 2 / 64 rounded         */
#define BV_LN2_64_LO 3.623510646634843e-19     /* ln 2 / 64 - BV_LN2_64_HI */
#define BV_64_LN2 92.33248261689366            /* 64 / ln 2                 */
BV_HD double bv_expneg_tab(double e, const double *tab) {
    const double kd = rint(e * BV_64_LN2);
    double r = fma(-kd, BV_LN2_64_HI, e);
    r = fma(-kd, BV_LN2_64_LO, r);
    const double t = -r;
    double q = fma(t, 1.0 / 120.0, 1.0 / 24.0);
    q = fma(t, q, 1.0 / 6.0);
    q = fma(t, q, 0.5);
    const double pm1 = fma(t * t, q, t);
    const int n = -(int)kd;
    const double th = tab[n & 63], tl = tab[BV_EXP_TAB + (n & 63)];
    return ldexp(th + fma(th, pm1, tl), n >> 6);
}

/* x^nu K_nu(x) as bv_xnu_knu_cheb, with e^-x by the table-driven
   exponential above (tab: its 2 x 64 entries) instead of exp. */
BV_HD double bv_xnu_knu_cheb_t(double x, const double *cheb, double nu, const double *tab) {
    if (!(x >= 0x1p-24)) return bv_xnu_knu_slow(x, nu);
    if (x >= 0x1p10) return 0.0;
    const double g = bv_cheb_g(x, cheb);
    if (x < 1.0) return g; /* the table holds x^nu K_nu(x) there */
    return g * bv_expneg_tab(x, tab);
}

/* General-nu entry at the scaled distance x = d * inv_scale (the reciprocal
   of beta * dist_scale, as the closed forms scale; within 2 ulp of the
   reference's two divisions): sigma^2 at 0, else con x^nu K_nu(x), from the
   evaluation's table cheb (kp.cheb or a copy of it). */
BV_HD double bv_matern_general_x(double x, const bv_kparams &kp, const double *cheb,
                                 const double *tab) {
    if (x == 0.0) return kp.sigma2;
    return kp.con * bv_xnu_knu_cheb_t(x, cheb, kp.nu, tab);
}

/* T_j = 2^(j/64) rounded to nearest and L_j = 2^(j/64) - T_j rounded, from
 * 50-digit values (mpmath: T = float(mp.power(2, mp.mpf(j)/64)),
 * L = float(mp.power(2, mp.mpf(j)/64) - T)). */
#define BV_EXP_TAB_HI { BV_EXP_TAB_HI_LIST }
#define BV_EXP_TAB_LO { BV_EXP_TAB_LO_LIST }
#define BV_EXP_TAB_HI_LIST \
    1.0, 1.0108892860517005, 1.0218971486541166, \
    1.0330248790212284, 1.0442737824274138, 1.0556451783605572, \
    1.0671404006768237, 1.0787607977571199, 1.0905077326652577, \
    1.102382583307841, 1.1143867425958924, 1.1265216186082418, \
    1.1387886347566916, 1.1511892299529827, 1.1637248587775775, \
    1.1763969916502812, 1.189207115002721, 1.202156731452703, \
    1.215247359980469, 1.22848053610687, 1.241857812073484, \
    1.255380757024691, 1.2690509571917332, 1.2828700160787783, \
    1.2968395546510096, 1.3109612115247644, 1.3252366431597413, \
    1.339667524053303, 1.3542555469368927, 1.3690024229745905, \
    1.383909881963832, 1.3989796725383112, 1.4142135623730951, \
    1.42961333839197, 1.4451808069770467, 1.460917794180647, \
    1.4768261459394993, 1.4929077282912648, 1.5091644275934228, \
    1.5255981507445384, 1.5422108254079407, 1.559004400237837, \
    1.5759808451078865, 1.593142151342267, 1.6104903319492543, \
    1.6280274218573478, 1.645755478153965, 1.6636765803267364, \
    1.681792830507429, 1.7001063537185235, 1.718619298122478, \
    1.7373338352737062, 1.7562521603732995, 1.7753764925265212, \
    1.7947090750031072, 1.8142521755003989, 1.8340080864093424, \
    1.8539791250833855, 1.8741676341103, 1.8945759815869656, \
    1.9152065613971474, 1.9360617934922943, 1.9571441241754002, \
    1.978456026387951
#define BV_EXP_TAB_LO_LIST \
    0.0, -1.5234778603368577e-17, 5.109225028973444e-17, \
    7.600838874027088e-18, 8.551889705537965e-17, 1.759325738772092e-18, \
    -7.899853966841582e-17, -6.656660436056593e-17, -3.046782079812471e-17, \
    5.2660368715706944e-17, 1.0410278456845571e-16, 5.165856758795457e-17, \
    8.912812676025408e-17, 3.250710218863827e-17, 3.8292048369240935e-17, \
    5.554203254218079e-17, 3.982015231465646e-17, 6.644981499252301e-17, \
    -7.712630692681488e-17, -1.89878163130253e-17, 4.658027591836937e-17, \
    -6.7113898212968784e-18, 2.667932131342186e-18, 1.713594918243561e-17, \
    2.5382502794888315e-17, -7.181536135519454e-17, -2.8587312100388614e-17, \
    8.927282594831732e-17, 7.70094837980299e-17, 9.593797919118849e-17, \
    -6.770511658794786e-17, -9.614213209051323e-17, -9.667293313452913e-17, \
    -1.2031642489053655e-17, -3.0237581349939873e-17, -5.600377186075216e-17, \
    -3.483994556892796e-17, 1.4192920154284036e-17, -1.016455327754295e-16, \
    -1.1024941712342561e-16, 7.949834809697621e-17, 3.7812070533575275e-17, \
    -1.0136916471278304e-17, -1.0094406542311964e-16, 2.4707192569797888e-17, \
    -6.712955084707084e-17, -1.0125679913674773e-16, 5.8909926967131e-17, \
    8.199010020581497e-17, -8.0237193703977e-18, -1.851380418263111e-17, \
    3.164389299292957e-17, 2.960140695448873e-17, 6.429731796556572e-17, \
    1.8227458427912087e-17, -9.969531538920349e-17, 3.283107224245627e-17, \
    9.761887490727594e-17, -6.122763413004143e-17, 3.4034035352165297e-17, \
    -1.0619946056195963e-16, 1.0332385960676326e-16, 8.960767791036668e-17, \
    4.0388753109278167e-17
BV_HD double bv_exp_tab_entry(int i) {
    const double hi[BV_EXP_TAB] = BV_EXP_TAB_HI;
    const double lo[BV_EXP_TAB] = BV_EXP_TAB_LO;
    return i < BV_EXP_TAB ? hi[i] : lo[i - BV_EXP_TAB];
}

/* One covariance entry between locations (x1,y1,z1) and (x2,y2,z2). */
template <int KIND>
BV_HD double bv_cov(double x1, double y1, double z1, double x2, double y2, double z2,
                    const bv_kparams &kp) {
    if (KIND == BV_KIND_SPACETIME) {
        /* core_dcmg_spacetime_matern (src/ckernel.cpp:1421-1458): spatial
           distance uses z as well, no dist_scale */
        const double d = bv_dist_euclid(x1, y1, z1, x2, y2, z2, 1);
        const double e = d / kp.beta;
        const double dt = fabs(z1 - z2);
        const double base = pow(dt, 2.0 * kp.nu_t) / kp.beta_t + 1.0;
        const double e2 = pow(base, kp.sep / 2.0);
        const double e3 = e / e2;
        const double e4 = pow(base, kp.sep + kp.aux);
        if (e == 0.0) return kp.sigma2 / e4;
        return kp.con * pow(e3, kp.nu) * bv_besselk(kp.nu, e3) / e4;
    }
    double d;
    if (kp.metric == 1) d = bv_dist_earth(x1, y1, x2, y2);
    else d = bv_dist_euclid(x1, y1, z1, x2, y2, z2, kp.use_z);
    if (KIND == BV_KIND_NU05 || KIND == BV_KIND_NU15 || KIND == BV_KIND_NU25)
        return bv_matern_closed<KIND>(d, kp);
    double e = d / kp.beta;
    e = e / kp.dist_scale;
    if (e == 0.0) return kp.sigma2;
    return kp.con * bv_xnu_knu_cheb(e, kp.cheb, kp.nu);
}

/* bv_cov for the kinds that need K_nu (GENERAL, SPACETIME), with the
   nu-only work taken from a table filled by bv_knu_tab_fill(kp.nu, ...). */
template <int KIND>
BV_HD double bv_cov_tab(double x1, double y1, double z1, double x2, double y2, double z2,
                        const bv_kparams &kp, const bv_knu_tab &t) {
    if (KIND == BV_KIND_SPACETIME) {
        const double d = bv_dist_euclid(x1, y1, z1, x2, y2, z2, 1);
        const double e = d / kp.beta;
        const double dt = fabs(z1 - z2);
        const double base = pow(dt, 2.0 * kp.nu_t) / kp.beta_t + 1.0;
        const double e2 = pow(base, kp.sep / 2.0);
        const double e3 = e / e2;
        const double e4 = pow(base, kp.sep + kp.aux);
        if (e == 0.0) return kp.sigma2 / e4;
        return kp.con * bv_xnu_knu_tab(e3, t) / e4;
    }
    double d;
    if (kp.metric == 1) d = bv_dist_earth(x1, y1, x2, y2);
    else d = bv_dist_euclid(x1, y1, z1, x2, y2, z2, kp.use_z);
    double e = d / kp.beta;
    e = e / kp.dist_scale;
    if (e == 0.0) return kp.sigma2;
    return kp.con * bv_xnu_knu_tab(e, t);
}

#endif /* BV_MATERN_H */

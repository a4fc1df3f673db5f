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
 * g(x) = x^nu K_nu(x) e^x is smooth on every half-octave: on
 * I(e, h) = 2^e [lo_h, hi_h), lo_0 = 1, hi_0 = lo_1 = sqrt 2, hi_1 = 2, for
 * e in [BV_CHEB_EMIN, BV_CHEB_EMAX), it is kept as a Chebyshev series of
 * BV_CHEB_N terms in t = (u - mid_h) / hw_h, u = x 2^-e (so t in [-1, 1]),
 * whose coefficients come from the accurate routine (bv_besselk_scaled) at
 * the Chebyshev points of each half-octave (bv_cheb_node_x; interpolation in
 * those points, coefficients by the discrete cosine sums, bv_cheb_build_kernel).  An entry then costs the
 * Clenshaw recurrence and one exp: f = g e^-x.  Interpolation error (against
 * 30-digit mpmath, nu in [0.01, 2.7], every half-octave): below 5e-15
 * relative; below 2^EMIN the accurate routine runs, from 2^EMAX on
 * f = 0 (e^-x is below the smallest double). */
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

/* The table's input: g = x^nu e^x K_nu(x) at node k of half-octave j. */
BV_HD double bv_cheb_node_g(double nu, int j, int k) {
    const double x = bv_cheb_node_x(j, k);
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

/* x^nu K_nu(x), x > 0, from the table. */
BV_HD double bv_xnu_knu_cheb(double x, const double *cheb, double nu) {
    if (!(x >= 0x1p-24)) return bv_xnu_knu_slow(x, nu);
    if (x >= 0x1p10) return 0.0;
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
    const double t2 = 2.0 * t;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = BV_CHEB_N - 1; k >= 1; --k) {
        const double b0 = t2 * b1 - b2 + c[k];
        b2 = b1;
        b1 = b0;
    }
    const double g = t * b1 - b2 + c[0];
    return g * exp(-x);
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

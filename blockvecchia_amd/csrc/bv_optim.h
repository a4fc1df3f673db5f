/*
 * bv_optim.h -- derivative-free maximisation under box bounds for the
 * driver's outer loop (the role NLopt's NLOPT_LN_BOBYQA plays in the
 * reference, src/test_Xvecchia_batch.cpp:700-725).
 *
 * The method follows the structure of BOBYQA (M. J. D. Powell, "The BOBYQA
 * algorithm for bound constrained optimization without derivatives", DAMTP
 * 2009/NA06): quadratic models of the objective built by interpolation,
 * trust-region steps kept inside the box, and a lower radius rho that
 * decreases from rho_beg to rho_end.  Differences, chosen for a short and
 * robust implementation: the model is the fully determined quadratic on
 * (n+1)(n+2)/2 points (n is 3, 4 or 7 here) instead of Powell's minimum-
 * Frobenius-norm updates on 2n+1 points; the trust-region subproblem is
 * solved by projected gradient iterations on the model, compared with the
 * projected Newton step; when the set is poorly poised or spread far
 * beyond the trust region, it is rebuilt as the structured initial design
 * around the best point.  The iterates therefore differ
 * from NLopt's; the maximiser agrees to the stopping tolerance.
 *
 * Stopping: rho <= xtol_rel * max(|x_best|_inf, 1e-3), or maxeval
 * evaluations.  Coordinates with lb == ub are fixed.  Non-finite objective
 * values (a failed factorization) are treated as rejected steps.
 */
#ifndef BV_OPTIM_H
#define BV_OPTIM_H

#include <math.h>
#include <algorithm>
#include <functional>
#include <numeric>
#include <vector>

namespace bvopt {

enum { BVOPT_XTOL = 1, BVOPT_MAXEVAL = 2 };

struct Result {
    std::vector<double> x;
    double f = -INFINITY; /* maximum found */
    int evals = 0;
    int status = 0;
};

/* Solve A x = b (dense, row-major n x n) by Gaussian elimination with partial
   pivoting; returns false when a pivot is below tol * max |A|. */
inline bool solve_dense(std::vector<double> A, std::vector<double> &b, int n, double tol) {
    double amax = 0.0;
    for (double v : A) amax = std::max(amax, fabs(v));
    if (!(amax > 0.0)) return false;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
        if (!(fabs(A[p * n + k]) > tol * amax)) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / A[k * n + k];
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
        b[k] = s / A[k * n + k];
    }
    return true;
}

/* Maximise f over lb <= x <= ub from x0. */
inline Result maximize(const std::function<double(const std::vector<double> &)> &fobj,
                       std::vector<double> x0, const std::vector<double> &lb,
                       const std::vector<double> &ub, double xtol_rel, int maxeval,
                       double rho_beg = -1.0) {
    const int N = (int)x0.size();
    for (int i = 0; i < N; ++i) x0[i] = std::min(std::max(x0[i], lb[i]), ub[i]);
    std::vector<int> fr;
    for (int i = 0; i < N; ++i)
        if (ub[i] > lb[i]) fr.push_back(i);
    const int n = (int)fr.size();
    Result res;
    res.x = x0;
    auto full = [&](const std::vector<double> &z) {
        std::vector<double> x = x0;
        for (int k = 0; k < n; ++k) x[fr[k]] = z[k];
        return x;
    };
    /* minimise F = -f; non-finite -> +inf */
    auto F = [&](const std::vector<double> &z) {
        ++res.evals;
        const double v = fobj(full(z));
        return std::isfinite(v) ? -v : INFINITY;
    };
    if (n == 0 || maxeval <= 1) {
        /* one evaluation at the (clipped) start: the reference's --perf */
        res.f = fobj(x0);
        res.evals = 1;
        res.status = BVOPT_MAXEVAL;
        return res;
    }
    std::vector<double> zl(n), zu(n), z0(n);
    double span = INFINITY, xinf = 0.0;
    for (int k = 0; k < n; ++k) {
        zl[k] = lb[fr[k]];
        zu[k] = ub[fr[k]];
        z0[k] = x0[fr[k]];
        span = std::min(span, zu[k] - zl[k]);
        xinf = std::max(xinf, fabs(z0[k]));
    }
    double rho = rho_beg > 0 ? rho_beg : std::max(0.1, 0.1 * xinf);
    rho = std::min(rho, span / 3.0);
    double delta = rho;
    const int p = (n + 1) * (n + 2) / 2;

    /* interpolation set */
    std::vector<std::vector<double>> Y;
    std::vector<double> FY;
    auto inbox = [&](int k, double v) { return v >= zl[k] && v <= zu[k]; };
    /* the structured design around c with step h (h <= span / 3): c,
       c + d_k e_k, c - d_k e_k (or c + 2 d_k e_k at a bound), and
       c + d_a e_a + d_b e_b; poised for the full quadratic by construction */
    /* centre and step of the last rebuild: rebuilding the same design again
       would reproduce the same model and the same rejected step */
    std::vector<double> last_c;
    double last_h = -1.0;
    auto rebuild = [&](const std::vector<double> c, double fc, double h) {
        /* c by value: callers pass points of Y, which the next line replaces */
        last_c = c;
        last_h = h;
        Y.assign(1, c);
        FY.assign(1, fc);
        std::vector<double> d(n);
        for (int k = 0; k < n && res.evals < maxeval; ++k) {
            d[k] = inbox(k, c[k] + h) ? h : -h;
            std::vector<double> z = c;
            z[k] += d[k];
            Y.push_back(z);
            FY.push_back(F(z));
        }
        for (int k = 0; k < n && res.evals < maxeval; ++k) {
            std::vector<double> z = c;
            z[k] = inbox(k, c[k] - d[k]) ? c[k] - d[k] : c[k] + 2.0 * d[k];
            Y.push_back(z);
            FY.push_back(F(z));
        }
        for (int a = 0; a < n && res.evals < maxeval; ++a)
            for (int b = a + 1; b < n && res.evals < maxeval; ++b) {
                std::vector<double> z = c;
                z[a] += d[a];
                z[b] += d[b];
                Y.push_back(z);
                FY.push_back(F(z));
            }
    };
    rebuild(z0, F(z0), rho);
    auto best_index = [&]() {
        int bi = 0;
        for (int i = 1; i < (int)FY.size(); ++i)
            if (FY[i] < FY[bi]) bi = i;
        return bi;
    };
    auto dist = [&](const std::vector<double> &a, const std::vector<double> &b) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += (a[k] - b[k]) * (a[k] - b[k]);
        return sqrt(s);
    };
    /* the set is spread beyond 2 r around c (the design itself reaches 2 h
       next to a bound), and rebuilding around c would not repeat the last
       rebuild */
    auto spread_out = [&](const std::vector<double> &c, double fard, double r) {
        return fard > 2.0 * r * (1.0 + 1e-9) && !(last_h == rho && last_c == c);
    };
    auto project = [&](const std::vector<double> &c, std::vector<double> &s) {
        for (int k = 0; k < n; ++k) s[k] = std::min(std::max(s[k], zl[k] - c[k]), zu[k] - c[k]);
        double nrm = 0.0;
        for (int k = 0; k < n; ++k) nrm += s[k] * s[k];
        nrm = sqrt(nrm);
        if (nrm > delta)
            for (int k = 0; k < n; ++k) s[k] *= delta / nrm; /* 0 is feasible: stays in box */
    };
    while (res.evals < maxeval) {
        const int bi = best_index();
        const std::vector<double> c = Y[bi];
        const double fc = FY[bi];
        double cinf = 0.0;
        for (int k = 0; k < n; ++k) cinf = std::max(cinf, fabs(c[k]));
        const double rho_end = xtol_rel * std::max(cinf, 1e-3);
        if (rho <= rho_end) { res.status = BVOPT_XTOL; break; }
        /* the model from the points with finite values (scaled by rho) */
        int far = -1;
        double fard = -1.0;
        for (int i = 0; i < (int)Y.size(); ++i) {
            const double dd = std::isfinite(FY[i]) ? dist(Y[i], c) : INFINITY;
            if (i != bi && dd > fard) { fard = dd; far = i; }
        }
        std::vector<double> M((size_t)p * p), rhs(p);
        bool ok = (int)Y.size() == p;
        for (int i = 0; i < p && ok; ++i) {
            if (!std::isfinite(FY[i])) { ok = false; break; }
            int col = 0;
            std::vector<double> u(n);
            for (int k = 0; k < n; ++k) u[k] = (Y[i][k] - c[k]) / rho;
            M[(size_t)i * p + col++] = 1.0;
            for (int k = 0; k < n; ++k) M[(size_t)i * p + col++] = u[k];
            for (int k = 0; k < n; ++k) M[(size_t)i * p + col++] = 0.5 * u[k] * u[k];
            for (int a = 0; a < n; ++a)
                for (int b = a + 1; b < n; ++b) M[(size_t)i * p + col++] = u[a] * u[b];
            rhs[i] = FY[i] - fc;
        }
        if (ok) ok = solve_dense(M, rhs, p, 1e-13);
        if (!ok) {
            /* poorly poised, incomplete, or a failed evaluation in the set;
               when the design just built here gives no model (failed
               evaluations in it), a smaller one */
            if (last_h == rho && last_c == c) {
                rho = std::max(0.1 * rho, 0.5 * rho_end);
                delta = std::max(0.5 * delta, rho);
                if (rho <= rho_end) { res.status = BVOPT_XTOL; break; }
            }
            rebuild(c, fc, rho);
            continue;
        }
        std::vector<double> g(n), H((size_t)n * n);
        {
            int col = 1;
            for (int k = 0; k < n; ++k) g[k] = rhs[col++] / rho;
            for (int k = 0; k < n; ++k) H[(size_t)k * n + k] = rhs[col++] / (rho * rho);
            for (int a = 0; a < n; ++a)
                for (int b = a + 1; b < n; ++b) {
                    H[(size_t)a * n + b] = H[(size_t)b * n + a] = rhs[col++] / (rho * rho);
                }
        }
        auto q = [&](const std::vector<double> &s) {
            double v = 0.0;
            for (int a = 0; a < n; ++a) {
                double hs = 0.0;
                for (int b = 0; b < n; ++b) hs += H[(size_t)a * n + b] * s[b];
                v += g[a] * s[a] + 0.5 * s[a] * hs;
            }
            return v;
        };
        /* trust-region step: projected gradient on the model, and the
           projected Newton step */
        std::vector<double> s(n, 0.0), best_s(n, 0.0);
        double hn = 0.0;
        for (double v : H) hn += v * v;
        hn = sqrt(hn);
        double gn = 0.0;
        for (double v : g) gn += v * v;
        gn = sqrt(gn);
        double step = hn > 0 ? 1.0 / hn : (gn > 0 ? delta / gn : 0.0);
        double qs = 0.0;
        for (int it = 0; it < 400 && step > 0; ++it) {
            std::vector<double> t(n);
            for (int a = 0; a < n; ++a) {
                double gr = g[a];
                for (int b = 0; b < n; ++b) gr += H[(size_t)a * n + b] * s[b];
                t[a] = s[a] - step * gr;
            }
            project(c, t);
            const double qt = q(t);
            if (qt < qs - 1e-15 * (fabs(qs) + 1e-300)) { s = t; qs = qt; }
            else { step *= 0.5; if (step * gn < 1e-12 * delta) break; }
        }
        best_s = s;
        {
            std::vector<double> sn(g.size());
            for (int k = 0; k < n; ++k) sn[k] = -g[k];
            if (solve_dense(H, sn, n, 1e-14)) {
                project(c, sn);
                if (q(sn) < qs) { best_s = sn; qs = q(sn); }
            }
        }
        const double sn2 = sqrt(std::inner_product(best_s.begin(), best_s.end(), best_s.begin(), 0.0));
        if (sn2 < 0.5 * rho || !(qs < 0.0)) {
            /* no useful step at this radius: repair geometry, else shrink rho */
            if (far >= 0 && spread_out(c, fard, rho)) {
                rebuild(c, fc, rho);
            } else {
                rho = std::max(0.1 * rho, 0.5 * rho_end);
                delta = std::max(0.5 * delta, rho);
            }
            continue;
        }
        std::vector<double> z(n);
        for (int k = 0; k < n; ++k) z[k] = std::min(std::max(c[k] + best_s[k], zl[k]), zu[k]);
        const double fz = F(z);
        const double ratio = std::isfinite(fz) ? (fc - fz) / (-qs) : -INFINITY;
        if (ratio <= 0.1) delta = std::max(0.5 * sn2, rho);
        else if (ratio <= 0.7) delta = std::max(0.5 * delta, sn2);
        else delta = std::max(0.5 * delta, 2.0 * sn2);
        if (std::isfinite(fz)) {
            /* the new point replaces the point farthest from the new best */
            const std::vector<double> &nb = (fz < fc) ? z : c;
            int j = -1;
            double jd = -1.0;
            for (int i = 0; i < (int)Y.size(); ++i) {
                if (i == bi && fz >= fc) continue;
                const double dd = std::isfinite(FY[i]) ? dist(Y[i], nb) : INFINITY;
                if (dd > jd) { jd = dd; j = i; }
            }
            Y[j] = z;
            FY[j] = fz;
        }
        if (ratio < 0.1 && res.evals < maxeval) {
            const int b2 = best_index();
            if (far >= 0 && spread_out(Y[b2], fard, delta)) {
                rebuild(Y[b2], FY[b2], rho);
            }
            else if (delta <= rho) {
                rho = std::max(0.1 * rho, 0.5 * rho_end);
                delta = std::max(0.5 * delta, rho);
            }
        }
    }
    if (res.status == 0) res.status = BVOPT_MAXEVAL;
    const int bi = best_index();
    res.x = full(Y[bi]);
    res.f = -FY[bi];
    return res;
}

} /* namespace bvopt */

#endif /* BV_OPTIM_H */

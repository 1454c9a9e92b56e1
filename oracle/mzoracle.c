/*
 * mzoracle.c -- CPU restatement of mizuba's conjunction-detection path.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle / CPU baseline); the
 * product path lives in mizuba_amd/ and never links this file.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the mizuba source tree, version 0.0.12). Choices the reference leaves to a
 * third-party library are fixed here and documented at the point of use:
 *   - heyoka expression order: sums are evaluated left to right in ascending
 *     index order; products as written ((c * binom) * a^k).
 *   - Boost.Math toms748_solve: restated from Alefeld, Potra & Shi,
 *     "Algorithm 748: Enclosing Zeros of Continuous Functions" (ACM TOMS 21,
 *     1995) with the eps_tolerance(53 bits) stopping rule and 100-evaluation
 *     limit used by poly_utils.cpp:226-247.
 *   - heyoka fex_check: interval-arithmetic Horner enclosure on [0, h].
 *   - heyoka poly_rtscc / poly_translate_1: reverse, Taylor shift by 1 via
 *     binomial sums, count sign changes ignoring zeros (Descartes).
 *   - tbb::parallel_sort is unstable; ties are broken by object index
 *     (Morton sort) and by the full lexicographic record (pair / conj sort).
 * Build with -ffp-contract=off so that no FMA contraction changes rounding.
 */
#include "mzoracle.h"

#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MZO_MAX_OP1 64

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void set_err(char *buf, size_t len, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, len, fmt, ap);
    va_end(ap);
}

/* Exact binomial coefficient as a double (boost::math::binomial_coefficient
 * is exact for the small arguments used here). */
static double binom_d(unsigned n, unsigned k)
{
    if (k > n) return 0.0;
    if (k > n - k) k = n - k;
    unsigned long long r = 1;
    for (unsigned i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return (double)r;
}

/* Powers a^0..a^order with the halving scheme of conjunctions_jit.cpp:83-95
 * (a^i = a^(i/2) * a^(i/2) for even i, a^(i/2+1) * a^(i/2) for odd i). */
static void halving_pows(double a, uint32_t order, double *p)
{
    p[0] = 1.0;
    if (order >= 1) p[1] = a;
    for (uint32_t i = 2; i <= order; ++i) {
        if (i % 2 == 0) {
            p[i] = p[i / 2] * p[i / 2];
        } else {
            p[i] = p[i / 2 + 1] * p[i / 2];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* JIT kernels of conjunctions_jit.cpp, restated                            */
/* ------------------------------------------------------------------------ */

/* c'_i = sum_{k=i}^{n} (c_k * C(k, k-i)) * a^(k-i)   (conjunctions_jit.cpp:68-115). */
void mzo_poly_translate(const double *cfs, uint32_t order, uint32_t stride, double a, double *out, uint32_t out_stride)
{
    double ap[MZO_MAX_OP1];
    double tmp[MZO_MAX_OP1];
    halving_pows(a, order, ap);
    for (uint32_t i = 0; i <= order; ++i) {
        double s = 0.0;
        for (uint32_t k = i; k <= order; ++k) {
            const double term = (cfs[(size_t)k * stride] * binom_d(k, k - i)) * ap[k - i];
            s = (k == i) ? term : s + term;
        }
        tmp[i] = s;
    }
    for (uint32_t i = 0; i <= order; ++i) out[(size_t)i * out_stride] = tmp[i];
}

/* NaN-propagating min/max (conjunctions_jit.cpp:142-146). */
static double hy_min(double a, double b) { return (b == b) ? ((b < a) ? b : a) : b; }
static double hy_max(double a, double b) { return (b == b) ? ((a < b) ? b : a) : b; }

/* Pairwise reduction (conjunctions_jit.cpp:148-172). */
static double pairwise_reduce(double *v, uint32_t n, double (*f)(double, double))
{
    while (n != 1) {
        uint32_t m = 0;
        for (uint32_t i = 0; i < n; i += 2) {
            if (i + 1 == n) {
                v[m++] = v[i];
            } else {
                v[m++] = f(v[i], v[i + 1]);
            }
        }
        n = m;
    }
    return v[0];
}

/* Cargo-Shisha enclosure over [lb, ub] (conjunctions_jit.cpp:179-267).
 * skip_translate reproduces the translation by a literal zero that heyoka
 * folds away in cs_enc (conjunctions_jit.cpp:317-320). */
static void cs_enclosure_impl(const double *cfs, uint32_t order, uint32_t stride, double lb, double ub,
                              int skip_translate, double *out_min, double *out_max)
{
    double t[MZO_MAX_OP1], dp[MZO_MAX_OP1], b[MZO_MAX_OP1], bmin[MZO_MAX_OP1], bmax[MZO_MAX_OP1];
    const double delta = ub - lb;
    if (skip_translate) {
        for (uint32_t i = 0; i <= order; ++i) t[i] = cfs[(size_t)i * stride];
    } else {
        mzo_poly_translate(cfs, order, stride, lb, t, 1);
    }
    halving_pows(delta, order, dp);
    for (uint32_t i = 0; i <= order; ++i) t[i] = t[i] * dp[i];
    for (uint32_t i = 0; i <= order; ++i) t[i] = t[i] / binom_d(order, i);
    /* difference table: row_r[k] = row_{r-1}[k] + row_{r-1}[k+1]; b_r = row_r[0] */
    b[0] = t[0];
    for (uint32_t r = 1; r <= order; ++r) {
        for (uint32_t k = 0; k + r <= order; ++k) t[k] = t[k] + t[k + 1];
        b[r] = t[0];
    }
    for (uint32_t i = 0; i <= order; ++i) bmin[i] = bmax[i] = b[i];
    *out_min = pairwise_reduce(bmin, order + 1, hy_min);
    *out_max = pairwise_reduce(bmax, order + 1, hy_max);
}

void mzo_cs_enclosure(const double *cfs, uint32_t order, uint32_t stride, double lb, double ub, double *out_min,
                      double *out_max)
{
    cs_enclosure_impl(cfs, order, stride, lb, ub, 0, out_min, out_max);
}

/* Chebyshev nodes on [0, h], ascending (conjunctions_jit.cpp:345-361). */
void mzo_cheby_nodes(uint32_t order, double h, double *nodes)
{
    const uint32_t op1 = order + 1;
    const double pi = 3.141592653589793238462643383279502884;
    for (uint32_t i = 0; i < op1; ++i) {
        const double k = (cos((2. * i + 1.) / (2. * op1) * pi) + 1.) / 2.;
        nodes[op1 - 1 - i] = k * h;
    }
}

/* Horner evaluation with stride (poly_utils.hpp:41-51). */
static double horner_eval(const double *p, uint32_t order, double x, uint32_t stride)
{
    double r = p[(size_t)order * stride];
    for (uint32_t o = 1; o <= order; ++o) r = p[(size_t)(order - o) * stride] + r * x;
    return r;
}

/* Bjorck-Pereyra divided differences and Newton->monomial conversion
 * (poly_utils.cpp:467-535). */
void mzo_vm_interp(uint32_t order, const double *alpha, const double *f, double *out)
{
    double c[MZO_MAX_OP1];
    for (uint32_t j = 0; j <= order; ++j) c[j] = f[j];
    for (uint32_t k = 0; k < order; ++k) {
        for (uint32_t j = order; j > k; --j) c[j] = (c[j] - c[j - 1]) / (alpha[j] - alpha[j - k - 1]);
    }
    for (uint32_t idx = 0; idx < order; ++idx) {
        const uint32_t k = order - idx - 1;
        for (uint32_t j = k; j < order; ++j) c[j] = c[j] - alpha[k] * c[j + 1];
    }
    for (uint32_t j = 0; j <= order; ++j) out[j] = c[j];
}

/* ------------------------------------------------------------------------ */
/* root finding (poly_utils.cpp:150-460)                                    */
/* ------------------------------------------------------------------------ */

static double poly_eval_1(const double *a, double x, uint32_t n)
{
    double r = a[n] * (double)n;
    for (uint32_t i = 1; i < n; ++i) r = a[n - i] * (double)(n - i) + r * x;
    return r;
}

static int sgn_d(double v) { return (0.0 < v) - (v < 0.0); }

/* Taylor shift by 1: out_i = sum_{k=i}^{n} c_k * C(k, i) (heyoka poly_translate_1). */
static void translate_1(const double *c, uint32_t n, double *out)
{
    double tmp[MZO_MAX_OP1];
    for (uint32_t i = 0; i <= n; ++i) {
        double s = c[i];
        for (uint32_t k = i + 1; k <= n; ++k) s = s + c[k] * binom_d(k, i);
        tmp[i] = s;
    }
    for (uint32_t i = 0; i <= n; ++i) out[i] = tmp[i];
}

/* Descartes sign-change count ignoring zero coefficients. */
static uint32_t count_sign_changes(const double *c, uint32_t n)
{
    uint32_t cnt = 0;
    int last = sgn_d(c[0]);
    for (uint32_t i = 1; i <= n; ++i) {
        const int cur = sgn_d(c[i]);
        cnt += (cur * last == -1);
        if (cur != 0) last = cur;
    }
    return cnt;
}

/* heyoka poly_rtscc: reverse into r, shift r by 1 into s, count sign changes of s. */
static uint32_t rtscc(const double *q, uint32_t n, double *r, double *s)
{
    for (uint32_t i = 0; i <= n; ++i) r[i] = q[n - i];
    translate_1(r, n, s);
    return count_sign_changes(s, n);
}

/* heyoka fex_check: interval Horner enclosure of the polynomial on [0, h];
 * returns 1 when the enclosure excludes zero. */
static int fex_check(const double *p, uint32_t n, double h)
{
    double lo = p[n], hi = p[n];
    for (uint32_t o = 1; o <= n; ++o) {
        const double a = lo * h, b = hi * h;
        double nlo = a < b ? a : b, nhi = a < b ? b : a;
        nlo = nlo < 0.0 ? nlo : 0.0;
        nhi = nhi > 0.0 ? nhi : 0.0;
        lo = nlo + p[n - o];
        hi = nhi + p[n - o];
    }
    return (lo > 0.0) || (hi < 0.0);
}

/* ---- TOMS 748 restatement ------------------------------------------------ */

typedef struct {
    const double *poly;
    uint32_t n;
    uint32_t nevals;
} t748_fn;

static double t748_eval(t748_fn *f, double x)
{
    ++f->nevals;
    return horner_eval(f->poly, f->n, x, 1);
}

static int bsign(double x) { return (x > 0) - (x < 0); }

static int t748_tol(double a, double b)
{
    const double eps = 4.0 * 2.220446049250313080847263336181640625e-16;
    const double m = fabs(a) < fabs(b) ? fabs(a) : fabs(b);
    return fabs(a - b) <= eps * m;
}

static double t748_secant(double a, double b, double fa, double fb)
{
    const double tol = 2.220446049250313080847263336181640625e-16 * 5;
    const double c = a - (fa / (fb - fa)) * (b - a);
    if ((c <= a + fabs(a) * tol) || (c >= b - fabs(b) * tol)) return (a + b) / 2;
    return c;
}

static double t748_safe_div(double num, double den, double r)
{
    if (fabs(den) < 1) {
        if (fabs(den * 1.7976931348623157e308) <= fabs(num)) return r;
    }
    return num / den;
}

static double t748_quadratic(double a, double b, double d, double fa, double fb, double fd, unsigned count)
{
    double B = t748_safe_div(fb - fa, b - a, 1.7976931348623157e308);
    double A = t748_safe_div(fd - fb, d - b, 1.7976931348623157e308);
    A = t748_safe_div(A - B, d - a, 0.0);
    if (A == 0) return t748_secant(a, b, fa, fb);
    double c = (bsign(A) * bsign(fa) > 0) ? a : b;
    for (unsigned i = 1; i <= count; ++i) {
        c -= t748_safe_div(fa + (B + A * (c - b)) * (c - a), B + A * (2 * c - a - b), 1 + c - a);
    }
    if ((c <= a) || (c >= b)) c = t748_secant(a, b, fa, fb);
    return c;
}

static double t748_cubic(double a, double b, double d, double e, double fa, double fb, double fd, double fe)
{
    const double q11 = (d - e) * fd / (fe - fd);
    const double q21 = (b - d) * fb / (fd - fb);
    const double q31 = (a - b) * fa / (fb - fa);
    const double d21 = (b - d) * fd / (fd - fb);
    const double d31 = (a - b) * fb / (fb - fa);
    const double q22 = (d21 - q11) * fb / (fe - fb);
    const double q32 = (d31 - q21) * fa / (fd - fa);
    const double d32 = (d31 - q21) * fd / (fd - fa);
    const double q33 = (d32 - q22) * fa / (fe - fa);
    double c = q31 + q32 + q33 + a;
    if ((c <= a) || (c >= b)) c = t748_quadratic(a, b, d, fa, fb, fd, 3);
    return c;
}

static void t748_bracket(t748_fn *f, double *a, double *b, double c, double *fa, double *fb, double *d, double *fd)
{
    const double tol = 2.220446049250313080847263336181640625e-16 * 2;
    if ((*b - *a) < 2 * tol * *a) {
        c = *a + (*b - *a) / 2;
    } else if (c <= *a + fabs(*a) * tol) {
        c = *a + fabs(*a) * tol;
    } else if (c >= *b - fabs(*b) * tol) {
        c = *b - fabs(*b) * tol;
    }
    const double fc = t748_eval(f, c);
    if (fc == 0) {
        *a = c;
        *fa = 0;
        *d = 0;
        *fd = 0;
        return;
    }
    if (bsign(*fa) * bsign(fc) < 0) {
        *d = *b;
        *fd = *fb;
        *b = c;
        *fb = fc;
    } else {
        *d = *a;
        *fd = *fa;
        *a = c;
        *fa = fc;
    }
}

static int t748_prof(double fa, double fb, double fd, double fe)
{
    const double md = 2.2250738585072014e-308 * 32;
    return (fabs(fa - fb) < md) || (fabs(fa - fd) < md) || (fabs(fa - fe) < md) || (fabs(fb - fd) < md)
           || (fabs(fb - fe) < md) || (fabs(fd - fe) < md);
}

/* Returns 0 on success, EDOM if [a, b] does not bracket a root or is out of
 * order; *iters receives the number of function evaluations. */
static int toms748(t748_fn *f, double ax, double bx, unsigned max_iter, double *ra, double *rb, unsigned *iters)
{
    double a = ax, b = bx;
    if (max_iter <= 2) {
        *ra = a;
        *rb = b;
        *iters = max_iter;
        return 0;
    }
    double fa = t748_eval(f, a), fb = t748_eval(f, b);
    unsigned count = max_iter - 2;
    const unsigned count0 = count;
    double c, u, fu, a0, b0, d, fd, e, fe;
    const double mu = 0.5;
    if (a >= b) {
        *ra = *rb = NAN;
        *iters = 2;
        return EDOM;
    }
    if (t748_tol(a, b) || (fa == 0) || (fb == 0)) {
        if (fa == 0) {
            b = a;
        } else if (fb == 0) {
            a = b;
        }
        *ra = a;
        *rb = b;
        *iters = 2;
        return 0;
    }
    if (bsign(fa) * bsign(fb) > 0) {
        *ra = *rb = NAN;
        *iters = 2;
        return EDOM;
    }
    fe = e = fd = 1e5;
    d = 0;
    if (fa != 0) {
        c = t748_secant(a, b, fa, fb);
        t748_bracket(f, &a, &b, c, &fa, &fb, &d, &fd);
        --count;
        if (count && (fa != 0) && !t748_tol(a, b)) {
            c = t748_quadratic(a, b, d, fa, fb, fd, 2);
            e = d;
            fe = fd;
            t748_bracket(f, &a, &b, c, &fa, &fb, &d, &fd);
            --count;
        }
    }
    while (count && (fa != 0) && !t748_tol(a, b)) {
        a0 = a;
        b0 = b;
        if (t748_prof(fa, fb, fd, fe)) {
            c = t748_quadratic(a, b, d, fa, fb, fd, 2);
        } else {
            c = t748_cubic(a, b, d, e, fa, fb, fd, fe);
        }
        e = d;
        fe = fd;
        t748_bracket(f, &a, &b, c, &fa, &fb, &d, &fd);
        if ((0 == --count) || (fa == 0) || t748_tol(a, b)) break;
        if (t748_prof(fa, fb, fd, fe)) {
            c = t748_quadratic(a, b, d, fa, fb, fd, 3);
        } else {
            c = t748_cubic(a, b, d, e, fa, fb, fd, fe);
        }
        t748_bracket(f, &a, &b, c, &fa, &fb, &d, &fd);
        if ((0 == --count) || (fa == 0) || t748_tol(a, b)) break;
        if (fabs(fa) < fabs(fb)) {
            u = a;
            fu = fa;
        } else {
            u = b;
            fu = fb;
        }
        c = u - 2 * (fu / (fb - fa)) * (b - a);
        if (fabs(c - u) > (b - a) / 2) c = a + (b - a) / 2;
        e = d;
        fe = fd;
        t748_bracket(f, &a, &b, c, &fa, &fb, &d, &fd);
        if ((0 == --count) || (fa == 0) || t748_tol(a, b)) break;
        if ((b - a) < mu * (b0 - a0)) continue;
        e = d;
        fe = fd;
        t748_bracket(f, &a, &b, a + (b - a) / 2, &fa, &fb, &d, &fd);
        --count;
    }
    if (fa == 0) {
        b = a;
    } else if (fb == 0) {
        a = b;
    }
    *ra = a;
    *rb = b;
    *iters = (count0 - count) + 2;
    return 0;
}

/* bracketed_root_find (poly_utils.cpp:213-268): search [lb, ub) by moving ub
 * one ulp towards lb, at most 100 evaluations, result = bracket midpoint. */
int mzo_bracketed_root(const double *poly, uint32_t order, double lb, double ub, double *root, uint32_t *n_iter)
{
    if (isfinite(lb) && isfinite(ub) && ub > lb) ub = nextafter(ub, lb);
    t748_fn f = {poly, order, 0};
    double ra, rb;
    unsigned it;
    const int e = toms748(&f, lb, ub, 100, &ra, &rb, &it);
    if (n_iter) *n_iter = it;
    *root = (ra / 2) + (rb / 2);
    if (e != 0) {
        *root = 0;
        return e;
    }
    if (it < 100) return 0;
    return -1;
}

typedef struct {
    double lb, ub;
    double p[MZO_MAX_OP1];
} wl_entry;

/* run_poly_root_finding (poly_utils.cpp:272-460). The work list is LIFO, the
 * isolating intervals are refined in insertion order. */
static int root_finding_impl(const double *poly, uint32_t order, double rf_int, int direction, double *roots,
                             uint32_t *n_roots, int *fex, uint32_t i, uint32_t j, char *err, size_t errlen)
{
    *n_roots = 0;
    *fex = 0;
    if (fex_check(poly, order, rf_int)) {
        *fex = 1;
        return MZO_OK;
    }
    wl_entry wl[251 + 2];
    size_t nwl = 0;
    double isol_lb[MZO_MAX_OP1 + 2], isol_ub[MZO_MAX_OP1 + 2];
    size_t nisol = 0;
    double q[MZO_MAX_OP1], r[MZO_MAX_OP1], s[MZO_MAX_OP1];

#define ADD_ROOT(rt)                                                                                                   \
    do {                                                                                                               \
        const double root_ = (rt);                                                                                     \
        if (!isfinite(root_)) {                                                                                        \
            set_err(err, errlen,                                                                                       \
                    "Polynomial root finding produced a non-finite root of %.17g involving objects %u and %u", root_,  \
                    i, j);                                                                                             \
            return MZO_EINVAL;                                                                                         \
        }                                                                                                              \
        int accept_ = 1;                                                                                               \
        if (direction != 0) {                                                                                          \
            const double der_ = poly_eval_1(poly, root_, order);                                                       \
            if (!isfinite(der_)) {                                                                                     \
                set_err(err, errlen,                                                                                   \
                        "Polynomial root finding produced the root %.17g with nonfinite derivative %.17g involving "   \
                        "objects %u and %u",                                                                           \
                        root_, der_, i, j);                                                                            \
                return MZO_EINVAL;                                                                                     \
            }                                                                                                          \
            if (sgn_d(der_) != direction) accept_ = 0;                                                                 \
        }                                                                                                              \
        if (accept_) {                                                                                                 \
            if (*n_roots >= 2 * order + 2) {                                                                           \
                set_err(err, errlen, "Too many roots detected for objects %u and %u", i, j);                          \
                return MZO_EINVAL;                                                                                     \
            }                                                                                                          \
            roots[(*n_roots)++] = root_;                                                                               \
        }                                                                                                              \
    } while (0)

    /* rescale [0, rf_int) -> [0, 1) */
    {
        double cf = 1.0;
        for (uint32_t k = 0; k <= order; ++k) {
            wl[0].p[k] = cf * poly[k];
            cf *= rf_int;
        }
        wl[0].lb = 0;
        wl[0].ub = 1;
        nwl = 1;
    }
    do {
        const double lb = wl[nwl - 1].lb, ub = wl[nwl - 1].ub;
        memcpy(q, wl[nwl - 1].p, sizeof(double) * (order + 1));
        --nwl;
        if (q[0] == 0) {
            int allf = 1;
            for (uint32_t k = 1; k <= order; ++k) allf = allf && isfinite(q[k]);
            if (allf) ADD_ROOT(lb * rf_int);
        }
        const uint32_t nsc = rtscc(q, order, r, s);
        if (nsc == 1) {
            isol_lb[nisol] = lb;
            isol_ub[nisol] = ub;
            ++nisol;
        } else if (nsc > 1) {
            /* 2^n q(x/2) into r, then its shift by 1 into s */
            double cf = 1.0;
            for (uint32_t k = 0; k <= order; ++k) {
                r[order - k] = cf * q[order - k];
                cf *= 2;
            }
            translate_1(r, order, s);
            const double mid = (lb / 2) + (ub / 2);
            wl[nwl].lb = lb;
            wl[nwl].ub = mid;
            memcpy(wl[nwl].p, r, sizeof(double) * (order + 1));
            ++nwl;
            wl[nwl].lb = mid;
            wl[nwl].ub = ub;
            memcpy(wl[nwl].p, s, sizeof(double) * (order + 1));
            ++nwl;
        }
        if (nwl > 250 || nisol > order) {
            set_err(err, errlen,
                    "The polynomial root isolation algorithm failed: the working list size is %zu and the number of "
                    "isolating intervals is %zu",
                    nwl, nisol);
            return MZO_EINVAL;
        }
    } while (nwl != 0);

    if (nisol) {
        double cf = 1.0;
        for (uint32_t k = 0; k <= order; ++k) {
            q[k] = cf * poly[k];
            cf *= rf_int;
        }
        for (size_t k = 0; k < nisol; ++k) {
            double rt;
            const int cflag = mzo_bracketed_root(q, order, isol_lb[k], isol_ub[k], &rt, NULL);
            if (cflag == 0) {
                ADD_ROOT(rt * rf_int);
            } else if (cflag == -1) {
                set_err(err, errlen, "Polynomial root finding failed due to too many iterations");
                return MZO_EINVAL;
            } else {
                set_err(err, errlen, "Polynomial root finding returned a nonzero errno with code '%d'", cflag);
                return MZO_EINVAL;
            }
        }
    }
#undef ADD_ROOT
    return MZO_OK;
}

int mzo_root_finding(const double *poly, uint32_t order, double rf_int, int direction, double *roots,
                     uint32_t *n_roots, int *fex)
{
    char err[256];
    return root_finding_impl(poly, order, rf_int, direction, roots, n_roots, fex, 0, 1, err, sizeof(err));
}

/* ------------------------------------------------------------------------ */
/* polyjectory access                                                       */
/* ------------------------------------------------------------------------ */

static const double *pj_traj(const mzo_pj *pj, uint32_t obj) { return pj->traj + pj->traj_off[obj]; }
static const double *pj_time(const mzo_pj *pj, uint32_t obj) { return pj->time + pj->time_off[obj]; }

/* std::upper_bound / std::lower_bound over [first, last). */
static const double *upper_bound_d(const double *first, const double *last, double v)
{
    size_t n = (size_t)(last - first);
    while (n > 0) {
        const size_t h = n / 2;
        if (!(v < first[h])) {
            first += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return first;
}

static const double *lower_bound_d(const double *first, const double *last, double v)
{
    size_t n = (size_t)(last - first);
    while (n > 0) {
        const size_t h = n / 2;
        if (first[h] < v) {
            first += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return first;
}

/* ------------------------------------------------------------------------ */
/* stage 0: step time range (conjunctions.cpp:373-392)                     */
/* ------------------------------------------------------------------------ */

int mzo_cd_begin_end(double maxT, uint64_t cd_idx, double conj_det_interval, uint64_t n_cd_steps, double *cbegin,
                     double *cend)
{
    const double b = conj_det_interval * (double)cd_idx;
    const double e = (cd_idx == n_cd_steps - 1) ? maxT : conj_det_interval * (double)(cd_idx + 1);
    *cbegin = b;
    *cend = e;
    if (!isfinite(b) || !isfinite(e) || !(e > b) || b < 0 || e > maxT) return MZO_EINVAL;
    return MZO_OK;
}

/* ------------------------------------------------------------------------ */
/* stage 1: AABBs (detect_conjunctions_aabbs.cpp:64-252, 300-398)           */
/* ------------------------------------------------------------------------ */

static int object_aabb_impl(const mzo_pj *pj, uint32_t obj, double cbegin, double cend, double conj_thresh,
                            float lb[4], float ub[4], char *err, size_t errlen)
{
    for (int d = 0; d < 4; ++d) {
        lb[d] = INFINITY;
        ub[d] = -INFINITY;
    }
    const uint64_t nsteps = pj->nsteps[obj];
    if (nsteps == 0) return MZO_OK;
    const double *t_begin = pj_time(pj, obj);
    const double *t_end = t_begin + (nsteps + 1);
    if (t_begin[0] >= cend || t_begin[nsteps] <= cbegin) return MZO_OK;
    const double radius = conj_thresh / 2;
    const uint32_t order = pj->order, op1 = order + 1;
    const double *ts_begin = upper_bound_d(t_begin + 1, t_end, cbegin);
    const double *ts_end = lower_bound_d(ts_begin, t_end, cend);
    ts_end += (ts_end != t_end);
    static const int chans[4] = {0, 1, 2, 6};
    for (const double *it = ts_begin; it != ts_end; ++it) {
        const double ss_start = *(it - 1);
        const double ev_lb = cbegin > ss_start ? cbegin : ss_start;
        const double ev_ub = cend < *it ? cend : *it;
        const double h_lb = ev_lb - ss_start;
        const double h_ub = ev_ub - ss_start;
        const size_t ss_idx = (size_t)(it - (t_begin + 1));
        const double *cf = pj_traj(pj, obj) + ss_idx * (size_t)op1 * 7u;
        for (int d = 0; d < 4; ++d) {
            double emin, emax;
            cs_enclosure_impl(cf + chans[d], order, 7, h_lb, h_ub, 0, &emin, &emax);
            emin = emin - radius;
            emax = emax + radius;
            const float fl = nextafterf((float)emin, -INFINITY);
            if (!isfinite(fl)) {
                set_err(err, errlen,
                        "The computation of the bounding box for the object at index %u produced the non-finite "
                        "lower bound %g",
                        obj, (double)fl);
                return MZO_EINVAL;
            }
            const float fu = nextafterf((float)emax, INFINITY);
            if (!isfinite(fu)) {
                set_err(err, errlen,
                        "The computation of the bounding box for the object at index %u produced the non-finite "
                        "upper bound %g",
                        obj, (double)fu);
                return MZO_EINVAL;
            }
            lb[d] = lb[d] < fl ? lb[d] : fl;
            ub[d] = ub[d] > fu ? ub[d] : fu;
        }
    }
    return MZO_OK;
}

int mzo_object_aabb(const mzo_pj *pj, uint32_t obj, double cbegin, double cend, double conj_thresh, float lb[4],
                    float ub[4])
{
    char err[256];
    return object_aabb_impl(pj, obj, cbegin, cend, conj_thresh, lb, ub, err, sizeof(err));
}

/* ------------------------------------------------------------------------ */
/* stage 2: Morton codes + sort (detect_conjunctions_morton.cpp:70-224)     */
/* ------------------------------------------------------------------------ */

uint64_t mzo_disc_single_coord(float x, float min, float max)
{
    const float isize = max - min;
    float rx = (x - min) / isize;
    rx = rx >= 0.f ? rx : 0.f;
    rx *= (float)((uint64_t)1 << 16);
    const uint64_t r = (uint64_t)rx;
    const uint64_t m = ((uint64_t)1 << 16) - 1u;
    return r < m ? r : m;
}

/* bit b of field d -> bit 4b + d (mortonND_LUT.h:262-299, field 0 = LSB). */
uint64_t mzo_morton_encode4(uint64_t n0, uint64_t n1, uint64_t n2, uint64_t n3)
{
    const uint64_t f[4] = {n0, n1, n2, n3};
    uint64_t code = 0;
    for (unsigned b = 0; b < 16; ++b) {
        for (unsigned d = 0; d < 4; ++d) code |= ((f[d] >> b) & 1u) << (4 * b + d);
    }
    return code;
}

typedef struct {
    uint64_t code;
    uint32_t inf;
    uint32_t idx;
} sort_item;

static int cmp_sort_item(const void *pa, const void *pb)
{
    const sort_item *a = (const sort_item *)pa, *b = (const sort_item *)pb;
    if (a->inf != b->inf) return a->inf < b->inf ? -1 : 1;
    if (a->code != b->code) return a->code < b->code ? -1 : 1;
    return (a->idx > b->idx) - (a->idx < b->idx);
}

/* ------------------------------------------------------------------------ */
/* stage 3: BVH (detect_conjunctions_bvh.cpp:267-699)                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    mzo_bvh_node *v;
    size_t n, cap;
} node_vec;

static int node_vec_resize(node_vec *nv, size_t n)
{
    if (n > nv->cap) {
        size_t nc = nv->cap ? nv->cap : 64;
        while (nc < n) nc *= 2;
        mzo_bvh_node *p = (mzo_bvh_node *)realloc(nv->v, nc * sizeof(mzo_bvh_node));
        if (!p) return MZO_ENOMEM;
        nv->v = p;
        nv->cap = nc;
    }
    nv->n = n;
    return MZO_OK;
}

/* Level-by-level top-down build. srt_aabbs is [N+1][8]. Restates the split
 * search including the ++split_idx loop of :426-441. A tree over n objects has
 * at most 2n - 1 nodes and at most n nodes per level. */
static int build_bvh(const float *srt_aabbs, const uint64_t *srt_mcodes, uint32_t tot_n, node_vec *tree)
{
    uint32_t n_objs = 0;
    while (n_objs < tot_n && !isinf(srt_aabbs[(size_t)n_objs * 8])) ++n_objs;
    tree->n = 0;
    if (n_objs == 0) return MZO_OK;

    const size_t max_nodes = 2 * (size_t)n_objs;
    uint32_t *split = (uint32_t *)malloc(sizeof(uint32_t) * max_nodes);
    uint32_t *nn_level = (uint32_t *)malloc(sizeof(uint32_t) * max_nodes);
    uint32_t *nc = (uint32_t *)malloc(sizeof(uint32_t) * n_objs);
    uint32_t *nolc = (uint32_t *)malloc(sizeof(uint32_t) * n_objs);
    uint32_t *ps = (uint32_t *)malloc(sizeof(uint32_t) * n_objs);
    int st = MZO_OK;
    if (!split || !nn_level || !nc || !nolc || !ps || node_vec_resize(tree, 1)) {
        st = MZO_ENOMEM;
        goto done;
    }
    mzo_bvh_node *root = &tree->v[0];
    root->begin = 0;
    root->end = n_objs;
    root->left = root->right = -1;
    for (int d = 0; d < 4; ++d) {
        root->lb[d] = INFINITY;
        root->ub[d] = -INFINITY;
    }
    split[0] = 0;
    nn_level[0] = 0;

    uint32_t cur_n_nodes = 1;
    while (cur_n_nodes != 0) {
        const size_t cur_tree_size = tree->n;
        const size_t n_begin = cur_tree_size - cur_n_nodes, n_end = cur_tree_size;
        uint32_t n_leaf = 0;
        /* step 1: leaf / split detection (:353-503) */
        for (size_t ni = n_begin; ni < n_end; ++ni) {
            mzo_bvh_node *cur = &tree->v[ni];
            int is_leaf = 0;
            size_t split_pos = 0;
            const uint64_t *mb = srt_mcodes + cur->begin;
            const size_t len = (size_t)(cur->end - cur->begin);
            if (len > 1u && split[ni] < 64u) {
                for (;;) {
                    const uint64_t mask = (uint64_t)1 << (63u - split[ni]);
                    /* lower_bound of bit value 1 over the (sorted) bit view */
                    size_t lo = 0, n = len;
                    while (n > 0) {
                        const size_t h = n / 2;
                        if ((mb[lo + h] & mask) == 0) {
                            lo += h + 1;
                            n -= h + 1;
                        } else {
                            n = h;
                        }
                    }
                    split_pos = lo;
                    if (split_pos != 0 && split_pos != len) break;
                    ++split[ni];
                    if (split[ni] == 64u) {
                        is_leaf = 1;
                        break;
                    }
                }
            } else {
                is_leaf = 1;
            }
            const size_t li = ni - n_begin;
            if (is_leaf) {
                nc[li] = 0;
                nolc[li] = 0;
                ++n_leaf;
                for (uint32_t o = cur->begin; o != cur->end; ++o) {
                    for (int d = 0; d < 4; ++d) {
                        const float l = srt_aabbs[(size_t)o * 8 + d], u = srt_aabbs[(size_t)o * 8 + 4 + d];
                        cur->lb[d] = cur->lb[d] < l ? cur->lb[d] : l;
                        cur->ub[d] = cur->ub[d] > u ? cur->ub[d] : u;
                    }
                }
            } else {
                nc[li] = 2;
                nolc[li] = (uint32_t)split_pos;
            }
        }
        /* steps 2-3: grow the tree, inclusive scan of the child counts (:505-536) */
        const uint32_t nn_next = cur_n_nodes * 2u - n_leaf * 2u;
        if (node_vec_resize(tree, cur_tree_size + nn_next)) {
            st = MZO_ENOMEM;
            goto done;
        }
        uint32_t acc = 0;
        for (uint32_t k = 0; k < cur_n_nodes; ++k) {
            acc += nc[k];
            ps[k] = acc;
        }
        /* step 4: children pointers and initial child set-up (:538-616) */
        for (size_t ni = n_begin; ni < n_end; ++ni) {
            mzo_bvh_node *cur = &tree->v[ni];
            const size_t li = ni - n_begin;
            nn_level[ni] = cur_n_nodes;
            if (nc[li] != 0) {
                const size_t lc = cur_tree_size + ps[li] - 2u;
                cur->left = (int32_t)lc;
                cur->right = (int32_t)(lc + 1);
                mzo_bvh_node *l = &tree->v[lc], *r = &tree->v[lc + 1];
                l->begin = cur->begin;
                l->end = cur->begin + nolc[li];
                r->begin = cur->begin + nolc[li];
                r->end = cur->end;
                l->left = l->right = r->left = r->right = -1;
                for (int d = 0; d < 4; ++d) {
                    l->lb[d] = r->lb[d] = INFINITY;
                    l->ub[d] = r->ub[d] = -INFINITY;
                }
                nn_level[lc] = nn_level[lc + 1] = 0;
                split[lc] = split[lc + 1] = split[ni] + 1u;
            }
        }
        cur_n_nodes = nn_next;
    }
    /* bottom-up refit, level by level (:629-691) */
    {
        size_t n_begin = tree->n - nn_level[tree->n - 1];
        size_t n_end = tree->n;
        if (n_begin != 0) {
            size_t new_end = n_begin;
            n_begin -= nn_level[n_begin - 1];
            n_end = new_end;
            for (;;) {
                for (size_t ni = n_begin; ni < n_end; ++ni) {
                    mzo_bvh_node *cur = &tree->v[ni];
                    if (cur->left != -1) {
                        const mzo_bvh_node *l = &tree->v[cur->left], *r = &tree->v[cur->right];
                        for (int d = 0; d < 4; ++d) {
                            cur->lb[d] = l->lb[d] < r->lb[d] ? l->lb[d] : r->lb[d];
                            cur->ub[d] = l->ub[d] > r->ub[d] ? l->ub[d] : r->ub[d];
                        }
                    }
                }
                if (n_begin == 0) break;
                new_end = n_begin;
                n_begin -= nn_level[n_begin - 1];
                n_end = new_end;
            }
        }
    }
done:
    free(split);
    free(nn_level);
    free(nc);
    free(nolc);
    free(ps);
    return st;
}

/* ------------------------------------------------------------------------ */
/* stage 4: broad phase (detect_conjunctions_broad_phase.cpp:153-339)       */
/* ------------------------------------------------------------------------ */

typedef struct {
    mzo_aabb_collision *v;
    size_t n, cap;
} bp_vec;

static int bp_push(bp_vec *bv, uint32_t i, uint32_t j)
{
    if (bv->n == bv->cap) {
        const size_t nc = bv->cap ? bv->cap * 2 : 64;
        mzo_aabb_collision *p = (mzo_aabb_collision *)realloc(bv->v, nc * sizeof(*p));
        if (!p) return MZO_ENOMEM;
        bv->v = p;
        bv->cap = nc;
    }
    bv->v[bv->n].i = i;
    bv->v[bv->n].j = j;
    ++bv->n;
    return MZO_OK;
}

static int cmp_bp(const void *pa, const void *pb)
{
    const mzo_aabb_collision *a = (const mzo_aabb_collision *)pa, *b = (const mzo_aabb_collision *)pb;
    if (a->i != b->i) return a->i < b->i ? -1 : 1;
    return (a->j > b->j) - (a->j < b->j);
}

static int broad_phase(const node_vec *tree, const uint32_t *vidx, const int32_t *otypes, const float *srt_aabbs,
                       bp_vec *out)
{
    out->n = 0;
    if (tree->n == 0) return MZO_OK;
    const uint32_t n_objs = tree->v[0].end - tree->v[0].begin;
    int32_t stack[130];
    for (uint32_t o = 0; o < n_objs; ++o) {
        const uint32_t orig = vidx[o];
        const int32_t ot = otypes[orig];
        const float *bx = srt_aabbs + (size_t)o * 8;
        size_t sp = 0;
        stack[sp++] = 0;
        do {
            const mzo_bvh_node *nd = &tree->v[stack[--sp]];
            const int overlap = (bx[4] >= nd->lb[0] && bx[0] <= nd->ub[0]) && (bx[5] >= nd->lb[1] && bx[1] <= nd->ub[1])
                                && (bx[6] >= nd->lb[2] && bx[2] <= nd->ub[2])
                                && (bx[7] >= nd->lb[3] && bx[3] <= nd->ub[3]);
            if (overlap) {
                if (nd->left == -1) {
                    for (uint32_t k = nd->begin; k != nd->end; ++k) {
                        const uint32_t oi = vidx[k];
                        if (orig >= oi) continue;
                        if (ot + otypes[oi] <= 3) {
                            if (bp_push(out, orig, oi)) return MZO_ENOMEM;
                        }
                    }
                } else {
                    stack[sp++] = nd->left;
                    stack[sp++] = nd->right;
                }
            }
        } while (sp != 0);
    }
    qsort(out->v, out->n, sizeof(mzo_aabb_collision), cmp_bp);
    return MZO_OK;
}

/* ------------------------------------------------------------------------ */
/* stage 5: narrow phase (detect_conjunctions_narrow_phase.cpp:167-769)     */
/* ------------------------------------------------------------------------ */

typedef struct {
    mzo_conj *v;
    size_t n, cap;
} conj_vec;

static int conj_push(conj_vec *cv, const mzo_conj *c)
{
    if (cv->n == cv->cap) {
        const size_t nc = cv->cap ? cv->cap * 2 : 16;
        mzo_conj *p = (mzo_conj *)realloc(cv->v, nc * sizeof(*p));
        if (!p) return MZO_ENOMEM;
        cv->v = p;
        cv->cap = nc;
    }
    cv->v[cv->n++] = *c;
    return MZO_OK;
}

static int cmp_d(double a, double b) { return (a > b) - (a < b); }

/* (tca, then lexicographic over all fields) (narrow_phase.cpp:751-757). */
static int cmp_conj(const void *pa, const void *pb)
{
    const mzo_conj *a = (const mzo_conj *)pa, *b = (const mzo_conj *)pb;
    if (a->tca != b->tca) return a->tca < b->tca ? -1 : 1;
    if (a->i != b->i) return a->i < b->i ? -1 : 1;
    if (a->j != b->j) return a->j < b->j ? -1 : 1;
    int c;
    if ((c = cmp_d(a->tca, b->tca))) return c;
    if ((c = cmp_d(a->dca, b->dca))) return c;
    for (int k = 0; k < 3; ++k)
        if ((c = cmp_d(a->ri[k], b->ri[k]))) return c;
    for (int k = 0; k < 3; ++k)
        if ((c = cmp_d(a->vi[k], b->vi[k]))) return c;
    for (int k = 0; k < 3; ++k)
        if ((c = cmp_d(a->rj[k], b->rj[k]))) return c;
    for (int k = 0; k < 3; ++k)
        if ((c = cmp_d(a->vj[k], b->vj[k]))) return c;
    return 0;
}

static void make_conj(mzo_conj *c, uint32_t i, uint32_t j, double lb_rf, double t, const double *pi, const double *pj,
                      uint32_t order)
{
    c->i = i;
    c->j = j;
    for (int k = 0; k < 3; ++k) {
        c->ri[k] = horner_eval(pi + k, order, t, 7);
        c->vi[k] = horner_eval(pi + 3 + k, order, t, 7);
        c->rj[k] = horner_eval(pj + k, order, t, 7);
        c->vj[k] = horner_eval(pj + 3 + k, order, t, 7);
    }
    const double dx = c->ri[0] - c->rj[0], dy = c->ri[1] - c->rj[1], dz = c->ri[2] - c->rj[2];
    c->dca = sqrt((dx * dx) + (dy * dy) + (dz * dz));
    c->tca = lb_rf + t;
}

static int narrow_phase(const mzo_pj *pj, const mzo_aabb_collision *bp, size_t n_bp, double cbegin, double cend,
                        double conj_thresh, conj_vec *out, uint64_t *rep, char *err, size_t errlen)
{
    const uint32_t order = pj->order, op1 = order + 1;
    const double thresh2 = conj_thresh * conj_thresh;
    double pti[MZO_MAX_OP1 * 7], ptj[MZO_MAX_OP1 * 7];
    double nodes[MZO_MAX_OP1], d2v[MZO_MAX_OP1], ptd2[MZO_MAX_OP1], ptd2p[MZO_MAX_OP1], roots[2 * MZO_MAX_OP1 + 2];
    out->n = 0;
    for (size_t b = 0; b < n_bp; ++b) {
        const uint32_t i = bp[b].i, j = bp[b].j;
        const uint64_t nsi = pj->nsteps[i], nsj = pj->nsteps[j];
        const double *tbi = pj_time(pj, i), *tbj = pj_time(pj, j);
        const double *tei = tbi + (nsi + 1), *tej = tbj + (nsj + 1);
        const double *tsbi = upper_bound_d(tbi + 1, tei, cbegin);
        const double *tsei = lower_bound_d(tsbi, tei, cend);
        tsei += (tsei != tei);
        const double *tsbj = upper_bound_d(tbj + 1, tej, cbegin);
        const double *tsej = lower_bound_d(tsbj, tej, cend);
        tsej += (tsej != tej);
        for (const double *iti = tsbi, *itj = tsbj; iti != tsei && itj != tsej;) {
            const double ts_i = *(iti - 1), ts_j = *(itj - 1);
            const double lb_i = cbegin > ts_i ? cbegin : ts_i;
            const double ub_i = cend < *iti ? cend : *iti;
            const double lb_j = cbegin > ts_j ? cbegin : ts_j;
            const double ub_j = cend < *itj ? cend : *itj;
            if (!(ub_i > lb_j && lb_i < ub_j)) goto next;
            ++rep[MZO_NP_CANDIDATES];
            {
                const double lb_rf = lb_i > lb_j ? lb_i : lb_j;
                const double ub_rf = ub_i < ub_j ? ub_i : ub_j;
                const double delta_i = lb_rf - ts_i, delta_j = lb_rf - ts_j;
                const double rf_int = ub_rf - lb_rf;
                if (!isfinite(delta_i) || !isfinite(delta_j) || !isfinite(rf_int) || delta_i < 0 || delta_j < 0
                    || rf_int < 0) {
                    set_err(err, errlen,
                            "During the narrow phase collision detection of objects %u and %u, an invalid time "
                            "interval for polynomial root finding was generated",
                            i, j);
                    return MZO_EINVAL;
                }
                const double *pi = pj_traj(pj, i) + (size_t)(iti - (tbi + 1)) * op1 * 7u;
                const double *pjp = pj_traj(pj, j) + (size_t)(itj - (tbj + 1)) * op1 * 7u;
                if (delta_i != 0) {
                    for (int ch = 0; ch < 7; ++ch) mzo_poly_translate(pi + ch, order, 7, delta_i, pti + ch, 7);
                    pi = pti;
                }
                if (delta_j != 0) {
                    for (int ch = 0; ch < 7; ++ch) mzo_poly_translate(pjp + ch, order, 7, delta_j, ptj + ch, 7);
                    pjp = ptj;
                }
                /* dist^2 interpolant at the Chebyshev nodes of [0, rf_int] (:95-152) */
                mzo_cheby_nodes(order, rf_int, nodes);
                for (uint32_t k = 0; k < op1; ++k) {
                    const double xi = horner_eval(pi + 0, order, nodes[k], 7);
                    const double yi = horner_eval(pi + 1, order, nodes[k], 7);
                    const double zi = horner_eval(pi + 2, order, nodes[k], 7);
                    const double xj = horner_eval(pjp + 0, order, nodes[k], 7);
                    const double yj = horner_eval(pjp + 1, order, nodes[k], 7);
                    const double zj = horner_eval(pjp + 2, order, nodes[k], 7);
                    const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
                    d2v[k] = dx * dx + dy * dy + dz * dz;
                }
                mzo_vm_interp(order, nodes, d2v, ptd2);
                double e_lo, e_hi;
                cs_enclosure_impl(ptd2, order, 1, 0.0, rf_int, 1, &e_lo, &e_hi);
                if (!isfinite(e_lo) || !isfinite(e_hi)) {
                    set_err(err, errlen, "Non-finite value(s) detected during conjunction tracking for objects %u and %u",
                            i, j);
                    return MZO_EINVAL;
                }
                if (e_lo < thresh2) {
                    for (uint32_t k = 0; k < order; ++k) ptd2p[k] = (double)(k + 1u) * ptd2[k + 1u];
                    ptd2p[order] = 0;
                    uint32_t nr = 0;
                    int fex = 0;
                    const int st = root_finding_impl(ptd2p, order, rf_int, 1, roots, &nr, &fex, i, j, err, errlen);
                    if (st != MZO_OK) return st;
                    ++rep[MZO_NP_POLY_ROOTS];
                    rep[MZO_NP_FEX_CHECK] += (uint64_t)fex;
                    rep[MZO_NP_POLY_NO_ROOTS] += (nr == 0);
                    rep[MZO_NP_DIST_MINIMA] += nr;
                    for (uint32_t r = 0; r < nr; ++r) {
                        const double cd2 = horner_eval(ptd2, order, roots[r], 1);
                        if (!isfinite(cd2)) {
                            set_err(err, errlen,
                                    "A non-finite conjunction distance square of %.17g was computed for the objects at "
                                    "indices %u and %u",
                                    cd2, i, j);
                            return MZO_EINVAL;
                        }
                        if (cd2 < thresh2) {
                            mzo_conj c;
                            make_conj(&c, i, j, lb_rf, roots[r], pi, pjp, order);
                            if (conj_push(out, &c)) return MZO_ENOMEM;
                        } else {
                            ++rep[MZO_NP_DISCARDED_MINIMA];
                        }
                    }
                    /* boundary minima (:629-714) */
                    if ((iti + 1 == tei && *iti == ub_rf) || (itj + 1 == tej && *itj == ub_rf)) {
                        const double tmc = nextafter(rf_int, -1.);
                        const double dv = horner_eval(ptd2p, order, tmc, 1);
                        if (!isfinite(dv)) {
                            set_err(err, errlen,
                                    "An invalid value of %.17g was computed for the derivative of the distance square "
                                    "function at the upper boundary of a trajectory",
                                    dv);
                            return MZO_EINVAL;
                        }
                        if (dv < 0) {
                            const double d2 = horner_eval(ptd2, order, tmc, 1);
                            if (!isfinite(d2)) {
                                set_err(err, errlen,
                                        "An invalid value of %.17g was computed for a candidate minimum of the distance "
                                        "square function at the upper boundary of a trajectory",
                                        d2);
                                return MZO_EINVAL;
                            }
                            if (d2 < thresh2) {
                                mzo_conj c;
                                make_conj(&c, i, j, lb_rf, tmc, pi, pjp, order);
                                if (conj_push(out, &c)) return MZO_ENOMEM;
                            }
                        }
                    }
                    if ((iti == tbi + 1 && lb_rf == *tbi) || (itj == tbj + 1 && lb_rf == *tbj)) {
                        const double dv = horner_eval(ptd2p, order, 0., 1);
                        if (!isfinite(dv)) {
                            set_err(err, errlen,
                                    "An invalid value of %.17g was computed for the derivative of the distance square "
                                    "function at the lower boundary of a trajectory",
                                    dv);
                            return MZO_EINVAL;
                        }
                        if (dv > 0) {
                            const double d2 = horner_eval(ptd2, order, 0., 1);
                            if (!isfinite(d2)) {
                                set_err(err, errlen,
                                        "An invalid value of %.17g was computed for a candidate minimum of the distance "
                                        "square function at the lower boundary of a trajectory",
                                        d2);
                                return MZO_EINVAL;
                            }
                            if (d2 < thresh2) {
                                mzo_conj c;
                                make_conj(&c, i, j, lb_rf, 0., pi, pjp, order);
                                if (conj_push(out, &c)) return MZO_ENOMEM;
                            }
                        }
                    }
                } else {
                    ++rep[MZO_NP_DIST2_CHECK];
                }
            }
        next : {
            const int inc_i = (*iti <= *itj), inc_j = (*itj <= *iti);
            iti += inc_i;
            itj += inc_j;
        }
        }
    }
    qsort(out->v, out->n, sizeof(mzo_conj), cmp_conj);
    return MZO_OK;
}

/* ------------------------------------------------------------------------ */
/* orchestration (conjunctions.cpp:263-360, detect_conjunctions.cpp:55-537) */
/* ------------------------------------------------------------------------ */

typedef struct {
    node_vec tree;
    bp_vec bp;
    conj_vec conj;
    int status;
    char err[512];
} step_out;

void mzo_free_result(mzo_result *r)
{
    if (!r) return;
    free(r->aabbs);
    free(r->srt_aabbs);
    free(r->mcodes);
    free(r->srt_mcodes);
    free(r->srt_idx);
    free(r->cd_end_times);
    free(r->trees);
    free(r->tree_offsets);
    free(r->bp);
    free(r->bp_offsets);
    free(r->conj);
    memset(r, 0, sizeof(*r));
}

int mzo_conjunctions(const mzo_pj *pj, double conj_thresh, double conj_det_interval, const int32_t *otypes_in,
                     uint64_t step_begin, uint64_t step_end, int n_threads, mzo_result *out)
{
    memset(out, 0, sizeof(*out));
    const uint32_t N = pj->n_objs;
    out->n_objs = N;
    if (pj->order + 1 > MZO_MAX_OP1 || pj->order < 2) {
        out->status = MZO_EINVAL;
        set_err(out->err, sizeof(out->err), "Unsupported polynomial order %u", pj->order);
        return out->status;
    }
    if (!isfinite(conj_thresh) || conj_thresh <= 0) {
        out->status = MZO_EINVAL;
        set_err(out->err, sizeof(out->err),
                "The conjunction threshold must be finite and positive, but instead a value of %.17g was provided",
                conj_thresh);
        return out->status;
    }
    if (!isfinite(conj_thresh * conj_thresh)) {
        out->status = MZO_EINVAL;
        set_err(out->err, sizeof(out->err),
                "A conjunction threshold of %.17g is too large and results in an overflow error", conj_thresh);
        return out->status;
    }
    if (!isfinite(conj_det_interval) || conj_det_interval <= 0) {
        out->status = MZO_EINVAL;
        set_err(out->err, sizeof(out->err),
                "The conjunction detection interval must be finite and positive, but instead a value of %.17g was "
                "provided",
                conj_det_interval);
        return out->status;
    }
    int32_t *otypes = (int32_t *)malloc(sizeof(int32_t) * (N ? N : 1));
    if (!otypes) {
        out->status = MZO_ENOMEM;
        return out->status;
    }
    int skip_cd = 1;
    if (otypes_in) {
        for (uint32_t k = 0; k < N; ++k) {
            const int32_t v = otypes_in[k];
            if (v != 1 && v != 2 && v != 4) {
                free(otypes);
                out->status = MZO_EINVAL;
                set_err(out->err, sizeof(out->err),
                        "The value of an object type must be one of [1, 2, 4], but a value of %d was detected instead",
                        v);
                return out->status;
            }
            skip_cd = skip_cd && (v != 1);
            otypes[k] = v;
        }
    } else {
        for (uint32_t k = 0; k < N; ++k) otypes[k] = 1;
        skip_cd = 0;
    }
    out->skip_cd = skip_cd;
    const double nsteps_d = ceil(pj->maxT / conj_det_interval);
    if (!(nsteps_d >= 0) || nsteps_d > 1e15) {
        free(otypes);
        out->status = MZO_EINVAL;
        set_err(out->err, sizeof(out->err), "Invalid number of conjunction steps");
        return out->status;
    }
    const uint64_t C = (uint64_t)nsteps_d;
    out->n_cd_steps = C;
    if (step_end == 0 || step_end > C) step_end = C;
    if (step_begin > step_end) step_begin = step_end;
    out->step_begin = step_begin;
    out->step_end = step_end;
    const uint64_t ns = step_end - step_begin;
    const size_t box_row = (size_t)(N + 1) * 8;

    out->aabbs = (float *)malloc(sizeof(float) * box_row * (ns ? ns : 1));
    out->srt_aabbs = (float *)malloc(sizeof(float) * box_row * (ns ? ns : 1));
    out->mcodes = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)N * (ns ? ns : 1));
    out->srt_mcodes = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)N * (ns ? ns : 1));
    out->srt_idx = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)N * (ns ? ns : 1));
    out->cd_end_times = (double *)malloc(sizeof(double) * (ns ? ns : 1));
    out->tree_offsets = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (ns ? ns : 1));
    out->bp_offsets = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (ns ? ns : 1));
    step_out *so = (step_out *)calloc(ns ? ns : 1, sizeof(step_out));
    if (!out->aabbs || !out->srt_aabbs || !out->mcodes || !out->srt_mcodes || !out->srt_idx || !out->cd_end_times
        || !out->tree_offsets || !out->bp_offsets || !so) {
        free(otypes);
        free(so);
        mzo_free_result(out);
        out->status = MZO_ENOMEM;
        return MZO_ENOMEM;
    }
    double st_aabb = 0, st_morton = 0, st_bvh = 0, st_broad = 0, st_narrow = 0;
    uint64_t rep[MZO_NP_COUNT] = {0};

#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif

    /* conjunction steps in parallel (detect_conjunctions.cpp:313-424) */
#pragma omp parallel num_threads(n_threads) reduction(+ : st_aabb, st_morton, st_bvh, st_broad, st_narrow)
    {
        sort_item *items = (sort_item *)malloc(sizeof(sort_item) * (N ? N : 1));
        uint64_t lrep[MZO_NP_COUNT] = {0};
#pragma omp for schedule(dynamic, 1)
        for (uint64_t s = 0; s < ns; ++s) {
            const uint64_t cd_idx = step_begin + s;
            step_out *o = &so[s];
            float *aabbs = out->aabbs + box_row * s;
            float *srt_aabbs = out->srt_aabbs + box_row * s;
            uint64_t *mcodes = out->mcodes + (size_t)N * s;
            uint64_t *srt_mcodes = out->srt_mcodes + (size_t)N * s;
            uint32_t *vidx = out->srt_idx + (size_t)N * s;
            double cb, ce;
            if (!items) {
                o->status = MZO_ENOMEM;
                continue;
            }
            if (mzo_cd_begin_end(pj->maxT, cd_idx, conj_det_interval, C, &cb, &ce) != MZO_OK) {
                o->status = MZO_EINVAL;
                set_err(o->err, sizeof(o->err), "Invalid conjunction step time range: [%.17g, %.17g)", cb, ce);
                continue;
            }
            double t0 = now_s();
            /* AABBs + global box */
            float glb[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, gub[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
            for (uint32_t ob = 0; ob < N && o->status == MZO_OK; ++ob) {
                float lb[4], ub[4];
                if (object_aabb_impl(pj, ob, cb, ce, conj_thresh, lb, ub, o->err, sizeof(o->err)) != MZO_OK) {
                    o->status = MZO_EINVAL;
                    break;
                }
                for (int d = 0; d < 4; ++d) {
                    aabbs[(size_t)ob * 8 + d] = lb[d];
                    aabbs[(size_t)ob * 8 + 4 + d] = ub[d];
                    glb[d] = glb[d] < lb[d] ? glb[d] : lb[d];
                    gub[d] = gub[d] > ub[d] ? gub[d] : ub[d];
                }
            }
            if (o->status != MZO_OK) continue;
            for (int d = 0; d < 4; ++d) {
                if (isfinite(glb[d]) && !isfinite(gub[d] - glb[d])) {
                    o->status = MZO_EINVAL;
                    set_err(o->err, sizeof(o->err), "A global bounding box with non-finite size was generated");
                }
                aabbs[(size_t)N * 8 + d] = glb[d];
                aabbs[(size_t)N * 8 + 4 + d] = gub[d];
            }
            if (o->status != MZO_OK) continue;
            out->cd_end_times[s] = ce;
            double t1 = now_s();
            st_aabb += t1 - t0;
            /* Morton codes, sort, gather */
            for (uint32_t ob = 0; ob < N; ++ob) {
                const float *bx = aabbs + (size_t)ob * 8;
                if (isfinite(bx[0])) {
                    uint64_t q[4];
                    for (int d = 0; d < 4; ++d) {
                        const float ctr = bx[d] / 2 + bx[4 + d] / 2;
                        q[d] = mzo_disc_single_coord(ctr, glb[d], gub[d]);
                    }
                    mcodes[ob] = mzo_morton_encode4(q[0], q[1], q[2], q[3]);
                } else {
                    mcodes[ob] = ~(uint64_t)0;
                }
                items[ob].code = mcodes[ob];
                items[ob].inf = isinf(bx[0]) ? 1u : 0u;
                items[ob].idx = ob;
            }
            qsort(items, N, sizeof(sort_item), cmp_sort_item);
            for (uint32_t p = 0; p < N; ++p) {
                const uint32_t si = items[p].idx;
                vidx[p] = si;
                memcpy(srt_aabbs + (size_t)p * 8, aabbs + (size_t)si * 8, sizeof(float) * 8);
                srt_mcodes[p] = mcodes[si];
            }
            memcpy(srt_aabbs + (size_t)N * 8, aabbs + (size_t)N * 8, sizeof(float) * 8);
            double t2 = now_s();
            st_morton += t2 - t1;
            /* BVH */
            if (build_bvh(srt_aabbs, srt_mcodes, N, &o->tree) != MZO_OK) {
                o->status = MZO_ENOMEM;
                continue;
            }
            double t3 = now_s();
            st_bvh += t3 - t2;
            if (!skip_cd) {
                if (broad_phase(&o->tree, vidx, otypes, srt_aabbs, &o->bp) != MZO_OK) {
                    o->status = MZO_ENOMEM;
                    continue;
                }
                double t4 = now_s();
                st_broad += t4 - t3;
                const int st = narrow_phase(pj, o->bp.v, o->bp.n, cb, ce, conj_thresh, &o->conj, lrep, o->err,
                                            sizeof(o->err));
                if (st != MZO_OK) {
                    o->status = st;
                    continue;
                }
                st_narrow += now_s() - t4;
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < MZO_NP_COUNT; ++k) rep[k] += lrep[k];
        }
        free(items);
    }
    free(otypes);

    /* gather per-step outputs in step order (the writer thread of :184-242) */
    size_t tot_nodes = 0, tot_bp = 0, tot_conj = 0;
    for (uint64_t s = 0; s < ns; ++s) {
        if (so[s].status != MZO_OK && out->status == MZO_OK) {
            out->status = so[s].status;
            memcpy(out->err, so[s].err, sizeof(out->err));
        }
        out->tree_offsets[2 * s] = tot_nodes;
        out->tree_offsets[2 * s + 1] = so[s].tree.n;
        tot_nodes += so[s].tree.n;
        out->bp_offsets[2 * s] = tot_bp;
        out->bp_offsets[2 * s + 1] = so[s].bp.n;
        tot_bp += so[s].bp.n;
        tot_conj += so[s].conj.n;
    }
    out->trees = (mzo_bvh_node *)malloc(sizeof(mzo_bvh_node) * (tot_nodes ? tot_nodes : 1));
    out->bp = (mzo_aabb_collision *)malloc(sizeof(mzo_aabb_collision) * (tot_bp ? tot_bp : 1));
    out->conj = (mzo_conj *)malloc(sizeof(mzo_conj) * (tot_conj ? tot_conj : 1));
    if (!out->trees || !out->bp || !out->conj) {
        if (out->status == MZO_OK) out->status = MZO_ENOMEM;
    } else {
        size_t a = 0, b = 0, c = 0;
        for (uint64_t s = 0; s < ns; ++s) {
            if (so[s].tree.n) memcpy(out->trees + a, so[s].tree.v, sizeof(mzo_bvh_node) * so[s].tree.n);
            a += so[s].tree.n;
            if (so[s].bp.n) memcpy(out->bp + b, so[s].bp.v, sizeof(mzo_aabb_collision) * so[s].bp.n);
            b += so[s].bp.n;
            if (so[s].conj.n) memcpy(out->conj + c, so[s].conj.v, sizeof(mzo_conj) * so[s].conj.n);
            c += so[s].conj.n;
        }
        out->n_conj = tot_conj;
    }
    for (uint64_t s = 0; s < ns; ++s) {
        free(so[s].tree.v);
        free(so[s].bp.v);
        free(so[s].conj.v);
    }
    free(so);
    memcpy(out->np_report, rep, sizeof(rep));
    out->stage_seconds[0] = st_aabb;
    out->stage_seconds[1] = st_morton;
    out->stage_seconds[2] = st_bvh;
    out->stage_seconds[3] = st_broad;
    out->stage_seconds[4] = st_narrow;
    return out->status;
}

/* ------------------------------------------------------------------------ */
/* polyjectory evaluation (polyjectory.cpp:1094-1366)                        */
/* ------------------------------------------------------------------------ */

int mzo_state_eval(const mzo_pj *pj, uint64_t obj, double tm, double out[7])
{
    /* polyjectory.cpp:1097-1100: the time is checked first */
    if (!isfinite(tm)) return MZO_EINVAL;
    /* polyjectory::operator[] (:1009-1014) */
    if (obj >= pj->n_objs) return MZO_ERANGE;
    const uint64_t nsteps = pj->nsteps[obj];
    const double *t = pj->time + pj->time_off[obj];
    /* no data, before the start or at/after the end: NaN (:1108-1120) */
    if (nsteps == 0 || tm < t[0] || tm >= t[nsteps]) {
        for (int ch = 0; ch < 7; ++ch) out[ch] = NAN;
        return MZO_OK;
    }
    /* first end time > tm among t[1..nsteps] (upper_bound, :1136-1138) */
    uint64_t lo = 1, hi = nsteps;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (t[mid] > tm) {
            hi = mid;
        } else {
            lo = mid + 1;
        }
    }
    const uint64_t step = lo - 1;
    const double h = tm - t[step]; /* :1141 */
    const uint32_t order = pj->order;
    const double *c = pj->traj + pj->traj_off[obj] + step * (uint64_t)(order + 1) * 7u;
    /* horner_eval with stride 7 (detail/poly_utils.hpp:41-51) */
    for (int ch = 0; ch < 7; ++ch) {
        double r = c[(uint64_t)order * 7u + ch];
        for (uint32_t o = 1; o <= order; ++o) r = c[(uint64_t)(order - o) * 7u + ch] + r * h;
        out[ch] = r;
    }
    return MZO_OK;
}

int mzo_state_meval(const mzo_pj *pj, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                    uint64_t n_times, double *out, uint64_t *bad)
{
    for (uint64_t r = 0; r < n_rows; ++r) {
        const uint64_t obj = sel ? sel[r] : r;
        for (uint64_t k = 0; k < n_times; ++k) {
            const uint64_t e = r * n_times + k;
            double t;
            switch (time_mode) {
                case MZO_TIME_SCALAR:
                    t = tm[0];
                    break;
                case MZO_TIME_PER_OBJECT:
                    t = tm[r];
                    break;
                case MZO_TIME_SHARED:
                    t = tm[k];
                    break;
                default:
                    t = tm[e];
                    break;
            }
            const int st = mzo_state_eval(pj, obj, t, out + e * 7u);
            if (st != MZO_OK) {
                if (bad) *bad = e;
                return st;
            }
        }
    }
    return MZO_OK;
}

// Narrow phase for gfx950 (detect_conjunctions_narrow_phase.cpp:167-769).
//
// One lane per broad-phase pair of any step of the batch. For every pair of
// overlapping trajectory steps the lane translates the position polynomials to
// the common interval start, evaluates them at the order+1 Chebyshev nodes of
// [0, rf_int], interpolates dist^2 with Bjorck-Pereyra, screens with a
// Cargo-Shisha enclosure and, when the enclosure reaches below thresh^2,
// isolates the minima (roots of d(dist^2)/dt with positive second derivative)
// with the Vincent-Collins-Akritas bisection and refines them with TOMS 748.
// All of it in FP64 with separately rounded operations, in the operation
// order of the CPU restatement, so TCA/DCA match it bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

// Order range of the kernels compiled into this object; a standalone compile
// builds every order and the host-side entry points.
#ifndef MZH_NARROW_LO
#define MZH_NARROW_LO 2
#define MZH_NARROW_HI 20
#define MZH_NARROW_HOST
#endif

namespace mzh
{

namespace
{

constexpr BinomTable kBinomN = make_binom_table();
__constant__ BinomTable c_binom_np = kBinomN;

// Work list of the root isolation. A lane first isolates with a private list
// of kWlCap entries; a candidate whose list outgrows it is isolated again by
// the slow pass with a list of kWlSlowCap entries in global memory, so the
// reference's limit (more than 250 pending intervals is an error,
// poly_utils.cpp:417-421) is the only one that can fail a screening. With
// the bound checked after every split, 250 entries plus the two children of
// the last split never exceed kWlSlowCap.
constexpr int kWlCap = 48;
constexpr int kWlSlowCap = 251;
// find_minima results
constexpr int kFmOk = 0, kFmError = 1, kFmListFull = 2;

// The problem of global pair index g: the largest p with pre[p] <= g.
__device__ inline uint32_t find_prob64(const uint64_t *__restrict__ pre, uint32_t np, uint64_t g)
{
    uint32_t lo = 0, n = np;
    while (n > 1) {
        const uint32_t h = n / 2;
        if (pre[lo + h] <= g) {
            lo += h;
            n -= h;
        } else {
            n = h;
        }
    }
    return lo;
}

// Offset of the report copy this wave adds to (see kReportSlots).
__device__ inline uint32_t report_slot()
{
    return ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kReportSlots - 1u)) * kReportStride;
}

// Add v[0..K) of every active lane to np_report[k0..k0+K) of the lane's problem
// p. The lanes of a wave nearly always belong to one problem: the wave's
// sums are then added by one lane. Call from all lanes of the wave.
template <int K>
__device__ inline void add_prob_report(const Prob *__restrict__ P, uint32_t p, bool active, int k0,
                                       const unsigned long long *v)
{
    const uint64_t act = __ballot(active);
    if (act == 0) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint32_t p0 = __shfl(p, leader, 64);
    const bool same = __ballot(active && p != p0) == 0;
    const uint32_t slot = report_slot() + (uint32_t)k0;
    if (same) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const unsigned long long s = wave_reduce_add_u64(active ? v[k] : 0ull);
            if ((int)__lane_id() == leader && s) atomicAdd(&P[p0].np_report[slot + k], s);
        }
    } else if (active) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (v[k]) atomicAdd(&P[p].np_report[slot + k], v[k]);
    }
}
// capacity of the per-interval root list (isolated roots plus roots found
// exactly at a work-list interval's lower end)
template <int ORDER>
constexpr int kMaxRoots = 2 * ORDER + 2;

template <int ORDER>
__device__ inline void halving_pows_d(double a, double *p)
{
    p[0] = 1.0;
    p[1] = a;
#pragma unroll
    for (int i = 2; i <= ORDER; ++i) p[i] = (i % 2 == 0) ? p[i / 2] * p[i / 2] : p[i / 2 + 1] * p[i / 2];
}

// p(x) -> p(x + a) for one channel of an AoS trajectory step (stride 7)
// (conjunctions_jit.cpp:68-115).
template <int ORDER>
__device__ inline void translate7(const double *__restrict__ src, double a, double *out)
{
    constexpr int OP1 = ORDER + 1;
    double c[OP1];
#pragma unroll
    for (int k = 0; k < OP1; ++k) c[k] = src[k * 7];
    if (a == 0.0) {
#pragma unroll
        for (int k = 0; k < OP1; ++k) out[k] = c[k];
        return;
    }
    double ap[OP1];
    halving_pows_d<ORDER>(a, ap);
#pragma unroll
    for (int i = 0; i < OP1; ++i) {
        double s = (c[i] * c_binom_np.v[i][0]) * ap[0];
#pragma unroll
        for (int k = i + 1; k < OP1; ++k) s = s + (c[k] * c_binom_np.v[k][k - i]) * ap[k - i];
        out[i] = s;
    }
}

template <int ORDER>
__device__ inline double horner(const double *p, double x)
{
    double r = p[ORDER];
#pragma unroll
    for (int o = 1; o <= ORDER; ++o) r = p[ORDER - o] + r * x;
    return r;
}

__device__ inline double hy_min_np(double a, double b) { return (b == b) ? ((b < a) ? b : a) : b; }
__device__ inline double hy_max_np(double a, double b) { return (b == b) ? ((a < b) ? b : a) : b; }

template <int N>
__device__ inline double pwr_min(double *v)
{
    if constexpr (N == 1) {
        return v[0];
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) v[i] = hy_min_np(v[2 * i], v[2 * i + 1]);
        if constexpr (N % 2) v[N / 2] = v[N - 1];
        return pwr_min<(N + 1) / 2>(v);
    }
}

template <int N>
__device__ inline double pwr_max(double *v)
{
    if constexpr (N == 1) {
        return v[0];
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) v[i] = hy_max_np(v[2 * i], v[2 * i + 1]);
        if constexpr (N % 2) v[N / 2] = v[N - 1];
        return pwr_max<(N + 1) / 2>(v);
    }
}

// Cargo-Shisha enclosure of p over [0, h] (cs_enc, conjunctions_jit.cpp:302-323).
template <int ORDER>
__device__ inline void cs_enc0(const double *p, double h, double &lo, double &hi)
{
    constexpr int OP1 = ORDER + 1;
    double t[OP1], dp[OP1], bmin[OP1], bmax[OP1];
    halving_pows_d<ORDER>(h, dp);
#pragma unroll
    for (int i = 0; i < OP1; ++i) t[i] = (p[i] * dp[i]) / c_binom_np.v[ORDER][i];
    bmin[0] = bmax[0] = t[0];
#pragma unroll
    for (int r = 1; r <= ORDER; ++r) {
#pragma unroll
        for (int k = 0; k + r <= ORDER; ++k) t[k] = t[k] + t[k + 1];
        bmin[r] = bmax[r] = t[0];
    }
    lo = pwr_min<OP1>(bmin);
    hi = pwr_max<OP1>(bmax);
}

// ---- root isolation + refinement (poly_utils.cpp:150-460) -----------------

__device__ inline int sgn_d(double v) { return (0.0 < v) - (v < 0.0); }

template <int ORDER>
__device__ inline void translate1(const double *c, double *out)
{
    constexpr int OP1 = ORDER + 1;
    double tmp[OP1];
#pragma unroll
    for (int i = 0; i < OP1; ++i) {
        double s = c[i];
#pragma unroll
        for (int k = i + 1; k < OP1; ++k) s = s + c[k] * c_binom_np.v[k][i];
        tmp[i] = s;
    }
#pragma unroll
    for (int i = 0; i < OP1; ++i) out[i] = tmp[i];
}

template <int ORDER>
__device__ inline uint32_t rtscc(const double *q)
{
    constexpr int OP1 = ORDER + 1;
    double r[OP1], s[OP1];
#pragma unroll
    for (int i = 0; i < OP1; ++i) r[i] = q[ORDER - i];
    translate1<ORDER>(r, s);
    uint32_t cnt = 0;
    int last = sgn_d(s[0]);
#pragma unroll
    for (int i = 1; i < OP1; ++i) {
        const int cur = sgn_d(s[i]);
        cnt += (cur * last == -1) ? 1u : 0u;
        if (cur != 0) last = cur;
    }
    return cnt;
}

template <int ORDER>
__device__ inline bool fex_excludes(const double *p, double h)
{
    double lo = p[ORDER], hi = p[ORDER];
#pragma unroll
    for (int o = 1; o <= ORDER; ++o) {
        const double a = lo * h, b = hi * h;
        double nlo = a < b ? a : b, nhi = a < b ? b : a;
        nlo = nlo < 0.0 ? nlo : 0.0;
        nhi = nhi > 0.0 ? nhi : 0.0;
        lo = nlo + p[ORDER - o];
        hi = nhi + p[ORDER - o];
    }
    return (lo > 0.0) || (hi < 0.0);
}

template <int ORDER>
__device__ inline double poly_eval_1(const double *a, double x)
{
    double r = a[ORDER] * (double)ORDER;
#pragma unroll
    for (int i = 1; i < ORDER; ++i) r = a[ORDER - i] * (double)(ORDER - i) + r * x;
    return r;
}

// TOMS 748 (Alefeld, Potra & Shi 1995), eps tolerance 4 * 2^-52, same
// step sequence as the CPU restatement.
constexpr double kEps = 2.220446049250313080847263336181640625e-16;

__device__ inline int bsign(double x) { return (x > 0) - (x < 0); }

__device__ inline bool t_tol(double a, double b)
{
    const double m = fabs(a) < fabs(b) ? fabs(a) : fabs(b);
    return fabs(a - b) <= (4.0 * kEps) * m;
}

__device__ inline double t_secant(double a, double b, double fa, double fb)
{
    const double tol = kEps * 5;
    const double c = a - (fa / (fb - fa)) * (b - a);
    if ((c <= a + fabs(a) * tol) || (c >= b - fabs(b) * tol)) return (a + b) / 2;
    return c;
}

__device__ inline double t_safe_div(double num, double den, double r)
{
    if (fabs(den) < 1) {
        if (fabs(den * 1.7976931348623157e308) <= fabs(num)) return r;
    }
    return num / den;
}

__device__ inline double t_quadratic(double a, double b, double d, double fa, double fb, double fd, int count)
{
    const double B = t_safe_div(fb - fa, b - a, 1.7976931348623157e308);
    double A = t_safe_div(fd - fb, d - b, 1.7976931348623157e308);
    A = t_safe_div(A - B, d - a, 0.0);
    if (A == 0) return t_secant(a, b, fa, fb);
    double c = (bsign(A) * bsign(fa) > 0) ? a : b;
    for (int i = 1; i <= count; ++i) {
        c -= t_safe_div(fa + (B + A * (c - b)) * (c - a), B + A * (2 * c - a - b), 1 + c - a);
    }
    if ((c <= a) || (c >= b)) c = t_secant(a, b, fa, fb);
    return c;
}

__device__ inline double t_cubic(double a, double b, double d, double e, double fa, double fb, double fd, double fe)
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
    if ((c <= a) || (c >= b)) c = t_quadratic(a, b, d, fa, fb, fd, 3);
    return c;
}

__device__ inline bool t_prof(double fa, double fb, double fd, double fe)
{
    const double md = 2.2250738585072014e-308 * 32;
    return (fabs(fa - fb) < md) || (fabs(fa - fd) < md) || (fabs(fa - fe) < md) || (fabs(fb - fd) < md)
           || (fabs(fb - fe) < md) || (fabs(fd - fe) < md);
}

template <int ORDER>
struct T748 {
    const double *q;
    __device__ inline double f(double x) const { return horner<ORDER>(q, x); }

    __device__ inline void bracket(double &a, double &b, double c, double &fa, double &fb, double &d, double &fd) const
    {
        const double tol = kEps * 2;
        if ((b - a) < 2 * tol * a) {
            c = a + (b - a) / 2;
        } else if (c <= a + fabs(a) * tol) {
            c = a + fabs(a) * tol;
        } else if (c >= b - fabs(b) * tol) {
            c = b - fabs(b) * tol;
        }
        const double fc = f(c);
        if (fc == 0) {
            a = c;
            fa = 0;
            d = 0;
            fd = 0;
            return;
        }
        if (bsign(fa) * bsign(fc) < 0) {
            d = b;
            fd = fb;
            b = c;
            fb = fc;
        } else {
            d = a;
            fd = fa;
            a = c;
            fa = fc;
        }
    }

    // returns 0 on success, 33 (EDOM) on a non-bracketing interval; iters =
    // number of function evaluations
    __device__ inline int solve(double a, double b, unsigned max_iter, double &ra, double &rb, unsigned &iters) const
    {
        double fa = f(a), fb = f(b);
        unsigned count = max_iter - 2;
        const unsigned count0 = count;
        if (a >= b) {
            iters = 2;
            return 33;
        }
        if (t_tol(a, b) || (fa == 0) || (fb == 0)) {
            if (fa == 0) {
                b = a;
            } else if (fb == 0) {
                a = b;
            }
            ra = a;
            rb = b;
            iters = 2;
            return 0;
        }
        if (bsign(fa) * bsign(fb) > 0) {
            iters = 2;
            return 33;
        }
        double c, u, fu, a0, b0, d = 0, fd = 1e5, e = 1e5, fe = 1e5;
        if (fa != 0) {
            c = t_secant(a, b, fa, fb);
            bracket(a, b, c, fa, fb, d, fd);
            --count;
            if (count && (fa != 0) && !t_tol(a, b)) {
                c = t_quadratic(a, b, d, fa, fb, fd, 2);
                e = d;
                fe = fd;
                bracket(a, b, c, fa, fb, d, fd);
                --count;
            }
        }
        while (count && (fa != 0) && !t_tol(a, b)) {
            a0 = a;
            b0 = b;
            c = t_prof(fa, fb, fd, fe) ? t_quadratic(a, b, d, fa, fb, fd, 2) : t_cubic(a, b, d, e, fa, fb, fd, fe);
            e = d;
            fe = fd;
            bracket(a, b, c, fa, fb, d, fd);
            if ((0 == --count) || (fa == 0) || t_tol(a, b)) break;
            c = t_prof(fa, fb, fd, fe) ? t_quadratic(a, b, d, fa, fb, fd, 3) : t_cubic(a, b, d, e, fa, fb, fd, fe);
            bracket(a, b, c, fa, fb, d, fd);
            if ((0 == --count) || (fa == 0) || t_tol(a, b)) break;
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
            bracket(a, b, c, fa, fb, d, fd);
            if ((0 == --count) || (fa == 0) || t_tol(a, b)) break;
            if ((b - a) < 0.5 * (b0 - a0)) continue;
            e = d;
            fe = fd;
            bracket(a, b, a + (b - a) / 2, fa, fb, d, fd);
            --count;
        }
        if (fa == 0) {
            b = a;
        } else if (fb == 0) {
            a = b;
        }
        ra = a;
        rb = b;
        iters = (count0 - count) + 2;
        return 0;
    }
};

template <int ORDER>
struct WLEntry {
    double lb, ub;
    double p[ORDER + 1];
};
static_assert(sizeof(WLEntry<8>) == narrow_wl_entry_bytes(8), "work-list entry size");
static_assert(sizeof(WLEntry<20>) == narrow_wl_entry_bytes(20), "work-list entry size");

// run_poly_root_finding with direction +1 (poly_utils.cpp:272-460), with the
// work list wl of `cap` entries. Returns kFmOk, kFmError (reported to err) or
// kFmListFull when a split would take the list beyond cap (nothing reported:
// the caller runs it again with a larger list).
template <int ORDER>
__device__ int find_minima(const double *poly, double rf_int, double *roots, int &nroots, bool &fex, uint32_t i,
                           uint32_t j, DevError *err, WLEntry<ORDER> *wl, int cap)
{
    constexpr int OP1 = ORDER + 1;
    nroots = 0;
    fex = false;
    if (fex_excludes<ORDER>(poly, rf_int)) {
        fex = true;
        return kFmOk;
    }
    double iso_lb[ORDER + 1], iso_ub[ORDER + 1];
    int nwl = 0, niso = 0;

    auto add_root = [&](double rt) -> bool {
        if (!is_finite_d(rt)) {
            report_error(err, kErrNpRoot, i, j, rt);
            return false;
        }
        const double der = poly_eval_1<ORDER>(poly, rt);
        if (!is_finite_d(der)) {
            report_error(err, kErrNpRootDer, i, j, rt);
            return false;
        }
        if (sgn_d(der) == 1) {
            if (nroots >= kMaxRoots<ORDER>) {
                report_error(err, kErrNpIsolation, (uint32_t)nwl, (uint32_t)niso, 0.0);
                return false;
            }
            roots[nroots++] = rt;
        }
        return true;
    };

    {
        double cf = 1.0;
#pragma unroll
        for (int k = 0; k < OP1; ++k) {
            wl[0].p[k] = cf * poly[k];
            cf *= rf_int;
        }
        wl[0].lb = 0;
        wl[0].ub = 1;
        nwl = 1;
    }
    do {
        double q[OP1];
        const double lb = wl[nwl - 1].lb, ub = wl[nwl - 1].ub;
#pragma unroll
        for (int k = 0; k < OP1; ++k) q[k] = wl[nwl - 1].p[k];
        --nwl;
        if (q[0] == 0) {
            bool allf = true;
#pragma unroll
            for (int k = 1; k < OP1; ++k) allf = allf && is_finite_d(q[k]);
            if (allf && !add_root(lb * rf_int)) return kFmError;
        }
        const uint32_t nsc = rtscc<ORDER>(q);
        if (nsc == 1) {
            if (niso >= ORDER) {
                report_error(err, kErrNpIsolation, (uint32_t)nwl, (uint32_t)(niso + 1), 0.0);
                return kFmError;
            }
            iso_lb[niso] = lb;
            iso_ub[niso] = ub;
            ++niso;
        } else if (nsc > 1) {
            if (nwl + 2 > cap) return kFmListFull;
            double r[OP1], s[OP1];
            double cf = 1.0;
#pragma unroll
            for (int k = 0; k < OP1; ++k) {
                r[ORDER - k] = cf * q[ORDER - k];
                cf *= 2;
            }
            translate1<ORDER>(r, s);
            const double mid = (lb / 2) + (ub / 2);
            wl[nwl].lb = lb;
            wl[nwl].ub = mid;
#pragma unroll
            for (int k = 0; k < OP1; ++k) wl[nwl].p[k] = r[k];
            ++nwl;
            wl[nwl].lb = mid;
            wl[nwl].ub = ub;
#pragma unroll
            for (int k = 0; k < OP1; ++k) wl[nwl].p[k] = s[k];
            ++nwl;
        }
        if (nwl > 250) {
            report_error(err, kErrNpIsolation, (uint32_t)nwl, (uint32_t)niso, 0.0);
            return kFmError;
        }
    } while (nwl != 0);

    if (niso) {
        double q[OP1];
        double cf = 1.0;
#pragma unroll
        for (int k = 0; k < OP1; ++k) {
            q[k] = cf * poly[k];
            cf *= rf_int;
        }
        const T748<ORDER> solver{q};
        for (int k = 0; k < niso; ++k) {
            double lb = iso_lb[k], ub = iso_ub[k];
            if (is_finite_d(lb) && is_finite_d(ub) && ub > lb) ub = next_toward_d(ub, lb);
            double ra = 0, rb = 0;
            unsigned it = 0;
            const int e = solver.solve(lb, ub, 100, ra, rb, it);
            if (e != 0) {
                report_error(err, kErrNpErrno, i, j, (double)e);
                return kFmError;
            }
            if (it >= 100) {
                report_error(err, kErrNpIters, i, j, 0.0);
                return kFmError;
            }
            if (!add_root(((ra / 2) + (rb / 2)) * rf_int)) return kFmError;
        }
    }
    return kFmOk;
}

// Bjorck-Pereyra interpolation in place: f -> monomial coefficients
// (poly_utils.cpp:467-535).
template <int ORDER>
__device__ inline void vm_interp(const double *alpha, double *c)
{
#pragma unroll
    for (int k = 0; k < ORDER; ++k) {
#pragma unroll
        for (int j = ORDER; j > k; --j) c[j] = (c[j] - c[j - 1]) / (alpha[j] - alpha[j - k - 1]);
    }
#pragma unroll
    for (int idx = 0; idx < ORDER; ++idx) {
        const int k = ORDER - idx - 1;
#pragma unroll
        for (int j = k; j < ORDER; ++j) c[j] = c[j] - alpha[k] * c[j + 1];
    }
}

template <int ORDER>
__device__ inline void emit_conj(const NarrowArgs &a, uint32_t gstep, uint32_t i, uint32_t j, double lb_rf, double t,
                                 const double *Pi, double di, const double *Pj, double dj)
{
    constexpr int OP1 = ORDER + 1;
    mzh_conj rec;
    rec.i = i;
    rec.j = j;
    double tp[OP1];
#pragma unroll 1
    for (int ch = 0; ch < 3; ++ch) {
        translate7<ORDER>(Pi + ch, di, tp);
        rec.ri[ch] = horner<ORDER>(tp, t);
        translate7<ORDER>(Pi + 3 + ch, di, tp);
        rec.vi[ch] = horner<ORDER>(tp, t);
        translate7<ORDER>(Pj + ch, dj, tp);
        rec.rj[ch] = horner<ORDER>(tp, t);
        translate7<ORDER>(Pj + 3 + ch, dj, tp);
        rec.vj[ch] = horner<ORDER>(tp, t);
    }
    const double dx = rec.ri[0] - rec.rj[0], dy = rec.ri[1] - rec.rj[1], dz = rec.ri[2] - rec.rj[2];
    rec.dca = sqrt((dx * dx) + (dy * dy) + (dz * dz));
    rec.tca = lb_rf + t;
    const unsigned long long slot = atomicAdd(a.conj_count, 1ull);
    if (slot < a.conj_cap) {
        a.conj[slot] = rec;
        a.conj_step[slot] = gstep;
    }
}

// A step-pair whose Cargo-Shisha enclosure of dist^2 reaches below thresh^2:
// everything the root finding and the conjunction records need.
template <int ORDER>
struct NpCand {
    double d2[ORDER + 1];  // dist^2 interpolant on [0, rf_int]
    double lb_rf, rf_int, di, dj;
    uint32_t c, i, j, flags;  // flags: bit 0 trajectory end reached, bit 1 trajectory begin
    uint64_t ki, kj;          // trajectory-step indices of i and j
    uint32_t p, pad;          // problem
};

constexpr uint32_t kCandEnd = 1u, kCandBegin = 2u;

// p(x) -> p(x + a) for one channel stored contiguously (the cf4 layout).
template <int ORDER>
__device__ inline void translate1s(const double *__restrict__ src, double a, double *out)
{
    constexpr int OP1 = ORDER + 1;
    double c[OP1];
#pragma unroll
    for (int k = 0; k < OP1; ++k) c[k] = src[k];
    if (a == 0.0) {
#pragma unroll
        for (int k = 0; k < OP1; ++k) out[k] = c[k];
        return;
    }
    double ap[OP1];
    halving_pows_d<ORDER>(a, ap);
#pragma unroll
    for (int i = 0; i < OP1; ++i) {
        double s = (c[i] * c_binom_np.v[i][0]) * ap[0];
#pragma unroll
        for (int k = i + 1; k < OP1; ++k) s = s + (c[k] * c_binom_np.v[k][k - i]) * ap[k - i];
        out[i] = s;
    }
}

// Mixed-precision pre-screen of a broad-phase pair (i, j) over its
// conjunction step, from the two objects' models (k_aabbs): true when no
// conjunction of the FP64 narrow phase can come from any of the pair's step
// pairs. Mixed mode owes only the FP64 conjunction list, and every conjunction
// (a minimum, or a boundary value, of a step pair's dist^2 interpolant p
// below thresh^2) needs p < thresh^2 somewhere in the step. Over the step,
// with s = t - tm in [-w, w], r_i(t) - r_j(t) is within R_i + R_j of
// d0 + v s + a s^2 / 2 (the differences of the models' position, velocity and
// acceleration), hence within R_i + R_j + |a| w^2 / 2 of the segment d0 + v s,
// whose least norm is |d0 x v| / |v| (a bound for the whole line), or an end
// point's when the line's closest point is clearly outside; and |p - F| <= E
// for the true dist^2 F (the problem's bound, Prob::np_e), plus a relative
// 1e-9 of the magnitudes for the FP64 path's rounding. Non-finite values
// never reject.
__device__ inline bool np_pair_far(const float4 *mi, const float4 *mj, double w, double thresh, double E)
{
    // {p0.x, p0.y, p0.z, v0.x}, {v0.y, v0.z, a0.x | a0.y << 16, a0.z | R << 16} (a0 and R bfloat16)
    const float4 ai = mi[0], bi = mi[1], aj = mj[0], bj = mj[1];
    const uint32_t zi = __float_as_uint(bi.z), wi = __float_as_uint(bi.w);
    const uint32_t zj = __float_as_uint(bj.z), wj = __float_as_uint(bj.w);
    const double R = (double)bf16_f(wi >> 16) + (double)bf16_f(wj >> 16);
    if (!(R < 1e300) || !(E < 1e300)) return false;
    const double d0[3] = {(double)ai.x - (double)aj.x, (double)ai.y - (double)aj.y, (double)ai.z - (double)aj.z};
    const double v[3] = {(double)ai.w - (double)aj.w, (double)bi.x - (double)bj.x, (double)bi.y - (double)bj.y};
    const double ax = (double)bf16_f(zi & 0xffffu) - (double)bf16_f(zj & 0xffffu);
    const double ay = (double)bf16_f(zi >> 16) - (double)bf16_f(zj >> 16);
    const double az = (double)bf16_f(wi & 0xffffu) - (double)bf16_f(wj & 0xffffu);
    const double an = sqrt(ax * ax + ay * ay + az * az);
    const double dd = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2];
    const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double dv = d0[0] * v[0] + d0[1] * v[1] + d0[2] * v[2];
    const double cx = d0[1] * v[2] - d0[2] * v[1];
    const double cy = d0[2] * v[0] - d0[0] * v[2];
    const double cz = d0[0] * v[1] - d0[1] * v[0];
    const double dn = sqrt(dd), vn = sqrt(vv);
    double m;
    const double s0 = vv > 0.0 ? -dv / vv : 0.0;
    const double slack = 1e-9 * (w + fabs(s0));
    if (vv > 0.0 && s0 < -w - slack) {
        const double x = d0[0] - v[0] * w, y = d0[1] - v[1] * w, z = d0[2] - v[2] * w;
        m = sqrt(x * x + y * y + z * z);
    } else if (vv > 0.0 && s0 > w + slack) {
        const double x = d0[0] + v[0] * w, y = d0[1] + v[1] * w, z = d0[2] + v[2] * w;
        m = sqrt(x * x + y * y + z * z);
    } else if (vv > 0.0) {
        m = sqrt(cx * cx + cy * cy + cz * cz) / vn;
    } else {
        m = dn;
    }
    const double curv = 0.5 * an * w * w;
    const double mag = dn + vn * w + curv + R;
    const double m_lb = m - curv - R - 1e-12 * mag;
    const double lim = thresh * thresh + E + 1e-9 * mag * mag + 1e-9;
    return m_lb > 0.0 && m_lb * m_lb * (1.0 - 1e-12) > lim;
}

// Narrow phase, part 1a: one lane per broad-phase pair. Merge-walk of the
// trajectory steps of i and j over the conjunction step
// (narrow_phase.cpp:357-420); every temporally overlapping pair of steps is
// appended to the step-pair list (wave-aggregated atomics). The step pairs of
// a mixed-precision problem's pair that np_pair_far rules out are counted as
// candidates rejected by the distance check and not appended.
__global__ __launch_bounds__(256) void k_np_enum(NarrowArgs a)
{
    const uint64_t g0 = a.pair_pre[a.p_first], g1 = a.pair_pre[a.p_end];
    const uint64_t g = g0 + a.pair_begin + (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t g_end = g0 + a.pair_end < g1 ? g0 + a.pair_end : g1;
    uint64_t kbi = 0, kbj = 0, kei = 0, kej = 0;
    uint32_t i = 0, j = 0, c = 0, p = 0;
    const double *ti = nullptr, *tj = nullptr;
    double cb = 0, ce = 0;
    bool far = false;
    const bool live = g < g_end;
    if (live) {
        p = find_prob64(a.pair_pre, a.np, g);
        const Prob &q = a.P[p];
        const uint64_t l = g - a.pair_pre[p];
        c = q.pair_c[l];  // the pair's conjunction step, recorded when the pair was placed
        i = q.pairs[l].i;
        j = q.pairs[l].j;
        cb = q.cd_begin[c];
        ce = q.cd_end[c];
        ti = q.pj.time + q.pj.time_off[i];
        tj = q.pj.time + q.pj.time_off[j];
        // overlapping trajectory steps of both objects, as found by the AABB stage
        const uint2 ri = q.kbke[(size_t)c * q.N + i], rj = q.kbke[(size_t)c * q.N + j];
        kbi = ri.x;
        kei = ri.y;
        kbj = rj.x;
        kej = rj.y;
        if (q.mode == MZH_MODE_MIXED) {
            const float4 *mi = q.nprec + ((size_t)c * q.N + i) * 2u, *mj = q.nprec + ((size_t)c * q.N + j) * 2u;
            const double tm = cb + 0.5 * (ce - cb);
            far = np_pair_far(mi, mj, fmax(tm - cb, ce - tm), q.thresh, q.np_e);
        }
    }
    // the merge-walk visits step pairs in order; the first kEnumBuf
    // overlapping ones wait in registers while the block's offsets are
    // found, the rare rest is found by walking again
    auto walk = [&](auto &&emit) {
        for (uint64_t ki = kbi, kj = kbj; ki < kei && kj < kej;) {
            const double tsi = ti[ki], tei = ti[ki + 1], tsj = tj[kj], tej = tj[kj + 1];
            const double lb_i = cb > tsi ? cb : tsi, ub_i = ce < tei ? ce : tei;
            const double lb_j = cb > tsj ? cb : tsj, ub_j = ce < tej ? ce : tej;
            if (ub_i > lb_j && lb_i < ub_j) emit(ki, kj);
            const bool inc_i = tei <= tej, inc_j = tej <= tei;
            ki += inc_i ? 1 : 0;
            kj += inc_j ? 1 : 0;
        }
    };
    constexpr uint32_t kEnumBuf = 4;
    uint32_t bki[kEnumBuf] = {}, bkj[kEnumBuf] = {};
    uint32_t cnt = 0;
    const uint64_t ni = kei - kbi, nj = kej - kbj;
    if (far && (ni == 1u || nj == 1u)) {
        // a pair the pre-screen rules out is only counted: both objects'
        // pieces cover the whole step (their models exist), so every piece
        // of one overlaps the other's single piece
        cnt = (uint32_t)(ni > nj ? ni : nj);
    } else if (far && ni == 2u && nj == 2u) {
        // three overlapping step pairs, or two when the inner boundaries meet
        cnt = ti[kbi + 1u] == tj[kbj + 1u] ? 2u : 3u;
    } else {
        walk([&](uint64_t ki, uint64_t kj) {
#pragma unroll
            for (uint32_t m = 0; m < kEnumBuf; ++m)
                if (m == cnt) {
                    bki[m] = (uint32_t)ki;
                    bkj[m] = (uint32_t)kj;
                }
            ++cnt;
        });
    }
    {
        // candidates, and those the pre-screen rejects
        const unsigned long long v[2] = {cnt, far ? cnt : 0u};
        add_prob_report<2>(a.P, p, live, 0, v);
    }
    if (far) cnt = 0;
    __shared__ uint32_t s_w[17];
    __shared__ unsigned long long s_base;
    uint32_t total;
    const uint32_t ex = block_excl_scan_u32(cnt, s_w, total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(a.steps_count, (unsigned long long)total) : 0ull;
    __syncthreads();
    const uint64_t slot = s_base + ex;
    StepPair *out = static_cast<StepPair *>(a.steps);
#pragma unroll
    for (uint32_t m = 0; m < kEnumBuf; ++m)
        if (m < cnt && slot + m < a.steps_cap) out[slot + m] = StepPair{p, i, j, c, bki[m], bkj[m]};
    if (cnt > kEnumBuf) {
        uint32_t m = 0;
        walk([&](uint64_t ki, uint64_t kj) {
            if (m >= kEnumBuf && slot + m < a.steps_cap) out[slot + m] = StepPair{p, i, j, c, (uint32_t)ki, (uint32_t)kj};
            ++m;
        });
    }
}

// Narrow phase, part 1b: one lane per overlapping step-pair. Root-finding
// interval and translations (narrow_phase.cpp:404-488), dist^2
// interpolation (:95-152) and the Cargo-Shisha screen (:502-513); survivors
// go to the candidate queue.
template <int ORDER>
__global__ __launch_bounds__(256) void k_np_screen(NarrowArgs a)
{
    constexpr int OP1 = ORDER + 1;
    const unsigned long long avail = *a.steps_count;
    const uint64_t nsp = avail < a.steps_cap ? avail : a.steps_cap;
    const uint64_t q = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    unsigned long long n_rej = 0;
    bool keep = false;
    uint32_t flags = 0, i = 0, j = 0, c = 0, p = 0;
    uint64_t ki = 0, kj = 0;
    double lb_rf = 0, rf_int = 0, di = 0, dj = 0;
    double d2[OP1];
    const bool live = q < nsp;
    if (live) {
        const StepPair sp = static_cast<const StepPair *>(a.steps)[q];
        p = sp.p;
        i = sp.i;
        j = sp.j;
        c = sp.c;
        ki = sp.ki;
        kj = sp.kj;
        const Prob &pb = a.P[p];
        const PjView &pj = pb.pj;
        const double cb = pb.cd_begin[c], ce = pb.cd_end[c];
        const double thresh2 = pb.thresh * pb.thresh;
        const uint64_t nsi = pj.nsteps[i], nsj = pj.nsteps[j];
        const double *ti = pj.time + pj.time_off[i];
        const double *tj = pj.time + pj.time_off[j];
        const double tsi = ti[ki], tei = ti[ki + 1], tsj = tj[kj], tej = tj[kj + 1];
        const double lb_i = cb > tsi ? cb : tsi, ub_i = ce < tei ? ce : tei;
        const double lb_j = cb > tsj ? cb : tsj, ub_j = ce < tej ? ce : tej;
        lb_rf = lb_i > lb_j ? lb_i : lb_j;
        const double ub_rf = ub_i < ub_j ? ub_i : ub_j;
        di = lb_rf - tsi;
        dj = lb_rf - tsj;
        rf_int = ub_rf - lb_rf;
        bool ok = true;
        if (!is_finite_d(di) || !is_finite_d(dj) || !is_finite_d(rf_int) || di < 0 || dj < 0 || rf_int < 0) {
            report_error(pb.err, kErrNpInterval, i, j, rf_int);
            ok = false;
        }
        if (ok) {
            // positions from the step-major copy: channel ch of a step at cf4[(s * 4 + ch) * OP1]
            const double *Qi = pj.cf4 + (pj.step0[i] + ki) * 4u * OP1;
            const double *Qj = pj.cf4 + (pj.step0[j] + kj) * 4u * OP1;
            double nodes[OP1];
#pragma unroll
            for (int m = 0; m < OP1; ++m) nodes[OP1 - 1 - m] = a.cheby[m] * rf_int;
#pragma unroll 1
            for (int ch = 0; ch < 3; ++ch) {
                double tp[OP1], vi[OP1];
                translate1s<ORDER>(Qi + ch * OP1, di, tp);
#pragma unroll
                for (int m = 0; m < OP1; ++m) vi[m] = horner<ORDER>(tp, nodes[m]);
                translate1s<ORDER>(Qj + ch * OP1, dj, tp);
#pragma unroll
                for (int m = 0; m < OP1; ++m) {
                    const double dx = vi[m] - horner<ORDER>(tp, nodes[m]);
                    d2[m] = (ch == 0) ? dx * dx : d2[m] + dx * dx;
                }
            }
            vm_interp<ORDER>(nodes, d2);
            double elo, ehi;
            cs_enc0<ORDER>(d2, rf_int, elo, ehi);
            if (!is_finite_d(elo) || !is_finite_d(ehi)) {
                report_error(pb.err, kErrNpDist2, i, j, elo);
            } else if (elo < thresh2) {
                flags = 0x80000000u;
                if ((ki + 1 == nsi && tei == ub_rf) || (kj + 1 == nsj && tej == ub_rf)) flags |= kCandEnd;
                if ((ki == 0 && lb_rf == ti[0]) || (kj == 0 && lb_rf == tj[0])) flags |= kCandBegin;
                keep = true;
            } else {
                ++n_rej;
            }
        }
    }
    add_prob_report<1>(a.P, p, live, 1, &n_rej);
    // block-aggregated append to the candidate queue
    __shared__ uint32_t s_w[17];
    __shared__ unsigned long long s_base;
    uint32_t total;
    const uint32_t ex = block_excl_scan_u32(keep ? 1u : 0u, s_w, total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(a.cand_count, (unsigned long long)total) : 0ull;
    __syncthreads();
    if (keep) {
        const unsigned long long slot = s_base + ex;
        if (slot < a.cand_cap) {
            flags &= 0x7fffffffu;
            NpCand<ORDER> &cd = static_cast<NpCand<ORDER> *>(a.cand)[slot];
#pragma unroll
                    for (int k = 0; k < OP1; ++k) cd.d2[k] = d2[k];
                    cd.lb_rf = lb_rf;
                    cd.rf_int = rf_int;
                    cd.di = di;
                    cd.dj = dj;
                    cd.c = c;
                    cd.i = i;
                    cd.j = j;
                    cd.flags = flags;
                    cd.ki = ki;
                    cd.kj = kj;
                    cd.p = p;
                    cd.pad = 0;
        }
    }
}

// Minima of dist^2 for one queued candidate (narrow_phase.cpp:519-605) and
// its boundary minima (:629-714), with the root-isolation work list wl of cap
// entries. rep accumulates np_report[2..6]. Returns kFmListFull, before
// anything is counted or emitted, when the work list is too small.
// The candidate's problem is cd.p; rep accumulates for the problem rep_p and
// is added to it (and restarted) when the candidate belongs to another one.
template <int ORDER>
__device__ inline int np_root_cand(const NarrowArgs &a, uint64_t q, unsigned long long *rep, uint32_t &rep_p,
                                   WLEntry<ORDER> *wl, int cap)
{
    constexpr int OP1 = ORDER + 1;
    const NpCand<ORDER> &cd = static_cast<const NpCand<ORDER> *>(a.cand)[q];
    const Prob &pb = a.P[cd.p];
    if (cd.p != rep_p) {
        if (rep_p != 0xffffffffu) {
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (rep[k]) atomicAdd(&a.P[rep_p].np_report[report_slot() + 2 + k], rep[k]);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) rep[k] = 0;
        rep_p = cd.p;
    }
    const double thresh2 = pb.thresh * pb.thresh;
    DevError *err = pb.err;
    double d2[OP1];
#pragma unroll
    for (int k = 0; k < OP1; ++k) d2[k] = cd.d2[k];
    const double rf_int = cd.rf_int, lb_rf = cd.lb_rf;
    const uint32_t i = cd.i, j = cd.j, flags = cd.flags;
    const uint32_t gstep = pb.step0 + cd.c;
    const double *Pi = pb.pj.traj + (pb.pj.step0[i] + cd.ki) * (uint64_t)OP1 * 7u;
    const double *Pj = pb.pj.traj + (pb.pj.step0[j] + cd.kj) * (uint64_t)OP1 * 7u;
    double d2p[OP1];
#pragma unroll
    for (int k = 0; k < ORDER; ++k) d2p[k] = (double)(k + 1) * d2[k + 1];
    d2p[ORDER] = 0;
    double roots[kMaxRoots<ORDER>];
    int nr = 0;
    bool fex = false;
    const int fm = find_minima<ORDER>(d2p, rf_int, roots, nr, fex, i, j, err, wl, cap);
    if (fm != kFmOk) return fm;
    ++rep[0];
    rep[1] += fex ? 1 : 0;
    rep[2] += (nr == 0) ? 1 : 0;
    rep[3] += (unsigned long long)nr;
    for (int r = 0; r < nr; ++r) {
        const double cd2 = horner<ORDER>(d2, roots[r]);
        if (!is_finite_d(cd2)) {
            report_error(err, kErrNpConjDist2, i, j, cd2);
            return kFmError;
        }
        if (cd2 < thresh2) {
            emit_conj<ORDER>(a, gstep, i, j, lb_rf, roots[r], Pi, cd.di, Pj, cd.dj);
        } else {
            ++rep[4];
        }
    }
    if (flags & kCandEnd) {
        const double tmc = next_toward_d(rf_int, -1.0);
        const double dv = horner<ORDER>(d2p, tmc);
        if (!is_finite_d(dv)) {
            report_error(err, kErrNpBndDerUb, i, j, dv);
            return kFmError;
        }
        if (dv < 0) {
            const double dd = horner<ORDER>(d2, tmc);
            if (!is_finite_d(dd)) {
                report_error(err, kErrNpBndDist2Ub, i, j, dd);
                return kFmError;
            }
            if (dd < thresh2) emit_conj<ORDER>(a, gstep, i, j, lb_rf, tmc, Pi, cd.di, Pj, cd.dj);
        }
    }
    if (flags & kCandBegin) {
        const double dv = horner<ORDER>(d2p, 0.0);
        if (!is_finite_d(dv)) {
            report_error(err, kErrNpBndDerLb, i, j, dv);
            return kFmError;
        }
        if (dv > 0) {
            const double dd = horner<ORDER>(d2, 0.0);
            if (!is_finite_d(dd)) {
                report_error(err, kErrNpBndDist2Lb, i, j, dd);
                return kFmError;
            }
            if (dd < thresh2) emit_conj<ORDER>(a, gstep, i, j, lb_rf, 0.0, Pi, cd.di, Pj, cd.dj);
        }
    }
    return kFmOk;
}

template <int ORDER>
__device__ inline void np_flush_report(const NarrowArgs &a, const unsigned long long *rep, uint32_t rep_p)
{
    add_prob_report<5>(a.P, rep_p, rep_p != 0xffffffffu, 2, rep);
}

// Narrow phase, part 2: one lane per queued candidate (grid-stride over the
// queue, whose length is read on the device), with a private work list of
// a.wl_fast_cap (<= kWlCap) entries. A candidate that needs a longer list is
// handed to the slow queue.
template <int ORDER>
__global__ __launch_bounds__(256) void k_np_roots(NarrowArgs a)
{
    const unsigned long long avail = *a.cand_count;
    const uint64_t n = avail < a.cand_cap ? avail : a.cand_cap;
    unsigned long long rep[5] = {0, 0, 0, 0, 0};
    uint32_t rep_p = 0xffffffffu;
    WLEntry<ORDER> wl[kWlCap];
    const int cap = a.wl_fast_cap < kWlCap ? a.wl_fast_cap : kWlCap;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t q = (uint64_t)blockIdx.x * 256u + threadIdx.x; q < n; q += stride) {
        if (np_root_cand<ORDER>(a, q, rep, rep_p, wl, cap) == kFmListFull) {
            const unsigned long long s = atomicAdd(a.slow_count, 1ull);
            if (s < a.slow_cap) a.slow[s] = q;
        }
    }
    np_flush_report<ORDER>(a, rep, rep_p);
}

// Narrow phase, part 2b: the candidates of the slow queue, each lane with its
// own work list of kWlSlowCap entries in a.wl_slow (gridDim.x * 256 lists).
template <int ORDER>
__global__ __launch_bounds__(256) void k_np_roots_slow(NarrowArgs a)
{
    const unsigned long long avail = *a.slow_count;
    const uint64_t n = avail < a.slow_cap ? avail : a.slow_cap;
    unsigned long long rep[5] = {0, 0, 0, 0, 0};
    uint32_t rep_p = 0xffffffffu;
    const uint64_t lane = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    WLEntry<ORDER> *wl = static_cast<WLEntry<ORDER> *>(a.wl_slow) + lane * kWlSlowCap;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t s = lane; s < n; s += stride) (void)np_root_cand<ORDER>(a, a.slow[s], rep, rep_p, wl, kWlSlowCap);
    np_flush_report<ORDER>(a, rep, rep_p);
}

// End of a narrow-phase pass: the queue lengths folded into peak, the queue
// counters cleared for the next pass.
__global__ void k_np_pass_end(unsigned long long *steps_count, unsigned long long *cand_count,
                              unsigned long long *slow_count, unsigned long long *peak)
{
    if (threadIdx.x != 0) return;
    peak[0] = peak[0] > *steps_count ? peak[0] : *steps_count;
    peak[1] = peak[1] > *cand_count ? peak[1] : *cand_count;
    peak[2] = peak[2] > *slow_count ? peak[2] : *slow_count;
    *steps_count = 0;
    *cand_count = 0;
    *slow_count = 0;
}

// Before the first pass: the problems' narrow-phase counters, the
// conjunction count, the queue counters and their peaks start at zero.
__global__ void k_np_reset(const Prob *__restrict__ P, uint32_t np, unsigned long long *conj_count,
                           unsigned long long *steps_count, unsigned long long *cand_count,
                           unsigned long long *slow_count, unsigned long long *peak)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < np * kReportWords) P[t / kReportWords].np_report[t % kReportWords] = 0ull;
    if (t == 0) {
        *conj_count = 0;
        *steps_count = 0;
        *cand_count = 0;
        *slow_count = 0;
        peak[0] = peak[1] = peak[2] = 0;
    }
}

// Root finding of free-standing polynomials (run_poly_root_finding,
// poly_utils.cpp:272-460, direction +1): one lane per polynomial, the
// private list first; a polynomial whose list outgrows it is done again by
// k_rootfind_slow with a kWlSlowCap-entry list in global memory.
// Outputs as documented with RootFindArgs.
template <int ORDER>
__device__ inline int rootfind_one(const RootFindArgs &a, uint64_t k, WLEntry<ORDER> *wl, int cap, uint32_t slow)
{
    const double rf_int = a.rf_int[k];
    if (!is_finite_d(rf_int) || !(rf_int > 0)) {
        report_error(a.err, kErrRfInterval, (uint32_t)k, 0u, rf_int);
        a.n_roots[k] = 0xfffffffeu;
        return kFmError;
    }
    double poly[ORDER + 1];
#pragma unroll
    for (int m = 0; m <= ORDER; ++m) poly[m] = a.polys[k * (ORDER + 1) + m];
    double roots[kMaxRoots<ORDER>];
    int nr = 0;
    bool fex = false;
    const int fm = find_minima<ORDER>(poly, rf_int, roots, nr, fex, (uint32_t)k, 0u, a.err, wl, cap);
    if (fm == kFmOk) {
        for (int r = 0; r < kMaxRoots<ORDER>; ++r) a.roots[k * kMaxRoots<ORDER> + r] = r < nr ? roots[r] : 0.0;
        a.n_roots[k] = (uint32_t)nr;
        if (a.flags) a.flags[k] = (fex ? 1u : 0u) | (slow << 1);
    } else if (fm == kFmError) {
        a.n_roots[k] = 0xfffffffeu;
    }
    return fm;
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_rootfind(RootFindArgs a)
{
    const uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (k >= a.n) return;
    WLEntry<ORDER> wl[kWlCap];
    const int cap = a.wl_fast_cap < kWlCap ? a.wl_fast_cap : kWlCap;
    if (rootfind_one<ORDER>(a, k, wl, cap, 0u) == kFmListFull) {
        a.n_roots[k] = 0xffffffffu;
        a.slow[atomicAdd(a.slow_count, 1ull)] = k;
    }
}

// zero the scratch counters of a root-finding launch
__global__ void k_rootfind_clear(unsigned long long *slow_count, DevError *err)
{
    if (threadIdx.x == 0) {
        *slow_count = 0ull;
        err->code = 0;
        err->i = err->j = 0u;
        err->value = 0.0;
    }
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_rootfind_slow(RootFindArgs a)
{
    const uint64_t n = *a.slow_count;
    const uint64_t lane = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    WLEntry<ORDER> *wl = static_cast<WLEntry<ORDER> *>(a.wl_slow) + lane * kWlSlowCap;
    for (uint64_t s = lane; s < n; s += (uint64_t)gridDim.x * 256u)
        (void)rootfind_one<ORDER>(a, a.slow[s], wl, kWlSlowCap, 1u);
}

#ifdef MZH_NARROW_HOST
// ---- conjunction ordering ---------------------------------------------------
// Records are appended in any order; the output is grouped by conjunction
// step and each step sorted by (tca, then every field lexicographically), the
// order of detect_conjunctions_narrow_phase.cpp:751-757. A merge sort over
// compact keys (step, tca, i, j, input index): 1024-key tiles sorted in LDS,
// then passes that merge runs of doubling width, where each key finds its
// place in the sibling run by binary search. O(M log M) whatever the number
// of conjunctions per step.

// strict weak order (tca, then all fields lexicographically)
__device__ inline bool conj_less(const mzh_conj &x, const mzh_conj &y)
{
    if (x.tca != y.tca) return x.tca < y.tca;
    if (x.i != y.i) return x.i < y.i;
    if (x.j != y.j) return x.j < y.j;
    if (x.dca != y.dca) return x.dca < y.dca;
    for (int k = 0; k < 3; ++k)
        if (x.ri[k] != y.ri[k]) return x.ri[k] < y.ri[k];
    for (int k = 0; k < 3; ++k)
        if (x.vi[k] != y.vi[k]) return x.vi[k] < y.vi[k];
    for (int k = 0; k < 3; ++k)
        if (x.rj[k] != y.rj[k]) return x.rj[k] < y.rj[k];
    for (int k = 0; k < 3; ++k)
        if (x.vj[k] != y.vj[k]) return x.vj[k] < y.vj[k];
    return false;
}

constexpr uint32_t kConjPad = 0xffffffffu;  // step of the padding keys of a partial tile
constexpr uint32_t kConjTile = 1024;

// Total order of the keys: (step, tca, i, j), then the remaining fields of
// the records, then the input index. tca >= 0, so its bits order as integers
// (tca + 0.0 turns a -0.0 into +0.0).
__device__ inline bool ck_less(const ConjKey &x, const ConjKey &y, const mzh_conj *__restrict__ rec)
{
    if (x.step != y.step) return x.step < y.step;
    if (x.step == kConjPad) return false;
    if (x.tca != y.tca) return x.tca < y.tca;
    if (x.ij != y.ij) return x.ij < y.ij;
    if (x.idx == y.idx) return false;
    const mzh_conj &a = rec[x.idx], &b = rec[y.idx];
    if (conj_less(a, b)) return true;
    if (conj_less(b, a)) return false;
    return x.idx < y.idx;
}

__device__ inline uint64_t dev_count(uint64_t n, const unsigned long long *n_dev)
{
    if (!n_dev) return n;
    const uint64_t m = *n_dev;
    return m < n ? m : n;
}

__global__ __launch_bounds__(256) void k_conj_keys(const mzh_conj *__restrict__ in, const uint32_t *__restrict__ step,
                                                   uint64_t n, const unsigned long long *n_dev,
                                                   ConjKey *__restrict__ keys)
{
    n = dev_count(n, n_dev);
    const uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const mzh_conj &r = in[k];
    ConjKey ck;
    ck.step = step[k];
    ck.idx = (uint32_t)k;
    ck.tca = (uint64_t)__double_as_longlong(r.tca + 0.0);
    ck.ij = ((uint64_t)r.i << 32) | r.j;
    keys[k] = ck;
}

// one workgroup sorts one tile of kConjTile keys in place (bitonic network in LDS)
__global__ __launch_bounds__(256) void k_conj_tiles(ConjKey *__restrict__ keys, uint64_t n,
                                                    const unsigned long long *n_dev, const mzh_conj *__restrict__ rec)
{
    __shared__ ConjKey s[kConjTile];
    n = dev_count(n, n_dev);
    const uint64_t base = (uint64_t)blockIdx.x * kConjTile;
    if (base >= n) return;
    for (uint32_t t = threadIdx.x; t < kConjTile; t += 256) {
        if (base + t < n) {
            s[t] = keys[base + t];
        } else {
            ConjKey pad;
            pad.step = kConjPad;
            pad.idx = kConjPad;
            pad.tca = ~0ull;
            pad.ij = ~0ull;
            s[t] = pad;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= kConjTile; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < kConjTile / 2; t += 256) {
                const uint32_t a = 2u * t - (t & (j - 1u)), b = a + j;
                const bool up = (a & k) == 0;
                if (ck_less(s[b], s[a], rec) == up) {
                    const ConjKey x = s[a];
                    s[a] = s[b];
                    s[b] = x;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t t = threadIdx.x; t < kConjTile; t += 256)
        if (base + t < n) keys[base + t] = s[t];
}

// merge the sorted runs of width w pairwise: in -> out
__global__ __launch_bounds__(256) void k_conj_merge(const ConjKey *__restrict__ in, ConjKey *__restrict__ out,
                                                    uint64_t n, const unsigned long long *n_dev, uint64_t w,
                                                    const mzh_conj *__restrict__ rec)
{
    n = dev_count(n, n_dev);
    const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (e >= n) return;
    const uint64_t r = e / w, base = (r & ~1ull) * w;
    const uint64_t sb = (r ^ 1ull) * w;
    const ConjKey me = in[e];
    if (sb >= n) {
        out[e] = me;  // no sibling run
        return;
    }
    const uint64_t se = sb + w < n ? sb + w : n;
    // number of keys of the sibling run below me (all keys are distinct)
    uint64_t lo = sb, len = se - sb;
    while (len > 0) {
        const uint64_t h = len / 2;
        if (ck_less(in[lo + h], me, rec)) {
            lo += h + 1;
            len -= h + 1;
        } else {
            len = h;
        }
    }
    out[base + (e - r * w) + (lo - sb)] = me;
}

// Every problem's range of the sorted keys (its batch steps are
// [step0, step0 + C)), counts[2..3], and kFlagConj if conj_out is too small.
// One thread per problem.
__global__ void k_conj_bounds(const Prob *__restrict__ P, uint32_t np, const ConjKey *__restrict__ keys,
                              uint64_t n, const unsigned long long *n_dev)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    n = dev_count(n, n_dev);
    const Prob &q = P[p];
    auto lower = [&](uint32_t st) {
        uint64_t lo = 0, len = n;
        while (len > 0) {
            const uint64_t h = len / 2;
            if (keys[lo + h].step < st) {
                lo += h + 1;
                len -= h + 1;
            } else {
                len = h;
            }
        }
        return lo;
    };
    const uint64_t b = lower(q.step0), e = lower(q.step0 + q.C);
    q.counts[2] = e - b;
    q.counts[3] = b;
    if (e - b > q.conj_out_cap) atomicOr(q.flags, kFlagConj);
}

// The sorted records into their problems' outputs (thread per record).
__global__ __launch_bounds__(256) void k_conj_gather(const Prob *__restrict__ P, uint32_t np,
                                                     const uint32_t *__restrict__ step_pre,
                                                     const ConjKey *__restrict__ keys, const mzh_conj *__restrict__ in,
                                                     uint64_t n, const unsigned long long *n_dev)
{
    n = dev_count(n, n_dev);
    const uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const ConjKey key = keys[k];
    uint32_t lo = 0, m = np;
    while (m > 1) {
        const uint32_t h = m / 2;
        if (step_pre[lo + h] <= key.step) {
            lo += h;
            m -= h;
        } else {
            m = h;
        }
    }
    const Prob &q = P[lo];
    const uint64_t at = k - q.counts[3];
    if (at < q.conj_out_cap) q.conj_out[at] = in[key.idx];
}
#endif // MZH_NARROW_HOST

} // namespace

template <int ORDER>
hipError_t launch_narrow_order(const NarrowArgs &a, hipStream_t stream)
{
    const uint64_t blocks = (a.pair_end - a.pair_begin + 255) / 256;
    if (blocks == 0) return hipSuccess;
    // the pass's pairs (fewer when the problems placed fewer: the kernel reads
    // the batch's count)
    hipLaunchKernelGGL(k_np_enum, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // one lane per step-pair; the list length is read on the device
    const uint64_t sblocks = (a.steps_cap + 255) / 256;
    hipLaunchKernelGGL(k_np_screen<ORDER>, dim3((unsigned)sblocks), dim3(256), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the queue length is only known on the device: a fixed grid strides over
    // it. Two workgroups per CU (512 on the MI355X) measured best: a lane then
    // takes several candidates, which evens out their uneven root-finding work,
    // and fewer workgroups add their report counters atomically (4096: 4.48 ms
    // for the narrow phase of the 30 000-object day, 1024: 4.10, 512: 4.06, 256: 4.19)
    const uint64_t qblocks = std::min<uint64_t>((a.cand_cap + 255) / 256, a.roots_wgs ? a.roots_wgs : 512);
    hipLaunchKernelGGL(k_np_roots<ORDER>, dim3((unsigned)qblocks), dim3(256), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the candidates whose root isolation outgrew the private work list
    hipLaunchKernelGGL(k_np_roots_slow<ORDER>, dim3(kNarrowSlowLanes / 256), dim3(256), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_np_pass_end, dim3(1), dim3(64), 0, stream, a.steps_count, a.cand_count, a.slow_count, a.peak);
    return hipGetLastError();
}

template <int ORDER>
hipError_t launch_rootfind_order(const RootFindArgs &a, hipStream_t stream)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rootfind_clear, dim3(1), dim3(64), 0, stream, a.slow_count, a.err);
    hipLaunchKernelGGL(k_rootfind<ORDER>, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rootfind_slow<ORDER>, dim3(kNarrowSlowLanes / 256), dim3(256), 0, stream, a);
    return hipGetLastError();
}

// Explicit instantiations for the orders compiled into this object (the build
// splits the order range over several objects so that they compile in parallel).
#define MZH_INST(N)                                                                                                    \
    template hipError_t launch_narrow_order<N>(const NarrowArgs &, hipStream_t);                                      \
    template hipError_t launch_rootfind_order<N>(const RootFindArgs &, hipStream_t);
#if MZH_NARROW_LO <= 2 && 2 <= MZH_NARROW_HI
MZH_INST(2)
#endif
#if MZH_NARROW_LO <= 3 && 3 <= MZH_NARROW_HI
MZH_INST(3)
#endif
#if MZH_NARROW_LO <= 4 && 4 <= MZH_NARROW_HI
MZH_INST(4)
#endif
#if MZH_NARROW_LO <= 5 && 5 <= MZH_NARROW_HI
MZH_INST(5)
#endif
#if MZH_NARROW_LO <= 6 && 6 <= MZH_NARROW_HI
MZH_INST(6)
#endif
#if MZH_NARROW_LO <= 7 && 7 <= MZH_NARROW_HI
MZH_INST(7)
#endif
#if MZH_NARROW_LO <= 8 && 8 <= MZH_NARROW_HI
MZH_INST(8)
#endif
#if MZH_NARROW_LO <= 9 && 9 <= MZH_NARROW_HI
MZH_INST(9)
#endif
#if MZH_NARROW_LO <= 10 && 10 <= MZH_NARROW_HI
MZH_INST(10)
#endif
#if MZH_NARROW_LO <= 11 && 11 <= MZH_NARROW_HI
MZH_INST(11)
#endif
#if MZH_NARROW_LO <= 12 && 12 <= MZH_NARROW_HI
MZH_INST(12)
#endif
#if MZH_NARROW_LO <= 13 && 13 <= MZH_NARROW_HI
MZH_INST(13)
#endif
#if MZH_NARROW_LO <= 14 && 14 <= MZH_NARROW_HI
MZH_INST(14)
#endif
#if MZH_NARROW_LO <= 15 && 15 <= MZH_NARROW_HI
MZH_INST(15)
#endif
#if MZH_NARROW_LO <= 16 && 16 <= MZH_NARROW_HI
MZH_INST(16)
#endif
#if MZH_NARROW_LO <= 17 && 17 <= MZH_NARROW_HI
MZH_INST(17)
#endif
#if MZH_NARROW_LO <= 18 && 18 <= MZH_NARROW_HI
MZH_INST(18)
#endif
#if MZH_NARROW_LO <= 19 && 19 <= MZH_NARROW_HI
MZH_INST(19)
#endif
#if MZH_NARROW_LO <= 20 && 20 <= MZH_NARROW_HI
MZH_INST(20)
#endif
#undef MZH_INST

#ifdef MZH_NARROW_HOST
size_t narrow_cand_bytes(uint32_t order)
{
    switch (order) {
        case 2:
            return sizeof(NpCand<2>);
        case 3:
            return sizeof(NpCand<3>);
        case 4:
            return sizeof(NpCand<4>);
        case 5:
            return sizeof(NpCand<5>);
        case 6:
            return sizeof(NpCand<6>);
        case 7:
            return sizeof(NpCand<7>);
        case 8:
            return sizeof(NpCand<8>);
        case 9:
            return sizeof(NpCand<9>);
        case 10:
            return sizeof(NpCand<10>);
        case 11:
            return sizeof(NpCand<11>);
        case 12:
            return sizeof(NpCand<12>);
        case 13:
            return sizeof(NpCand<13>);
        case 14:
            return sizeof(NpCand<14>);
        case 15:
            return sizeof(NpCand<15>);
        case 16:
            return sizeof(NpCand<16>);
        case 17:
            return sizeof(NpCand<17>);
        case 18:
            return sizeof(NpCand<18>);
        case 19:
            return sizeof(NpCand<19>);
        case 20:
            return sizeof(NpCand<20>);
        default:
            return 0;
    }
}

template <int ORDER>
hipError_t dispatch_rootfind(uint32_t order, const RootFindArgs &a, hipStream_t stream)
{
    if (order == ORDER) return launch_rootfind_order<ORDER>(a, stream);
    if constexpr (ORDER < kMaxOrder) {
        return dispatch_rootfind<ORDER + 1>(order, a, stream);
    } else {
        return hipErrorInvalidValue;
    }
}

hipError_t launch_rootfind(uint32_t order, const RootFindArgs &a, hipStream_t stream)
{
    return dispatch_rootfind<kMinOrder>(order, a, stream);
}

hipError_t launch_narrow(const NarrowArgs &a, hipStream_t stream)
{
    switch (a.order) {
        case 2:
            return launch_narrow_order<2>(a, stream);
        case 3:
            return launch_narrow_order<3>(a, stream);
        case 4:
            return launch_narrow_order<4>(a, stream);
        case 5:
            return launch_narrow_order<5>(a, stream);
        case 6:
            return launch_narrow_order<6>(a, stream);
        case 7:
            return launch_narrow_order<7>(a, stream);
        case 8:
            return launch_narrow_order<8>(a, stream);
        case 9:
            return launch_narrow_order<9>(a, stream);
        case 10:
            return launch_narrow_order<10>(a, stream);
        case 11:
            return launch_narrow_order<11>(a, stream);
        case 12:
            return launch_narrow_order<12>(a, stream);
        case 13:
            return launch_narrow_order<13>(a, stream);
        case 14:
            return launch_narrow_order<14>(a, stream);
        case 15:
            return launch_narrow_order<15>(a, stream);
        case 16:
            return launch_narrow_order<16>(a, stream);
        case 17:
            return launch_narrow_order<17>(a, stream);
        case 18:
            return launch_narrow_order<18>(a, stream);
        case 19:
            return launch_narrow_order<19>(a, stream);
        case 20:
            return launch_narrow_order<20>(a, stream);
        default:
            return hipErrorInvalidValue;
    }
}

// Conjunction ordering used by the orchestrator (see k_conj_keys ..
// k_conj_gather): sized by the capacity n, bounded on the device by n_dev.
hipError_t launch_conj_sort(const Prob *P, uint32_t np, const uint32_t *step_pre, const mzh_conj *in,
                            const uint32_t *in_step, uint64_t n, const unsigned long long *n_dev, ConjKey *keys0,
                            ConjKey *keys1, hipStream_t stream)
{
    if (n >= 0xffffffffull) return hipErrorInvalidValue;
    if (n > 0) {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_conj_keys, dim3(blocks), dim3(256), 0, stream, in, in_step, n, n_dev, keys0);
        hipLaunchKernelGGL(k_conj_tiles, dim3((unsigned)((n + kConjTile - 1) / kConjTile)), dim3(256), 0, stream,
                           keys0, n, n_dev, in);
    }
    ConjKey *a = keys0, *b = keys1;
    for (uint64_t w = kConjTile; w < n; w *= 2) {
        hipLaunchKernelGGL(k_conj_merge, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, b, n, n_dev, w,
                           in);
        ConjKey *t = a;
        a = b;
        b = t;
    }
    hipLaunchKernelGGL(k_conj_bounds, dim3((np + 255) / 256), dim3(256), 0, stream, P, np, a, n, n_dev);
    if (n > 0)
        hipLaunchKernelGGL(k_conj_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, P, np, step_pre, a,
                           in, n, n_dev);
    return hipGetLastError();
}

hipError_t launch_narrow_reset(const NarrowArgs &a, hipStream_t stream)
{
    hipLaunchKernelGGL(k_np_reset, dim3((a.np * kReportWords + 255) / 256), dim3(256), 0, stream, a.P, a.np, a.conj_count,
                       a.steps_count, a.cand_count, a.slow_count, a.peak);
    return hipGetLastError();
}
#endif // MZH_NARROW_HOST

} // namespace mzh

// SGP4 on the GPU: propagation of general perturbations element sets (GPEs)
// and the construction of polyjectories from them, the upstream producer of
// the screening path (mizuba/make_sgp4_polyjectory.cpp:478-913).
//
// The propagator is the near-earth SGP4 model of Spacetrack Report #3 (Hoots &
// Roehrich, 1980) with the corrections of Vallado, Crawford, Hujsak & Kelso,
// "Revisiting Spacetrack Report #3" (AIAA 2006-6753), WGS-72 constants, in
// FP64. Deep-space element sets (orbital period of 225 minutes or more) are
// recognised and reported, not propagated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace mzh
{

namespace
{

// WGS-72
constexpr double kMuE = 398600.8;         // km^3 / s^2
constexpr double kRe = 6378.135;          // km
constexpr double kJ2 = 0.001082616;
constexpr double kJ3 = -0.00000253881;
constexpr double kJ4 = -0.00000165597;
constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ inline double xke() { return 60.0 / sqrt(kRe * kRe * kRe / kMuE); }

// Initialise the near-earth model from the mean elements of a GPE (the
// initialisation of the Report #3 / 2006 formulation: recovery of the
// original mean motion and semimajor axis, atmospheric drag and gravity
// coefficients). Returns 0, or the model's error code.
__device__ int sgp4_init(const mzh_gpe &g, Sgp4Sat &s)
{
    const double x2o3 = 2.0 / 3.0;
    const double ke = xke();
    const double j3oj2 = kJ3 / kJ2;
    s.bstar = g.bstar;
    s.ecco = g.e0;
    s.argpo = g.omega0;
    s.inclo = g.i0;
    s.mo = g.m0;
    s.nodeo = g.node0;
    s.error = 0;
    s.deep = 0;
    s.isimp = 0;
    const double no_kozai = g.n0;
    // recovery of the original (un-Kozai'd) mean motion
    const double eccsq = s.ecco * s.ecco;
    const double omeosq = 1.0 - eccsq;
    const double rteosq = sqrt(omeosq);
    const double cosio = cos(s.inclo);
    const double cosio2 = cosio * cosio;
    const double ak = pow(ke / no_kozai, x2o3);
    const double d1 = 0.75 * kJ2 * (3.0 * cosio2 - 1.0) / (rteosq * omeosq);
    double del = d1 / (ak * ak);
    const double adel = ak * (1.0 - del * del - del * (1.0 / 3.0 + 134.0 * del * del / 81.0));
    del = d1 / (adel * adel);
    s.no_unkozai = no_kozai / (1.0 + del);
    const double ao = pow(ke / s.no_unkozai, x2o3);
    const double sinio = sin(s.inclo);
    const double po = ao * omeosq;
    const double con42 = 1.0 - 5.0 * cosio2;
    s.con41 = -con42 - cosio2 - cosio2;
    const double posq = po * po;
    const double rp = ao * (1.0 - s.ecco);
    if (omeosq < 0.0 && s.no_unkozai < 0.0) return 0;
    // deep space: period of 225 minutes or more
    if (kTwoPi / s.no_unkozai >= 225.0) {
        s.deep = 1;
        return 0;
    }
    if (rp < 220.0 / kRe + 1.0) s.isimp = 1;
    double sfour = 78.0 / kRe + 1.0;
    const double qzms2ttemp = (120.0 - 78.0) / kRe;
    double qzms24 = qzms2ttemp * qzms2ttemp * qzms2ttemp * qzms2ttemp;
    const double perige = (rp - 1.0) * kRe;
    if (perige < 156.0) {
        sfour = perige - 78.0;
        if (perige < 98.0) sfour = 20.0;
        const double qzms24temp = (120.0 - sfour) / kRe;
        qzms24 = qzms24temp * qzms24temp * qzms24temp * qzms24temp;
        sfour = sfour / kRe + 1.0;
    }
    const double pinvsq = 1.0 / posq;
    const double tsi = 1.0 / (ao - sfour);
    s.eta = ao * s.ecco * tsi;
    const double etasq = s.eta * s.eta;
    const double eeta = s.ecco * s.eta;
    const double psisq = fabs(1.0 - etasq);
    const double coef = qzms24 * pow(tsi, 4.0);
    const double coef1 = coef / pow(psisq, 3.5);
    const double cc2 = coef1 * s.no_unkozai
                       * (ao * (1.0 + 1.5 * etasq + eeta * (4.0 + etasq))
                          + 0.375 * kJ2 * tsi / psisq * s.con41 * (8.0 + 3.0 * etasq * (8.0 + etasq)));
    s.cc1 = s.bstar * cc2;
    double cc3 = 0.0;
    if (s.ecco > 1.0e-4) cc3 = -2.0 * coef * tsi * j3oj2 * s.no_unkozai * sinio / s.ecco;
    s.x1mth2 = 1.0 - cosio2;
    s.cc4 = 2.0 * s.no_unkozai * coef1 * ao * omeosq
            * (s.eta * (2.0 + 0.5 * etasq) + s.ecco * (0.5 + 2.0 * etasq)
               - kJ2 * tsi / (ao * psisq)
                     * (-3.0 * s.con41 * (1.0 - 2.0 * eeta + etasq * (1.5 - 0.5 * eeta))
                        + 0.75 * s.x1mth2 * (2.0 * etasq - eeta * (1.0 + etasq)) * cos(2.0 * s.argpo)));
    s.cc5 = 2.0 * coef1 * ao * omeosq * (1.0 + 2.75 * (etasq + eeta) + eeta * etasq);
    const double cosio4 = cosio2 * cosio2;
    const double temp1 = 1.5 * kJ2 * pinvsq * s.no_unkozai;
    const double temp2 = 0.5 * temp1 * kJ2 * pinvsq;
    const double temp3 = -0.46875 * kJ4 * pinvsq * pinvsq * s.no_unkozai;
    s.mdot = s.no_unkozai + 0.5 * temp1 * rteosq * s.con41
             + 0.0625 * temp2 * rteosq * (13.0 - 78.0 * cosio2 + 137.0 * cosio4);
    s.argpdot = -0.5 * temp1 * con42 + 0.0625 * temp2 * (7.0 - 114.0 * cosio2 + 395.0 * cosio4)
                + temp3 * (3.0 - 36.0 * cosio2 + 49.0 * cosio4);
    const double xhdot1 = -temp1 * cosio;
    s.nodedot = xhdot1 + (0.5 * temp2 * (4.0 - 19.0 * cosio2) + 2.0 * temp3 * (3.0 - 7.0 * cosio2)) * cosio;
    s.omgcof = s.bstar * cc3 * cos(s.argpo);
    s.xmcof = 0.0;
    if (s.ecco > 1.0e-4) s.xmcof = -x2o3 * coef * s.bstar / eeta;
    s.nodecf = 3.5 * omeosq * xhdot1 * s.cc1;
    s.t2cof = 1.5 * s.cc1;
    // the 1 + cos(i) divisor is kept away from zero for retrograde equatorial orbits
    if (fabs(cosio + 1.0) > 1.5e-12)
        s.xlcof = -0.25 * j3oj2 * sinio * (3.0 + 5.0 * cosio) / (1.0 + cosio);
    else
        s.xlcof = -0.25 * j3oj2 * sinio * (3.0 + 5.0 * cosio) / 1.5e-12;
    s.aycof = -0.5 * j3oj2 * sinio;
    const double delmotemp = 1.0 + s.eta * cos(s.mo);
    s.delmo = delmotemp * delmotemp * delmotemp;
    s.sinmao = sin(s.mo);
    s.x7thm1 = 7.0 * cosio2 - 1.0;
    if (s.isimp != 1) {
        const double cc1sq = s.cc1 * s.cc1;
        s.d2 = 4.0 * ao * tsi * cc1sq;
        const double temp = s.d2 * tsi * s.cc1 / 3.0;
        s.d3 = (17.0 * ao + sfour) * temp;
        s.d4 = 0.5 * temp * ao * tsi * (221.0 * ao + 31.0 * sfour) * s.cc1;
        s.t3cof = s.d2 + 2.0 * cc1sq;
        s.t4cof = 0.25 * (3.0 * s.d3 + s.cc1 * (12.0 * s.d2 + 10.0 * cc1sq));
        s.t5cof = 0.2 * (3.0 * s.d4 + 12.0 * s.cc1 * s.d3 + 6.0 * s.d2 * s.d2 + 15.0 * cc1sq * (2.0 * s.d2 + cc1sq));
    }
    return 0;
}

} // namespace

// Position (km) and velocity (km/s) in TEME at tsince minutes from the epoch.
// Returns 0, or the model's error code: 1 mean eccentricity out of range, 2
// negative mean motion, 4 negative semi-latus rectum, 6 decayed (radius below
// one Earth radius).
// SMOOTH = false follows the published code exactly: its solution of Kepler's
// equation stops once a Newton correction is below 1e-12 rad and keeps the
// sine and cosine from before that correction, so the state jumps by up to
// about 1e-12 rad (some 1e-8 km) where the iteration count changes. SMOOTH =
// true takes them at the converged anomaly instead: the model without those
// jumps, for interpolation.
template <bool SMOOTH>
__device__ int sgp4_prop(const Sgp4Sat &s, double t, double r[3], double v[3])
{
    const double x2o3 = 2.0 / 3.0;
    const double ke = xke();
    const double vkmpersec = kRe * ke / 60.0;
    // secular gravity and atmospheric drag
    const double xmdf = s.mo + s.mdot * t;
    const double argpdf = s.argpo + s.argpdot * t;
    const double nodedf = s.nodeo + s.nodedot * t;
    double argpm = argpdf;
    double mm = xmdf;
    const double t2 = t * t;
    double nodem = nodedf + s.nodecf * t2;
    double tempa = 1.0 - s.cc1 * t;
    double tempe = s.bstar * s.cc4 * t;
    double templ = s.t2cof * t2;
    if (s.isimp != 1) {
        const double delomg = s.omgcof * t;
        const double delmtemp = 1.0 + s.eta * cos(xmdf);
        const double delm = s.xmcof * (delmtemp * delmtemp * delmtemp - s.delmo);
        const double temp = delomg + delm;
        mm = xmdf + temp;
        argpm = argpdf - temp;
        const double t3 = t2 * t;
        const double t4 = t3 * t;
        tempa = tempa - s.d2 * t2 - s.d3 * t3 - s.d4 * t4;
        tempe = tempe + s.bstar * s.cc5 * (sin(mm) - s.sinmao);
        templ = templ + s.t3cof * t3 + t4 * (s.t4cof + t * s.t5cof);
    }
    double nm = s.no_unkozai;
    double em = s.ecco;
    const double inclm = s.inclo;
    if (nm <= 0.0) return 2;
    const double am = pow(ke / nm, x2o3) * tempa * tempa;
    nm = ke / pow(am, 1.5);
    em = em - tempe;
    if (em >= 1.0 || em < -0.001) return 1;
    if (em < 1.0e-6) em = 1.0e-6;
    mm = mm + s.no_unkozai * templ;
    double xlm = mm + argpm + nodem;
    nodem = fmod(nodem, kTwoPi);
    argpm = fmod(argpm, kTwoPi);
    xlm = fmod(xlm, kTwoPi);
    mm = fmod(xlm - argpm - nodem, kTwoPi);
    const double sinim = sin(inclm), cosim = cos(inclm);
    // long-period periodics
    const double ep = em, xincp = inclm, argpp = argpm, nodep = nodem, mp = mm;
    const double axnl = ep * cos(argpp);
    double temp = 1.0 / (am * (1.0 - ep * ep));
    const double aynl = ep * sin(argpp) + temp * s.aycof;
    const double xl = mp + argpp + nodep + temp * s.xlcof * axnl;
    // Kepler's equation for the eccentric longitude
    const double u = fmod(xl - nodep, kTwoPi);
    double eo1 = u, tem5 = 9999.9, sineo1 = 0.0, coseo1 = 0.0;
    for (int ktr = 1; fabs(tem5) >= 1.0e-12 && ktr <= 10; ++ktr) {
        sineo1 = sin(eo1);
        coseo1 = cos(eo1);
        tem5 = 1.0 - coseo1 * axnl - sineo1 * aynl;
        tem5 = (u - aynl * coseo1 + axnl * sineo1 - eo1) / tem5;
        if (fabs(tem5) >= 0.95) tem5 = tem5 > 0.0 ? 0.95 : -0.95;
        eo1 = eo1 + tem5;
    }
    if (SMOOTH) {
        sineo1 = sin(eo1);
        coseo1 = cos(eo1);
    }
    // short-period periodics
    const double ecose = axnl * coseo1 + aynl * sineo1;
    const double esine = axnl * sineo1 - aynl * coseo1;
    const double el2 = axnl * axnl + aynl * aynl;
    const double pl = am * (1.0 - el2);
    if (pl < 0.0) return 4;
    const double rl = am * (1.0 - ecose);
    const double rdotl = sqrt(am) * esine / rl;
    const double rvdotl = sqrt(pl) / rl;
    const double betal = sqrt(1.0 - el2);
    temp = esine / (1.0 + betal);
    const double sinu = am / rl * (sineo1 - aynl - axnl * temp);
    const double cosu = am / rl * (coseo1 - axnl + aynl * temp);
    double su = atan2(sinu, cosu);
    const double sin2u = (cosu + cosu) * sinu;
    const double cos2u = 1.0 - 2.0 * sinu * sinu;
    temp = 1.0 / pl;
    const double temp1 = 0.5 * kJ2 * temp;
    const double temp2 = temp1 * temp;
    const double mrt = rl * (1.0 - 1.5 * temp2 * betal * s.con41) + 0.5 * temp1 * s.x1mth2 * cos2u;
    su = su - 0.25 * temp2 * s.x7thm1 * sin2u;
    const double xnode = nodep + 1.5 * temp2 * cosim * sin2u;
    const double xinc = xincp + 1.5 * temp2 * cosim * sinim * cos2u;
    const double mvt = rdotl - nm * temp1 * s.x1mth2 * sin2u / ke;
    const double rvdot = rvdotl + nm * temp1 * (s.x1mth2 * cos2u + 1.5 * s.con41) / ke;
    // orientation vectors
    const double sinsu = sin(su), cossu = cos(su), snod = sin(xnode), cnod = cos(xnode), sini = sin(xinc),
                 cosi = cos(xinc);
    const double xmx = -snod * cosi, xmy = cnod * cosi;
    const double ux = xmx * sinsu + cnod * cossu, uy = xmy * sinsu + snod * cossu, uz = sini * sinsu;
    const double vx = xmx * cossu - cnod * sinsu, vy = xmy * cossu - snod * sinsu, vz = sini * cossu;
    const double mr = mrt * kRe;
    r[0] = mr * ux;
    r[1] = mr * uy;
    r[2] = mr * uz;
    v[0] = (mvt * ux + rvdot * vx) * vkmpersec;
    v[1] = (mvt * uy + rvdot * vy) * vkmpersec;
    v[2] = (mvt * uz + rvdot * vz) * vkmpersec;
    if (mrt < 1.0) return 6;
    return 0;
}

// The model of a GPE, checked as the published initialisation does by a
// propagation to the epoch: 0, an SGP4 error code, or deep space.
__device__ int sgp4_setup(const mzh_gpe &g, Sgp4Sat &s)
{
    int st = sgp4_init(g, s);
    if (st == 0 && s.deep) return kSgp4DeepSpace;
    if (st == 0) {
        double r[3], v[3];
        st = sgp4_prop<false>(s, 0.0, r, v);
    }
    return st;
}

namespace
{

// One lane per (GPE, time): state at tsince[k] minutes from the GPE's epoch.
__global__ __launch_bounds__(256) void k_sgp4_propagate(const mzh_gpe *__restrict__ gpes, uint64_t n_gpes,
                                                        const double *__restrict__ tsince, uint64_t n_times,
                                                        double *__restrict__ out, int32_t *__restrict__ err)
{
    const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (e >= n_gpes * n_times) return;
    const uint64_t g = e / n_times;
    Sgp4Sat s;
    int st = sgp4_setup(gpes[g], s);
    double r[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    if (st == 0) st = sgp4_prop<false>(s, tsince[e], r, v);
    for (int i = 0; i < 3; ++i) {
        out[e * 6 + i] = r[i];
        out[e * 6 + 3 + i] = v[i];
    }
    err[e] = st;
}

// Bjorck-Pereyra: values f at the nodes x -> monomial coefficients, in place
// (the Newton divided differences, then their expansion).
template <int N>
__device__ inline void bp_interp(const double *x, double *c)
{
    for (int k = 0; k < N; ++k)
        for (int j = N; j > k; --j) c[j] = (c[j] - c[j - 1]) / (x[j] - x[j - k - 1]);
    for (int k = N - 1; k >= 0; --k)
        for (int j = k; j < N; ++j) c[j] = c[j] - x[k] * c[j + 1];
}

template <int N>
__device__ inline double horner_n(const double *c, double x)
{
    double r = c[N];
    for (int k = N - 1; k >= 0; --k) r = c[k] + r * x;
    return r;
}

constexpr double kMinStep = 5.0 / 86400.0;  // days: shorter pieces are accepted as they are
constexpr double kErrPos = 1e-7;             // km: interpolation error of the position
constexpr int kMaxPolys = 20;                // polynomials per nominal step before giving up
constexpr int kStepsPerPeriod = 20;          // nominal steps per orbit of a circular orbit

// One segment per lane: the model, the nominal step and the number of steps.
// The nominal step is a twentieth of the orbital period divided by the ratio
// of the angular rate at perigee to the mean motion, (1 + e)^2 / (1 - e^2)^1.5,
// so that eccentric orbits are sampled as densely at perigee as circular
// ones: with order-8 polynomials that keeps the interpolation error of
// positions and velocities near 1e-9 km and 2e-12 km/s, as close to SGP4 as
// the published code's own stopping rule for Kepler's equation. The last
// step absorbs a remainder shorter than a fifth of a step.
__global__ __launch_bounds__(256) void k_sgp4_seg_init(Sgp4Gen a)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n_segs) return;
    Sgp4Seg sg = a.segs[i];
    Sgp4Sat s;
    int st = sgp4_setup(a.gpes[sg.gpe], s);
    uint32_t n = 1;
    double h0 = 0.0;
    if (st == 0) {
        const double ec = fmin(fmax(s.ecco, 0.0), 0.9);
        const double perigee_rate = (1.0 + ec) * (1.0 + ec) / ((1.0 - ec * ec) * sqrt(1.0 - ec * ec));
        h0 = kTwoPi / s.no_unkozai / 1440.0 / (kStepsPerPeriod * perigee_rate);
        const double q = (sg.t_e - sg.t_b) / h0 - 0.2;
        if (!(h0 > 0.0) || !is_finite_d(h0) || !(q < 4.0e9))
            st = MZH_SGP4_NONFINITE;
        else if (q > 1.0)
            n = (uint32_t)ceil(q);
    }
    sg.h0 = h0;
    sg.n_steps = n;
    sg.status = st;
    a.segs[i] = sg;
    a.sats[i] = s;
}

// Nominal step k of a segment: [b, e), days from the polyjectory's epoch.
__device__ inline void nominal_step(const Sgp4Seg &sg, uint32_t k, double &b, double &e)
{
    b = k == 0 ? sg.t_b : sg.t_b + (double)k * sg.h0;
    e = k + 1u == sg.n_steps ? sg.t_e : sg.t_b + (double)(k + 1u) * sg.h0;
}

// The segment holding nominal step `step` (binary search on first steps).
__device__ inline uint32_t seg_of_step(const uint64_t *seg_step0, uint32_t n_segs, uint64_t step)
{
    uint32_t lo = 0, hi = n_segs;  // seg_step0[lo] <= step < seg_step0[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (seg_step0[mid] <= step)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Interpolate nominal step [b, e) of a model whose epoch is eoff (days from
// the polyjectory's epoch), as make_sgp4_polyjectory.cpp:478-702 does: the
// state at the step's beginning is checked for errors, reentry and exit
// (:838-855); the step is then interpolated at the Chebyshev nodes and halved
// while the position is 1e-7 km or more off SGP4 at one of the midpoints of
// [0, x0], [x0, x1], ..., [x7, x8] (:376-456), pieces shorter than 5 s being
// accepted as they are; more than 20 pieces with more to do is status 14. The pieces are visited in time order, without a stack: piece (d, i) is
// [b + (e - b) i / 2^d, b + (e - b) (i + 1) / 2^d).
// Returns the number of pieces (> 0), or -status. With WRITE, piece j (for j <
// max_out) goes to traj[j] / time[j + 1], and time[0] gets b when first.
template <int ORDER, bool WRITE>
__device__ int interp_step(const Sgp4Sat &sat, double eoff, double b, double e, double reentry_r2, double exit_r2,
                           double *__restrict__ traj, double *__restrict__ time, bool first, int max_out)
{
    constexpr int OP1 = ORDER + 1;
    double kfrac[OP1];
    for (int i = 0; i < OP1; ++i) kfrac[OP1 - 1 - i] = (cos((2.0 * i + 1.0) / (2.0 * OP1) * (kTwoPi / 2.0)) + 1.0) / 2.0;
    double r[3], v[3];
    int st = sgp4_prop<true>(sat, (b - eoff) * 1440.0, r, v);
    if (st != 0) return -st;
    const double rr2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (rr2 < reentry_r2) return -MZH_SGP4_REENTRY;
    if (rr2 >= exit_r2) return -MZH_SGP4_EXIT;
    const double dur0 = e - b;
    if (!is_finite_d(dur0)) return -MZH_SGP4_NONFINITE;
    if (WRITE && first) time[0] = b;
    int npolys = 0;
    uint32_t d = 0, i = 0;
    for (;;) {
        const double scale = ldexp(1.0, -(int)d);
        const double pb = i == 0 ? b : b + dur0 * ((double)i * scale);
        const double pe = (i + 1u) == (1u << d) ? e : b + dur0 * ((double)(i + 1u) * scale);
        const double dur = pe - pb;
        const double base = pb - eoff;
        double x[OP1], c[7][OP1];
        for (int k = 0; k < OP1; ++k) {
            x[k] = kfrac[k] * dur;
            st = sgp4_prop<true>(sat, (base + x[k]) * 1440.0, r, v);
            if (st != 0) return -st;
            for (int ch = 0; ch < 3; ++ch) {
                c[ch][k] = r[ch];
                c[ch + 3][k] = v[ch];
            }
            c[6][k] = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        }
        for (int ch = 0; ch < 7; ++ch) bp_interp<ORDER>(x, c[ch]);
        bool accept = dur < kMinStep;
        if (!accept) {
            double ep2 = 0.0;
            for (int k = 0; k < OP1; ++k) {
                const double lo = k == 0 ? 0.0 : x[k - 1], xm = 0.5 * lo + 0.5 * x[k];
                st = sgp4_prop<true>(sat, (base + xm) * 1440.0, r, v);
                if (st != 0) return -st;
                double p2 = 0.0;
                for (int ch = 0; ch < 3; ++ch) {
                    const double dp = horner_n<ORDER>(c[ch], xm) - r[ch];
                    p2 += dp * dp;
                }
                if (!is_finite_d(p2)) return -MZH_SGP4_NONFINITE;
                ep2 = fmax(ep2, p2);
            }
            accept = ep2 < kErrPos * kErrPos;
        }
        if (accept) {
            for (int ch = 0; ch < 7; ++ch)
                for (int k = 0; k < OP1; ++k)
                    if (!is_finite_d(c[ch][k])) return -MZH_SGP4_NONFINITE;
            if (WRITE && npolys < max_out) {
                double *out = traj + (size_t)npolys * OP1 * 7u;
                for (int k = 0; k < OP1; ++k)
                    for (int ch = 0; ch < 7; ++ch) out[k * 7 + ch] = c[ch][k];
                time[npolys + 1] = pe;
            }
            ++npolys;
            // next piece in time order: up while a right child, then right
            while (d > 0 && (i & 1u)) {
                i >>= 1;
                --d;
            }
            if (d == 0) break;
            ++i;
            if (npolys > kMaxPolys) return -MZH_SGP4_INTERPOLATION;
        } else {
            if (d >= 30u || npolys > kMaxPolys) return -MZH_SGP4_INTERPOLATION;
            ++d;
            i <<= 1;
        }
    }
    return npolys;
}

// One lane per nominal step: its number of polynomials, or -status.
template <int ORDER>
__global__ __launch_bounds__(256) void k_sgp4_count(Sgp4Gen a)
{
    const uint64_t step = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (step >= a.n_steps) return;
    const uint32_t si = seg_of_step(a.seg_step0, a.n_segs, step);
    const Sgp4Seg sg = a.segs[si];
    int res;
    if (sg.status != 0) {
        res = -sg.status;
    } else {
        double b, e;
        nominal_step(sg, (uint32_t)(step - a.seg_step0[si]), b, e);
        res = interp_step<ORDER, false>(a.sats[si], sg.eoff, b, e, a.reentry_r2, a.exit_r2, nullptr, nullptr, false,
                                        0);
    }
    a.step_res[step] = res;
}

// One lane per satellite: its steps in time order up to the first failing one
// (whose status becomes the satellite's, make_sgp4_polyjectory.cpp:1376-1381);
// each kept step's first polynomial index, ~0u for the others.
__global__ __launch_bounds__(256) void k_sgp4_scan(Sgp4Gen a)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= a.n_sats) return;
    uint64_t acc = 0;
    int32_t status = 0;
    double t_last = 0.0;
    for (uint32_t si = a.sat_seg0[s]; si < a.sat_seg0[s + 1u]; ++si) {
        const Sgp4Seg sg = a.segs[si];
        for (uint32_t k = 0; k < sg.n_steps; ++k) {
            const uint64_t step = a.seg_step0[si] + k;
            const int32_t res = status == 0 ? a.step_res[step] : 0;
            if (status == 0 && res < 0) status = -res;
            if (status != 0) {
                a.step_out[step] = ~0u;
                continue;
            }
            a.step_out[step] = (uint32_t)acc;
            acc += (uint64_t)res;
            double b, e;
            nominal_step(sg, k, b, e);
            t_last = e;
        }
    }
    a.sat_npolys[s] = acc;
    a.sat_status[s] = status;
    a.sat_tlast[s] = t_last;
}

// One lane per kept nominal step: interpolate again (the same arithmetic as
// k_sgp4_count) and write the polynomials and their end times.
template <int ORDER>
__global__ __launch_bounds__(256) void k_sgp4_write(Sgp4Gen a)
{
    constexpr int OP1 = ORDER + 1;
    const uint64_t step = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (step >= a.n_steps) return;
    const uint32_t out = a.step_out[step];
    if (out == ~0u) return;
    const int32_t cnt = a.step_res[step];
    if (cnt <= 0) return;
    const uint32_t si = seg_of_step(a.seg_step0, a.n_segs, step);
    const Sgp4Seg sg = a.segs[si];
    double b, e;
    nominal_step(sg, (uint32_t)(step - a.seg_step0[si]), b, e);
    double *traj = a.traj + (a.sat_poly0[sg.sat] + out) * (uint64_t)(OP1 * 7);
    double *time = a.time + a.sat_time0[sg.sat] + out;
    interp_step<ORDER, true>(a.sats[si], sg.eoff, b, e, a.reentry_r2, a.exit_r2, traj, time, out == 0u, cnt);
}

} // namespace

hipError_t launch_sgp4_seg_init(const Sgp4Gen &g, hipStream_t stream)
{
    if (g.n_segs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgp4_seg_init, dim3((g.n_segs + 255u) / 256u), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_sgp4_count(const Sgp4Gen &g, uint32_t order, hipStream_t stream)
{
    if (g.n_steps == 0) return hipSuccess;
    if (order != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sgp4_count<8>, dim3((unsigned)((g.n_steps + 255u) / 256u)), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_sgp4_scan(const Sgp4Gen &g, hipStream_t stream)
{
    if (g.n_sats == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgp4_scan, dim3((g.n_sats + 255u) / 256u), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_sgp4_write(const Sgp4Gen &g, uint32_t order, hipStream_t stream)
{
    if (g.n_steps == 0) return hipSuccess;
    if (order != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sgp4_write<8>, dim3((unsigned)((g.n_steps + 255u) / 256u)), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_sgp4_propagate(const mzh_gpe *gpes, uint64_t n_gpes, const double *tsince, uint64_t n_times,
                                 double *out, int32_t *err, hipStream_t stream)
{
    const uint64_t n = n_gpes * n_times;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgp4_propagate, dim3((unsigned)((n + 255u) / 256u)), dim3(256), 0, stream, gpes, n_gpes,
                       tsince, n_times, out, err);
    return hipGetLastError();
}

} // namespace mzh

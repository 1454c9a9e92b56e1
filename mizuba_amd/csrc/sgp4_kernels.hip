// SGP4 on the GPU: propagation of general perturbations element sets (GPEs)
// and the construction of polyjectories from them, the upstream producer of
// the screening path (mizuba/make_sgp4_polyjectory.cpp:478-913).
//
// The propagator is SGP4 / SDP4 of Spacetrack Report #3 (Hoots & Roehrich,
// 1980) with the corrections of Vallado, Crawford, Hujsak & Kelso,
// "Revisiting Spacetrack Report #3" (AIAA 2006-6753), WGS-72 constants, AFSPC
// operation mode ('a'), in FP64: the near-earth model, and for element sets
// with an orbital period of 225 minutes or more the deep-space lunar-solar
// and resonance terms.
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

constexpr double kPi = kTwoPi / 2.0;
constexpr double kRptim = 4.37526908801129966e-3;  // Earth's rotation, rad/min

// Deep-space initialisation (Hoots & Roehrich 1980, Spacetrack Report #3,
// with the corrections of Vallado et al. 2006): the Greenwich sidereal angle
// at epoch (the AFSPC formula of opsmode 'a'), the lunar and solar terms of
// the orbit's geometry, their periodic coefficients, the secular rates, and
// the geopotential resonance terms of one-day (synchronous) and half-day
// (e >= 0.5) orbits. epoch: days from 1949 December 31 0h.
__device__ void sdp4_init(double epoch, double eccsq, Sgp4Sat &s)
{
    Sdp4Terms &d = s.ds;
    // sidereal angle at epoch
    {
        const double ts70 = epoch - 7305.0;
        const double ds70 = floor(ts70 + 1.0e-8);
        const double tfrac = ts70 - ds70;
        const double c1 = 1.72027916940703639e-2, thgr70 = 1.7321343856509374, fk5r = 5.07551419432269442e-15;
        double g = fmod(thgr70 + c1 * ds70 + (c1 + kTwoPi) * tfrac + ts70 * ts70 * fk5r, kTwoPi);
        if (g < 0.0) g += kTwoPi;
        d.gsto = g;
    }
    // lunar and solar terms at epoch
    const double zes = 0.01675, zel = 0.05490, zns = 1.19459e-5, znl = 1.5835218e-4;
    const double nm = s.no_unkozai, em = s.ecco;
    const double snodm = sin(s.nodeo), cnodm = cos(s.nodeo);
    const double sinomm = sin(s.argpo), cosomm = cos(s.argpo);
    const double sinim = sin(s.inclo), cosim = cos(s.inclo);
    const double emsq = em * em, betasq = 1.0 - emsq, rtemsq = sqrt(betasq);
    d.peo = d.pinco = d.plo = d.pgho = d.pho = 0.0;
    const double day = epoch + 18261.5;
    const double xnodce = fmod(4.5236020 - 9.2422029e-4 * day, kTwoPi);
    const double stem = sin(xnodce), ctem = cos(xnodce);
    const double zcosil = 0.91375164 - 0.03568096 * ctem;
    const double zsinil = sqrt(1.0 - zcosil * zcosil);
    const double zsinhl = 0.089683511 * stem / zsinil;
    const double zcoshl = sqrt(1.0 - zsinhl * zsinhl);
    const double gam = 5.8351514 + 0.0019443680 * day;
    double zx = 0.39785416 * stem / zsinil;
    const double zy = zcoshl * ctem + 0.91744867 * zsinhl * stem;
    zx = atan2(zx, zy);
    zx = gam + zx - xnodce;
    const double zcosgl = cos(zx), zsingl = sin(zx);
    // the sun first (k = 0), then the moon (k = 1)
    double S[2][7], Z[2][12];  // s1..s7; z1 z2 z3 z11 z12 z13 z21 z22 z23 z31 z32 z33
    double zcosg = 0.1945905, zsing = -0.98088458, zcosi = 0.91744867, zsini = 0.39785416;
    double zcosh = cnodm, zsinh = snodm, cc = 2.9864797e-6;
    const double xnoi = 1.0 / nm;
    for (int k = 0; k < 2; ++k) {
        const double a1 = zcosg * zcosh + zsing * zcosi * zsinh;
        const double a3 = -zsing * zcosh + zcosg * zcosi * zsinh;
        const double a7 = -zcosg * zsinh + zsing * zcosi * zcosh;
        const double a8 = zsing * zsini;
        const double a9 = zsing * zsinh + zcosg * zcosi * zcosh;
        const double a10 = zcosg * zsini;
        const double a2 = cosim * a7 + sinim * a8;
        const double a4 = cosim * a9 + sinim * a10;
        const double a5 = -sinim * a7 + cosim * a8;
        const double a6 = -sinim * a9 + cosim * a10;
        const double x1 = a1 * cosomm + a2 * sinomm;
        const double x2 = a3 * cosomm + a4 * sinomm;
        const double x3 = -a1 * sinomm + a2 * cosomm;
        const double x4 = -a3 * sinomm + a4 * cosomm;
        const double x5 = a5 * sinomm, x6 = a6 * sinomm, x7 = a5 * cosomm, x8 = a6 * cosomm;
        double *z = Z[k], *sv = S[k];
        z[9] = 12.0 * x1 * x1 - 3.0 * x3 * x3;
        z[10] = 24.0 * x1 * x2 - 6.0 * x3 * x4;
        z[11] = 12.0 * x2 * x2 - 3.0 * x4 * x4;
        z[0] = 3.0 * (a1 * a1 + a2 * a2) + z[9] * emsq;
        z[1] = 6.0 * (a1 * a3 + a2 * a4) + z[10] * emsq;
        z[2] = 3.0 * (a3 * a3 + a4 * a4) + z[11] * emsq;
        z[3] = -6.0 * a1 * a5 + emsq * (-24.0 * x1 * x7 - 6.0 * x3 * x5);
        z[4] = -6.0 * (a1 * a6 + a3 * a5) + emsq * (-24.0 * (x2 * x7 + x1 * x8) - 6.0 * (x3 * x6 + x4 * x5));
        z[5] = -6.0 * a3 * a6 + emsq * (-24.0 * x2 * x8 - 6.0 * x4 * x6);
        z[6] = 6.0 * a2 * a5 + emsq * (24.0 * x1 * x5 - 6.0 * x3 * x7);
        z[7] = 6.0 * (a4 * a5 + a2 * a6) + emsq * (24.0 * (x2 * x5 + x1 * x6) - 6.0 * (x4 * x7 + x3 * x8));
        z[8] = 6.0 * a4 * a6 + emsq * (24.0 * x2 * x6 - 6.0 * x4 * x8);
        z[0] = z[0] + z[0] + betasq * z[9];
        z[1] = z[1] + z[1] + betasq * z[10];
        z[2] = z[2] + z[2] + betasq * z[11];
        sv[2] = cc * xnoi;
        sv[1] = -0.5 * sv[2] / rtemsq;
        sv[3] = sv[2] * rtemsq;
        sv[0] = -15.0 * em * sv[3];
        sv[4] = x1 * x3 + x2 * x4;
        sv[5] = x2 * x3 + x1 * x4;
        sv[6] = x2 * x4 - x1 * x3;
        // the moon's geometry for the second pass
        zcosg = zcosgl;
        zsing = zsingl;
        zcosi = zcosil;
        zsini = zsinil;
        zcosh = zcoshl * cnodm + zsinhl * snodm;
        zsinh = snodm * zcoshl - cnodm * zsinhl;
        cc = 4.7968065e-7;
    }
    d.zmol = fmod(4.7199672 + 0.22997150 * day - gam, kTwoPi);
    d.zmos = fmod(6.2565837 + 0.017201977 * day, kTwoPi);
    const double *ss = S[0], *sz = Z[0], *sl = S[1], *zl = Z[1];
    // solar periodic coefficients
    d.se2 = 2.0 * ss[0] * ss[5];
    d.se3 = 2.0 * ss[0] * ss[6];
    d.si2 = 2.0 * ss[1] * sz[4];
    d.si3 = 2.0 * ss[1] * (sz[5] - sz[3]);
    d.sl2 = -2.0 * ss[2] * sz[1];
    d.sl3 = -2.0 * ss[2] * (sz[2] - sz[0]);
    d.sl4 = -2.0 * ss[2] * (-21.0 - 9.0 * emsq) * zes;
    d.sgh2 = 2.0 * ss[3] * sz[10];
    d.sgh3 = 2.0 * ss[3] * (sz[11] - sz[9]);
    d.sgh4 = -18.0 * ss[3] * zes;
    d.sh2 = -2.0 * ss[1] * sz[7];
    d.sh3 = -2.0 * ss[1] * (sz[8] - sz[6]);
    // lunar periodic coefficients
    d.ee2 = 2.0 * sl[0] * sl[5];
    d.e3 = 2.0 * sl[0] * sl[6];
    d.xi2 = 2.0 * sl[1] * zl[4];
    d.xi3 = 2.0 * sl[1] * (zl[5] - zl[3]);
    d.xl2 = -2.0 * sl[2] * zl[1];
    d.xl3 = -2.0 * sl[2] * (zl[2] - zl[0]);
    d.xl4 = -2.0 * sl[2] * (-21.0 - 9.0 * emsq) * zel;
    d.xgh2 = 2.0 * sl[3] * zl[10];
    d.xgh3 = 2.0 * sl[3] * (zl[11] - zl[9]);
    d.xgh4 = -18.0 * sl[3] * zel;
    d.xh2 = -2.0 * sl[1] * zl[7];
    d.xh3 = -2.0 * sl[1] * (zl[8] - zl[6]);

    // secular rates from the sun and the moon
    const bool polar_flat = s.inclo < 5.2359877e-2 || s.inclo > kPi - 5.2359877e-2;
    const double ses = ss[0] * zns * ss[4];
    const double sis = ss[1] * zns * (sz[3] + sz[5]);
    const double sls = -zns * ss[2] * (sz[0] + sz[2] - 14.0 - 6.0 * emsq);
    const double sghs = ss[3] * zns * (sz[9] + sz[11] - 6.0);
    double shs = -zns * ss[1] * (sz[6] + sz[8]);
    if (polar_flat) shs = 0.0;
    if (sinim != 0.0) shs = shs / sinim;
    const double sgs = sghs - cosim * shs;
    d.dedt = ses + sl[0] * znl * sl[4];
    d.didt = sis + sl[1] * znl * (zl[3] + zl[5]);
    d.dmdt = sls - znl * sl[2] * (zl[0] + zl[2] - 14.0 - 6.0 * emsq);
    const double sghl = sl[3] * znl * (zl[9] + zl[11] - 6.0);
    double shll = -znl * sl[1] * (zl[6] + zl[8]);
    if (polar_flat) shll = 0.0;
    d.domdt = sgs + sghl;
    d.dnodt = shs;
    if (sinim != 0.0) {
        d.domdt = d.domdt - cosim / sinim * shll;
        d.dnodt = d.dnodt + shll / sinim;
    }

    // geopotential resonances
    d.irez = 0;
    if (nm < 0.0052359877 && nm > 0.0034906585) d.irez = 1;
    if (nm >= 8.26e-3 && nm <= 9.24e-3 && em >= 0.5) d.irez = 2;
    d.d2201 = d.d2211 = d.d3210 = d.d3222 = d.d4410 = d.d4422 = 0.0;
    d.d5220 = d.d5232 = d.d5421 = d.d5433 = 0.0;
    d.del1 = d.del2 = d.del3 = d.xfact = d.xlamo = 0.0;
    if (d.irez == 0) return;
    const double theta = d.gsto;
    const double aonv = pow(nm / xke(), 2.0 / 3.0);
    if (d.irez == 2) {
        // half-day orbits
        const double cosisq = cosim * cosim;
        const double e = s.ecco, esq = eccsq, eoc = e * esq;
        const double g201 = -0.306 - (e - 0.64) * 0.440;
        double g211, g310, g322, g410, g422, g520, g521, g532, g533;
        if (e <= 0.65) {
            g211 = 3.616 - 13.2470 * e + 16.2900 * esq;
            g310 = -19.302 + 117.3900 * e - 228.4190 * esq + 156.5910 * eoc;
            g322 = -18.9068 + 109.7927 * e - 214.6334 * esq + 146.5816 * eoc;
            g410 = -41.122 + 242.6940 * e - 471.0940 * esq + 313.9530 * eoc;
            g422 = -146.407 + 841.8800 * e - 1629.014 * esq + 1083.4350 * eoc;
            g520 = -532.114 + 3017.977 * e - 5740.032 * esq + 3708.2760 * eoc;
        } else {
            g211 = -72.099 + 331.819 * e - 508.738 * esq + 266.724 * eoc;
            g310 = -346.844 + 1582.851 * e - 2415.925 * esq + 1246.113 * eoc;
            g322 = -342.585 + 1554.908 * e - 2366.899 * esq + 1215.972 * eoc;
            g410 = -1052.797 + 4758.686 * e - 7193.992 * esq + 3651.957 * eoc;
            g422 = -3581.690 + 16178.110 * e - 24462.770 * esq + 12422.520 * eoc;
            if (e > 0.715)
                g520 = -5149.66 + 29936.92 * e - 54087.36 * esq + 31324.56 * eoc;
            else
                g520 = 1464.74 - 4664.75 * e + 3763.64 * esq;
        }
        if (e < 0.7) {
            g533 = -919.22770 + 4988.6100 * e - 9064.7700 * esq + 5542.21 * eoc;
            g521 = -822.71072 + 4568.6173 * e - 8491.4146 * esq + 5337.524 * eoc;
            g532 = -853.66600 + 4690.2500 * e - 8624.7700 * esq + 5341.4 * eoc;
        } else {
            g533 = -37995.780 + 161616.52 * e - 229838.20 * esq + 109377.94 * eoc;
            g521 = -51752.104 + 218913.95 * e - 309468.16 * esq + 146349.42 * eoc;
            g532 = -40023.880 + 170470.89 * e - 242699.48 * esq + 115605.82 * eoc;
        }
        const double sini2 = sinim * sinim;
        const double f220 = 0.75 * (1.0 + 2.0 * cosim + cosisq);
        const double f221 = 1.5 * sini2;
        const double f321 = 1.875 * sinim * (1.0 - 2.0 * cosim - 3.0 * cosisq);
        const double f322 = -1.875 * sinim * (1.0 + 2.0 * cosim - 3.0 * cosisq);
        const double f441 = 35.0 * sini2 * f220;
        const double f442 = 39.3750 * sini2 * sini2;
        const double f522 = 9.84375 * sinim
                            * (sini2 * (1.0 - 2.0 * cosim - 5.0 * cosisq) + 0.33333333 * (-2.0 + 4.0 * cosim + 6.0 * cosisq));
        const double f523 = sinim * (4.92187512 * sini2 * (-2.0 - 4.0 * cosim + 10.0 * cosisq)
                                     + 6.56250012 * (1.0 + 2.0 * cosim - 3.0 * cosisq));
        const double f542 = 29.53125 * sinim * (2.0 - 8.0 * cosim + cosisq * (-12.0 + 8.0 * cosim + 10.0 * cosisq));
        const double f543 = 29.53125 * sinim * (-2.0 - 8.0 * cosim + cosisq * (12.0 + 8.0 * cosim - 10.0 * cosisq));
        const double xno2 = nm * nm, ainv2 = aonv * aonv;
        double temp1 = 3.0 * xno2 * ainv2;
        double temp = temp1 * 1.7891679e-6;
        d.d2201 = temp * f220 * g201;
        d.d2211 = temp * f221 * g211;
        temp1 = temp1 * aonv;
        temp = temp1 * 3.7393792e-7;
        d.d3210 = temp * f321 * g310;
        d.d3222 = temp * f322 * g322;
        temp1 = temp1 * aonv;
        temp = 2.0 * temp1 * 7.3636953e-9;
        d.d4410 = temp * f441 * g410;
        d.d4422 = temp * f442 * g422;
        temp1 = temp1 * aonv;
        temp = temp1 * 1.1428639e-7;
        d.d5220 = temp * f522 * g520;
        d.d5232 = temp * f523 * g532;
        temp = 2.0 * temp1 * 2.1765803e-9;
        d.d5421 = temp * f542 * g521;
        d.d5433 = temp * f543 * g533;
        d.xlamo = fmod(s.mo + s.nodeo + s.nodeo - theta - theta, kTwoPi);
        d.xfact = s.mdot + d.dmdt + 2.0 * (s.nodedot + d.dnodt - kRptim) - s.no_unkozai;
    } else {
        // one-day (synchronous) orbits
        const double g200 = 1.0 + emsq * (-2.5 + 0.8125 * emsq);
        const double g310 = 1.0 + 2.0 * emsq;
        const double g300 = 1.0 + emsq * (-6.0 + 6.60937 * emsq);
        const double f220 = 0.75 * (1.0 + cosim) * (1.0 + cosim);
        const double f311 = 0.9375 * sinim * sinim * (1.0 + 3.0 * cosim) - 0.75 * (1.0 + cosim);
        double f330 = 1.0 + cosim;
        f330 = 1.875 * f330 * f330 * f330;
        const double del1 = 3.0 * nm * nm * aonv * aonv;
        d.del2 = 2.0 * del1 * f220 * g200 * 1.7891679e-6;
        d.del3 = 3.0 * del1 * f330 * g300 * 2.2123015e-7 * aonv;
        d.del1 = del1 * f311 * g310 * 2.1460748e-6 * aonv;
        d.xlamo = fmod(s.mo + s.nodeo + s.argpo - theta, kTwoPi);
        d.xfact = s.mdot + (s.argpdot + s.nodedot) - kRptim + d.dmdt + d.domdt + d.dnodt - s.no_unkozai;
    }
}

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
    // deep space: period of 225 minutes or more
    if (kTwoPi / s.no_unkozai >= 225.0) {
        s.deep = 1;
        s.isimp = 1;
        sdp4_init((g.epoch_jd1 - 2433281.5) + g.epoch_jd2, eccsq, s);
    }
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

// Deep-space secular effects at t minutes: the lunar-solar secular rates,
// then, for resonant orbits, the resonance terms integrated from epoch with
// the fixed 720-minute Euler-Maclaurin steps of the published model (the
// same steps from epoch at every call, so the result does not depend on
// earlier calls). Returns false for a resonant orbit at a time that is not
// finite or beyond 5e7 minutes (about 95 years) of its epoch, where the
// integration would not end in reasonable time.
__device__ bool sdp4_secular(const Sgp4Sat &s, double t, double &em, double &argpm, double &inclm, double &mm,
                             double &nodem, double &nm)
{
    const Sdp4Terms &d = s.ds;
    const double theta = fmod(d.gsto + t * kRptim, kTwoPi);
    em = em + d.dedt * t;
    inclm = inclm + d.didt * t;
    argpm = argpm + d.domdt * t;
    nodem = nodem + d.dnodt * t;
    mm = mm + d.dmdt * t;
    if (d.irez == 0) return true;
    if (!(fabs(t) <= 5.0e7)) return false;
    const double step = t > 0.0 ? 720.0 : -720.0, step2 = 259200.0;
    double atime = 0.0, xni = s.no_unkozai, xli = d.xlamo, ft = 0.0;
    double xndt, xldot, xnddt;
    for (;;) {
        if (d.irez != 2) {
            const double fasx2 = 0.13130908, fasx4 = 2.8843198, fasx6 = 0.37448087;
            xndt = d.del1 * sin(xli - fasx2) + d.del2 * sin(2.0 * (xli - fasx4)) + d.del3 * sin(3.0 * (xli - fasx6));
            xldot = xni + d.xfact;
            xnddt = d.del1 * cos(xli - fasx2) + 2.0 * d.del2 * cos(2.0 * (xli - fasx4))
                    + 3.0 * d.del3 * cos(3.0 * (xli - fasx6));
            xnddt = xnddt * xldot;
        } else {
            const double g22 = 5.7686396, g32 = 0.95240898, g44 = 1.8014998, g52 = 1.0508330, g54 = 4.4108898;
            const double xomi = s.argpo + s.argpdot * atime;
            const double x2omi = xomi + xomi, x2li = xli + xli;
            xndt = d.d2201 * sin(x2omi + xli - g22) + d.d2211 * sin(xli - g22) + d.d3210 * sin(xomi + xli - g32)
                   + d.d3222 * sin(-xomi + xli - g32) + d.d4410 * sin(x2omi + x2li - g44) + d.d4422 * sin(x2li - g44)
                   + d.d5220 * sin(xomi + xli - g52) + d.d5232 * sin(-xomi + xli - g52)
                   + d.d5421 * sin(xomi + x2li - g54) + d.d5433 * sin(-xomi + x2li - g54);
            xldot = xni + d.xfact;
            xnddt = d.d2201 * cos(x2omi + xli - g22) + d.d2211 * cos(xli - g22) + d.d3210 * cos(xomi + xli - g32)
                    + d.d3222 * cos(-xomi + xli - g32) + d.d5220 * cos(xomi + xli - g52)
                    + d.d5232 * cos(-xomi + xli - g52)
                    + 2.0 * (d.d4410 * cos(x2omi + x2li - g44) + d.d4422 * cos(x2li - g44)
                             + d.d5421 * cos(xomi + x2li - g54) + d.d5433 * cos(-xomi + x2li - g54));
            xnddt = xnddt * xldot;
        }
        if (fabs(t - atime) >= 720.0) {
            xli = xli + xldot * step + xndt * step2;
            xni = xni + xndt * step + xnddt * step2;
            atime = atime + step;
        } else {
            ft = t - atime;
            break;
        }
    }
    nm = xni + xndt * ft + xnddt * ft * ft * 0.5;
    const double xl = xli + xldot * ft + xndt * ft * ft * 0.5;
    if (d.irez != 1)
        mm = xl - 2.0 * nodem + 2.0 * theta;
    else
        mm = xl - nodem - argpm + theta;
    const double dndt = nm - s.no_unkozai;
    nm = s.no_unkozai + dndt;
    return true;
}

// Deep-space lunar-solar periodics at t minutes, applied to the elements
// unless `at_init` (the published initialisation's own propagation to epoch,
// which leaves them out). Inclinations below 0.2 rad use Lyddane's form, with
// the AFSPC (opsmode 'a') quadrant fixes of the node.
__device__ void sdp4_periodics(const Sdp4Terms &d, double t, bool at_init, double &ep, double &inclp, double &nodep,
                               double &argpp, double &mp)
{
    const double zns = 1.19459e-5, zes = 0.01675, znl = 1.5835218e-4, zel = 0.05490;
    double zm = at_init ? d.zmos : d.zmos + zns * t;
    double zf = zm + 2.0 * zes * sin(zm);
    double sinzf = sin(zf);
    double f2 = 0.5 * sinzf * sinzf - 0.25;
    double f3 = -0.5 * sinzf * cos(zf);
    const double ses = d.se2 * f2 + d.se3 * f3;
    const double sis = d.si2 * f2 + d.si3 * f3;
    const double sls = d.sl2 * f2 + d.sl3 * f3 + d.sl4 * sinzf;
    const double sghs = d.sgh2 * f2 + d.sgh3 * f3 + d.sgh4 * sinzf;
    const double shs = d.sh2 * f2 + d.sh3 * f3;
    zm = at_init ? d.zmol : d.zmol + znl * t;
    zf = zm + 2.0 * zel * sin(zm);
    sinzf = sin(zf);
    f2 = 0.5 * sinzf * sinzf - 0.25;
    f3 = -0.5 * sinzf * cos(zf);
    const double sel = d.ee2 * f2 + d.e3 * f3;
    const double sil = d.xi2 * f2 + d.xi3 * f3;
    const double sll = d.xl2 * f2 + d.xl3 * f3 + d.xl4 * sinzf;
    const double sghl = d.xgh2 * f2 + d.xgh3 * f3 + d.xgh4 * sinzf;
    const double shll = d.xh2 * f2 + d.xh3 * f3;
    if (at_init) return;
    const double pe = (ses + sel) - d.peo;
    const double pinc = (sis + sil) - d.pinco;
    const double pl = (sls + sll) - d.plo;
    double pgh = (sghs + sghl) - d.pgho;
    double ph = (shs + shll) - d.pho;
    inclp = inclp + pinc;
    ep = ep + pe;
    const double sinip = sin(inclp), cosip = cos(inclp);
    if (inclp >= 0.2) {
        ph = ph / sinip;
        pgh = pgh - cosip * ph;
        argpp = argpp + pgh;
        nodep = nodep + ph;
        mp = mp + pl;
    } else {
        const double sinop = sin(nodep), cosop = cos(nodep);
        double alfdp = sinip * sinop, betdp = sinip * cosop;
        const double dalf = ph * cosop + pinc * cosip * sinop;
        const double dbet = -ph * sinop + pinc * cosip * cosop;
        alfdp = alfdp + dalf;
        betdp = betdp + dbet;
        nodep = fmod(nodep, kTwoPi);
        if (nodep < 0.0) nodep = nodep + kTwoPi;
        double xls = mp + argpp + cosip * nodep;
        const double dls = pl + pgh - pinc * nodep * sinip;
        xls = xls + dls;
        const double xnoh = nodep;
        nodep = atan2(alfdp, betdp);
        if (nodep < 0.0) nodep = nodep + kTwoPi;
        if (fabs(xnoh - nodep) > kPi) {
            if (nodep < xnoh)
                nodep = nodep + kTwoPi;
            else
                nodep = nodep - kTwoPi;
        }
        mp = mp + pl;
        argpp = xls - mp - cosip * nodep;
    }
}

} // namespace

// Position (km) and velocity (km/s) in TEME at tsince minutes from the epoch.
// Returns 0, or the model's error code: 1 mean eccentricity out of range, 2
// negative mean motion, 3 perturbed eccentricity out of range (deep space), 4
// negative semi-latus rectum, 6 decayed (radius below one Earth radius); 10
// for a resonant deep-space orbit at a time beyond 5e7 minutes of its epoch.
// at_init leaves out the deep-space periodics, as the published
// initialisation's propagation to epoch does.
// SMOOTH = false follows the published code exactly: its solution of Kepler's
// equation stops once a Newton correction is below 1e-12 rad and keeps the
// sine and cosine from before that correction, so the state jumps by up to
// about 1e-12 rad (some 1e-8 km) where the iteration count changes. SMOOTH =
// true takes them at the converged anomaly instead: the model without those
// jumps, for interpolation.
// DEEP selects the model at compile time (the element set's own `deep` flag
// must match), so that near-earth code carries none of the deep-space terms.
template <bool SMOOTH, bool DEEP>
__device__ __attribute__((always_inline)) int sgp4_prop(const Sgp4Sat &s, double t, double r[3], double v[3], bool at_init = false)
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
    double inclm = s.inclo;
    if constexpr (DEEP) {
        if (!sdp4_secular(s, t, em, argpm, inclm, mm, nodem, nm)) return MZH_SGP4_NONFINITE;
    }
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
    // lunar-solar periodics (deep space)
    double ep = em, xincp = inclm, argpp = argpm, nodep = nodem, mp = mm;
    double sinip, cosip, aycof = s.aycof, xlcof = s.xlcof;
    if constexpr (DEEP) {
        sdp4_periodics(s.ds, t, at_init, ep, xincp, nodep, argpp, mp);
        if (xincp < 0.0) {
            xincp = -xincp;
            nodep = nodep + kPi;
            argpp = argpp - kPi;
        }
        if (ep < 0.0 || ep > 1.0) return 3;
        sinip = sin(xincp);
        cosip = cos(xincp);
        aycof = -0.5 * (kJ3 / kJ2) * sinip;
        if (fabs(cosip + 1.0) > 1.5e-12)
            xlcof = -0.25 * (kJ3 / kJ2) * sinip * (3.0 + 5.0 * cosip) / (1.0 + cosip);
        else
            xlcof = -0.25 * (kJ3 / kJ2) * sinip * (3.0 + 5.0 * cosip) / 1.5e-12;
    } else {
        sinip = sin(inclm);
        cosip = cos(inclm);
    }
    // long-period periodics
    const double axnl = ep * cos(argpp);
    double temp = 1.0 / (am * (1.0 - ep * ep));
    const double aynl = ep * sin(argpp) + temp * aycof;
    const double xl = mp + argpp + nodep + temp * xlcof * axnl;
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
    double con41 = s.con41, x1mth2 = s.x1mth2, x7thm1 = s.x7thm1;
    if constexpr (DEEP) {
        const double cosisq = cosip * cosip;
        con41 = 3.0 * cosisq - 1.0;
        x1mth2 = 1.0 - cosisq;
        x7thm1 = 7.0 * cosisq - 1.0;
    }
    const double mrt = rl * (1.0 - 1.5 * temp2 * betal * con41) + 0.5 * temp1 * x1mth2 * cos2u;
    su = su - 0.25 * temp2 * x7thm1 * sin2u;
    const double xnode = nodep + 1.5 * temp2 * cosip * sin2u;
    const double xinc = xincp + 1.5 * temp2 * cosip * sinip * cos2u;
    const double mvt = rdotl - nm * temp1 * x1mth2 * sin2u / ke;
    const double rvdot = rvdotl + nm * temp1 * (x1mth2 * cos2u + 1.5 * con41) / ke;
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
// propagation to the epoch: 0 or an SGP4 error code.
__device__ int sgp4_setup(const mzh_gpe &g, Sgp4Sat &s)
{
    int st = sgp4_init(g, s);
    if (st == 0) {
        double r[3], v[3];
        st = s.deep ? sgp4_prop<false, true>(s, 0.0, r, v, true) : sgp4_prop<false, false>(s, 0.0, r, v, true);
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
    if (st == 0) st = s.deep ? sgp4_prop<false, true>(s, tsince[e], r, v) : sgp4_prop<false, false>(s, tsince[e], r, v);
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
    sg.deep = s.deep;
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
// max_out) goes to traj[j * 7 (order + 1) + (k * 7 + ch) * cstride] for
// coefficient k of channel ch, its end time to time[j + 1], and time[0] gets b
// when first.
template <int ORDER, bool WRITE, bool DEEP>
__device__ __attribute__((always_inline)) int interp_step(const Sgp4Sat &sat, double eoff, double b, double e, double reentry_r2, double exit_r2,
                           double *__restrict__ traj, double *__restrict__ time, bool first, int max_out,
                           uint64_t cstride = 1)
{
    constexpr int OP1 = ORDER + 1;
    double kfrac[OP1];
    for (int i = 0; i < OP1; ++i) kfrac[OP1 - 1 - i] = (cos((2.0 * i + 1.0) / (2.0 * OP1) * (kTwoPi / 2.0)) + 1.0) / 2.0;
    double r[3], v[3];
    int st = sgp4_prop<true, DEEP>(sat, (b - eoff) * 1440.0, r, v);
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
            st = sgp4_prop<true, DEEP>(sat, (base + x[k]) * 1440.0, r, v);
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
                st = sgp4_prop<true, DEEP>(sat, (base + xm) * 1440.0, r, v);
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
                    for (int ch = 0; ch < 7; ++ch) out[(k * 7 + ch) * cstride] = c[ch][k];
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
// The near-earth instantiation (DEEP = false) takes the steps of near-earth
// segments and of failed ones, the deep-space one the others.
__device__ inline bool takes(const Sgp4Seg &sg, bool deep) { return deep == (sg.status == 0 && sg.deep != 0); }

template <int ORDER, bool DEEP, bool STAGE>
__global__ __launch_bounds__(256) void k_sgp4_count(Sgp4Gen a)
{
    const uint64_t step = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (step >= a.n_steps) return;
    const uint32_t si = seg_of_step(a.seg_step0, a.n_segs, step);
    const Sgp4Seg sg = a.segs[si];
    if (!takes(sg, DEEP)) return;
    int res;
    if (sg.status != 0) {
        res = -sg.status;
    } else {
        double b, e;
        nominal_step(sg, (uint32_t)(step - a.seg_step0[si]), b, e);
        if constexpr (STAGE)
            // the first piece goes to the staging slot of the step
            res = interp_step<ORDER, true, DEEP>(a.sats[si], sg.eoff, b, e, a.reentry_r2, a.exit_r2, a.stage + step,
                                                 a.stage_t + step, false, 1, a.n_steps);
        else
            res = interp_step<ORDER, false, DEEP>(a.sats[si], sg.eoff, b, e, a.reentry_r2, a.exit_r2, nullptr,
                                                  nullptr, false, 0);
    }
    a.step_res[step] = res;
}

// 64 steps per workgroup: the staged polynomials of the kept single-piece
// steps go through an LDS tile from the coefficient-major staging array to
// their records in the polyjectory, both sides read and written contiguously;
// their end times (and a satellite's start time) go along.
template <int ORDER>
__global__ __launch_bounds__(256) void k_sgp4_copy(Sgp4Gen a)
{
    constexpr uint32_t NC = 7u * (ORDER + 1u);
    __shared__ double tile[NC * 65u];
    __shared__ uint64_t dst[64];
    const uint64_t s0 = (uint64_t)blockIdx.x * 64u;
    const uint32_t tid = threadIdx.x;
    if (tid < 64u) {
        const uint64_t step = s0 + tid;
        uint64_t d = ~0ull;
        if (step < a.n_steps && a.step_res[step] == 1) {
            const uint32_t out = a.step_out[step];
            if (out != ~0u) {
                const uint32_t si = seg_of_step(a.seg_step0, a.n_segs, step);
                const Sgp4Seg sg = a.segs[si];
                d = (a.sat_poly0[sg.sat] + out) * NC;
                double *time = a.time + a.sat_time0[sg.sat] + out;
                time[1] = a.stage_t[step + 1u];
                if (out == 0u) {
                    double b, e;
                    nominal_step(sg, (uint32_t)(step - a.seg_step0[si]), b, e);
                    time[0] = b;
                }
            }
        }
        dst[tid] = d;
    }
    for (uint32_t i = tid; i < NC * 64u; i += 256u) {
        const uint32_t k = i / 64u, j = i % 64u;
        if (s0 + j < a.n_steps) tile[k * 65u + j] = a.stage[k * a.n_steps + s0 + j];
    }
    __syncthreads();
    for (uint32_t i = tid; i < NC * 64u; i += 256u) {
        const uint32_t j = i / NC, k = i % NC;
        const uint64_t d = dst[j];
        if (d != ~0ull) a.traj[d + k] = tile[k * 65u + j];
    }
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
            if (res > 1 || !a.stage) a.redo[atomicAdd(a.redo_count, 1ull)] = (uint32_t)step;
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
template <int ORDER, bool DEEP>
__global__ __launch_bounds__(256) void k_sgp4_write(Sgp4Gen a, uint64_t n_redo)
{
    constexpr int OP1 = ORDER + 1;
    const uint64_t r = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (r >= n_redo) return;
    const uint32_t step = a.redo[r];
    const uint32_t out = a.step_out[step];
    const int32_t cnt = a.step_res[step];
    if (out == ~0u || cnt <= 0) return;
    const uint32_t si = seg_of_step(a.seg_step0, a.n_segs, step);
    const Sgp4Seg sg = a.segs[si];
    if (!takes(sg, DEEP)) return;
    double b, e;
    nominal_step(sg, (uint32_t)(step - a.seg_step0[si]), b, e);
    double *traj = a.traj + (a.sat_poly0[sg.sat] + out) * (uint64_t)(OP1 * 7);
    double *time = a.time + a.sat_time0[sg.sat] + out;
    interp_step<ORDER, true, DEEP>(a.sats[si], sg.eoff, b, e, a.reentry_r2, a.exit_r2, traj, time, out == 0u, cnt);
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
    const dim3 grid((unsigned)((g.n_steps + 255u) / 256u));
    if (g.stage) {
        hipLaunchKernelGGL((k_sgp4_count<8, false, true>), grid, dim3(256), 0, stream, g);
        if (g.any_deep) hipLaunchKernelGGL((k_sgp4_count<8, true, true>), grid, dim3(256), 0, stream, g);
    } else {
        hipLaunchKernelGGL((k_sgp4_count<8, false, false>), grid, dim3(256), 0, stream, g);
        if (g.any_deep) hipLaunchKernelGGL((k_sgp4_count<8, true, false>), grid, dim3(256), 0, stream, g);
    }
    return hipGetLastError();
}

hipError_t launch_sgp4_scan(const Sgp4Gen &g, hipStream_t stream)
{
    if (g.n_sats == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgp4_scan, dim3((g.n_sats + 255u) / 256u), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_sgp4_write(const Sgp4Gen &g, uint32_t order, uint64_t n_redo, hipStream_t stream)
{
    if (g.n_steps == 0) return hipSuccess;
    if (order != 8) return hipErrorInvalidValue;
    if (g.stage) {
        const uint64_t blocks = (g.n_steps + 63u) / 64u;
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_sgp4_copy<8>), dim3((unsigned)blocks), dim3(256), 0, stream, g);
    }
    if (n_redo) {
        const dim3 grid((unsigned)((n_redo + 255u) / 256u));
        hipLaunchKernelGGL((k_sgp4_write<8, false>), grid, dim3(256), 0, stream, g, n_redo);
        if (g.any_deep) hipLaunchKernelGGL((k_sgp4_write<8, true>), grid, dim3(256), 0, stream, g, n_redo);
    }
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

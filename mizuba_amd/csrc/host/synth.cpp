#include "synth.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace mzh_host
{

namespace
{

constexpr double kMu = 398600.8;    // km^3/s^2 (WGS-72)
constexpr double kRe = 6378.135;    // km (WGS-72)
constexpr double kJ2 = 0.001082616; // WGS-72
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDay = 86400.0;

// splitmix64: independent, reproducible stream per object
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform(double a, double b) { return a + (b - a) * ((double)(next() >> 11) * 0x1.0p-53); }
};

struct Rates {
    double n, raan_dot, argp_dot;
};

Rates secular_rates(const Elements &el)
{
    const double n0 = std::sqrt(kMu / (el.a * el.a * el.a));
    const double p = el.a * (1 - el.e * el.e);
    const double k = 1.5 * kJ2 * (kRe / p) * (kRe / p) * n0;
    const double ci = std::cos(el.inc), si = std::sin(el.inc);
    Rates r;
    r.raan_dot = -k * ci;
    r.argp_dot = k * (2.0 - 2.5 * si * si);
    r.n = n0 + 0.5 * k * std::sqrt(1 - el.e * el.e) * (3 * ci * ci - 1);
    return r;
}

// Bjorck-Pereyra: values f at nodes x -> monomial coefficients (in place).
void bp_interp(const double *x, double *c, int n)
{
    for (int k = 0; k < n; ++k)
        for (int j = n; j > k; --j) c[j] = (c[j] - c[j - 1]) / (x[j] - x[j - k - 1]);
    for (int k = n - 1; k >= 0; --k)
        for (int j = k; j < n; ++j) c[j] = c[j] - x[k] * c[j + 1];
}

double horner(const double *c, int n, double x)
{
    double r = c[n];
    for (int k = n - 1; k >= 0; --k) r = c[k] + r * x;
    return r;
}

} // namespace

void synth_state(const Elements &el, double t_days, double out[7])
{
    const Rates rt = secular_rates(el);
    const double t = t_days * kDay;
    const double M = el.m0 + rt.n * t;
    const double raan = el.raan + rt.raan_dot * t;
    const double argp = el.argp + rt.argp_dot * t;
    const double e = el.e;
    double E = M + e * std::sin(M);
    for (int it = 0; it < 30; ++it) {
        const double f = E - e * std::sin(E) - M;
        const double dE = f / (1 - e * std::cos(E));
        E -= dE;
        if (std::fabs(dE) < 1e-15) break;
    }
    const double cE = std::cos(E), sE = std::sin(E);
    const double sq = std::sqrt(1 - e * e);
    const double rr = el.a * (1 - e * cE);
    const double xp = el.a * (cE - e), yp = el.a * sq * sE;
    const double edot = rt.n / (1 - e * cE);
    const double vxp = -el.a * sE * edot, vyp = el.a * sq * cE * edot;
    const double cO = std::cos(raan), sO = std::sin(raan), cw = std::cos(argp), sw = std::sin(argp);
    const double ci = std::cos(el.inc), si = std::sin(el.inc);
    const double r11 = cO * cw - sO * sw * ci, r12 = -cO * sw - sO * cw * ci;
    const double r21 = sO * cw + cO * sw * ci, r22 = -sO * sw + cO * cw * ci;
    const double r31 = sw * si, r32 = cw * si;
    out[0] = r11 * xp + r12 * yp;
    out[1] = r21 * xp + r22 * yp;
    out[2] = r31 * xp + r32 * yp;
    out[3] = r11 * vxp + r12 * vyp;
    out[4] = r21 * vxp + r22 * vyp;
    out[5] = r31 * vxp + r32 * vyp;
    out[6] = rr;
}

SynthOut make_synthetic(const SynthParams &p)
{
    const uint32_t N = p.n_objs;
    const int n = (int)p.order, op1 = n + 1;
    SynthOut out;
    out.order = p.order;
    out.elements.resize(N);
    // element sampling (deterministic per object index)
    for (uint32_t o = 0; o < N; ++o) {
        Rng rng{p.seed * 0x100000001b3ull + o};
        Elements el;
        if (o < p.n_shell) {
            // dense shells: 5 altitudes, 72 evenly phased planes, i = 53 or 97.6 deg
            const uint32_t shell = o % 5, plane = (o / 5) % 72, slot = o / 360;
            const uint32_t per_plane = std::max<uint32_t>(1, (p.n_shell + 359) / 360);
            const double alt = 540.0 + 10.0 * shell;
            el.e = rng.uniform(0, 1e-4);
            el.a = (kRe + alt) / (1 - el.e);
            el.inc = ((plane % 2) ? 97.6 : 53.0) * kPi / 180;
            el.raan = 2 * kPi * plane / 72.0;
            el.argp = rng.uniform(0, 2 * kPi);
            el.m0 = 2 * kPi * slot / per_plane + 2 * kPi * shell / (5.0 * per_plane);
            el.bstar = rng.uniform(1e-5, 1e-4);
        } else {
            const bool high = rng.uniform(0, 1) < p.frac_high;
            const double hp = high ? rng.uniform(1200, 2000) : rng.uniform(400, 1200);
            el.e = high ? rng.uniform(0, 0.05) : rng.uniform(0, 0.01);
            el.a = (kRe + hp) / (1 - el.e);
            el.inc = rng.uniform(0, 100) * kPi / 180;
            el.raan = rng.uniform(0, 2 * kPi);
            el.argp = rng.uniform(0, 2 * kPi);
            el.m0 = rng.uniform(0, 2 * kPi);
            el.bstar = rng.uniform(1e-5, 1e-4);
        }
        out.elements[o] = el;
    }
    // Chebyshev node fractions on [0, 1], ascending
    std::vector<double> kfrac(op1);
    for (int i = 0; i < op1; ++i) kfrac[op1 - 1 - i] = (std::cos((2. * i + 1.) / (2. * op1) * kPi) + 1.) / 2.;

    std::vector<std::vector<double>> traj_o(N), time_o(N);
#pragma omp parallel for schedule(dynamic, 16)
    for (long long oi = 0; oi < (long long)N; ++oi) {
        const Elements &el = out.elements[oi];
        const double T_days = 2 * kPi / secular_rates(el).n / kDay;
        const double h0 = T_days / 8;
        const double hmin = 5.0 / kDay;
        std::vector<double> &tr = traj_o[oi], &tm = time_o[oi];
        // one allocation per object: growing these vectors step by step from
        // many threads at once serialises on the allocator's memory maps
        const size_t est_steps = (size_t)std::ceil(p.days / h0) + 16;
        tr.reserve(est_steps * 7 * (size_t)op1);
        tm.reserve(est_steps + 1);
        double t0 = 0;
        tm.push_back(0.0);
        std::vector<double> x(op1), st(7), c(7 * (size_t)op1);
        while (t0 < p.days) {
            // a remainder shorter than a fifth of a step is absorbed by the last
            // step instead of becoming an ill-conditioned sliver
            double h = (p.days - t0 < 1.2 * h0) ? p.days - t0 : h0;
            for (;;) {
                for (int k = 0; k < op1; ++k) x[k] = kfrac[k] * h;
                for (int k = 0; k < op1; ++k) {
                    synth_state(el, t0 + x[k], st.data());
                    for (int ch = 0; ch < 7; ++ch) c[(size_t)ch * op1 + k] = st[ch];
                }
                for (int ch = 0; ch < 7; ++ch) bp_interp(x.data(), &c[(size_t)ch * op1], n);
                double err = 0;
                for (int k = 0; k + 1 < op1; ++k) {
                    const double xm = 0.5 * (x[k] + x[k + 1]);
                    synth_state(el, t0 + xm, st.data());
                    for (int ch = 0; ch < 3; ++ch)
                        err = std::max(err, std::fabs(horner(&c[(size_t)ch * op1], n, xm) - st[ch]));
                }
                if (err < 1e-7 || h <= hmin) break;
                h *= 0.5;
            }
            for (int k = 0; k < op1; ++k)
                for (int ch = 0; ch < 7; ++ch) tr.push_back(c[(size_t)ch * op1 + k]);
            const double t1 = (p.days - t0 <= h) ? p.days : t0 + h;
            tm.push_back(t1);
            t0 = t1;
        }
    }
    uint64_t toff = 0, moff = 0;
    out.traj_off.resize(N);
    out.nsteps.resize(N);
    out.time_off.resize(N);
    for (uint32_t o = 0; o < N; ++o) {
        out.traj_off[o] = toff;
        out.nsteps[o] = traj_o[o].size() / (7ull * op1);
        out.time_off[o] = moff;
        toff += traj_o[o].size();
        moff += time_o[o].size();
    }
    out.traj.resize(toff);
    out.time.resize(moff);
#pragma omp parallel for schedule(static)
    for (long long oi = 0; oi < (long long)N; ++oi) {
        std::copy(traj_o[oi].begin(), traj_o[oi].end(), out.traj.begin() + out.traj_off[oi]);
        std::copy(time_o[oi].begin(), time_o[oi].end(), out.time.begin() + out.time_off[oi]);
    }
    out.maxT = p.days;
    return out;
}

} // namespace mzh_host

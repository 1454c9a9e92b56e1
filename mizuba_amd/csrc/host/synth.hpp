// Seeded synthetic polyjectory generator (test and benchmark inputs).
//
// Orbits are Keplerian with the secular J2 drift of the node, the argument of
// perigee and the mean anomaly (WGS-72 constants, the ones SGP4 uses). Each
// trajectory is cut into steps of about a eighth of the orbital period and
// every step holds order-n monomial polynomials in (t - t_step) [days] for the
// 7 channels x, y, z [km], vx, vy, vz [km/s], r [km], fitted through the n+1
// Chebyshev nodes of the step with Bjorck-Pereyra interpolation; a step is
// halved until the position error at the node midpoints is below 1e-7 km.
// This mirrors the layout and step rule of make_sgp4_polyjectory
// (mizuba/make_sgp4_polyjectory.cpp:478-702) without its SGP4 dependency.
#pragma once

#include <cstdint>
#include <vector>

namespace mzh_host
{

struct SynthParams {
    uint32_t n_objs = 1000;
    double days = 1.0;
    uint64_t seed = 42;
    double frac_high = 0.0;   // fraction drawn from the high (1200-2000 km, e < 0.05) population
    uint32_t order = 8;
    // dense shells: n_shell objects on near-circular shells (0 = none)
    uint32_t n_shell = 0;
};

struct Elements {
    double a, e, inc, raan, argp, m0, bstar;
};

struct SynthOut {
    uint32_t order = 8;
    std::vector<double> traj;      // [S][op1][7]
    std::vector<double> time;      // per object nsteps + 1
    std::vector<uint64_t> traj_off;
    std::vector<uint64_t> nsteps;
    std::vector<uint64_t> time_off;
    std::vector<Elements> elements;
    double maxT = 0;
};

SynthOut make_synthetic(const SynthParams &p);

// Exact state of an element set at t [days] (x y z vx vy vz r).
void synth_state(const Elements &el, double t_days, double out[7]);

} // namespace mzh_host

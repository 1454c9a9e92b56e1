// TEST INFRASTRUCTURE ONLY. A C entry point over the reference's vendored copy
// of Vallado's SGP4 (mizuba/detail/sgp4/SGP4.cpp, compiled from where it lies
// by oracle/Makefile into oracle/_ref/), used by the tests as the checker of
// the GPU SGP4 propagation. It initialises each GPE the way
// make_sgp4_polyjectory does (make_sgp4_polyjectory.cpp:722-758: WGS-72, opsmode
// 'a', epoch in days since 1949 Dec 31 0h UTC, zero ndot / nddot) and
// propagates it to the given minutes since its epoch.
#include <cstdint>

#include "SGP4.h"

extern "C" {

// gpe: {norad_id, epoch_jd1, epoch_jd2, n0, e0, i0, node0, omega0, m0, bstar}
// (make_sgp4_polyjectory.hpp:35-47) as 10 doubles-sized fields; tsince[k] in
// minutes from the epoch; out[k] = {x, y, z, vx, vy, vz} (km, km/s, TEME);
// err[k] = the SGP4 error code (0: none). Returns the sgp4init error code.
int ref_sgp4_propagate(const double *gpe_fields, const double *tsince, int64_t n, double *out, int *err)
{
    const double jd_sub = 2433281.5;
    const double epoch = (gpe_fields[1] - jd_sub) + gpe_fields[2];
    elsetrec sat;
    SGP4Funcs::sgp4init(wgs72, 'a', "", epoch, gpe_fields[9], 0., 0., gpe_fields[4], gpe_fields[7], gpe_fields[5],
                        gpe_fields[8], gpe_fields[3], gpe_fields[6], sat);
    if (sat.error != 0) return sat.error;
    for (int64_t k = 0; k < n; ++k) {
        double r[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        SGP4Funcs::sgp4(sat, tsince[k], r, v);
        err[k] = sat.error;
        for (int i = 0; i < 3; ++i) {
            out[k * 6 + i] = r[i];
            out[k * 6 + 3 + i] = v[i];
        }
        sat.error = 0;
    }
    return 0;
}

// 1 when the GPE is a deep-space one (orbital period of 225 minutes or more).
int ref_sgp4_is_deep_space(const double *gpe_fields)
{
    const double jd_sub = 2433281.5;
    const double epoch = (gpe_fields[1] - jd_sub) + gpe_fields[2];
    elsetrec sat;
    SGP4Funcs::sgp4init(wgs72, 'a', "", epoch, gpe_fields[9], 0., 0., gpe_fields[4], gpe_fields[7], gpe_fields[5],
                        gpe_fields[8], gpe_fields[3], gpe_fields[6], sat);
    return sat.method == 'd' ? 1 : 0;
}
}

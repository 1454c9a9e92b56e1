// Host build of the AABB stage's polynomial range enclosures
// (mizuba_amd/csrc/mzh_enclosure.hpp: Cargo-Shisha in FP64, and in FP32 with
// its a-priori rounding margin), driven by tests/test_enclosure_host.py.
#include <cstdint>
#include <cstring>

#include "../mizuba_amd/csrc/mzh_enclosure.hpp"

namespace
{

constexpr mzh::BinomTable kB = mzh::make_binom_table();

template <int O>
int run(int order, uint64_t n, const double *cf, const double *lb, const double *ub, double *e64, float *e32)
{
    if (order == O) {
        int same = 0;
        for (uint64_t k = 0; k < n; ++k) {
            mzh::cs_enclosure_fp64<O>(cf + k * (O + 1), lb[k], ub[k], e64[2 * k], e64[2 * k + 1], kB);
            mzh::cs_enclosure_fp32<O>(cf + k * (O + 1), lb[k], ub[k], e32[2 * k], e32[2 * k + 1], kB);
            // the kernels' float copy of the coefficients (mixed-precision
            // AABBs) must give the same bits
            float cff[O + 1], a, b;
            for (int i = 0; i <= O; ++i) cff[i] = (float)cf[k * (O + 1) + i];
            mzh::cs_enclosure_fp32<O, 1, float>(cff, lb[k], ub[k], a, b, kB);
            same |= std::memcmp(&a, &e32[2 * k], sizeof(float)) | std::memcmp(&b, &e32[2 * k + 1], sizeof(float));
        }
        return same ? 2 : 0;
    }
    if constexpr (O < mzh::kMaxOrder) {
        return run<O + 1>(order, n, cf, lb, ub, e64, e32);
    } else {
        return 1;
    }
}

} // namespace

// n polynomials of the given order (coefficients in increasing degree, order + 1
// per polynomial) over [lb[k], ub[k]]: FP64 enclosures into e64, FP32 ones
// into e32 (pairs min, max). Returns 1 for an order without a build, 2 when
// the FP32 enclosure from the float-rounded coefficients differs in any bit
// from the one from the doubles.
extern "C" int enc_batch(int order, uint64_t n, const double *cf, const double *lb, const double *ub, double *e64,
                         float *e32)
{
    return run<mzh::kMinOrder>(order, n, cf, lb, ub, e64, e32);
}

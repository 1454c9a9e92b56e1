// Polynomial range enclosures of the AABB stage (Cargo-Shisha, FP64 and FP32
// with a rigorous margin) and the constants they use. Host and device code:
// the kernels include this through mzh_device.hpp, and a CPU test compiles it
// on its own to check the FP32 enclosure against the FP64 one.
//
// Compiled with -ffp-contract=off on both sides: every multiply and add is
// rounded separately.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MZH_HD __host__ __device__
#else
#define MZH_HD
#endif

namespace mzh
{

constexpr int kMaxOrder = 20;  // highest polynomial order with a compiled kernel
constexpr int kMinOrder = 2;   // polyjectory.cpp:434-438 rejects order < 2

// Binomial coefficients C(n, k) for n, k <= kMaxOrder, exact in double.
struct BinomTable {
    double v[kMaxOrder + 1][kMaxOrder + 1];
};

constexpr BinomTable make_binom_table()
{
    BinomTable t{};
    for (int n = 0; n <= kMaxOrder; ++n) {
        for (int k = 0; k <= kMaxOrder; ++k) {
            if (k > n) {
                t.v[n][k] = 0.0;
            } else {
                unsigned long long r = 1;
                const int kk = (k > n - k) ? n - k : k;
                for (int i = 1; i <= kk; ++i) r = r * (unsigned long long)(n - kk + i) / (unsigned long long)i;
                t.v[n][k] = (double)r;
            }
        }
    }
    return t;
}

// a^0..a^ORDER with the halving scheme of conjunctions_jit.cpp:83-95.
template <int ORDER, typename T>
MZH_HD inline void halving_pows(T a, T *p)
{
    p[0] = T(1);
    if constexpr (ORDER >= 1) p[1] = a;
#pragma unroll
    for (int i = 2; i <= ORDER; ++i) p[i] = (i % 2 == 0) ? p[i / 2] * p[i / 2] : p[i / 2 + 1] * p[i / 2];
}

MZH_HD inline double hy_min(double a, double b) { return (b == b) ? ((b < a) ? b : a) : b; }
MZH_HD inline double hy_max(double a, double b) { return (b == b) ? ((a < b) ? b : a) : b; }

template <int N>
MZH_HD inline double pw_min(double *v)
{
    if constexpr (N == 1) {
        return v[0];
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) v[i] = hy_min(v[2 * i], v[2 * i + 1]);
        if constexpr (N % 2) v[N / 2] = v[N - 1];
        return pw_min<(N + 1) / 2>(v);
    }
}

template <int N>
MZH_HD inline double pw_max(double *v)
{
    if constexpr (N == 1) {
        return v[0];
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) v[i] = hy_max(v[2 * i], v[2 * i + 1]);
        if constexpr (N % 2) v[N / 2] = v[N - 1];
        return pw_max<(N + 1) / 2>(v);
    }
}

// FP64 Cargo-Shisha enclosure of one polynomial (contiguous coefficients)
// over [lb, ub]: translate by lb, scale by (ub - lb)^i, divide by C(n, i),
// difference table, NaN-propagating pairwise min/max.
template <int ORDER, int STRIDE = 1>
MZH_HD inline void cs_enclosure_fp64(const double *__restrict__ cf, double lb, double ub, double &emin,
                                     double &emax, const BinomTable &B)
{
    constexpr int OP1 = ORDER + 1;
    double c[OP1], t[OP1], ap[OP1], dp[OP1];
#pragma unroll
    for (int i = 0; i < OP1; ++i) c[i] = cf[i * STRIDE];
    halving_pows<ORDER>(lb, ap);
#pragma unroll
    for (int i = 0; i < OP1; ++i) {
        double s = (c[i] * B.v[i][0]) * ap[0];
#pragma unroll
        for (int k = i + 1; k < OP1; ++k) s = s + (c[k] * B.v[k][k - i]) * ap[k - i];
        t[i] = s;
    }
    halving_pows<ORDER>(ub - lb, dp);
#pragma unroll
    for (int i = 0; i < OP1; ++i) t[i] = t[i] * dp[i];
#pragma unroll
    for (int i = 0; i < OP1; ++i) t[i] = t[i] / B.v[ORDER][i];
    double bmin[OP1], bmax[OP1];
    bmin[0] = bmax[0] = t[0];
#pragma unroll
    for (int r = 1; r <= ORDER; ++r) {
#pragma unroll
        for (int k = 0; k + r <= ORDER; ++k) t[k] = t[k] + t[k + 1];
        bmin[r] = bmax[r] = t[0];
    }
    emin = pw_min<OP1>(bmin);
    emax = pw_max<OP1>(bmax);
}

// FP32 Cargo-Shisha enclosure with a rigorous margin: the same algorithm in
// single precision on coefficients rounded to float, widened by an a-priori
// bound on every rounding error made (coefficient rounding included), so that
// [emin, emax] contains the exact range of the FP64 polynomial and the
// result of the FP64 enclosure.
//
// Error model: with u = 2^-24 every float operation returns x(1 + d), |d| <= u.
// Bernstein coefficients b_r are sums with non-negative weights of
// t_i = c'_i (ub - lb)^i / C(n, i), so |fl(b_r) - b_r| <= gamma_K * B where
// B = sum_i |t_i| C(r, i) bounds the magnitude of the whole computation and K
// counts the rounding steps on any path (translation: 3 + order, scaling: order,
// division: 1, difference table: order). The same bound covers the FP64
// enclosure's own (smaller) errors, so the float interval [b - e, b + e]
// contains the FP64 Bernstein coefficient. The margin is evaluated from
// magnitudes accumulated alongside the values and inflated by a factor 2.
// The coefficients may come as doubles or already rounded to float (CT =
// float): each is used as (float)cf[i], so both give the same result.
template <int ORDER, int STRIDE = 1, typename CT = double>
MZH_HD inline void cs_enclosure_fp32(const CT *__restrict__ cf, double lb_d, double ub_d, float &emin,
                                     float &emax, const BinomTable &B)
{
    constexpr int OP1 = ORDER + 1;
    const float lb = (float)lb_d;
    const float dl = (float)(ub_d - lb_d);
    float t[OP1], m[OP1], ap[OP1], dp[OP1];
    halving_pows<ORDER>(lb, ap);
    halving_pows<ORDER>(dl, dp);
    float cfl[OP1];
#pragma unroll
    for (int i = 0; i < OP1; ++i) cfl[i] = (float)cf[i * STRIDE];
#pragma unroll
    for (int i = 0; i < OP1; ++i) {
        float s = cfl[i], ms = __builtin_fabsf(cfl[i]);
#pragma unroll
        for (int k = i + 1; k < OP1; ++k) {
            const float w = (float)B.v[k][k - i] * ap[k - i];
            s = s + cfl[k] * w;
            ms = ms + __builtin_fabsf(cfl[k]) * __builtin_fabsf(w);
        }
        const float sc = dp[i] / (float)B.v[ORDER][i];
        t[i] = s * sc;
        m[i] = ms * __builtin_fabsf(sc);
    }
    float bmin = t[0], bmax = t[0];
    float mag = m[0];
#pragma unroll
    for (int r = 1; r <= ORDER; ++r) {
#pragma unroll
        for (int k = 0; k + r <= ORDER; ++k) {
            t[k] = t[k] + t[k + 1];
            m[k] = m[k] + m[k + 1];
        }
        // NaN-propagating min/max, as in the FP64 enclosure
        bmin = (t[0] == t[0]) ? (t[0] < bmin ? t[0] : bmin) : t[0];
        bmax = (t[0] == t[0]) ? (bmax < t[0] ? t[0] : bmax) : t[0];
        mag = m[0] > mag ? m[0] : mag;
    }
    // Rounding steps on any path, relative to the magnitude sum: coefficient
    // and parameter conversion (3), powers of lb and dl (2 * order each, the
    // inputs' own rounding included), binomial products and the translation
    // sum (order + 2), scaling and division (2), difference table (order).
    // The FP64 enclosure's error is bounded by the same count at u = 2^-53 and
    // is absorbed by the factor 2.
    constexpr float kSteps = (float)(6 * ORDER + 8);
    constexpr float kU = 5.9604645e-08f;
    const float gamma = kSteps * kU / (1.0f - kSteps * kU);
    const float margin = 2.0f * gamma * mag + 1.0e-30f;
    emin = bmin - margin;
    emax = bmax + margin;
}

} // namespace mzh

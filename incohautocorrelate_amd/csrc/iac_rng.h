// iac_rng.h -- counter-based random numbers and the Poisson sampler of the
// simulator post-processing (ArrayOperators::ScalarPoissonSampling,
// IAC/ArrayOperators.h:25-55), for device and host.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) gives every (seed, frame, pixel) its own uniforms,
// so sampling is order free.  exp, log and cos are evaluated from their
// series (range reduction + Taylor / atanh series) with plain double
// arithmetic: no library transcendental is involved, so any IEEE double
// implementation without contraction produces the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace iac {

__host__ __device__ inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Two 53-bit uniforms in [0, 1) for (seed, frame, pixel, stream); stream 0 feeds
// the Poisson draw, stream 1 the white noise.
__host__ __device__ inline void uniform2(uint64_t seed, uint32_t frame, uint32_t pixel, uint32_t stream, double u[2])
{
    const uint32_t ctr[4] = {pixel, frame, 0x49414350u, stream};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    u[0] = ((double)(r[0] >> 5) * 67108864.0 + (double)(r[1] >> 6)) * (1.0 / 9007199254740992.0);
    u[1] = ((double)(r[2] >> 5) * 67108864.0 + (double)(r[3] >> 6)) * (1.0 / 9007199254740992.0);
}

// e^x: x = k ln2 + r (ln2 in a 32-bit head and a tail), |r| <= ln2/2, Taylor to r^13.
__host__ __device__ inline double series_exp(double x)
{
    if (x < -708.0) return 0.0;
    if (x > 709.0) return __builtin_inf();
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double log2e = 1.44269504088896338700e+00;
    const double k = floor(x * log2e + 0.5);
    const double r = (x - k * ln2_hi) - k * ln2_lo;
    const double inv_fact[13] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,
                                 1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,     1.0 / 120.0,
                                 1.0 / 24.0,        1.0 / 6.0,        1.0 / 2.0,       1.0,
                                 1.0};
    double p = 1.0 / 6227020800.0;
    for (int i = 0; i < 13; i++) p = p * r + inv_fact[i];
    return ldexp(p, (int)k);
}

// log x, x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1).
__host__ __device__ inline double series_log(double x)
{
    if (!(x > 0)) return -__builtin_inf();
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double t = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2) t = t * s2 + 1.0 / k;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return (e * ln2_hi + 2.0 * s * t) + e * ln2_lo;
}

// cos x for moderate |x|: x = n pi/2 + r, |r| <= pi/4, Taylor series of cos / sin.
__host__ __device__ inline double series_cos(double x)
{
    const double pio2_hi = 1.57079632673412561417e+00;
    const double pio2_lo = 6.07710050650619224932e-11;
    const double twoopi = 6.36619772367581382433e-01;
    const double n = floor(x * twoopi + 0.5);
    const double r = (x - n * pio2_hi) - n * pio2_lo;
    const double r2 = r * r;
    const double cf[8] = {-1.0 / 87178291200.0, 1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0,
                          -1.0 / 720.0,         1.0 / 24.0,        -1.0 / 2.0,       1.0};
    const double sf[8] = {-1.0 / 1307674368000.0, 1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0,
                          -1.0 / 5040.0,          1.0 / 120.0,        -1.0 / 6.0,        1.0};
    double c = 1.0 / 20922789888000.0;
    double sn = 1.0 / 355687428096000.0;
    for (int i = 0; i < 8; i++) c = c * r2 + cf[i];
    for (int i = 0; i < 8; i++) sn = sn * r2 + sf[i];
    sn = sn * r;
    const long q = ((long)n) & 3;
    if (q == 0) return c;
    if (q == 1) return -sn;
    if (q == 2) return -c;
    return sn;
}

// ScalarPoissonSampling (IAC/ArrayOperators.h:25-55): inversion for mean < 25,
// Box-Muller otherwise; u[0] is the first Drand() draw, u[1] the second.  The
// search stops after 4096 steps (a draw above the rounded CDF total).
__host__ __device__ inline uint32_t poisson_sample(double mean, const double u[2])
{
    double ret = 0;
    if (mean < 25) {
        int i = 0;
        double r = series_exp(-mean);
        double p = r;
        const double q = u[0];
        while (p < q && i < 4096) {
            i++;
            r *= mean / i;
            p += r;
        }
        ret = i;
    } else {
        const double a = 1 - u[0];
        const double b = 1 - u[1];
        ret = (mean + sqrt(mean) * sqrt(-2. * series_log(a)) * series_cos(2. * 3.14159265358979323846 * b));
        if (ret < 0) ret = 0;
    }
    return (uint32_t)ret;
}

}  // namespace iac

// synth_math.h - the wavelet arithmetic of the synthetic database generator, shared by the
// host generator (synth.cpp, g++) and the device generator (synth_gpu.hip, hipcc).
//
// exp, log, cos and sqrt are written here from +, -, *, / and exact scalings by powers of
// two only, evaluated in a fixed order (both translation units are built with
// -ffp-contract=off, and IEEE double +, -, *, / are correctly rounded on the host and on
// gfx950), so the two generators produce the same float for every sample.  The
// approximations are accurate to a few units in the last place of a double, far below the
// float rounding of the stored traces.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SYNTH_HD __host__ __device__
#else
#define SYNTH_HD
#endif

namespace synth_math {

constexpr double kPi = 3.1415926535898;        // global_parameters.f90:40
constexpr double kTwoPiFull = 6.283185307179586;
constexpr double kLn2Hi = 6.93147180369123816490e-01;   // 32 significant bits: k * hi exact
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;

SYNTH_HD inline double pow2i(int k) {
    // 2^k for -1022 <= k <= 1023, built from its exponent bits
    uint64_t bits = (uint64_t)(k + 1023) << 52;
    double d;
    memcpy(&d, &bits, sizeof d);
    return d;
}

SYNTH_HD inline int round_int(double t) { return t < 0.0 ? (int)(t - 0.5) : (int)(t + 0.5); }

// e^x for -700 < x < 700: x = k ln2 + r (Cody-Waite), |r| <= ln2/2, Taylor series of
// degree 14 by Horner
SYNTH_HD inline double exp_d(double x) {
    const int k = round_int(x * kInvLn2);
    const double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 87178291200.0;                      // 1/14!
    const double inv_fact[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0,
                                 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};
    for (int i = 0; i < 14; ++i) p = p * r + inv_fact[i];
    return p * pow2i(k);
}

// natural logarithm for 0 < y < 2^1023: y = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1) / (m + 1), series to s^23
SYNTH_HD inline double log_d(double y) {
    uint64_t bits;
    memcpy(&bits, &y, sizeof bits);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;   // m in [1, 2)
    double m;
    memcpy(&m, &bits, sizeof m);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 23.0;
    for (int j = 21; j >= 1; j -= 2) p = p * s2 + 1.0 / (double)j;
    const double lm = 2.0 * s * p;
    return e * kLn2Hi + (e * kLn2Lo + lm);
}

// sqrt for a >= 0 by Newton's iteration from an exponent-halving first guess
SYNTH_HD inline double sqrt_d(double a) {
    if (!(a > 0.0)) return 0.0;
    uint64_t bits;
    memcpy(&bits, &a, sizeof bits);
    bits = (bits >> 1) + 0x1ff8000000000000ull;
    double y;
    memcpy(&y, &bits, sizeof y);
    for (int i = 0; i < 7; ++i) y = 0.5 * (y + a / y);
    return y;
}

// sin and cos of |t| <= pi/4 (Taylor series to t^19 and t^18)
SYNTH_HD inline double sin_small(double t) {
    const double t2 = t * t;
    double p = -1.0 / 121645100408832000.0;              // -1/19!
    const double c[9] = {1.0 / 355687428096000.0, -1.0 / 1307674368000.0, 1.0 / 6227020800.0,
                         -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0, 1.0 / 120.0, -1.0 / 6.0, 1.0};
    for (int i = 0; i < 9; ++i) p = p * t2 + c[i];
    return p * t;
}
SYNTH_HD inline double cos_small(double t) {
    const double t2 = t * t;
    double p = -1.0 / 6402373705728000.0;                // -1/18!
    const double c[9] = {1.0 / 20922789888000.0, -1.0 / 87178291200.0, 1.0 / 479001600.0,
                         -1.0 / 3628800.0, 1.0 / 40320.0, -1.0 / 720.0, 1.0 / 24.0, -0.5, 1.0};
    for (int i = 0; i < 9; ++i) p = p * t2 + c[i];
    return p;
}

// cos(2 pi q): q = n/4 + f with |f| <= 1/8, cos(n pi/2 + 2 pi f)
SYNTH_HD inline double cos_turns(double q) {
    const int n = round_int(q * 4.0);
    const double f = q - n * 0.25;
    const double t = kTwoPiFull * f;
    switch (n & 3) {
    case 0: return cos_small(t);
    case 1: return -sin_small(t);
    case 2: return -cos_small(t);
    default: return sin_small(t);
    }
}

// counter-based hash (splitmix64 finaliser) -> uniform double in [0, 1)
SYNTH_HD inline uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
SYNTH_HD inline double unif(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = mix64(seed ^ mix64(a ^ mix64(b ^ mix64(c))));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// wavelet w of the trace of (point, variable): amplitude N(0,1) (Box-Muller) and arrival
// time U(100, 900) s; samples i0..i1 (|t - tau| <= 240 s) carry it
struct Wavelet {
    double amp, tau;
    int i0, i1;
};
SYNTH_HD inline Wavelet wavelet(uint64_t seed, int64_t point, int var, int w, int nt, double dt) {
    const double u1 = unif(seed, (uint64_t)point, (uint64_t)var, 4 * w + 0);
    const double u2 = unif(seed, (uint64_t)point, (uint64_t)var, 4 * w + 1);
    const double u3 = unif(seed, (uint64_t)point, (uint64_t)var, 4 * w + 2);
    Wavelet W;
    W.amp = sqrt_d(-2.0 * log_d(1.0 - u1)) * cos_turns(u2);
    W.tau = 100.0 + 800.0 * u3;
    // first sample at or after tau - 240 and last at or before tau + 240
    const double a = (W.tau - 240.0) / dt, b = (W.tau + 240.0) / dt;
    int i0 = (int)a;
    if ((double)i0 < a) i0 += 1;
    int i1 = (int)b;
    W.i0 = i0 < 0 ? 0 : i0;
    W.i1 = i1 > nt - 1 ? nt - 1 : i1;
    return W;
}

// contribution of wavelet W at sample i: amp * exp(-(x/40)^2) * cos(2 pi x/40), x = i dt - tau
SYNTH_HD inline double wavelet_sample(const Wavelet &W, int i, double dt) {
    const double q = (i * dt - W.tau) / 40.0;
    return W.amp * exp_d(-(q * q)) * cos_turns(q);
}

}  // namespace synth_math

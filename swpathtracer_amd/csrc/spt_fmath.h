// spt_fmath.h - single-precision sin/cos, exp, log and atan built from float
// operations only (+, -, *, /, rint and exponent-field bit manipulation), so
// a host compiler without FMA contraction and a gfx950 kernel compiled with
// -ffp-contract=off produce the same bits for the same input. No FP64.
//
// Method (Cody and Waite, "Software Manual for the Elementary Functions",
// 1980, the argument reductions also used by fdlibm):
//   sin/cos: j = rint(x * 2/pi), r = x - j * pi/2 carried as r_hi + r_lo with
//            pi/2 split into two 16-bit parts and a float tail (j * part is
//            exact for |j| < 2^8); odd/even polynomials on |r| <= pi/4 with a
//            first-order correction for r_lo; quadrant j mod 4.
//   exp:     k = rint(x / ln 2), r = x - k ln 2 (16-bit high part, float
//            tail), e^r = 1 + r + r^2 E(r) on |r| <= ln(2)/2, times 2^k in
//            two exact power-of-two scalings (subnormal results round once).
//   log:     x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f / (2 + f),
//            log(1 + f) = f - (f^2/2 - s (f^2/2 + R(s^2))), plus e ln 2 with a
//            12-bit high part.
//   atan:    breakpoints 7/16, 11/16, 19/16, 39/16 reduce to
//            atan(c) + atan(u), |u| <= 7/16, c in {0, 1/2, 1, 3/2, inf},
//            atan(c) as a float pair.
// The polynomial coefficients are least-squares fits on Chebyshev nodes of
// the relative error, rounded to float (gen_fmath.py re-derives them).
// Measured on dense grids against float64: at most 0.78 / 0.80 ulp (sin /
// cos on |x| <= 100), 0.87 (exp, normal results), 0.84 (log), 0.84 (atan);
// the correctly rounded float is returned for 90-97% of inputs.
#pragma once

#include <stdint.h>
#include <string.h>

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SPT_FM_HD __host__ __device__ inline
#else
#define SPT_FM_HD inline
#endif

namespace spt {
namespace fm {

SPT_FM_HD uint32_t bits(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(x);
#else
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
#endif
}
SPT_FM_HD float from_bits(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float x;
    memcpy(&x, &u, 4);
    return x;
#endif
}
// round to nearest, ties to even (the default rounding mode on both sides)
SPT_FM_HD float rint(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_rintf(x);
#else
    return std::rint(x);
#endif
}
// 2^k for k in [-126, 127], exact
SPT_FM_HD float pow2(int k) { return from_bits((uint32_t)(k + 127) << 23); }

}  // namespace fm

// sin and cos of x (|x| < 2^31 for a meaningful result; the reduction is
// accurate while |rint(x * 2/pi)| < 2^8).
SPT_FM_HD void fm_sincos(float x, float& s_out, float& c_out) {
    const float j = fm::rint(x * 0.6366197466850281f);
    const float r1 = x - j * 1.570770263671875f;          // pi/2, first 16 bits: exact
    const float t = j * 2.6063062250614166259765625e-05f;  // next 16 bits: exact
    const float rh = r1 - t;
    const float rl = ((r1 - rh) - t) - j * 6.077094383272197e-11f;
    const float z = rh * rh;
    const float ps = -0.1666666716337204f +
                     z * (0.008333329111337662f +
                          z * (-0.0001983929832931608f + z * 2.7180080905964132e-06f));
    const float sn = rh + ((rh * z) * ps + (rl - (0.5f * z) * rl));
    const float pc = 0.0416666679084301f +
                     z * (-0.0013888883404433727f +
                          z * (2.4799432139843702e-05f + z * -2.720371981013159e-07f));
    const float hz = 0.5f * z;
    const float w = 1.0f - hz;
    const float cs = w + (((1.0f - w) - hz) + ((z * z) * pc - rh * rl));
    const int q = (int)(int64_t)j & 3;
    s_out = (q == 0) ? sn : (q == 1) ? cs : (q == 2) ? -sn : -cs;
    c_out = (q == 0) ? cs : (q == 1) ? -sn : (q == 2) ? -cs : sn;
}
SPT_FM_HD float fm_sin(float x) {
    float s, c;
    fm_sincos(x, s, c);
    return s;
}
SPT_FM_HD float fm_cos(float x) {
    float s, c;
    fm_sincos(x, s, c);
    return c;
}

SPT_FM_HD float fm_exp(float x) {
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    if (!(x == x)) return x;
    const float k = fm::rint(x * 1.4426950216293335f);
    const float rh = x - k * 0.693145751953125f;  // ln 2, first 16 bits: exact
    const float rl = k * 1.428606765330187e-06f;
    const float r = rh - rl;
    const float c = (rh - r) - rl;  // rounding of the reduction
    const float p = r * r *
                    (0.5f + r * (0.1666666567325592f +
                                 r * (0.041666269302368164f +
                                      r * (0.00833342969417572f +
                                           r * (0.0013946021208539605f + r * 0.00019826495554298162f)))));
    const float y = 1.0f + (r + (p + c));
    const int ki = (int)k;
    const int k1 = ki / 2;  // truncating halves keep both scalings in range
    const int k2 = ki - k1;
    return (y * fm::pow2(k1)) * fm::pow2(k2 < -126 ? -126 : k2);
}

SPT_FM_HD float fm_log(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    uint32_t b = fm::bits(x);
    int e = 0;
    if (b < 0x00800000u) {  // subnormal: scale by 2^23
        x = x * 8388608.0f;
        b = fm::bits(x);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    float m = fm::from_bits((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.4142135381698608f) {
        m = m * 0.5f;
        ++e;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float R = z * (0.6666666269302368f +
                         z * (0.4000034034252167f + z * (0.2853710949420929f + z * 0.23585404455661774f)));
    const float hfsq = 0.5f * f * f;
    const float ef = (float)e;
    return ef * 0.693115234375f + (f - (hfsq - (s * (hfsq + R) + ef * 3.194618329871446e-05f)));
}

SPT_FM_HD float fm_atan(float x) {
    if (!(x == x)) return x;
    const float ax = (x < 0.0f) ? -x : x;
    float u, hi = 0.0f, lo = 0.0f;
    bool reduced = true;
    if (ax < 0.4375f) {
        u = ax;
        reduced = false;
    } else if (ax < 0.6875f) {  // c = 1/2
        u = (2.0f * ax - 1.0f) / (2.0f + ax);
        hi = 0.46364760398864746f;
        lo = 5.01215868808913e-09f;
    } else if (ax < 1.1875f) {  // c = 1
        u = (ax - 1.0f) / (ax + 1.0f);
        hi = 0.7853981852531433f;
        lo = -2.1855694143368964e-08f;
    } else if (ax < 2.4375f) {  // c = 3/2
        u = (ax - 1.5f) / (1.0f + 1.5f * ax);
        hi = 0.9827937483787537f;
        lo = -2.5131424052915463e-08f;
    } else {  // c = inf
        u = -1.0f / ax;
        hi = 1.5707963705062866f;
        lo = -4.371138828673793e-08f;
    }
    const float z = u * u;
    const float p =
        u * z *
        (-0.3333333134651184f +
         z * (0.19999836385250092f +
              z * (-0.14280132949352264f +
                   z * (0.11021947860717773f + z * (-0.08351728320121765f + z * 0.04491681605577469f)))));
    const float r = reduced ? hi + ((p + lo) + u) : u + p;
    return (x < 0.0f) ? -r : r;
}

}  // namespace spt

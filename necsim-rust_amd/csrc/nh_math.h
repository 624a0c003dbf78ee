// nh_math.h -- the MathsCore of the HIP Independent algorithm.
//
// necsim-rust's ReproducibleMathsCore (necsim/impls/no-std/src/cogs/maths/
// reproducible.rs:5-52) forwards to crate `libm` 0.2.8 (Cargo.lock:953-954),
// which restates the FreeBSD msun / musl double-precision routines.  The
// functions below restate the same published algorithms (e_log.c, e_exp.c,
// k_sin.c, k_cos.c, __rem_pio2.c) so that host and gfx950 device produce the
// same bits.  Everything here is compiled with -ffp-contract=off: a*b+c is
// never fused, like in Rust.
//
// Names carry an nh_ prefix so that clang never substitutes its own __ocml_*
// implementations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NH_HD __host__ __device__ __forceinline__

namespace nh {

NH_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
NH_HD double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }

// ---- Rust `as` casts: saturating, NaN -> 0 -------------------------------
NH_HD uint64_t as_u64(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)x;
}
NH_HD uint32_t as_u32(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)x;
}
NH_HD int64_t as_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
NH_HD int32_t as_i32(double x) {
    if (x != x) return 0;
    if (x >= 2147483648.0) return INT32_MAX;
    if (x < -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

// ---- exact IEEE operations -------------------------------------------------
// floor / round-half-away-from-zero / sqrt are exactly specified by IEEE 754,
// so the builtins (v_floor_f64, trunc-based round, the correctly rounded
// f64 sqrt expansion) are bit-identical to libm's.
NH_HD double floor_(double x) { return __builtin_floor(x); }
NH_HD double round_(double x) { return __builtin_round(x); }
NH_HD double sqrt_(double x) { return __builtin_sqrt(x); }

// ---- log: x = 2^k (1+f), log(1+f) = f - hfsq + s (hfsq + R(s^2)) ----------
NH_HD double ln(double x) {
    const double ln2_hi = 6.93147180369123816490e-01; /* 3fe62e42 fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10; /* 3dea39ef 35793c76 */
    const double Lg1 = 6.666666666666735130e-01;
    const double Lg2 = 3.999999999940941908e-01;
    const double Lg3 = 2.857142874366239149e-01;
    const double Lg4 = 2.222219843214978396e-01;
    const double Lg5 = 1.818357216161805012e-01;
    const double Lg6 = 1.531383769920937332e-01;
    const double Lg7 = 1.479819860511658591e-01;
    uint64_t ui = bits(x);
    uint32_t hx = (uint32_t)(ui >> 32);
    int32_t k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((ui << 1) == 0) return -1.0 / (x * x);
        if (hx >> 31) return (x - x) / 0.0;
        k -= 54;
        x *= from_bits(0x4350000000000000ull);
        ui = bits(x);
        hx = (uint32_t)(ui >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (ui << 32) == 0) {
        return 0.0;
    }
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int32_t)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    x = from_bits(((uint64_t)hx << 32) | (ui & 0xffffffffull));
    double f = x - 1.0;
    double hfsq = 0.5 * f * f;
    double s = f / (2.0 + f);
    double z = s * s;
    double w = z * z;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    double R = t2 + t1;
    double dk = (double)k;
    return s * (hfsq + R) + dk * ln2_lo - hfsq + f + dk * ln2_hi;
}

NH_HD double scalbn(double x, int n) {
    double y = x;
    if (n > 1023) {
        y *= from_bits(0x7fe0000000000000ull);
        n -= 1023;
        if (n > 1023) {
            y *= from_bits(0x7fe0000000000000ull);
            n -= 1023;
            if (n > 1023) n = 1023;
        }
    } else if (n < -1022) {
        y *= from_bits(0x0010000000000000ull) * from_bits(0x4340000000000000ull);
        n += 1022 - 53;
        if (n < -1022) {
            y *= from_bits(0x0010000000000000ull) * from_bits(0x4340000000000000ull);
            n += 1022 - 53;
            if (n < -1022) n = -1022;
        }
    }
    return y * from_bits((uint64_t)(0x3ff + n) << 52);
}

// ---- exp: x = k ln2 + r, exp(r) = 1 + r + r c / (2 - c) ---------------------
NH_HD double exp(double x) {
    const double ln2hi = 6.93147180369123816490e-01;
    const double ln2lo = 1.90821492927058770002e-10;
    const double invln2 = 1.44269504088896338700e+00;
    const double P1 = 1.66666666666666019037e-01;
    const double P2 = -2.77777777770155933842e-03;
    const double P3 = 6.61375632143793436117e-05;
    const double P4 = -1.65339022054652515390e-06;
    const double P5 = 4.13813679705723846039e-08;
    uint32_t hx = (uint32_t)(bits(x) >> 32);
    int32_t sign = (int32_t)(hx >> 31);
    hx &= 0x7fffffffu;
    double hi, lo;
    int32_t k;
    if (hx >= 0x4086232bu) {
        if (x != x) return x;
        if (x > 709.782712893383973096) return x * from_bits(0x7fe0000000000000ull);
        if (x < -745.13321910194110842) return 0.0;
    }
    if (hx > 0x3fd62e42u) {
        if (hx >= 0x3ff0a2b2u) k = as_i32(invln2 * x + (sign ? -0.5 : 0.5));
        else k = 1 - sign - sign;
        hi = x - (double)k * ln2hi;
        lo = (double)k * ln2lo;
        x = hi - lo;
    } else if (hx > 0x3e300000u) {
        k = 0;
        hi = x;
        lo = 0.0;
    } else {
        return 1.0 + x;
    }
    double xx = x * x;
    double c = x - xx * (P1 + xx * (P2 + xx * (P3 + xx * (P4 + xx * P5))));
    double y = 1.0 + (x * c / (2.0 - c) - lo + hi);
    return k == 0 ? y : scalbn(y, k);
}

// ---- sin / cos kernels on [-pi/4, pi/4] -----------------------------------
NH_HD double k_sin(double x, double y, int iy) {
    const double S1 = -1.66666666666666324348e-01;
    const double S2 = 8.33333333332248946124e-03;
    const double S3 = -1.98412698298579493134e-04;
    const double S4 = 2.75573137070700676789e-06;
    const double S5 = -2.50507602534068634195e-08;
    const double S6 = 1.58969099521155010221e-10;
    double z = x * x;
    double w = z * z;
    double r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6);
    double v = z * x;
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

NH_HD double k_cos(double x, double y) {
    const double C1 = 4.16666666666666019037e-02;
    const double C2 = -1.38888888888741095749e-03;
    const double C3 = 2.48015872894767294178e-05;
    const double C4 = -2.75573143513906633035e-07;
    const double C5 = 2.08757232129817482790e-09;
    const double C6 = -1.13596475577881948265e-11;
    double z = x * x;
    double w = z * z;
    double r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6));
    double hz = 0.5 * z;
    double ww = 1.0 - hz;
    return ww + (((1.0 - ww) - hz) + (z * r - x * y));
}

// ---- x - n pi/2 = y0 + y1 for |x| < 2^20 pi/2 ------------------------------
// The Independent algorithm only evaluates sin/cos at Box-Muller angles in
// (-2 pi, 0], far inside this range; larger arguments return n = 0 and NaN
// remainders (the Payne-Hanek path is not part of this library).
NH_HD int32_t rem_pio2_medium(double x, uint32_t ix, double &y0o, double &y1o) {
    const double toint = 1.5 / 2.220446049250313080847e-16;
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    const double pio2_2 = 6.07710050630396597660e-11;
    const double pio2_2t = 2.02226624879595063154e-21;
    const double pio2_3 = 2.02226624871116645580e-21;
    const double pio2_3t = 8.47842766036889956997e-32;
    double fn = x * invpio2 + toint - toint;
    int32_t n = as_i32(fn);
    double r = x - fn * pio2_1;
    double w = fn * pio2_1t;
    double y0 = r - w;
    int32_t ey = (int32_t)((bits(y0) >> 52) & 0x7ff);
    int32_t ex = (int32_t)(ix >> 20);
    if (ex - ey > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y0 = r - w;
        ey = (int32_t)((bits(y0) >> 52) & 0x7ff);
        if (ex - ey > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y0 = r - w;
        }
    }
    y0o = y0;
    y1o = (r - y0) - w;
    return n;
}

NH_HD int32_t rem_pio2(double x, double &y0, double &y1) {
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    uint64_t u = bits(x);
    int sign = (int)(u >> 63);
    uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    double z;
    if (ix <= 0x400f6a7au) {
        if ((ix & 0xfffffu) == 0x921fbu) return rem_pio2_medium(x, ix, y0, y1);
        if (ix <= 0x4002d97cu) {
            if (!sign) { z = x - pio2_1; y0 = z - pio2_1t; y1 = (z - y0) - pio2_1t; return 1; }
            z = x + pio2_1; y0 = z + pio2_1t; y1 = (z - y0) + pio2_1t; return -1;
        }
        if (!sign) { z = x - 2 * pio2_1; y0 = z - 2 * pio2_1t; y1 = (z - y0) - 2 * pio2_1t; return 2; }
        z = x + 2 * pio2_1; y0 = z + 2 * pio2_1t; y1 = (z - y0) + 2 * pio2_1t; return -2;
    }
    if (ix <= 0x401c463bu) {
        if (ix <= 0x4015fdbcu) {
            if (ix == 0x4012d97cu) return rem_pio2_medium(x, ix, y0, y1);
            if (!sign) { z = x - 3 * pio2_1; y0 = z - 3 * pio2_1t; y1 = (z - y0) - 3 * pio2_1t; return 3; }
            z = x + 3 * pio2_1; y0 = z + 3 * pio2_1t; y1 = (z - y0) + 3 * pio2_1t; return -3;
        }
        if (ix == 0x401921fbu) return rem_pio2_medium(x, ix, y0, y1);
        if (!sign) { z = x - 4 * pio2_1; y0 = z - 4 * pio2_1t; y1 = (z - y0) - 4 * pio2_1t; return 4; }
        z = x + 4 * pio2_1; y0 = z + 4 * pio2_1t; y1 = (z - y0) + 4 * pio2_1t; return -4;
    }
    if (ix < 0x413921fbu) return rem_pio2_medium(x, ix, y0, y1);
    y0 = y1 = x - x; // inf/NaN -> NaN; huge finite -> unsupported, NaN
    if (ix < 0x7ff00000u) y0 = y1 = __builtin_nan("");
    return 0;
}

// sin(x) and cos(x) of the same argument.  The reference calls M::sin and
// M::cos separately (necsim/core/src/cogs/rng.rs:188); both reduce x with the
// same __rem_pio2, so sharing the reduction gives the same bits.
NH_HD void sincos(double x, double &s, double &c) {
    uint32_t ix = (uint32_t)(bits(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        s = ix < 0x3e500000u ? x : k_sin(x, 0.0, 0);
        c = (ix < 0x3e46a09eu && as_i32(x) == 0) ? 1.0 : k_cos(x, 0.0);
        return;
    }
    if (ix >= 0x7ff00000u) {
        s = c = x - x;
        return;
    }
    double y0, y1;
    int32_t n = rem_pio2(x, y0, y1);
    double ks = k_sin(y0, y1, 1);
    double kc = k_cos(y0, y1);
    switch (n & 3) {
    case 0: s = ks; c = kc; break;
    case 1: s = kc; c = -ks; break;
    case 2: s = -ks; c = -kc; break;
    default: s = -kc; c = ks; break;
    }
}

// The same sin / cos as one straight-line sequence for the arguments where
// the case-by-case code has several paths: __rem_pio2's cases for
// |x| <= 9 pi / 4 become one arithmetic sequence with the multiple k of pi/2
// and its constants picked by selects (x + k pio2_1 is computed as
// x - (-k pio2_1), the same IEEE operation), |x| <= pi/4 takes y0 = x,
// y1 = 0, n = 0, and the two kernels run once; only the medium reduction
// (near multiples of pi/2, |x| > 9 pi / 4) is a branch.  The bits equal
// sincos's.  A lone lane (the serial tail of a coalescing simulation) then
// runs one basic block that the compiler interleaves with the logarithm of
// the same Box-Muller draw.  A full wave of angles takes every branch of the
// case-by-case code, so the pooled round kernel (whose execution passes run
// full waves) uses this form too: 5 % less time per launch on the
// almost-infinite scenario.
NH_HD void sincos_flat(double x, double &s, double &c) {
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    const uint64_t u = bits(x);
    const bool neg = (u >> 63) != 0;
    const uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    const bool small = ix <= 0x3fe921fbu;
    const bool medium = ix > 0x401c463bu || (ix <= 0x400f6a7au && (ix & 0xfffffu) == 0x921fbu) ||
                        ix == 0x4012d97cu || ix == 0x401921fbu;
    const int32_t k = ix <= 0x4002d97cu ? 1 : ix <= 0x400f6a7au ? 2 : ix <= 0x4015fdbcu ? 3 : 4;
    const double ck = k == 1 ? pio2_1 : k == 2 ? 2 * pio2_1 : k == 3 ? 3 * pio2_1 : 4 * pio2_1;
    const double tk = k == 1 ? pio2_1t : k == 2 ? 2 * pio2_1t : k == 3 ? 3 * pio2_1t : 4 * pio2_1t;
    const double cs = neg ? -ck : ck, ts = neg ? -tk : tk;
    const double z = x - cs;
    double y0 = z - ts;
    double y1 = (z - y0) - ts;
    int32_t n = neg ? -k : k;
    if (small) {
        y0 = x;
        y1 = 0.0;
        n = 0;
    } else if (medium && ix < 0x7ff00000u) {
        n = rem_pio2(x, y0, y1);
    }
    // k_sin (iy = 0 for the unreduced small argument) and k_cos
    const double S1 = -1.66666666666666324348e-01;
    const double S2 = 8.33333333332248946124e-03;
    const double S3 = -1.98412698298579493134e-04;
    const double S4 = 2.75573137070700676789e-06;
    const double S5 = -2.50507602534068634195e-08;
    const double S6 = 1.58969099521155010221e-10;
    const double zz = y0 * y0;
    const double w = zz * zz;
    const double r = S2 + zz * (S3 + zz * S4) + zz * w * (S5 + zz * S6);
    const double v = zz * y0;
    double ks = small ? y0 + v * (S1 + zz * r) : y0 - ((zz * (0.5 * y1 - v * r) - y1) - v * S1);
    double kc = k_cos(y0, y1);
    if (small) {
        if (ix < 0x3e500000u) ks = x;
        if (ix < 0x3e46a09eu && as_i32(x) == 0) kc = 1.0;
    }
    // quadrant n & 3: (s, c) = (ks, kc), (kc, -ks), (-ks, -kc), (-kc, ks)
    const bool odd = (n & 1) != 0;
    s = odd ? kc : ks;
    c = odd ? ks : kc;
    if (n & 2) s = -s;
    if ((n + 1) & 2) c = -c;
    if (ix >= 0x7ff00000u) s = c = x - x;
}

NH_HD double sin(double x) { double s, c; sincos(x, s, c); return s; }
NH_HD double cos(double x) { double s, c; sincos(x, s, c); return c; }

// PositiveF64::max_after (necsim/core/bond/src/positive_f64.rs:103-113)
NH_HD double max_after(double before, double value) {
    if (value > before || before != before) return value;
    if (before == __builtin_inf()) return __builtin_inf();
    return from_bits(bits(before) + 1);
}

// Repeated sequential sum `x + x + ... + x` (n terms, folded from 0.0), the
// value of the water-level event-rate sum over n lineages with a uniform rate
// (parallelisation/independent/monolithic/mod.rs:150-164), equal bit for bit
// to the fold but computed binade by binade.  While s and s + x stay in one
// binade [2^(e-1), 2^e) with ulp u, every addition adds the same multiple D
// of u: the multiple nearest to x, since s is a multiple of u.  (When x / u
// is a half-integer, ties go to the even neighbour; the increment is then
// even after at most one step, which keeps the parity of s / u and the
// increment fixed.)  So a run of steps inside a binade is one integer
// multiply-add; a step that reaches the next binade, and a step whose
// successor adds a different increment, are taken one at a time.  A fold of
// 12.6M terms takes milliseconds of host time per water-level round; this
// takes a few dozen iterations.
inline double repeated_sum_fold(double x, uint64_t n) {
    double s = 0.0;
    for (uint64_t j = 0; j < n; j++) s = s + x;
    return s;
}

inline double repeated_sum(double x, uint64_t n) {
    if (n == 0) return 0.0;
    if (!(x > 0.0) || x == __builtin_inf()) return repeated_sum_fold(x, n);
    const uint64_t two53 = 1ull << 53;
    double s = x; // 0.0 + x
    uint64_t k = 1;
    while (k < n) {
        int e = 0;
        (void)__builtin_frexp(s, &e);          // s in [2^(e-1), 2^e)
        if (e < -1021) e = -1021;              // subnormals share the ulp 2^-1074
        const double u = __builtin_ldexp(1.0, e - 53);
        const uint64_t S = (uint64_t)(s / u);  // exact: s is a multiple of u, below 2^53 u
        const uint64_t xi = (uint64_t)__builtin_floor(x / u);
        // the steps j = 1, 2, ... stay in the binade while S + (j-1) D + x / u < 2^53,
        // i.e. S + (j-1) D <= B
        const uint64_t B = two53 - 1 - xi;
        const double s1 = s + x;
        if (S > B) { // this step reaches the next binade
            s = s1;
            k++;
            continue;
        }
        const double d1 = s1 - s; // exact (x <= s, so s <= s1 <= 2s)
        if (d1 == 0.0) return s;  // x is below half an ulp: s no longer changes
        const uint64_t D = (uint64_t)(d1 / u);
        if (S + D <= B && (s1 + x) - s1 != d1) { // the tie's parity settles first
            s = s1;
            k++;
            continue;
        }
        uint64_t m = (B - S) / D + 1;
        if (m > n - k) m = n - k;
        s = (double)(S + m * D) * u; // at most 2^53 units: exact
        k += m;
    }
    return s;
}

} // namespace nh

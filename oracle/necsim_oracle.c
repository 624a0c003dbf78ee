/*
 * necsim_oracle.c -- CPU restatement of necsim-rust's Independent algorithm.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline), see
 * necsim_oracle.h.  Written as a literal restatement of the reference's Rust
 * code paths, in the same order of floating-point operations, so that it is a
 * bit-exact stand-in for the reference CPU Independent algorithm run with
 * ReproducibleMathsCore (crate libm 0.2.8) and WyHash.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (see Makefile).  Rust never
 * contracts a*b+c into an FMA and neither may this file.
 */
#include "necsim_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ======================================================================== */
/* Rust numeric semantics                                                   */
/* ======================================================================== */

static inline uint64_t bits_of(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double of_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* `f64 as u64`: saturating, NaN -> 0 */
static uint64_t f64_as_u64(double x) {
    if (!(x > 0.0)) return 0; /* NaN, <= 0 (incl. -0.0) */
    if (x >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)x;
}
/* `f64 as u32`: saturating, NaN -> 0 */
static uint32_t f64_as_u32(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 4294967296.0) return UINT32_MAX;
    return (uint32_t)x;
}
/* `f64 as i64`: saturating, NaN -> 0 */
static int64_t f64_as_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
/* `f64 as i32`: saturating, NaN -> 0 */
static int32_t f64_as_i32(double x) {
    if (x != x) return 0;
    if (x >= 2147483648.0) return INT32_MAX;
    if (x < -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

/* ======================================================================== */
/* MathsCore: ReproducibleMathsCore                                          */
/* necsim/impls/no-std/src/cogs/maths/reproducible.rs:5-52 calls crate       */
/* libm 0.2.8 (Cargo.lock:953-954).  That crate restates the FreeBSD msun /  */
/* musl algorithms (e_log.c, e_exp.c, k_sin.c, k_cos.c, __rem_pio2.c), which */
/* are restated below from their published form.                            */
/* ======================================================================== */

/* floor, round (half away from zero), sqrt: exact IEEE operations.
 * reproducible.rs:18-20 (floor), :38-40 (sqrt), :48-50 (round) */
double orc_floor(double x) { return floor(x); }
double orc_round(double x) { return round(x); }
double orc_sqrt(double x) { return sqrt(x); }

/* log (msun e_log.c method: x = 2^k (1+f), log(1+f) = f - hfsq + s(hfsq+R)) */
static const double ln2_hi = 6.93147180369123816490e-01; /* 3fe62e42 fee00000 */
static const double ln2_lo = 1.90821492927058770002e-10; /* 3dea39ef 35793c76 */
static const double Lg1 = 6.666666666666735130e-01;  /* 3FE55555 55555593 */
static const double Lg2 = 3.999999999940941908e-01;  /* 3FD99999 9997FA04 */
static const double Lg3 = 2.857142874366239149e-01;  /* 3FD24924 94229359 */
static const double Lg4 = 2.222219843214978396e-01;  /* 3FCC71C5 1D8E78AF */
static const double Lg5 = 1.818357216161805012e-01;  /* 3FC74664 96CB03DE */
static const double Lg6 = 1.531383769920937332e-01;  /* 3FC39A09 D078C69F */
static const double Lg7 = 1.479819860511658591e-01;  /* 3FC2F112 DF3E5244 */

double orc_ln(double x) {
    uint64_t ui = bits_of(x);
    uint32_t hx = (uint32_t)(ui >> 32);
    int32_t k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((ui << 1) == 0) return -1.0 / (x * x); /* log(+-0) = -inf */
        if (hx >> 31) return (x - x) / 0.0;         /* log(-#) = NaN */
        /* subnormal: scale up */
        k -= 54;
        x *= of_bits(0x4350000000000000ull); /* 2^54 */
        ui = bits_of(x);
        hx = (uint32_t)(ui >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (ui << 32) == 0) {
        return 0.0;
    }
    /* reduce x into [sqrt(2)/2, sqrt(2)] */
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int32_t)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    ui = ((uint64_t)hx << 32) | (ui & 0xffffffffull);
    x = of_bits(ui);

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

/* scalbn (musl) */
static double orc_scalbn(double x, int n) {
    double y = x;
    if (n > 1023) {
        y *= of_bits(0x7fe0000000000000ull);
        n -= 1023;
        if (n > 1023) {
            y *= of_bits(0x7fe0000000000000ull);
            n -= 1023;
            if (n > 1023) n = 1023;
        }
    } else if (n < -1022) {
        /* 0x1p-1022 * 0x1p53 */
        y *= of_bits(0x0010000000000000ull) * of_bits(0x4340000000000000ull);
        n += 1022 - 53;
        if (n < -1022) {
            y *= of_bits(0x0010000000000000ull) * of_bits(0x4340000000000000ull);
            n += 1022 - 53;
            if (n < -1022) n = -1022;
        }
    }
    return y * of_bits((uint64_t)(0x3ff + n) << 52);
}

/* exp (msun e_exp.c method: x = k ln2 + r, exp(r) = 1 + r + r c/(2-c)) */
static const double half_tab[2] = {0.5, -0.5};
static const double e_ln2hi = 6.93147180369123816490e-01;  /* 0x3fe62e42, 0xfee00000 */
static const double e_ln2lo = 1.90821492927058770002e-10;  /* 0x3dea39ef, 0x35793c76 */
static const double e_invln2 = 1.44269504088896338700e+00; /* 0x3ff71547, 0x652b82fe */
static const double P1 = 1.66666666666666019037e-01;  /* 0x3FC55555, 0x5555553E */
static const double P2 = -2.77777777770155933842e-03; /* 0xBF66C16C, 0x16BEBD93 */
static const double P3 = 6.61375632143793436117e-05;  /* 0x3F11566A, 0xAF25DE2C */
static const double P4 = -1.65339022054652515390e-06; /* 0xBEBBBD41, 0xC5D26BF1 */
static const double P5 = 4.13813679705723846039e-08;  /* 0x3E663769, 0x72BEA4D0 */

double orc_exp(double x) {
    double hi, lo, c, xx, y;
    int32_t k, sign;
    uint32_t hx = (uint32_t)(bits_of(x) >> 32);
    sign = (int32_t)(hx >> 31);
    hx &= 0x7fffffffu;

    if (hx >= 0x4086232bu) { /* |x| >= 708.39... */
        if (x != x) return x;
        if (x > 709.782712893383973096) return x * of_bits(0x7fe0000000000000ull);
        if (x < -708.39641853226410622) {
            if (x < -745.13321910194110842) return 0.0;
        }
    }
    if (hx > 0x3fd62e42u) {      /* |x| > 0.5 ln2 */
        if (hx >= 0x3ff0a2b2u) { /* |x| >= 1.5 ln2 */
            k = f64_as_i32(e_invln2 * x + half_tab[sign]);
        } else {
            k = 1 - sign - sign;
        }
        hi = x - (double)k * e_ln2hi;
        lo = (double)k * e_ln2lo;
        x = hi - lo;
    } else if (hx > 0x3e300000u) { /* |x| > 2^-28 */
        k = 0;
        hi = x;
        lo = 0.0;
    } else {
        return 1.0 + x;
    }
    xx = x * x;
    c = x - xx * (P1 + xx * (P2 + xx * (P3 + xx * (P4 + xx * P5))));
    y = 1.0 + (x * c / (2.0 - c) - lo + hi);
    if (k == 0) return y;
    return orc_scalbn(y, k);
}

/* k_sin / k_cos kernels on [-pi/4, pi/4] (msun k_sin.c, k_cos.c) */
static const double S1 = -1.66666666666666324348e-01; /* 0xBFC55555, 0x55555549 */
static const double S2 = 8.33333333332248946124e-03;  /* 0x3F811111, 0x1110F8A6 */
static const double S3 = -1.98412698298579493134e-04; /* 0xBF2A01A0, 0x19C161D5 */
static const double S4 = 2.75573137070700676789e-06;  /* 0x3EC71DE3, 0x57B1FE7D */
static const double S5 = -2.50507602534068634195e-08; /* 0xBE5AE5E6, 0x8A2B9CEB */
static const double S6 = 1.58969099521155010221e-10;  /* 0x3DE5D93A, 0x5ACFD57C */

static double orc_k_sin(double x, double y, int iy) {
    double z = x * x;
    double w = z * z;
    double r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6);
    double v = z * x;
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

static const double C1 = 4.16666666666666019037e-02;  /* 0x3FA55555, 0x5555554C */
static const double C2 = -1.38888888888741095749e-03; /* 0xBF56C16C, 0x16C15177 */
static const double C3 = 2.48015872894767294178e-05;  /* 0x3EFA01A0, 0x19CB1590 */
static const double C4 = -2.75573143513906633035e-07; /* 0xBE927E4F, 0x809C52AD */
static const double C5 = 2.08757232129817482790e-09;  /* 0x3E21EE9E, 0xBDB4B1C4 */
static const double C6 = -1.13596475577881948265e-11; /* 0xBDA8FAE9, 0xBE8838D4 */

static double orc_k_cos(double x, double y) {
    double z = x * x;
    double w = z * z;
    double r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6));
    double hz = 0.5 * z;
    double ww = 1.0 - hz;
    return ww + (((1.0 - ww) - hz) + (z * r - x * y));
}

/* __rem_pio2 (musl): x - n*pi/2 = y0 + y1.  Arguments with
 * |x| >= 2^20 * pi/2 need the Payne-Hanek path, which the Independent
 * algorithm never reaches (Box-Muller angles lie in (-2pi, 0]); the oracle
 * aborts on them instead of restating it. */
static const double toint = 1.5 / 2.220446049250313080847e-16; /* 1.5/EPSILON */
static const double invpio2 = 6.36619772367581382433e-01; /* 0x3FE45F30, 0x6DC9C883 */
static const double pio2_1 = 1.57079632673412561417e+00;  /* 0x3FF921FB, 0x54400000 */
static const double pio2_1t = 6.07710050650619224932e-11; /* 0x3DD0B461, 0x1A626331 */
static const double pio2_2 = 6.07710050630396597660e-11;  /* 0x3DD0B461, 0x1A600000 */
static const double pio2_2t = 2.02226624879595063154e-21; /* 0x3BA3198A, 0x2E037073 */
static const double pio2_3 = 2.02226624871116645580e-21;  /* 0x3BA3198A, 0x2E000000 */
static const double pio2_3t = 8.47842766036889956997e-32; /* 0x397B839A, 0x252049C1 */

static int32_t rem_pio2_medium(double x, uint32_t ix, double *y0o, double *y1o) {
    double fn = x * invpio2 + toint - toint;
    int32_t n = f64_as_i32(fn);
    double r = x - fn * pio2_1;
    double w = fn * pio2_1t;
    double y0 = r - w;
    int32_t ey = (int32_t)((bits_of(y0) >> 52) & 0x7ff);
    int32_t ex = (int32_t)(ix >> 20);
    if (ex - ey > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y0 = r - w;
        ey = (int32_t)((bits_of(y0) >> 52) & 0x7ff);
        if (ex - ey > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y0 = r - w;
        }
    }
    *y0o = y0;
    *y1o = (r - y0) - w;
    return n;
}

static int32_t orc_rem_pio2(double x, double *y0, double *y1) {
    uint64_t u = bits_of(x);
    int sign = (int)(u >> 63);
    uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    double z;
    if (ix <= 0x400f6a7au) {                  /* |x| ~<= 5pi/4 */
        if ((ix & 0xfffffu) == 0x921fbu)      /* |x| ~= pi/2 or 2pi/2 */
            return rem_pio2_medium(x, ix, y0, y1);
        if (ix <= 0x4002d97cu) {              /* |x| ~<= 3pi/4 */
            if (!sign) {
                z = x - pio2_1;
                *y0 = z - pio2_1t;
                *y1 = (z - *y0) - pio2_1t;
                return 1;
            } else {
                z = x + pio2_1;
                *y0 = z + pio2_1t;
                *y1 = (z - *y0) + pio2_1t;
                return -1;
            }
        } else {
            if (!sign) {
                z = x - 2 * pio2_1;
                *y0 = z - 2 * pio2_1t;
                *y1 = (z - *y0) - 2 * pio2_1t;
                return 2;
            } else {
                z = x + 2 * pio2_1;
                *y0 = z + 2 * pio2_1t;
                *y1 = (z - *y0) + 2 * pio2_1t;
                return -2;
            }
        }
    }
    if (ix <= 0x401c463bu) {         /* |x| ~<= 9pi/4 */
        if (ix <= 0x4015fdbcu) {     /* |x| ~<= 7pi/4 */
            if (ix == 0x4012d97cu)   /* |x| ~= 3pi/2 */
                return rem_pio2_medium(x, ix, y0, y1);
            if (!sign) {
                z = x - 3 * pio2_1;
                *y0 = z - 3 * pio2_1t;
                *y1 = (z - *y0) - 3 * pio2_1t;
                return 3;
            } else {
                z = x + 3 * pio2_1;
                *y0 = z + 3 * pio2_1t;
                *y1 = (z - *y0) + 3 * pio2_1t;
                return -3;
            }
        } else {
            if (ix == 0x401921fbu)   /* |x| ~= 4pi/2 */
                return rem_pio2_medium(x, ix, y0, y1);
            if (!sign) {
                z = x - 4 * pio2_1;
                *y0 = z - 4 * pio2_1t;
                *y1 = (z - *y0) - 4 * pio2_1t;
                return 4;
            } else {
                z = x + 4 * pio2_1;
                *y0 = z + 4 * pio2_1t;
                *y1 = (z - *y0) + 4 * pio2_1t;
                return -4;
            }
        }
    }
    if (ix < 0x413921fbu) return rem_pio2_medium(x, ix, y0, y1);
    if (ix >= 0x7ff00000u) {
        *y0 = *y1 = x - x;
        return 0;
    }
    fprintf(stderr, "necsim oracle: rem_pio2 of |x| >= 2^20*pi/2 is not restated\n");
    abort();
}

double orc_sin(double x) {
    uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    double y0, y1;
    if (ix <= 0x3fe921fbu) {       /* |x| ~< pi/4 */
        if (ix < 0x3e500000u) return x; /* |x| < 2^-26 */
        return orc_k_sin(x, 0.0, 0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    int32_t n = orc_rem_pio2(x, &y0, &y1);
    switch (n & 3) {
    case 0: return orc_k_sin(y0, y1, 1);
    case 1: return orc_k_cos(y0, y1);
    case 2: return -orc_k_sin(y0, y1, 1);
    default: return -orc_k_cos(y0, y1);
    }
}

double orc_cos(double x) {
    uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    double y0, y1;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e46a09eu) {  /* |x| < 2^-27 * sqrt(2) */
            if (f64_as_i32(x) == 0) return 1.0;
        }
        return orc_k_cos(x, 0.0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    int32_t n = orc_rem_pio2(x, &y0, &y1);
    switch (n & 3) {
    case 0: return orc_k_cos(y0, y1);
    case 1: return -orc_k_sin(y0, y1, 1);
    case 2: return -orc_k_cos(y0, y1);
    default: return orc_k_sin(y0, y1, 1);
    }
}

/* ------------------------------------------------------------------------ */
/* Test-only twins in fdlibm 5.3's evaluation order (e_log.c, k_sin.c,       */
/* k_cos.c), built from the SAME constants and the same pi/2 reduction as    */
/* orc_ln / orc_sin / orc_cos above.  They are not on the simulated path:    */
/* an independent fdlibm port (V8's Math.log / Math.sin / Math.cos) must     */
/* reproduce them bit for bit, which pins every constant and the reduction   */
/* of the restatement; only the musl evaluation order of the final           */
/* assembly (which the libm crate follows) remains pinned by the <= 1 ulp    */
/* checks alone.                                                             */
/* ------------------------------------------------------------------------ */
double orc_fdlibm_log(double x) {
    const double two54 = of_bits(0x4350000000000000ull);
    uint64_t ui = bits_of(x);
    int32_t hx = (int32_t)(ui >> 32);
    uint32_t lx = (uint32_t)ui;
    int32_t k = 0;
    if (hx < 0x00100000) {
        if (((hx & 0x7fffffff) | lx) == 0) return -1.0 / (x * x);
        if (hx < 0) return (x - x) / 0.0;
        k -= 54;
        x *= two54;
        hx = (int32_t)(bits_of(x) >> 32);
    }
    if (hx >= 0x7ff00000) return x + x;
    k += (hx >> 20) - 1023;
    hx &= 0x000fffff;
    int32_t i = (hx + 0x95f64) & 0x100000;
    x = of_bits(((uint64_t)(uint32_t)(hx | (i ^ 0x3ff00000)) << 32) | (bits_of(x) & 0xffffffffull));
    k += (i >> 20);
    double f = x - 1.0, dk, R, hfsq;
    if ((0x000fffff & (2 + hx)) < 3) { /* -2^-20 <= f < 2^-20 */
        if (f == 0.0) {
            if (k == 0) return 0.0;
            dk = (double)k;
            return dk * ln2_hi + dk * ln2_lo;
        }
        R = f * f * (0.5 - 0.33333333333333333 * f);
        if (k == 0) return f - R;
        dk = (double)k;
        return dk * ln2_hi - ((R - dk * ln2_lo) - f);
    }
    double s = f / (2.0 + f);
    dk = (double)k;
    double z = s * s;
    i = hx - 0x6147a;
    double w = z * z;
    int32_t j = 0x6b851 - hx;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    i |= j;
    R = t2 + t1;
    if (i > 0) {
        hfsq = 0.5 * f * f;
        if (k == 0) return f - (hfsq - s * (hfsq + R));
        return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
    }
    if (k == 0) return f - s * (f - R);
    return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

static double fdlibm_k_sin(double x, double y, int iy) {
    double z = x * x;
    double v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

static double fdlibm_k_cos(double x, double y) {
    uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3fd33333u) return 1.0 - (0.5 * z - (z * r - x * y)); /* |x| < 0.3 */
    double qx = ix > 0x3fe90000u ? 0.28125 : of_bits((uint64_t)(ix - 0x00200000u) << 32); /* x/4 */
    double iz = 0.5 * z - qx;
    double a = 1.0 - qx;
    return a - (iz - (z * r - x * y));
}

double orc_fdlibm_sin(double x) {
    uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    double y0, y1;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e400000u && (int)x == 0) return x;
        return fdlibm_k_sin(x, 0.0, 0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    int32_t n = orc_rem_pio2(x, &y0, &y1);
    switch (n & 3) {
    case 0: return fdlibm_k_sin(y0, y1, 1);
    case 1: return fdlibm_k_cos(y0, y1);
    case 2: return -fdlibm_k_sin(y0, y1, 1);
    default: return -fdlibm_k_cos(y0, y1);
    }
}

double orc_fdlibm_cos(double x) {
    uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    double y0, y1;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e400000u && (int)x == 0) return 1.0;
        return fdlibm_k_cos(x, 0.0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    int32_t n = orc_rem_pio2(x, &y0, &y1);
    switch (n & 3) {
    case 0: return fdlibm_k_cos(y0, y1);
    case 1: return -fdlibm_k_sin(y0, y1, 1);
    case 2: return -fdlibm_k_cos(y0, y1);
    default: return fdlibm_k_sin(y0, y1, 1);
    }
}

/* Vector evaluation for the pins: op 0 ln, 1 exp, 2 sin, 3 cos, 4 the fdlibm
 * log twin, 5 the fdlibm sin twin, 6 the fdlibm cos twin. */
int orc_maths_vec(int op, const double *in, double *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        double x = in[i];
        switch (op) {
        case 0: out[i] = orc_ln(x); break;
        case 1: out[i] = orc_exp(x); break;
        case 2: out[i] = orc_sin(x); break;
        case 3: out[i] = orc_cos(x); break;
        case 4: out[i] = orc_fdlibm_log(x); break;
        case 5: out[i] = orc_fdlibm_sin(x); break;
        case 6: out[i] = orc_fdlibm_cos(x); break;
        default: return -1;
        }
    }
    return 0;
}

/* ======================================================================== */
/* RNG: WyHash PrimeableRng and RngSampler                                    */
/* ======================================================================== */

/* necsim/impls/no-std/src/cogs/rng/wyhash.rs:9-12 */
#define WY_P0 0xa0761d6478bd642full
#define WY_P1 0xe7037ed1a0b428dbull
#define WY_P2 0x8ebc6af09c88c6e3ull
#define WY_P5 0xeb44accab455d165ull

/* necsim/core/src/cogs/rng.rs:33-63 (rand 0.7 seed_from_u64 via PCG32) then
 * WyHash::from_seed wyhash.rs:41-49 (seed = u64::from_le_bytes) */
uint64_t orc_seed_from_u64(uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull;
    const uint64_t INC = 11634580027462260723ull;
    uint8_t seed[8];
    for (int chunk = 0; chunk < 2; chunk++) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        uint32_t x = (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
        for (int b = 0; b < 4; b++) seed[chunk * 4 + b] = (uint8_t)(x >> (8 * b));
    }
    uint64_t s = 0;
    for (int b = 0; b < 8; b++) s |= (uint64_t)seed[b] << (8 * b);
    return s;
}

/* wyhash.rs:85-96 */
uint64_t orc_wymum(uint64_t a, uint64_t b) {
    unsigned __int128 r = (unsigned __int128)a * (unsigned __int128)b;
    a ^= (uint64_t)r;
    b ^= (uint64_t)(r >> 64);
    return a ^ b;
}

/* wyhash.rs:99-119 */
uint64_t orc_seahash_diffuse(uint64_t x) {
    x = x * 0x6eed0e9da4d94a4full;
    uint64_t a = x >> 32;
    uint64_t b = x >> 60;
    x ^= a >> b;
    x = x * 0x6eed0e9da4d94a4full;
    return x;
}

/* wyhash.rs:67-80 */
uint64_t orc_prime(uint64_t seed, uint64_t location_index, uint64_t time_index) {
    uint64_t l = orc_seahash_diffuse(location_index);
    uint64_t t = orc_seahash_diffuse(time_index);
    uint64_t hash = orc_wymum(((l << 32) | (l >> 32)) ^ (seed ^ WY_P0), ((t << 32) | (t >> 32)) ^ WY_P2);
    return orc_wymum(hash, 16ull ^ WY_P5);
}

/* wyhash.rs:53-63 */
uint64_t orc_sample_u64(uint64_t *state) {
    *state = *state + WY_P0;
    return orc_seahash_diffuse(orc_wymum(*state ^ WY_P1, *state));
}

/* rng.rs:75-81 */
double orc_uniform_closed_open(uint64_t *state) {
    return (double)(orc_sample_u64(state) >> 11) * of_bits(0x3CA0000000000000ull);
}

/* rng.rs:86-93 */
double orc_uniform_open_closed(uint64_t *state) {
    return (double)((orc_sample_u64(state) >> 11) + 1) * of_bits(0x3CA0000000000000ull);
}

/* rng.rs:132-144 */
static uint64_t orc_sample_index_u64(uint64_t *state, uint64_t length) {
    uint64_t index = f64_as_u64(orc_floor(orc_uniform_closed_open(state) * (double)length));
    return index < length - 1 ? index : length - 1;
}

/* rng.rs:180-197: Box-Muller, returns (z0*sigma+mu, z1*sigma+mu) */
static void orc_sample_2d_normal(uint64_t *state, double mu, double sigma, double *a, double *b) {
    double u0 = orc_uniform_open_closed(state);
    double u1 = orc_uniform_closed_open(state);
    double r = orc_sqrt(-2.0 * orc_ln(u0));
    double theta = -6.283185307179586 * u1; /* -core::f64::consts::TAU */
    double z0 = r * orc_sin(theta);
    double z1 = r * orc_cos(theta);
    *a = z0 * sigma + mu;
    *b = z1 * sigma + mu;
}

/* FNV-1a 64 (crate fnv 1.0.7, Cargo.lock:738-739; FnvHasher::default + write) */
uint64_t orc_fnv1a64(const uint8_t *bytes, size_t n) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; i++) {
        h ^= (uint64_t)bytes[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

/* Hash of SpeciationSample (necsim/impls/no-std/src/cogs/event_sampler/tracking.rs:69-75):
 * location (x, y as u32 via derive(Hash), landscape/location.rs:11-14), index
 * u32 (:47-51), sample_time bits u64, speciation_sample bits u64; each written
 * as native-endian (little-endian) bytes */
uint64_t orc_speciation_sample_hash(double u, double time, uint32_t x, uint32_t y, uint32_t i) {
    uint8_t buf[28];
    uint64_t tb = bits_of(time), ub = bits_of(u);
    for (int b = 0; b < 4; b++) buf[b] = (uint8_t)(x >> (8 * b));
    for (int b = 0; b < 4; b++) buf[4 + b] = (uint8_t)(y >> (8 * b));
    for (int b = 0; b < 4; b++) buf[8 + b] = (uint8_t)(i >> (8 * b));
    for (int b = 0; b < 8; b++) buf[12 + b] = (uint8_t)(tb >> (8 * b));
    for (int b = 0; b < 8; b++) buf[20 + b] = (uint8_t)(ub >> (8 * b));
    return orc_fnv1a64(buf, 28);
}

/* ======================================================================== */
/* Alias method: necsim/impls/no-std/src/alias/packed.rs                      */
/* ======================================================================== */

/* packed.rs:59-120 (create), with pop_overfull_underfull_pair_atomic :166-182 */
void orc_alias_create(const uint64_t *events, const double *weights, size_t n,
                      double *u, uint64_t *e, uint64_t *k) {
    double total = 0.0; /* NonNegativeF64 Sum = fold from zero, bond non_negative_f64.rs:276-280 */
    for (size_t i = 0; i < n; i++) total = total + weights[i];
    double nf = (double)n;
    for (size_t i = 0; i < n; i++) {
        u[i] = weights[i] * nf / total;
        e[i] = events[i];
        k[i] = events[i];
    }
    size_t *over = (size_t *)malloc(sizeof(size_t) * (n + 1));
    size_t *under = (size_t *)malloc(sizeof(size_t) * (n + 1));
    size_t no = 0, nu = 0;
    for (size_t i = 0; i < n; i++) if (u[i] > 1.0) over[no++] = i;
    for (size_t i = 0; i < n; i++) if (u[i] < 1.0) under[nu++] = i;
    while (no > 0 && nu > 0) {
        size_t oi = over[--no];
        size_t ui = under[--nu];
        u[oi] = u[oi] + u[ui] - 1.0;
        k[ui] = e[oi];
        if (u[oi] < 1.0) under[nu++] = oi;
        else if (u[oi] > 1.0) over[no++] = oi;
    }
    for (size_t j = 0; j < no; j++) u[over[j]] = 1.0;
    for (size_t j = 0; j < nu; j++) u[under[j]] = 1.0;
    free(over);
    free(under);
}

/* packed.rs:135-163 */
uint64_t orc_alias_sample(const double *u, const uint64_t *e, const uint64_t *k, size_t n,
                          double limit, double u01) {
    double f = u01 * limit * (double)n;
    uint64_t i = f64_as_u64(orc_floor(f));
    double y = f - (double)i;
    return y < u[i] ? e[i] : k[i];
}

/* ======================================================================== */
/* Scenario: habitat, dispersal sampler, tables                               */
/* ======================================================================== */

typedef struct {
    double u;
    uint64_t e, k;
} orc_atom;

struct orc_scenario {
    int kind;
    uint32_t width, height, deme;
    double sigma;
    /* in-memory */
    uint32_t *habitat;        /* W*H */
    uint64_t *injection;      /* W*H+1 prefix sums, in_memory.rs:151-160 */
    uint64_t *range_start, *range_end; /* W*H */
    double *self_p;           /* W*H, separable only */
    uint64_t *self_event;     /* W*H, separable only */
    orc_atom *atoms;
    size_t n_atoms;
};

static uint32_t habitat_at(const orc_scenario *s, uint32_t x, uint32_t y) {
    switch (s->kind) {
    case ORC_NON_SPATIAL: return s->deme;              /* non_spatial.rs:88-91 */
    case ORC_ALMOST_INFINITE_NORMAL: return 1;         /* almost_infinite.rs:62-65 */
    default: {
        /* in_memory.rs:60-69: .get(y*W + x).unwrap_or(0) */
        uint64_t idx = (uint64_t)y * s->width + x;
        if (idx >= (uint64_t)s->width * s->height) return 0;
        return s->habitat[idx];
    }
    }
}

uint32_t orc_deme_at(const orc_scenario *s, uint32_t x, uint32_t y) { return habitat_at(s, x, y); }

/* map_indexed_location_to_u64_injective */
uint64_t orc_location_id(const orc_scenario *s, uint32_t x, uint32_t y, uint32_t i) {
    switch (s->kind) {
    case ORC_NON_SPATIAL: /* non_spatial.rs:93-109, origin (0,0) */
        return ((uint64_t)y * (uint64_t)s->width + (uint64_t)x) * (uint64_t)s->deme + (uint64_t)i;
    case ORC_ALMOST_INFINITE_NORMAL: /* almost_infinite.rs:67-71 */
        return ((uint64_t)y << 32) | (uint64_t)x;
    default: { /* in_memory.rs:71-81 */
        uint64_t idx = (uint64_t)y * s->width + x;
        uint64_t base = idx < (uint64_t)s->width * s->height + 1 ? s->injection[idx] : 0;
        return base + (uint64_t)i;
    }
    }
}

/* packed_separable_alias/mod.rs:64-203 and packed_alias/mod.rs:58-118 */
static int build_in_memory(orc_scenario *s, const orc_scenario_desc *d) {
    size_t cells = (size_t)s->width * s->height;
    s->habitat = (uint32_t *)malloc(sizeof(uint32_t) * cells);
    memcpy(s->habitat, d->habitat, sizeof(uint32_t) * cells);
    s->injection = (uint64_t *)malloc(sizeof(uint64_t) * (cells + 1));
    uint64_t acc = 0;
    for (size_t c = 0; c < cells; c++) {
        s->injection[c] = acc;
        acc += s->habitat[c];
    }
    s->injection[cells] = acc;
    if (acc == 0) return -1; /* InMemoryHabitat::try_new -> None */

    s->range_start = (uint64_t *)calloc(cells, sizeof(uint64_t));
    s->range_end = (uint64_t *)calloc(cells, sizeof(uint64_t));
    int separable = s->kind == ORC_SPATIALLY_EXPLICIT;
    if (separable) {
        s->self_p = (double *)calloc(cells, sizeof(double));
        s->self_event = (uint64_t *)malloc(sizeof(uint64_t) * cells);
        for (size_t c = 0; c < cells; c++) s->self_event[c] = c; /* mod.rs:77-86 */
    }
    size_t cap = d->row_ptr[cells] + 16;
    s->atoms = (orc_atom *)malloc(sizeof(orc_atom) * cap);
    s->n_atoms = 0;

    uint64_t *ev = NULL;
    double *w = NULL, *au = NULL;
    uint64_t *ae = NULL, *ak = NULL;
    size_t wcap = 0;

    for (size_t row = 0; row < cells; row++) {
        /* contract (in_memory/contract.rs:13-69): dispersal only between habitat */
        uint64_t b = d->row_ptr[row], en = d->row_ptr[row + 1];
        size_t len = (size_t)(en - b);
        if (len + 1 > wcap) {
            wcap = (len + 1) * 2;
            ev = (uint64_t *)realloc(ev, sizeof(uint64_t) * wcap);
            w = (double *)realloc(w, sizeof(double) * wcap);
            au = (double *)realloc(au, sizeof(double) * wcap);
            ae = (uint64_t *)realloc(ae, sizeof(uint64_t) * wcap);
            ak = (uint64_t *)realloc(ak, sizeof(uint64_t) * wcap);
        }
        size_t nw = 0;
        double self_w = 0.0;
        for (uint64_t j = b; j < en; j++) {
            uint32_t col = d->col[j];
            double p = d->weight[j];
            if (p < 0.0 || (p > 0.0 && (s->habitat[row] == 0 || s->habitat[col] == 0))) return -2;
            double weight = p * (double)s->habitat[col];
            if (weight > 0.0) {
                ev[nw] = col;
                w[nw] = weight;
                nw++;
                if (col == row) self_w = weight;
            }
        }
        if (s->habitat[row] > 0 && nw == 0) return -3; /* NoDispersalFromHabitat */
        size_t from = s->n_atoms;
        if (nw == 0) {
            s->range_start[row] = from;
            s->range_end[row] = from;
            continue;
        }
        if (!separable) {
            orc_alias_create(ev, w, nw, au, ae, ak);
        } else {
            double total = 0.0;
            for (size_t j = 0; j < nw; j++) total = total + w[j];
            double self_prob = self_w / total;
            double self_u = self_w * (double)nw / total;
            uint64_t non_self = row;
            if (self_u < 1.0) {
                orc_alias_create(ev, w, nw, au, ae, ak);
                for (size_t j = 0; j < nw; j++) {
                    if (ae[j] == row || ak[j] == row) {
                        if (ae[j] == row) { /* flip: swap e/k, u = 1 - u */
                            uint64_t t = ae[j];
                            ae[j] = ak[j];
                            ak[j] = t;
                            au[j] = 1.0 - au[j];
                        }
                        non_self = ae[j];
                        size_t last = nw - 1;
                        double tu = au[j]; au[j] = au[last]; au[last] = tu;
                        uint64_t te = ae[j]; ae[j] = ae[last]; ae[last] = te;
                        uint64_t tk = ak[j]; ak[j] = ak[last]; ak[last] = tk;
                        break;
                    }
                }
            } else {
                /* retain(|(event, _)| event != row) */
                size_t m = 0;
                for (size_t j = 0; j < nw; j++) {
                    if (ev[j] != row) {
                        ev[m] = ev[j];
                        w[m] = w[j];
                        m++;
                    }
                }
                nw = m;
                if (nw > 0) orc_alias_create(ev, w, nw, au, ae, ak);
            }
            s->self_p[row] = self_prob;
            s->self_event[row] = non_self;
        }
        for (size_t j = 0; j < nw; j++) {
            s->atoms[s->n_atoms].u = au[j];
            s->atoms[s->n_atoms].e = ae[j];
            s->atoms[s->n_atoms].k = ak[j];
            s->n_atoms++;
        }
        s->range_start[row] = from;
        s->range_end[row] = s->n_atoms;
    }
    free(ev); free(w); free(au); free(ae); free(ak);
    return 0;
}

orc_scenario *orc_scenario_new(const orc_scenario_desc *d) {
    orc_scenario *s = (orc_scenario *)calloc(1, sizeof(orc_scenario));
    s->kind = d->kind;
    s->width = d->width;
    s->height = d->height;
    s->deme = d->deme;
    s->sigma = d->sigma;
    if (s->kind == ORC_SPATIALLY_EXPLICIT || s->kind == ORC_SPATIALLY_EXPLICIT_PACKED) {
        if (build_in_memory(s, d) != 0) {
            orc_scenario_free(s);
            return NULL;
        }
    }
    return s;
}

void orc_scenario_free(orc_scenario *s) {
    if (!s) return;
    free(s->habitat); free(s->injection); free(s->range_start); free(s->range_end);
    free(s->self_p); free(s->self_event); free(s->atoms);
    free(s);
}

size_t orc_scenario_tables(const orc_scenario *s, uint64_t *rs, uint64_t *re, double *sp,
                           uint64_t *se, double *au, uint64_t *ae, uint64_t *ak) {
    size_t cells = (size_t)s->width * s->height;
    if (rs) memcpy(rs, s->range_start, cells * 8);
    if (re) memcpy(re, s->range_end, cells * 8);
    if (sp && s->self_p) memcpy(sp, s->self_p, cells * 8);
    if (se && s->self_event) memcpy(se, s->self_event, cells * 8);
    for (size_t j = 0; j < s->n_atoms; j++) {
        if (au) au[j] = s->atoms[j].u;
        if (ae) ae[j] = s->atoms[j].e;
        if (ak) ak[j] = s->atoms[j].k;
    }
    return s->n_atoms;
}

static uint64_t alias_sample_atoms(const orc_atom *a, size_t n, double limit, uint64_t *state) {
    double f = orc_uniform_closed_open(state) * limit * (double)n;
    uint64_t i = f64_as_u64(orc_floor(f));
    double y = f - (double)i;
    return y < a[i].u ? a[i].e : a[i].k;
}

/* DispersalSampler::sample_dispersal_from_location for each scenario */
void orc_sample_dispersal(const orc_scenario *s, uint32_t x, uint32_t y, uint64_t *state,
                          uint32_t *tx, uint32_t *ty) {
    switch (s->kind) {
    case ORC_NON_SPATIAL: { /* dispersal_sampler/non_spatial.rs:41-70 */
        uint64_t W = s->width, H = s->height;
        uint64_t idx = orc_sample_index_u64(state, W * H);
        *tx = (uint32_t)(idx % W); /* origin.x (=0) wrapping_add */
        *ty = (uint32_t)(idx / W);
        return;
    }
    case ORC_ALMOST_INFINITE_NORMAL: { /* almost_infinite_normal.rs:52-69, habitat/almost_infinite.rs:98-117 */
        double dx, dy;
        orc_sample_2d_normal(state, 0.0, s->sigma, &dx, &dy);
        int64_t idx = f64_as_i64(orc_round(dx));
        int64_t idy = f64_as_i64(orc_round(dy));
        const int64_t WRAP = (int64_t)1 << 32;
        int64_t nx = ((int64_t)x + idx) % WRAP; /* Rust % on i64 = C % (sign of dividend) */
        int64_t ny = ((int64_t)y + idy) % WRAP;
        *tx = (uint32_t)((nx + WRAP) % WRAP);
        *ty = (uint32_t)((ny + WRAP) % WRAP);
        return;
    }
    case ORC_SPATIALLY_EXPLICIT: { /* packed_separable_alias/dispersal.rs:16-77 */
        uint64_t self_index = (uint64_t)y * s->width + x;
        double sp = s->self_p[self_index];
        uint64_t nse = s->self_event[self_index];
        uint64_t target;
        if (sp == 1.0) {
            target = self_index;
        } else if (sp == 0.0 || nse != self_index || orc_uniform_closed_open(state) >= sp) {
            target = alias_sample_atoms(s->atoms + s->range_start[self_index],
                                        (size_t)(s->range_end[self_index] - s->range_start[self_index]),
                                        1.0, state);
        } else {
            target = self_index;
        }
        *tx = (uint32_t)(target % s->width);
        *ty = (uint32_t)(target / s->width);
        return;
    }
    default: { /* packed_alias/dispersal.rs:16-52 */
        uint64_t self_index = (uint64_t)y * s->width + x;
        uint64_t target = alias_sample_atoms(s->atoms + s->range_start[self_index],
                                             (size_t)(s->range_end[self_index] - s->range_start[self_index]),
                                             1.0, state);
        *tx = (uint32_t)(target % s->width);
        *ty = (uint32_t)(target / s->width);
        return;
    }
    }
}

/* ======================================================================== */
/* Poisson event time sampler                                                */
/* necsim/impls/no-std/src/cogs/active_lineage_sampler/independent/          */
/*   event_time_sampler/poisson.rs:30-109                                    */
/* ======================================================================== */

#define INV_PHI_U64 0x9e3779b97f4a7c15ull

static double poisson_next(uint64_t seed, uint64_t loc_id, double time, double lambda,
                           double delta_t, uint64_t *state) {
    double lambda_per_step = lambda * delta_t;
    double no_event_probability_per_step = orc_exp(-lambda_per_step);
    uint64_t time_step = f64_as_u64(orc_floor(time / delta_t));
    double event_time;
    uint32_t event_index;
    for (;;) {
        *state = orc_prime(seed, loc_id, time_step);
        uint32_t number;
        if (no_event_probability_per_step > 0.0) {
            uint32_t poisson = 0;
            double prod = no_event_probability_per_step;
            double acc = no_event_probability_per_step;
            double u = orc_uniform_closed_open(state);
            while (u > acc && prod > 0.0) {
                poisson += 1;
                prod *= lambda_per_step / (double)poisson;
                acc += prod;
            }
            number = poisson;
        } else {
            double a, b;
            orc_sample_2d_normal(state, lambda_per_step, orc_sqrt(lambda_per_step), &a, &b);
            number = f64_as_u32(fmax(a, 0.0)); /* f64::max ignores NaN like fmax */
        }
        int found = 0;
        double best = 0.0;
        uint32_t best_index = 0;
        for (uint32_t e = 0; e < number; e++) {
            double t = ((double)time_step + orc_uniform_closed_open(state)) * delta_t;
            if (t > time) {
                if (!found || best > t) {
                    best = t;
                    best_index = e;
                    found = 1;
                }
            }
        }
        if (found) {
            event_time = best;
            event_index = best_index;
            break;
        }
        time_step += 1;
    }
    *state = orc_prime(seed, loc_id, time_step + INV_PHI_U64 * (uint64_t)(event_index + 1));
    return event_time;
}

double orc_poisson_next_event_time(uint64_t seed, uint64_t location_id, double time,
                                   double lambda, double delta_t, uint64_t *state_out) {
    uint64_t st = 0;
    double t = poisson_next(seed, location_id, time, lambda, delta_t, &st);
    if (state_out) *state_out = st;
    return t;
}

/* PositiveF64::max_after, necsim/core/bond/src/positive_f64.rs:103-113 */
static double max_after(double before, double value) {
    if (value > before || before != before) return value;
    if (isinf(before)) return INFINITY;
    return of_bits(bits_of(before) + 1);
}

/* ======================================================================== */
/* Origin samplers                                                           */
/* ======================================================================== */

typedef struct {
    /* base = 0.. ; percentage layer ; partition layer */
    uint64_t next_base;
    int pct_mode; /* 0: p <= 0 (empty), 1: p >= 1 (all), 2: skip sampling */
    double percentage, inv_geometric_sample_rate, quasi_random;
    uint32_t rank, size;
    int partition_started;
    int exhausted;
} pre_sampler;

/* inner iterator: OriginPreSampler::all().percentage(p)
 * origin_sampler/pre_sampler.rs:50-57, 71-101 */
static int pct_next(pre_sampler *p, uint64_t *out) {
    if (p->pct_mode == 0) return 0;
    if (p->pct_mode == 1) {
        *out = p->next_base++;
        return 1;
    }
    p->quasi_random += 6.180339887498949e-1;
    p->quasi_random -= orc_floor(p->quasi_random);
    uint64_t skip = f64_as_u64(orc_floor(orc_ln(p->quasi_random) * p->inv_geometric_sample_rate));
    /* self.nth(skip): skip elements then yield the next one of 0.. */
    p->next_base += skip;
    *out = p->next_base++;
    return 1;
}

/* .partition(partition): advance_by(rank) then step_by(size), pre_sampler.rs:103-114 */
static int pre_next(pre_sampler *p, uint64_t *out) {
    uint64_t v;
    if (!p->partition_started) {
        p->partition_started = 1;
        for (uint32_t r = 0; r < p->rank; r++)
            if (!pct_next(p, &v)) return 0;
        return pct_next(p, out);
    }
    /* step_by(size): skip size-1 then yield */
    for (uint32_t r = 0; r + 1 < p->size; r++)
        if (!pct_next(p, &v)) return 0;
    return pct_next(p, out);
}

static void pre_init(pre_sampler *p, double percentage, uint32_t rank, uint32_t size) {
    memset(p, 0, sizeof(*p));
    p->percentage = percentage;
    if (percentage <= 0.0) p->pct_mode = 0;
    else if (percentage >= 1.0) p->pct_mode = 1;
    else p->pct_mode = 2;
    p->inv_geometric_sample_rate = 1.0 / orc_ln(1.0 - percentage);
    p->quasi_random = 0.5;
    p->rank = rank;
    p->size = size ? size : 1;
}

size_t orc_sample_origins(const orc_scenario *s, double percentage, uint32_t rank, uint32_t size,
                          uint32_t ccx, uint32_t ccy, uint32_t radius, orc_lineage *out,
                          size_t cap, size_t *total) {
    pre_sampler p;
    pre_init(&p, percentage, rank, size);
    size_t n = 0;
    uint64_t next_index;
    if (s->kind == ORC_ALMOST_INFINITE_NORMAL) {
        /* origin_sampler/singleton_demes/circle.rs:52-156 */
        uint32_t ox = ccx - radius, oy = ccy - radius;
        uint64_t diameter = (uint64_t)radius * 2 + 1;
        uint64_t r2 = (uint64_t)radius * radius;
        uint64_t last_index = 0;
        /* location iterator over the sample extent, row-major (landscape/extent.rs:77-100) */
        uint64_t it = 0, it_end = diameter * diameter;
        while (pre_next(&p, &next_index)) {
            uint64_t diff = next_index - last_index;
            last_index = next_index + 1;
            int got = 0;
            while (it < it_end) {
                uint32_t lx = ox + (uint32_t)(it % diameter);
                uint32_t ly = oy + (uint32_t)(it / diameter);
                it++;
                uint32_t dx1 = lx - ccx, dx2 = ccx - lx;
                uint32_t dy1 = ly - ccy, dy2 = ccy - ly;
                uint64_t dx = dx1 < dx2 ? dx1 : dx2;
                uint64_t dy = dy1 < dy2 ? dy1 : dy2;
                if (dx * dx + dy * dy <= r2) {
                    if (diff == 0) {
                        if (n < cap) {
                            out[n].global_reference = ((uint64_t)ly << 32) | lx;
                            out[n].last_event_time = 0.0;
                            out[n].x = lx; out[n].y = ly; out[n].index = 0; out[n]._pad = 0;
                        }
                        n++;
                        got = 1;
                        break;
                    }
                    diff--;
                }
            }
            if (!got) break;
        }
    } else {
        /* origin_sampler/non_spatial.rs:89-112 and in_memory.rs:89-119: the
         * location iterator walks the full extent row-major; demes of size
         * habitat_at(loc) are consumed index by index */
        uint64_t W = s->width, H = s->height;
        uint64_t loc = 0, cells = W * H;
        uint64_t last_index = 0;
        uint32_t nli = 0;
        while (pre_next(&p, &next_index)) {
            uint64_t diff = next_index - last_index;
            last_index = next_index;
            int ok = 1;
            for (;;) {
                if (loc >= cells) { ok = 0; break; }
                uint32_t h = habitat_at(s, (uint32_t)(loc % W), (uint32_t)(loc / W));
                if ((uint64_t)nli + diff >= (uint64_t)h) {
                    diff -= (uint64_t)(h - nli);
                    nli = 0;
                    loc++;
                } else break;
            }
            if (!ok) break;
            nli += (uint32_t)diff;
            uint32_t lx = (uint32_t)(loc % W), ly = (uint32_t)(loc / W);
            /* IndependentActiveLineageSampler::resume_with_store_and_lineages
             * filters (independent/mod.rs:98-117) always pass here */
            if (n < cap) {
                out[n].global_reference = orc_location_id(s, lx, ly, nli);
                out[n].last_event_time = 0.0;
                out[n].x = lx; out[n].y = ly; out[n].index = nli; out[n]._pad = 0;
            }
            n++;
        }
    }
    if (total) *total = n;
    return n < cap ? n : cap;
}

/* ======================================================================== */
/* Event order and Biodiversity                                              */
/* ======================================================================== */

/* PackedEvent::cmp (necsim/core/src/event.rs:256-280) for speciation events:
 * (event_time, origin, target = origin, prior_time, global_reference) */
static int cmp_f64_total(double a, double b) {
    /* f64::total_cmp on non-negative values orders like the bit patterns */
    int64_t ia = (int64_t)bits_of(a), ib = (int64_t)bits_of(b);
    ia ^= (int64_t)(((uint64_t)(ia >> 63)) >> 1);
    ib ^= (int64_t)(((uint64_t)(ib >> 63)) >> 1);
    return ia < ib ? -1 : ia > ib;
}

static int cmp_event(const void *pa, const void *pb) {
    const orc_speciation_event *a = (const orc_speciation_event *)pa;
    const orc_speciation_event *b = (const orc_speciation_event *)pb;
    int c = cmp_f64_total(fabs(a->event_time), fabs(b->event_time));
    if (c) return c;
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    if (a->y != b->y) return a->y < b->y ? -1 : 1;
    if (a->index != b->index) return a->index < b->index ? -1 : 1;
    c = cmp_f64_total(fabs(a->prior_time), fabs(b->prior_time));
    if (c) return c;
    if (a->global_reference != b->global_reference)
        return a->global_reference < b->global_reference ? -1 : 1;
    return 0;
}

/* PackedEvent::cmp (event.rs:256-280) */
static int cmp_packed(const void *pa, const void *pb) {
    const orc_packed_event *a = (const orc_packed_event *)pa;
    const orc_packed_event *b = (const orc_packed_event *)pb;
    int c = cmp_f64_total(fabs(a->event_time), fabs(b->event_time));
    if (c) return c;
    if (a->ox != b->ox) return a->ox < b->ox ? -1 : 1;
    if (a->oy != b->oy) return a->oy < b->oy ? -1 : 1;
    if (a->oindex != b->oindex) return a->oindex < b->oindex ? -1 : 1;
    if (a->tx != b->tx) return a->tx < b->tx ? -1 : 1;
    if (a->ty != b->ty) return a->ty < b->ty ? -1 : 1;
    if (a->tindex != b->tindex) return a->tindex < b->tindex ? -1 : 1;
    c = cmp_f64_total(fabs(a->prior_time), fabs(b->prior_time));
    if (c) return c;
    if (a->global_reference != b->global_reference)
        return a->global_reference < b->global_reference ? -1 : 1;
    return 0;
}

static double negated(double x) { return of_bits(bits_of(x) | 0x8000000000000000ull); }

/* growing PackedEvent buffer of orc_simulate_packed */
typedef struct {
    orc_packed_event *b;
    size_t n, cap;
} packed_sink;

static void sink_push(packed_sink *k, orc_packed_event e) {
    if (k->n == k->cap) {
        k->cap = k->cap ? 2 * k->cap : 1024;
        k->b = (orc_packed_event *)realloc(k->b, k->cap * sizeof(orc_packed_event));
    }
    k->b[k->n++] = e;
}

uint64_t orc_biodiversity(const orc_speciation_event *ev, size_t n) {
    uint64_t count = 0;
    for (size_t j = 0; j < n; j++) {
        if (j > 0 && ev[j].x == ev[j - 1].x && ev[j].y == ev[j - 1].y &&
            ev[j].index == ev[j - 1].index && ev[j].event_time == ev[j - 1].event_time)
            continue;
        count++;
    }
    return count;
}

/* ======================================================================== */
/* The monolithic water-level loop                                           */
/* necsim/impls/no-std/src/parallelisation/independent/monolithic/mod.rs     */
/* ======================================================================== */

typedef struct {
    orc_lineage l;
    double next_event; /* pessimistic next event time */
} queued;

typedef struct {
    queued *buf;
    size_t cap, head, len;
} deque;

static void dq_init(deque *d, size_t cap) {
    d->cap = cap ? cap : 1;
    d->buf = (queued *)malloc(sizeof(queued) * d->cap);
    d->head = d->len = 0;
}
static void dq_push(deque *d, queued q) {
    d->buf[(d->head + d->len) % d->cap] = q;
    d->len++;
}
static int dq_pop(deque *d, queued *q) {
    if (!d->len) return 0;
    *q = d->buf[d->head];
    d->head = (d->head + 1) % d->cap;
    d->len--;
    return 1;
}

typedef struct {
    int some;
    double u, time;
    uint32_t x, y, i;
} spec_sample;

typedef struct {
    spec_sample *b;
    size_t cap;
} dedup_cache;

/* cache.rs:39-54 */
static int cache_insert(dedup_cache *c, const spec_sample *v) {
    if (c->cap == 0) return 1;
    uint64_t h = orc_speciation_sample_hash(v->u, v->time, v->x, v->y, v->i);
    spec_sample *bucket = &c->b[h % (uint64_t)c->cap];
    int insert = !(bucket->some && bucket->u == v->u && bucket->time == v->time &&
                   bucket->x == v->x && bucket->y == v->y && bucket->i == v->i);
    *bucket = *v;
    bucket->some = 1;
    return insert;
}

/* The loop of both entry points: speciation events go to events_out (40 B),
 * or, with a sink, every reported event goes there as a PackedEvent. */
static int simulate_impl(const orc_scenario *s, const orc_params *p, const orc_lineage *lineages,
                         size_t n, orc_speciation_event *events_out, packed_sink *sink,
                         orc_lineage *remaining_out, orc_outcome *outcome) {
    const int report_dispersal = sink && p->report_dispersal;
    const uint64_t seed = orc_seed_from_u64(p->seed);
    const double lambda = p->turnover_rate;
    const double nu = p->speciation_probability;
    const double min_event_time = p->min_event_time;

    deque slow, fast;
    dq_init(&slow, n);
    dq_init(&fast, n);
    for (size_t j = 0; j < n; j++) {
        queued q = {lineages[j], lineages[j].last_event_time};
        dq_push(&slow, q);
    }
    /* EventSlice::capacity, parallelisation/independent/mod.rs:64-81 (unused
     * for the sorted output, but it sets the water level) */
    size_t event_slice;
    if (p->event_slice_mode == ORC_SLICE_ABSOLUTE) event_slice = (size_t)p->event_slice_capacity;
    else {
        event_slice = (size_t)f64_as_u64((double)n * p->event_slice_factor);
        if (event_slice < 1) event_slice = 1;
    }
    /* DedupCache::construct, mod.rs:34-55 */
    dedup_cache cache;
    if (p->dedup_mode == ORC_DEDUP_ABSOLUTE) cache.cap = (size_t)p->dedup_capacity;
    else if (p->dedup_mode == ORC_DEDUP_RELATIVE) cache.cap = (size_t)f64_as_u64((double)n * p->dedup_factor);
    else cache.cap = 0;
    cache.b = (spec_sample *)calloc(cache.cap ? cache.cap : 1, sizeof(spec_sample));

    uint64_t total_steps = 0;
    double max_time = 0.0, level_time = 0.0;
    if (n > 0) {
        max_time = lineages[0].last_event_time;
        level_time = lineages[0].last_event_time;
        for (size_t j = 1; j < n; j++) {
            if (cmp_f64_total(lineages[j].last_event_time, max_time) > 0) max_time = lineages[j].last_event_time;
            if (cmp_f64_total(lineages[j].last_event_time, level_time) < 0) level_time = lineages[j].last_event_time;
        }
    }
    size_t n_events = 0;

    while (slow.len > 0 && (!p->has_pause || level_time < p->pause_before)) {
        /* the reported event rate per slow lineage, summed sequentially over
         * the slow deque: lambda with dispersal reporting (mod.rs:139-148),
         * lambda * nu for speciation only (mod.rs:150-164) */
        double total_event_rate = 0.0;
        const double rate = report_dispersal ? lambda : lambda * nu;
        for (size_t j = 0; j < slow.len; j++) total_event_rate = total_event_rate + rate;
        level_time = level_time + (double)event_slice / total_event_rate;
        if (p->has_pause && p->pause_before < level_time) level_time = p->pause_before;

        int have_active = 0;
        orc_lineage active = {0};
        int have_prev_next = 0;
        double prev_next = 0.0;
        spec_sample min_sample = {0};

        while (slow.len > 0 || have_active) {
            /* mod.rs:186-196 */
            int have_next = 0;
            queued q;
            orc_lineage next_slow = {0};
            while (dq_pop(&slow, &q)) {
                if (q.next_event < level_time) {
                    next_slow = q.l;
                    have_next = 1;
                    break;
                }
                dq_push(&fast, q);
            }
            /* replace_active_lineage (singular.rs:45-58) */
            int have_prev = have_active;
            orc_lineage prev_task = active;
            have_active = have_next;
            if (have_next) active = next_slow;
            double sampler_last_event_time = have_active ? active.last_event_time : 0.0;
            /* replace_min_speciation(None) */
            spec_sample prev_sample = min_sample;
            min_sample.some = 0;

            if (have_prev && have_prev_next) {
                int duplicate = prev_sample.some && !cache_insert(&cache, &prev_sample);
                if (!duplicate) {
                    queued r = {prev_task, prev_next};
                    if (prev_next < level_time) dq_push(&slow, r);
                    else dq_push(&fast, r);
                }
            }
            have_prev_next = 0;

            /* simulate_incremental_early_stop (necsim/core/src/simulation/mod.rs:53-134) */
            uint64_t steps = 0;
            for (;;) {
                if (!have_active) break;
                double before = active.last_event_time > min_event_time ? active.last_event_time : min_event_time;
                uint64_t loc_id = orc_location_id(s, active.x, active.y, active.index);
                uint64_t state;
                double et = poisson_next(seed, loc_id, before, lambda, p->delta_t, &state);
                double next_event_time = max_after(before, et);
                /* early stop closure, monolithic/mod.rs:229-237 */
                prev_next = next_event_time;
                have_prev_next = 1;
                if (steps >= p->step_slice || next_event_time >= level_time) break;
                sampler_last_event_time = next_event_time;
                orc_lineage chosen = active;
                have_active = 0;

                /* IndependentEventSampler (event_sampler/independent.rs:112-222) */
                double u_spec = orc_uniform_closed_open(&state);
                if (!(min_sample.some && min_sample.u <= u_spec)) {
                    min_sample.some = 1;
                    min_sample.u = u_spec;
                    min_sample.time = next_event_time;
                    min_sample.x = chosen.x;
                    min_sample.y = chosen.y;
                    min_sample.i = chosen.index;
                }
                if (u_spec < nu) {
                    if (sink) {
                        orc_packed_event e = {chosen.global_reference, negated(chosen.last_event_time),
                                              next_event_time, chosen.x, chosen.y, chosen.index,
                                              chosen.x, chosen.y, chosen.index, chosen.global_reference};
                        sink_push(sink, e);
                    } else {
                        orc_speciation_event *e = &events_out[n_events];
                        e->global_reference = chosen.global_reference;
                        e->prior_time = chosen.last_event_time;
                        e->event_time = next_event_time;
                        e->x = chosen.x;
                        e->y = chosen.y;
                        e->index = chosen.index;
                        e->_pad = 0;
                    }
                    n_events++;
                } else {
                    uint32_t tx, ty;
                    orc_sample_dispersal(s, chosen.x, chosen.y, &state, &tx, &ty);
                    /* CoalescenceRngSample::new then the coalescence sampler
                     * (coalescence_sampler/independent.rs:38-76) */
                    double u_c = orc_uniform_closed_open(&state);
                    uint32_t deme = habitat_at(s, tx, ty);
                    uint32_t idx;
                    if (s->kind == ORC_ALMOST_INFINITE_NORMAL) idx = 0; /* SingletonDemesHabitat */
                    else idx = f64_as_u32(orc_floor(u_c * (double)deme));
                    if (report_dispersal) {
                        /* DispersalEvent with LineageInteraction::Maybe
                         * (event_sampler/independent.rs:174-222, event.rs:140-146) */
                        orc_packed_event e = {chosen.global_reference, chosen.last_event_time,
                                              negated(next_event_time), chosen.x, chosen.y, chosen.index,
                                              tx, ty, idx, chosen.global_reference};
                        sink_push(sink, e);
                        n_events++;
                    }
                    active = chosen;
                    active.x = tx;
                    active.y = ty;
                    active.index = idx;
                    active.last_event_time = next_event_time;
                    have_active = 1;
                    sampler_last_event_time = next_event_time;
                }
                steps += 1;
            }
            double new_time = sampler_last_event_time > min_event_time ? sampler_last_event_time : min_event_time;
            total_steps += steps;
            if (cmp_f64_total(new_time, max_time) > 0) max_time = new_time;
        }
        /* swap slow and fast */
        deque t = slow;
        slow = fast;
        fast = t;
    }

    if (sink) qsort(sink->b, sink->n, sizeof(orc_packed_event), cmp_packed);
    else qsort(events_out, n_events, sizeof(orc_speciation_event), cmp_event);

    size_t n_rem = slow.len;
    if (remaining_out) {
        queued q;
        size_t j = 0;
        while (dq_pop(&slow, &q)) remaining_out[j++] = q.l;
    }
    outcome->paused = n_rem > 0;
    outcome->time = max_time;
    outcome->steps = total_steps;
    outcome->n_events = n_events;
    outcome->n_remaining = n_rem;

    free(slow.buf);
    free(fast.buf);
    free(cache.b);
    return 0;
}

int orc_simulate(const orc_scenario *s, const orc_params *p, const orc_lineage *lineages,
                 size_t n, orc_speciation_event *events_out, orc_lineage *remaining_out,
                 orc_outcome *outcome) {
    return simulate_impl(s, p, lineages, n, events_out, NULL, remaining_out, outcome);
}

int orc_simulate_packed(const orc_scenario *s, const orc_params *p, const orc_lineage *lineages,
                        size_t n, orc_packed_event **events_out, size_t *n_events, orc_lineage *remaining_out,
                        orc_outcome *outcome) {
    packed_sink sink = {NULL, 0, 0};
    int rc = simulate_impl(s, p, lineages, n, NULL, &sink, remaining_out, outcome);
    *events_out = sink.b;
    *n_events = sink.n;
    return rc;
}

void orc_free(void *ptr) { free(ptr); }

/* ======================================================================== */
/* IsolatedIndividuals CPU baseline (multi-threaded)                         */
/* rustcoalescence/algorithms/independent/src/launch.rs:89-100               */
/* ======================================================================== */

typedef struct {
    const orc_scenario *s;
    const orc_params *p;
    double percentage;
    uint32_t cx, cy, r, rank, size;
    uint64_t steps, events;
    double seconds;
} part_job;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void *part_run(void *arg) {
    part_job *j = (part_job *)arg;
    size_t total = 0;
    orc_sample_origins(j->s, j->percentage, j->rank, j->size, j->cx, j->cy, j->r, NULL, 0, &total);
    orc_lineage *l = (orc_lineage *)malloc(sizeof(orc_lineage) * (total ? total : 1));
    orc_sample_origins(j->s, j->percentage, j->rank, j->size, j->cx, j->cy, j->r, l, total, &total);
    orc_speciation_event *ev = (orc_speciation_event *)malloc(sizeof(orc_speciation_event) * (total ? total : 1));
    orc_outcome o;
    double t0 = now_s();
    orc_simulate(j->s, j->p, l, total, ev, NULL, &o);
    j->seconds = now_s() - t0;
    j->steps = o.steps;
    j->events = o.n_events;
    free(l);
    free(ev);
    return NULL;
}

uint64_t orc_simulate_partitioned(const orc_scenario *s, const orc_params *p, double percentage,
                                  uint32_t cx, uint32_t cy, uint32_t r, uint32_t threads,
                                  uint32_t partitions, double *seconds, uint64_t *n_events) {
    if (threads < 1) threads = 1;
    if (partitions < threads) partitions = threads;
    part_job *jobs = (part_job *)calloc(threads, sizeof(part_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (uint32_t t = 0; t < threads; t++) {
        jobs[t].s = s; jobs[t].p = p; jobs[t].percentage = percentage;
        jobs[t].cx = cx; jobs[t].cy = cy; jobs[t].r = r;
        jobs[t].rank = t; jobs[t].size = partitions;
        pthread_create(&tids[t], NULL, part_run, &jobs[t]);
    }
    uint64_t steps = 0, events = 0;
    double wall = 0.0;
    for (uint32_t t = 0; t < threads; t++) {
        pthread_join(tids[t], NULL);
        steps += jobs[t].steps;
        events += jobs[t].events;
        if (jobs[t].seconds > wall) wall = jobs[t].seconds;
    }
    free(jobs);
    free(tids);
    if (seconds) *seconds = wall;
    if (n_events) *n_events = events;
    return steps;
}

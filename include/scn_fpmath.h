/*
 * scn_fpmath.h -- deterministic FP64 arithmetic shared by the HIP engine and the CPU oracle.
 *
 * Everything the multigroup tracking loop needs beyond IEEE +, -, *, /, sqrt and floor lives
 * here, written with those operations only and evaluated in a fixed order, so a gfx950 kernel
 * and a gcc build produce the same bits for the same inputs (both sides are compiled with
 * -ffp-contract=off and without fast-math).
 *
 *  - scn_rng_*   : SCONE's 63-bit LCG (RandomNumbers/RNG_class.f90:40-44 constants,
 *                  get :147-168, skip :251-304, stride :310-316).  Brown's O(log k) skip.
 *  - scn_log     : natural log by the fdlibm e_log.c method (argument reduction to
 *                  [sqrt(2)/2, sqrt(2)), s = f/(2+f), log(1+f) = f - (hfsq - s*(hfsq + R))),
 *                  with R the Taylor series of 2*atanh(s)/s - 2 truncated after z^10
 *                  (coefficients 2/(2k+1)); <= 1 ulp against libm.
 *  - scn_sincos  : Cody-Waite reduction by a three-part pi/2, Taylor polynomials to x^19 / x^18.
 *  - scn_norm2   : the scaled two-norm Fortran NORM2 evaluates (gfortran's inline expansion:
 *                  scale/sum update per nonzero element, scale*sqrt(sum)).
 *  - scn_rotate  : genericProcedures.f90:1047-1084 rotateVector, including the single-precision
 *                  literal 1E-8 of its pole test and the final division by norm2.
 *
 * The transcendental functions are not correctly rounded; their bits therefore match glibc's
 * only most of the time.  What is exact is CPU == GPU.
 */
#ifndef SCN_FPMATH_H
#define SCN_FPMATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SCN_HD __host__ __device__ static inline
#else
#define SCN_HD static inline
#endif

/* ---- constants (universalVariables.f90, numPrecision.f90) ------------------------------- */
#define SCN_INF        9223372036854775808.0          /* 2**63, INF / INFINITY            */
#define SCN_SURF_TOL   1.0e-12                        /* SURF_TOL                          */
#define SCN_FP_REL_TOL 1.0e-7                         /* FP_REL_TOL                        */
#define SCN_NUDGE      1.0e-8
#define SCN_PI         0x1.921fb54442d18p+1           /* 4*atan(1)                          */
#define SCN_TWO_PI     0x1.921fb54442d18p+2
#define SCN_HUGE       1.7976931348623157e308         /* huge(1.0_defReal)                  */
/* single-precision literals of the reference, promoted to double */
#define SCN_ST_TOL      9.999999960041972e-13         /* transportOperatorST_class.f90:58 1.0E-12   */
#define SCN_ROT_POLE    9.99999993922529e-09          /* genericProcedures.f90:1069 1E-8            */
#define SCN_KAPPA_DEF   202.27000427246094            /* fissionMG_class.f90:64 202.27              */
#define SCN_CHI_TOL_F   0.009999999776482582          /* fissionMG_class.f90:268 0.01 (times FP_REL_TOL) */
#define SCN_MISS_TOL    (1.0 + 10.0 * 2.220446049250313e-16) /* box_class FP_MISS_TOL            */

SCN_HD uint64_t scn_dbits(double x) { uint64_t b; __builtin_memcpy(&b, &x, 8); return b; }
SCN_HD double scn_bitsd(uint64_t b) { double x; __builtin_memcpy(&x, &b, 8); return x; }

/* Fortran MIN/MAX/SIGN on doubles, written once so both builds make the same choice on ties. */
SCN_HD double scn_fmin(double a, double b) { return (b < a) ? b : a; }
SCN_HD double scn_fmax(double a, double b) { return (b > a) ? b : a; }
SCN_HD double scn_sign(double a, double b) { return copysign(fabs(a), b); }

/* floor(x / y) exactly as the IEEE division rounds it, mostly without the division.
 * t = x * y_inv with y_inv = fl(1 / y) is within 3.5e-16 |t| of fl(x / y); when t is further than
 * 4e-15 |t| from both integers around it, floor(t) = floor(fl(x / y)).  Otherwise (t close to an
 * integer, or not finite) the division decides.  Near zero the sign alone decides and the test is
 * relative, so small quotients never fall back. */
SCN_HD double scn_floor_div(double x, double y, double y_inv)
{
    const double t = x * y_inv;
    const double f = floor(t);
    const double tol = 4.0e-15 * fabs(t);
    if (t - f > tol && (f + 1.0) - t > tol) return f;
    return floor(x / y);
}

/* a / n for 0 <= a < 2^31 and n >= 1 given n_inv = fl(1 / n), without an integer division:
 * (a + 1/2) / n = q + (r + 1/2) / n lies at least 1 / (2n) inside (q, q + 1), and the product with
 * n_inv is off by at most a 2^-52 / n < 1 / (2n), so its truncation is q. */
SCN_HD int scn_idiv(int a, double n_inv) { return (int)(((double)a + 0.5) * n_inv); }

/* Order-independent score accumulation.  A score v (|v| < 2^63) becomes the 128-bit two's-
 * complement integer q = v * 2^64 truncated below 2^-64: hi = floor(v), lo = (v - floor(v)) 2^64.
 * It is added as four 32-bit limbs, each into its own 64-bit word, so a word holds 2^31 additions
 * without overflow and the four words are the same whatever the order of the additions.  The
 * value of a bin is read back by carrying the words into 128 bits and rounding once per half
 * (hi exactly, lo * 2^-64 then the sum): deterministic, within 2 ulp of the exact sum of the
 * truncated scores, each truncation at most 2^-64 absolute. */
typedef struct scn_fix { uint64_t lo; int64_t hi; } scn_fix;

SCN_HD scn_fix scn_fix_of(double v)
{
    scn_fix q;
    const double h = floor(v);
    q.hi = (int64_t)h;
    q.lo = (uint64_t)((v - h) * 18446744073709551616.0);
    return q;
}

/* the four limbs of q: w[0] += bits 0-31, w[1] += bits 32-63, w[2] += bits 64-95,
 * w[3] += bits 96-127 sign-extended (added as two's complement) */
SCN_HD void scn_fix_limbs(scn_fix q, uint64_t limb[4])
{
    limb[0] = q.lo & 0xFFFFFFFFULL;
    limb[1] = q.lo >> 32;
    limb[2] = (uint64_t)q.hi & 0xFFFFFFFFULL;
    limb[3] = (uint64_t)(q.hi >> 32);
}

SCN_HD double scn_fix_value(const uint64_t w[4])
{
    const uint64_t lo = w[0] + (w[1] << 32);
    const uint64_t carry = lo < w[0] ? 1ULL : 0ULL;
    const uint64_t hi = (w[1] >> 32) + w[2] + (w[3] << 32) + carry;
    return (double)(int64_t)hi + (double)lo * 0x1.0p-64;
}

/* ---- RNG ---------------------------------------------------------------------------------- */
#define SCN_LCG_G      2806196910506780709ULL
#define SCN_LCG_MASK   0x7FFFFFFFFFFFFFFFULL
#define SCN_STRIDE     152917LL

/* RNG_class.f90:147-168: s <- (g*s + 1) mod 2^63, u = real(s) * 2^-63 */
SCN_HD double scn_rng_get(uint64_t* s)
{
    uint64_t x = (SCN_LCG_G * (*s)) & SCN_LCG_MASK;
    x = (x + 1ULL) & SCN_LCG_MASK;
    *s = x;
    return (double)(int64_t)x * 0x1p-63;
}

/* RNG_class.f90:251-304 (Brown's algorithm); negative k wraps through the period 2^63 */
SCN_HD uint64_t scn_rng_skip(uint64_t s, int64_t k_in)
{
    uint64_t k;
    if (k_in >= 0) k = (uint64_t)k_in;
    else k = (uint64_t)(INT64_MAX - (-k_in) + 1);
    k &= SCN_LCG_MASK;
    uint64_t Gk = 1ULL, Ck = 0ULL, gsq = SCN_LCG_G, L = 1ULL;
    while (k > 0) {
        if (k & 1ULL) {
            Gk = (Gk * gsq) & SCN_LCG_MASK;
            Ck = (Ck * gsq) & SCN_LCG_MASK;
            Ck = (Ck + L) & SCN_LCG_MASK;
        }
        L = (L * (gsq + 1ULL)) & SCN_LCG_MASK;
        gsq = (gsq * gsq) & SCN_LCG_MASK;
        k >>= 1;
    }
    return (Gk * s + Ck) & SCN_LCG_MASK;
}

/* RNG_class.f90:310-316: skip by strideSize * n */
SCN_HD uint64_t scn_rng_stride(uint64_t s, int64_t n) { return scn_rng_skip(s, SCN_STRIDE * n); }

#ifdef __cplusplus
/* stride(s, n) = A_n s + C_n (mod 2^63): the maps of strides b * 256^k (k < 5, b < 256), built at
 * compile time; a stride by n < 2^40 applies the five maps of n's bytes (they are powers of one
 * map, so they commute).  The same states as scn_rng_stride, with five multiplies. */
struct scn_rng_stride_table {
    uint64_t a[5][256];
    uint64_t c[5][256];
};

constexpr scn_rng_stride_table scn_make_rng_stride_table()
{
    scn_rng_stride_table t{};
    uint64_t ja[63] = {}, jc[63] = {};   /* the generator advanced by 2^i steps */
    uint64_t a = SCN_LCG_G, c = 1ULL;
    for (int i = 0; i < 63; ++i) {
        ja[i] = a;
        jc[i] = c;
        c = (c * (a + 1ULL)) & SCN_LCG_MASK;
        a = (a * a) & SCN_LCG_MASK;
    }
    for (int k = 0; k < 5; ++k) {
        /* the map of a stride by 256^k histories: skip 152917 * 256^k steps */
        uint64_t steps = (uint64_t)SCN_STRIDE << (8 * k), ma = 1ULL, mc = 0ULL;
        for (int i = 0; i < 63; ++i)
            if ((steps >> i) & 1ULL) {
                mc = (ja[i] * mc + jc[i]) & SCN_LCG_MASK;
                ma = (ja[i] * ma) & SCN_LCG_MASK;
            }
        t.a[k][0] = 1ULL;
        t.c[k][0] = 0ULL;
        for (int b = 1; b < 256; ++b) {
            t.c[k][b] = (ma * t.c[k][b - 1] + mc) & SCN_LCG_MASK;
            t.a[k][b] = (ma * t.a[k][b - 1]) & SCN_LCG_MASK;
        }
    }
    return t;
}

/* the generator advanced by 4b steps, b < 256: skips the 4-draw attempts of a rejection loop */
struct scn_rng_skip4_table {
    uint64_t a[256];
    uint64_t c[256];
};

constexpr scn_rng_skip4_table scn_make_rng_skip4_table()
{
    scn_rng_skip4_table t{};
    uint64_t a4 = 1ULL, c4 = 0ULL;
    for (int i = 0; i < 4; ++i) {
        c4 = (SCN_LCG_G * c4 + 1ULL) & SCN_LCG_MASK;
        a4 = (SCN_LCG_G * a4) & SCN_LCG_MASK;
    }
    t.a[0] = 1ULL;
    t.c[0] = 0ULL;
    for (int b = 1; b < 256; ++b) {
        t.c[b] = (a4 * t.c[b - 1] + c4) & SCN_LCG_MASK;
        t.a[b] = (a4 * t.a[b - 1]) & SCN_LCG_MASK;
    }
    return t;
}

SCN_HD uint64_t scn_rng_stride_bytes(const scn_rng_stride_table& t, uint64_t s, int64_t n)
{
    if (n < 0 || n >= ((int64_t)1 << 40)) return scn_rng_stride(s, n);
    const uint64_t u = (uint64_t)n;
    for (int k = 0; k < 5; ++k) {
        const int b = (int)((u >> (8 * k)) & 255ULL);
        s = (t.a[k][b] * s + t.c[k][b]) & SCN_LCG_MASK;
    }
    return s;
}
#endif

/* ---- log --------------------------------------------------------------------------------- */
SCN_HD double scn_log(double x)
{
    const double LN2_HI = 0x1.62e42fee00000p-1;
    const double LN2_LO = 0x1.a39ef35793c76p-33;
    const double SQRT2 = 0x1.6a09e667f3bcdp+0;
    if (!(x > 0.0)) {
        if (x == 0.0) return -HUGE_VAL;
        return NAN;
    }
    uint64_t b = scn_dbits(x);
    if (b >= 0x7FF0000000000000ULL) return x; /* +inf */
    int k = 0;
    if (b < 0x0010000000000000ULL) { /* subnormal */
        x = x * 0x1p54;
        b = scn_dbits(x);
        k = -54;
    }
    k += (int)(b >> 52) - 1023;
    double m = scn_bitsd((b & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);
    if (m > SQRT2) {
        m = m * 0.5;
        k += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = z * (0x1.5555555555555p-1 + z * (0x1.999999999999ap-2 + z * (0x1.2492492492492p-2 +
               z * (0x1.c71c71c71c71cp-3 + z * (0x1.745d1745d1746p-3 + z * (0x1.3b13b13b13b14p-3 +
               z * (0x1.1111111111111p-3 + z * (0x1.e1e1e1e1e1e1ep-4 + z * (0x1.af286bca1af28p-4 +
               z * 0x1.8618618618618p-4)))))))));
    double hfsq = 0.5 * f * f;
    double dk = (double)k;
    return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

/* ---- sin / cos --------------------------------------------------------------------------- */
SCN_HD void scn_sincos(double x, double* sn, double* cs)
{
    const double INVPIO2 = 0x1.45f306dc9c883p-1;
    const double PIO2_1 = 0x1.921fb54400000p+0;
    const double PIO2_2 = 0x1.0b4611a600000p-34;
    const double PIO2_3 = 0x1.3198a2e037073p-69;
    double kd = floor(x * INVPIO2 + 0.5);
    double r = ((x - kd * PIO2_1) - kd * PIO2_2) - kd * PIO2_3;
    int q = ((int)(int64_t)kd) & 3;
    double z = r * r;
    double sp = -0x1.5555555555555p-3 + z * (0x1.1111111111111p-7 + z * (-0x1.a01a01a01a01ap-13 +
                z * (0x1.71de3a556c734p-19 + z * (-0x1.ae64567f544e4p-26 + z * (0x1.6124613a86d09p-33 +
                z * (-0x1.ae7f3e733b81fp-41 + z * (0x1.952c77030ad4ap-49 + z * -0x1.2f49b46814157p-57)))))));
    double s = r + r * z * sp;
    double cp = 0x1.5555555555555p-5 + z * (-0x1.6c16c16c16c17p-10 + z * (0x1.a01a01a01a01ap-16 +
                z * (-0x1.27e4fb7789f5cp-22 + z * (0x1.1eed8eff8d898p-29 + z * (-0x1.93974a8c07c9dp-37 +
                z * (0x1.ae7f3e733b81fp-45 + z * -0x1.6827863b97d97p-53))))));
    double c = 1.0 - (0.5 * z - z * z * cp);
    switch (q) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
    }
}

/* ---- Fortran NORM2 of a 3-vector --------------------------------------------------------- */
SCN_HD double scn_norm2(const double v[3])
{
    double result = 0.0, scale = 1.0;
    for (int i = 0; i < 3; ++i) {
        if (v[i] != 0.0) {
            double ax = fabs(v[i]);
            if (scale < ax) {
                double val = scale / ax;
                result = 1.0 + result * val * val;
                scale = ax;
            } else {
                double val = ax / scale;
                result = result + val * val;
            }
        }
    }
    return scale * sqrt(result);
}

/* ---- rotateVector (genericProcedures.f90:1047-1084) -------------------------------------- */
SCN_HD void scn_rotate(const double dir[3], double mu, double phi, double out[3])
{
    double sinPol, cosPol;
    scn_sincos(phi, &sinPol, &cosPol);
    double u = dir[0], v = dir[1], w = dir[2];
    double A = sqrt(scn_fmax(0.0, 1.0 - mu * mu));
    double B = sqrt(scn_fmax(0.0, 1.0 - w * w));
    double n[3];
    if (B > SCN_ROT_POLE) {
        n[0] = mu * u + A * (u * w * cosPol - v * sinPol) / B;
        n[1] = mu * v + A * (v * w * cosPol + u * sinPol) / B;
        n[2] = mu * w - A * B * cosPol;
    } else {
        B = sqrt(scn_fmax(0.0, 1.0 - v * v));
        n[0] = mu * u + A * (u * v * cosPol + w * sinPol) / B;
        n[1] = mu * v - A * B * cosPol;
        n[2] = mu * w + A * (v * w * cosPol - u * sinPol) / B;
    }
    double nn = scn_norm2(n);
    out[0] = n[0] / nn;
    out[1] = n[1] / nn;
    out[2] = n[2] / nn;
}

/* ---- canonical cross-history sum --------------------------------------------------------
 * Sums that feed back into the simulation (population weights, k-eff clerk scores) are taken
 * in one fixed order on both sides: consecutive chunks of SCN_DSUM_CHUNK values are summed
 * left to right starting from 0.0, then the chunk sums are combined pairwise level by level
 * (y[i] = y[2i] + y[2i+1]; an odd tail element is carried up unchanged).                      */
#define SCN_DSUM_CHUNK 32

#endif /* SCN_FPMATH_H */

/* hh_pow.h -- reproducible FP64 power functions shared by the HIP kernels and
 * the CPU oracle.
 *
 * hiten evaluates `r2**1.5`, `r2**2.5` (dynamics/rtbp.py:119-122, 230-231)
 * and the PI step-size factors `err**(-beta)`, `err_prev**alpha`,
 * `err**(-1/8)` (integrators/utils.py:242-247, 280) through numba, which
 * lowers a float exponent to llvm.pow, i.e. the C library `pow`.  The GPU has
 * no glibc, and ocml's pow differs from glibc's in the last bit, which the
 * chaotic CR3BP flow amplifies.  These routines are written only with IEEE
 * +, -, *, /, sqrt and fused multiply-add, so a host build and a device build
 * return the same bits:
 *
 *   hh_pow15(x), hh_pow25(x)  x^1.5 and x^2.5, correctly rounded: the square
 *                             root's remainder (one fma) gives a double-double
 *                             correction and the product is summed before one
 *                             final rounding.
 *   hh_pow(x, y)              x^y for y a double, via a table-driven
 *                             double-double log and exp (~2^-68 relative error
 *                             before the final rounding), i.e. correctly
 *                             rounded except in rare near-midpoint cases, the
 *                             same accuracy class as glibc's pow.
 *   hh_pow_t(x, y, tables)    hh_pow on tables passed by pointer (a kernel's
 *                             LDS copy of them); same bits.
 *
 * Contract: translation units including this header must be compiled with
 * -ffp-contract=off (no implicit fma), and fma() must be a true fused
 * multiply-add (gfx950 v_fma_f64; x86-64 vfmadd with -mfma).
 *
 * The includer defines HH_FN (function qualifiers) and HH_TABLE (table
 * storage) before including; see hh_tables.h.
 */
#ifndef HH_POW_H
#define HH_POW_H

#ifndef HH_FN
#error "define HH_FN before including hh_pow.h"
#endif

#include "hh_tables.h"

/* The correctly rounded square root; a device includer may supply a faster
 * correctly rounded one for the argument range it guarantees.              */
#ifndef HH_SQRT
#define HH_SQRT(x) __builtin_sqrt(x)
#endif

#ifdef __cplusplus
#define HH_U64 unsigned long long
#else
#include <stdint.h>
#define HH_U64 uint64_t
#endif

HH_FN HH_U64 hh_bits(double x) {
    HH_U64 u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

HH_FN double hh_from_bits(HH_U64 u) {
    double x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

/* Exact product a*b = p + e. */
#define HH_TWO_PROD(a, b, p, e) do { (p) = (a) * (b); (e) = __builtin_fma((a), (b), -(p)); } while (0)
/* Exact sum a+b = s + e (any magnitudes). */
#define HH_TWO_SUM(a, b, s, e) do { double _bb; (s) = (a) + (b); _bb = (s) - (a); \
    (e) = ((a) - ((s) - _bb)) + ((b) - _bb); } while (0)

/* Special values shared by hh_pow15/hh_pow25 (x <= 0, inf or NaN). */
HH_FN double hh_pow_pos_special(double x) {
    if (x == 0.0) return 0.0;                         /* +-0 ** 1.5 = +0 */
    if (x > 0.0) return x;                            /* +inf            */
    return __builtin_nan("");                         /* x < 0 or NaN    */
}

/* Correctly rounded x^1.5 and x^2.5.
 *
 * With s = RN(sqrt(x)) and the exact remainder e = x - s*s (one fma),
 * sqrt(x) = s + d with d = e / (sqrt(x) + s), so
 *   x * d   = (s * e / 2)     * (1 + O(2^-52))
 *   x^2 * d = (x * s * e / 2) * (1 + O(2^-52)).
 * Each correction is ~2^-53 of the result, so evaluating it to double
 * precision leaves an error of ~2^-104 relative before the single final
 * rounding of (p + (pe + correction)), where p + pe is the exact product of
 * the leading terms.  No division is needed.                              */
HH_FN double hh_pow15(double x) {
    if (!(x > 0.0) || x > 1.7976931348623157e308) return hh_pow_pos_special(x);
    double s = HH_SQRT(x);
    double e = __builtin_fma(-s, s, x);               /* x - s*s, exact    */
    double p, pe;
    HH_TWO_PROD(x, s, p, pe);
    double r = p + (pe + s * (0.5 * e));
    return (p > 1.7976931348623157e308) ? p : r;      /* overflow -> +inf   */
}

HH_FN double hh_pow25(double x) {
    if (!(x > 0.0) || x > 1.7976931348623157e308) return hh_pow_pos_special(x);
    double s = HH_SQRT(x);
    double e = __builtin_fma(-s, s, x);
    double q, qe;
    HH_TWO_PROD(x, x, q, qe);                         /* x^2 = q + qe      */
    double p, pe;
    HH_TWO_PROD(q, s, p, pe);
    double r = p + (pe + ((x * s) * (0.5 * e) + qe * s));
    return (p > 1.7976931348623157e308) ? p : r;
}

/* Both powers of one argument sharing the square root (the variational
 * field needs r^3 and r^5 of the same r^2).                                */
HH_FN void hh_pow15_25(double x, double* p15, double* p25) {
    if (!(x > 0.0) || x > 1.7976931348623157e308) {
        *p15 = hh_pow_pos_special(x);
        *p25 = *p15;
        return;
    }
    double s = HH_SQRT(x);
    double e = __builtin_fma(-s, s, x);
    double he = 0.5 * e;
    double p, pe;
    HH_TWO_PROD(x, s, p, pe);
    double r15 = p + (pe + s * he);
    *p15 = (p > 1.7976931348623157e308) ? p : r15;
    double q, qe;
    HH_TWO_PROD(x, x, q, qe);
    double p2, pe2;
    HH_TWO_PROD(q, s, p2, pe2);
    double r25 = p2 + (pe2 + (p * he + qe * s));
    *p25 = (p2 > 1.7976931348623157e308) ? p2 : r25;
}

/* x^y for finite y.  Special cases follow C99 pow for the cases a step-size
 * controller can produce (x = 0, inf, NaN, 1, negative).  The five tables
 * come in as pointers, so that a device caller can read copies of them from
 * LDS; hh_pow below reads them where they are defined.                      */
HH_FN double hh_pow_t(double x, double y, const double* invc, const double* log_hi,
                      const double* log_lo, const double* exp_hi, const double* exp_lo) {
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;               /* NaN                */
    if (x == 1.0) return 1.0;
    if (x == 0.0) return (y > 0.0) ? 0.0 : __builtin_inf();
    if (x < 0.0) return __builtin_nan("");            /* non-integer y only */
    if (x > 1.7976931348623157e308) return (y > 0.0) ? __builtin_inf() : 0.0;

    /* x = 2^k * m, m in [1, 2) */
    int k = 0;
    HH_U64 ux = hh_bits(x);
    if ((ux >> 52) == 0) {                            /* subnormal          */
        x = x * 18014398509481984.0;                  /* 2^54               */
        ux = hh_bits(x);
        k = -54;
    }
    k += (int)((ux >> 52) & 0x7ff) - 1023;
    double m = hh_from_bits((ux & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);

    /* m = c_i (1 + z), c_i = 1 + i/128, |z| < 2^-7, z = zh + zl */
    int i = (int)((ux >> 45) & 0x7f);
    double c = 1.0 + (double)i * 0.0078125;
    double dlt = m - c;                               /* exact (Sterbenz)   */
    double ic = invc[i];                       /* RN(1 / c_i)        */
    double zh = dlt * ic;
    /* c_i has 8 significant bits, so dlt - zh*c_i is exact in one fma      */
    double zl = __builtin_fma(-zh, c, dlt) * ic;

    /* log1p(z) = z - z^2/2 + ... ; leading part as a double-double */
    double sq, sqe;
    HH_TWO_PROD(zh, zh, sq, sqe);
    double a_hi, a_lo;
    HH_TWO_SUM(zh, -0.5 * sq, a_hi, a_lo);
    double w = zh;
    double tail = w * w * w * (1.0 / 3.0 - w * (0.25 - w * (0.2 - w * (1.0 / 6.0 - w * (1.0 / 7.0
                  - w * (0.125 - w * (1.0 / 9.0 - w * 0.1)))))));
    double small = (-0.5 * sqe + (zl - zl * zh)) + tail;

    /* L = k ln2 + log(c_i) + log1p(z) */
    double kd = (double)k;
    double s1, e1, s2, e2;
    HH_TWO_SUM(kd * HH_POW_LN2_HI, log_hi[i], s1, e1);
    HH_TWO_SUM(s1, a_hi, s2, e2);
    double l_lo = ((e1 + e2) + (kd * HH_POW_LN2_LO + log_lo[i])) + (a_lo + small);
    double l_hi, l_lo2;
    HH_TWO_SUM(s2, l_lo, l_hi, l_lo2);

    /* t = y * L */
    double t_hi, t_lo;
    HH_TWO_PROD(y, l_hi, t_hi, t_lo);
    t_lo = t_lo + y * l_lo2;
    if (t_hi > 709.8) return __builtin_inf();
    if (t_hi < -745.2) return 0.0;

    /* exp(t) = 2^(N/128) * exp(r),  |r| <= ln2/256 */
    double nd = (t_hi * HH_POW_INV_L128 + 6755399441055744.0) - 6755399441055744.0;
    int n = (int)nd;
    int j = n & 127;
    int e = (n - j) / 128;
    double r_hi = t_hi - nd * HH_POW_L128_HI;         /* exact              */
    double r_lo = t_lo - nd * HH_POW_L128_LO;
    double r, rr;
    HH_TWO_SUM(r_hi, r_lo, r, rr);
    double q = r * r * (0.5 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0 + r * (1.0 / 720.0)))))
               + (rr + rr * r);
    double e_hi, e_lo;
    HH_TWO_SUM(1.0, r, e_hi, e_lo);
    e_lo = e_lo + q;
    double p_hi, p_lo;
    HH_TWO_PROD(exp_hi[j], e_hi, p_hi, p_lo);
    p_lo = p_lo + (exp_hi[j] * e_lo + exp_lo[j] * e_hi);
    double res = p_hi + p_lo;
    /* scale by 2^e in two exact steps (keeps each factor a normal number) */
    int e1s = e / 2;
    int e2s = e - e1s;
    res = res * hh_from_bits((HH_U64)(e1s + 1023) << 52);
    res = res * hh_from_bits((HH_U64)(e2s + 1023) << 52);
    return res;
}

HH_FN double hh_pow(double x, double y) {
    return hh_pow_t(x, y, HH_POW_INVC, HH_POW_LOG_HI, HH_POW_LOG_LO, HH_POW_EXP_HI, HH_POW_EXP_LO);
}

#endif /* HH_POW_H */

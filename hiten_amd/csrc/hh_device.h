/* hh_device.h -- device-side building blocks of the CR3BP propagation
 * kernels for gfx950 (CDNA4, wave64).
 *
 * Every arithmetic expression here is written in the operation order of
 * hiten's numba kernels (fastmath=False), and the library is compiled with
 * -ffp-contract=off, so that a kernel's FP64 results are bit-identical to a
 * scalar CPU evaluation of the reference formulas:
 *   rhs6            dynamics/rtbp.py:65-74   (_crtbp_accel)
 *   Lane42::rhs     dynamics/rtbp.py:115-141, 227-242 (Jacobian entries and
 *                   the state part of _var_equations)
 *   column_rhs      dynamics/rtbp.py:219-225 (F @ Phi; the zero entries of F
 *                   contribute exactly +-0 to the reference's running sum, so
 *                   only the non-zero products are formed, starting the sum
 *                   from +0.0 exactly as the reference does)
 *   DOP853 stepping integrators/rk.py:1636-1708, 1790-2003, 2005-2102,
 *                   2443-2484; integrators/utils.py:13-296
 *
 * Two lane layouts share the stepping code through a policy class:
 *   Lane6   one trajectory per lane, the whole 6-d state in registers.
 *   Lane42  one trajectory per 8-lane group; lane b of the group owns one
 *           6-double block of the 42-d variational state (b = 0: the state
 *           x, b = 1..6: STM column b-1, b = 7: idle padding).  The state
 *           position is broadcast to the group with DPP moves at every stage,
 *           error norms are reduced across the group with an xor butterfly
 *           (ds_swizzle).
 */
#ifndef HH_DEVICE_H
#define HH_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HH_TABLE static __device__ const
#define HH_FN static __device__ __forceinline__

/* Correctly rounded FP64 square root.  For normal x >= 2^-767 (the range in
 * which the compiler's own expansion applies no scaling) this is that
 * expansion without its scaling and special-value fixups: the hardware
 * reciprocal square root refined by the coupled iteration g ~ sqrt(x),
 * h ~ 1/(2 sqrt(x)) and two residual corrections d = x - g*g (fma), so the
 * result is the same bits as __builtin_sqrt; other x take __builtin_sqrt.  */
static __device__ __forceinline__ bool hh_dev_sqrt_fast(double x) {
    return x >= 0x1p-767 && x <= 1.7976931348623157e308;
}
/* the expansion for hh_dev_sqrt_fast(x) only */
static __device__ __forceinline__ double hh_dev_sqrt_core(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    double g = x * y0;
    double h = y0 * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
static __device__ __forceinline__ double hh_dev_sqrt(double x) {
    if (!hh_dev_sqrt_fast(x)) return __builtin_sqrt(x);
    return hh_dev_sqrt_core(x);
}
/* hh_dev_sqrt with one wave-uniform test in place of a per-lane branch:
 * when every active lane's argument is in the fast range (the case of every
 * CR3BP distance away from a collision) the wave runs the expansion alone,
 * else every lane takes hh_dev_sqrt.  Same bits either way.              */
static __device__ __forceinline__ double hh_dev_sqrt_wave(double x) {
    if (__ballot(!hh_dev_sqrt_fast(x)) == 0) return hh_dev_sqrt_core(x);
    return hh_dev_sqrt(x);
}
#define HH_SQRT(x) hh_dev_sqrt(x)
#include "hh_pow.h"

/* hh_pow15_25 with one wave-uniform test in place of its per-lane branches
 * (special arguments, and the square root's range): when every active
 * lane's argument is in [2^-767, DBL_MAX] the wave runs hh_pow15_25's
 * expressions on the fast square root directly, else every lane takes
 * hh_pow15_25.  Same bits either way.                                     */
static __device__ __forceinline__ void hh_dev_pow15_25(double x, double* p15, double* p25) {
    if (__ballot(!hh_dev_sqrt_fast(x)) != 0) {
        hh_pow15_25(x, p15, p25);
        return;
    }
    const double s = hh_dev_sqrt_core(x);
    const double e = __builtin_fma(-s, s, x);
    const double he = 0.5 * e;
    double p, pe;
    HH_TWO_PROD(x, s, p, pe);
    const double r15 = p + (pe + s * he);
    *p15 = (p > 1.7976931348623157e308) ? p : r15;
    double q, qe;
    HH_TWO_PROD(x, x, q, qe);
    double p2, pe2;
    HH_TWO_PROD(q, s, p2, pe2);
    const double r25 = p2 + (pe2 + (p * he + qe * s));
    *p25 = (p2 > 1.7976931348623157e308) ? p2 : r25;
}

namespace hh {

/* hh_pow's five tables in one block.  The latency layouts (one trajectory
 * per wave, pair or 8-lane group) read them from an LDS copy that the
 * kernel fills at its start: their step-size factors sit on the
 * trajectory's dependent chain, and two rounds of table loads from L2
 * under a full GPU cost more than the rest of the pow.                   */
struct PowTab {
    double invc[128], log_hi[128], log_lo[128], exp_hi[128], exp_lo[128];
};

/* copy the tables into t with the workgroup's threads; a barrier must
 * follow before any thread reads t                                       */
__device__ __forceinline__ void pow_tab_fill(PowTab* t) {
    const int nt = (int)(blockDim.x * blockDim.y * blockDim.z);
    for (int i = (int)threadIdx.x; i < 128; i += nt) {
        t->invc[i] = HH_POW_INVC[i];
        t->log_hi[i] = HH_POW_LOG_HI[i];
        t->log_lo[i] = HH_POW_LOG_LO[i];
        t->exp_hi[i] = HH_POW_EXP_HI[i];
        t->exp_lo[i] = HH_POW_EXP_LO[i];
    }
}

/* hh_pow on the tables at t (same bits as hh_pow) */
__device__ __forceinline__ double pow_tab(const PowTab* t, double x, double y) {
    return hh_pow_t(x, y, t->invc, t->log_hi, t->log_lo, t->exp_hi, t->exp_lo);
}

constexpr int DOP_S = 12;     /* stages of one DOP853 step                 */
constexpr int DOP_K = 13;     /* stored stage derivatives (k[12] = f(y_high)) */
constexpr int DOP_SX = 16;    /* stages of the dense-output extension       */
constexpr int DOP_P = 7;      /* interpolant rows                           */

constexpr int ST_OK = 0;
constexpr int ST_MAXSTEPS = 1;
constexpr int ST_NONFINITE = 2;
constexpr int ST_CANCELLED = 3;   /* stopped by its caller; the result is not used */
constexpr int ST_DROPPED = 4;     /* lost by the persistent scheduler (ruled out by its
                                     queue capacity; reported rather than silent)     */

__device__ __forceinline__ double np_max(double a, double b) {
    /* numpy.maximum: NaN-propagating */
    if (a != a || b != b) return a + b;
    return a >= b ? a : b;
}

__device__ __forceinline__ bool isclose(double a, double b) {
    /* numpy.isclose defaults (rtol 1e-5, atol 1e-8), dynamics/base.py:421 */
    return __builtin_fabs(a - b) <= 1e-8 + 1e-5 * __builtin_fabs(b);
}

/* ds_swizzle in bitmask mode: lane reads ((lane & AND) | OR) ^ XOR within
 * its 32-lane half.  Used on the two halves of a double.                  */
template <int PATTERN>
__device__ __forceinline__ double swz(double v) {
    int lo = __double2loint(v);
    int hi = __double2hiint(v);
    lo = __builtin_amdgcn_ds_swizzle(lo, PATTERN);
    hi = __builtin_amdgcn_ds_swizzle(hi, PATTERN);
    return __hiloint2double(hi, lo);
}

/* value held by lane K of this lane's aligned 8-lane group */
template <int K>
__device__ __forceinline__ double group_get(double v) { return swz<0x18 | (K << 5)>(v); }

template <int X>
__device__ __forceinline__ double group_xor(double v) { return swz<0x1f | (X << 10)>(v); }

/* Broadcast lane 0 of each aligned 8-lane group to the group with two DPP
 * moves per dword: quad_perm [0,0,0,0] (lanes 4..7 then hold lane 4), then
 * row_shr:4 written only to banks 1 and 3 (lanes 4..7 and 12..15 of each
 * row take lanes 0..3 and 8..11).                                          */
__device__ __forceinline__ int bcast8_dw(int v) {
    const int t = __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false);
    return __builtin_amdgcn_update_dpp(t, t, 0x114, 0xF, 0xA, false);
}

__device__ __forceinline__ double bcast8(double v) {
    const int lo = bcast8_dw(__double2loint(v));
    const int hi = bcast8_dw(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

/* exact negation when bit d of m is set (x * -1.0 as the reference does) */
__device__ __forceinline__ double neg_if(double x, unsigned m, int d) {
    const long long b = (long long)((m >> d) & 1u) << 63;
    return __longlong_as_double(__double_as_longlong(x) ^ b);
}

/* ---------------------------------------------------------------------- */
/* CR3BP vector fields                                                      */
/* ---------------------------------------------------------------------- */

/* dynamics/rtbp.py:65-74 */
__device__ __forceinline__ void rhs6(const double (&s)[6], double mu, double (&o)[6]) {
    const double x = s[0], y = s[1], z = s[2], vx = s[3], vy = s[4], vz = s[5];
    const double xm = x + mu;
    const double r1 = __builtin_sqrt((xm * xm + y * y) + z * z);
    const double one_mu = 1.0 - mu;
    const double xr = x - one_mu;
    const double r2 = __builtin_sqrt((xr * xr + y * y) + z * z);
    const double r1c = r1 * (r1 * r1);
    const double r2c = r2 * (r2 * r2);
    const double xm1 = (x - 1.0) + mu;
    o[0] = vx;
    o[1] = vy;
    o[2] = vz;
    o[3] = ((2.0 * vy + x) - (one_mu * xm) / r1c) - (mu * xm1) / r2c;
    o[4] = ((-2.0 * vx + y) - (one_mu * y) / r1c) - (mu * y) / r2c;
    o[5] = ((-one_mu) * z) / r1c - (mu * z) / r2c;
}

/* One STM column p = Phi[:, j] -> (F @ Phi)[:, j] given the Hessian entries
 * of the Jacobian (rtbp.py:124-141).                                       */
__device__ __forceinline__ void column_rhs(double uxx, double uyy, double uzz, double uxy,
                                           double uxz, double uyz, const double (&p)[6],
                                           double (&o)[6]) {
    o[0] = 0.0 + p[3];
    o[1] = 0.0 + p[4];
    o[2] = 0.0 + p[5];
    o[3] = (((0.0 + uxx * p[0]) + uxy * p[1]) + uxz * p[2]) + 2.0 * p[4];
    o[4] = (((0.0 + uxy * p[0]) + uyy * p[1]) + uyz * p[2]) + (-2.0) * p[3];
    o[5] = ((0.0 + uxz * p[0]) + uyz * p[1]) + uzz * p[2];
}

/* ---------------------------------------------------------------------- */
/* lane layout policies                                                     */
/* ---------------------------------------------------------------------- */

/* One trajectory per lane, 6-d state. */
struct Lane6 {
    static constexpr int N = 6;
    double mu;
    unsigned flipbits;  /* bit d: negate component d (backward _DirectedSystem) */
    bool flip_on;       /* kernel-uniform: any flip at all                     */

    __device__ __forceinline__ void rhs(const double (&s)[6], double (&o)[6]) const {
        rhs6(s, mu, o);
        if (flip_on) {
#pragma unroll
            for (int d = 0; d < 6; ++d) o[d] = neg_if(o[d], flipbits, d);
        }
    }
    __device__ __forceinline__ double reduce(double v) const { return v; }
};

/* One trajectory per aligned group of 8 lanes, 42-d variational state.
 *
 * The variational field needs twelve quotients (rtbp.py:124-141 and
 * 233-242).  The STM-column lanes use only
 *   mu2/r5, mu/R5, mu2/r3, mu/R3, (mu2 xm)/r5, (mu xr)/R5
 * and the state lane only
 *   xm/r3, xr/R3, y/r3, y/R3, z/r3, z/R3,
 * so each lane forms the six it needs (operands selected per lane) instead of
 * all twelve.  Only x, y, z are broadcast from the state lane.             */
struct Lane42 {
    static constexpr int N = 42;
    double mu, mu2;
    unsigned flipbits;
    bool flip_on;
    int blk;            /* lane & 7 */

    __device__ __forceinline__ void rhs(const double (&s)[6], double (&o)[6]) const {
        const double x = bcast8(s[0]);
        const double y = bcast8(s[1]);
        const double z = bcast8(s[2]);
        const bool st = (blk == 0);
        const double xm = x + mu;
        const double xr = x - mu2;
        const double r2 = (xm * xm + y * y) + z * z;
        const double R2 = (xr * xr + y * y) + z * z;
        double r3, r5, R3, R5;
        hh_dev_pow15_25(r2, &r3, &r5);
        hh_dev_pow15_25(R2, &R3, &R5);
        const double q1 = (st ? xm : mu2) / (st ? r3 : r5);
        const double q2 = (st ? xr : mu) / (st ? R3 : R5);
        const double q3 = (st ? y : mu2) / r3;
        const double q4 = (st ? y : mu) / R3;
        const double q5 = (st ? z : mu2 * xm) / (st ? r3 : r5);
        const double q6 = (st ? z : mu * xr) / (st ? R3 : R5);
        if (st) {
            /* rtbp.py:233-242 */
            o[0] = s[3];
            o[1] = s[4];
            o[2] = s[5];
            o[3] = ((x - mu2 * q1) - mu * q2) + 2.0 * s[4];
            o[4] = ((y - mu2 * q3) - mu * q4) - 2.0 * s[3];
            o[5] = (-mu2 * q5) - mu * q6;
        } else {
            /* rtbp.py:124-141 */
            const double a53 = q1 * 3.0;
            const double b53 = q2 * 3.0;
            const double c3 = q3 + q4;
            const double gx = q5 + q6;
            const double uxx = ((1.0 + a53 * (xm * xm)) + b53 * (xr * xr)) - c3;
            const double uyy = ((1.0 + a53 * (y * y)) + b53 * (y * y)) - c3;
            const double uzz = ((0.0 + a53 * (z * z)) + b53 * (z * z)) - c3;
            const double uxy = (3.0 * y) * gx;
            const double uxz = (3.0 * z) * gx;
            const double uyz = ((3.0 * y) * z) * (q1 + q2);
            column_rhs(uxx, uyy, uzz, uxy, uxz, uyz, s, o);
        }
        if (flip_on) {
#pragma unroll
            for (int d = 0; d < 6; ++d) o[d] = neg_if(o[d], flipbits, d);
        }
    }
    /* canonical block tree: ((b0+b1)+(b2+b3)) + ((b4+b5)+(b6+b7)), b7 = 0 */
    __device__ __forceinline__ double reduce(double v) const {
        v = (blk == 7) ? 0.0 : v;
        v = v + group_xor<1>(v);
        v = v + group_xor<2>(v);
        v = v + group_xor<4>(v);
        return v;
    }
};

/* component index of lane-block b, entry d, in the reference's 42-vector
 * layout [Phi row-major (36) | x (6)]                                       */
__device__ __forceinline__ int comp42(int b, int d) { return b == 0 ? 36 + d : 6 * d + (b - 1); }

/* ---------------------------------------------------------------------- */
/* step control: integrators/utils.py                                       */
/* ---------------------------------------------------------------------- */

__device__ __forceinline__ double select_initial_step(double d0, double d1, double min_step,
                                                      double max_step) {
    double h = (d0 < 1.0e-5 || d1 < 1.0e-5) ? 1.0e-6 : (0.01 * d0) / d1;
    if (h > max_step) h = max_step;
    if (h < min_step) h = min_step;
    return h;
}

__device__ __forceinline__ double clamp_step(double h, double max_step, double min_step) {
    if (h > max_step) h = max_step;
    if (h < min_step) h = min_step;
    return h;
}

__device__ __forceinline__ double adjust_step_to_endpoint(double t, double h, double t_end) {
    return (t + h > t_end) ? __builtin_fabs(t_end - t) : h;
}

/* beta = 1/9, alpha = 0.4/9 as the reference computes them for order 8 */
__device__ __forceinline__ double pi_accept_factor(double err, double err_prev) {
    const double beta = 1.0 / (8.0 + 1.0);
    const double alpha = 0.4 * beta;
    double f;
    if (err == 0.0) f = 10.0;
    else if (err_prev < 0.0) f = 0.9 * hh_pow(err, -beta);
    else f = (0.9 * hh_pow(err, -beta)) * hh_pow(err_prev, alpha);
    if (!(f == f)) f = 10.0;
    if (f < 0.2) f = 0.2;
    if (f > 10.0) f = 10.0;
    return f;
}

__device__ __forceinline__ double pi_reject_factor(double err) {
    const double err_exp = 1.0 / 8.0;
    double f = (err <= 0.0) ? 0.2 : 0.9 * hh_pow(err, -err_exp);
    if (!(f == f)) f = 0.2;
    if (f < 0.2) f = 0.2;
    if (f > 10.0) f = 10.0;
    return f;
}

/* pi_accept_factor(err, err_prev) and pi_reject_factor(err) of an aligned
 * 8-lane group holding one trajectory (err, err_prev group-uniform; every
 * lane of the group active): entry b of the group forms one of the three
 * powers (b = 0: err^-beta, b = 1: err_prev^alpha, b >= 2: err^(-1/8)), so
 * a wave forms one pow per lane whatever its groups decide, where the two
 * factors formed in their branches take up to three pows one after the
 * other.  Same expressions on the same powers: same bits.                 */
__device__ __forceinline__ void group_step_factors(int b, double err, double err_prev,
                                                   double& f_acc, double& f_rej) {
    const double beta = 1.0 / (8.0 + 1.0);
    const double alpha = 0.4 * beta;
    const double err_exp = 1.0 / 8.0;
    const double px = (b == 1) ? err_prev : err;
    const double pe = (b == 0) ? -beta : ((b == 1) ? alpha : -err_exp);
    const double pw = hh_pow(px, pe);
    const double pb = group_get<0>(pw), pa = group_get<1>(pw), pr = group_get<2>(pw);
    double f;
    if (err == 0.0) f = 10.0;
    else if (err_prev < 0.0) f = 0.9 * pb;
    else f = (0.9 * pb) * pa;
    if (!(f == f)) f = 10.0;
    if (f < 0.2) f = 0.2;
    if (f > 10.0) f = 10.0;
    f_acc = f;
    double g = (err <= 0.0) ? 0.2 : 0.9 * pr;
    if (!(g == g)) g = 0.2;
    if (g < 0.2) g = 0.2;
    if (g > 10.0) g = 10.0;
    f_rej = g;
}

/* rk.py:2443-2448 */
template <class P>
__device__ __forceinline__ double initial_step(const P& p, const double (&y)[6],
                                               const double (&f0)[6], double rtol, double atol,
                                               double min_step, double max_step) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const double sc = atol + rtol * __builtin_fabs(y[d]);
        const double a = y[d] / sc;
        const double b = f0[d] / sc;
        s0 += a * a;
        s1 += b * b;
    }
    s0 = p.reduce(s0);
    s1 = p.reduce(s1);
    const double sn = __builtin_sqrt((double)P::N);
    const double d0 = __builtin_sqrt(s0) / sn;
    const double d1 = __builtin_sqrt(s1) / sn;
    return select_initial_step(d0, d1, min_step, max_step);
}

/* ---------------------------------------------------------------------- */
/* DOP853 step attempt (rk.py:1636-1699) + scaled error norm (2457-2467)   */
/* k[0] must hold f(y) on entry; k[1..12] and yh are written.               */
/* ---------------------------------------------------------------------- */
template <class P>
__device__ __forceinline__ double dop853_attempt(const P& p, const double (&y)[6],
                                                 double (&k)[DOP_K][6], double h, double rtol,
                                                 double atol, double (&yh)[6]) {
#pragma unroll
    for (int i = 1; i < DOP_S; ++i) {
        double ys[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) ys[d] = y[d];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double a = HH_DOP_A[i][j];
            if (a != 0.0) {
                const double ha = h * a;
#pragma unroll
                for (int d = 0; d < 6; ++d) ys[d] = ys[d] + ha * k[j][d];
            }
        }
        p.rhs(ys, k[i]);
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) yh[d] = y[d];
#pragma unroll
    for (int j = 0; j < DOP_S; ++j) {
        const double b = HH_DOP_B[j];
        if (b != 0.0) {
            const double hb = h * b;
#pragma unroll
            for (int d = 0; d < 6; ++d) yh[d] = yh[d] + hb * k[j][d];
        }
    }
    p.rhs(yh, k[DOP_S]);
    double e5[6], e3[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        e5[d] = 0.0;
        e3[d] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) {
        const double c5 = HH_DOP_E5[j];
        const double c3 = HH_DOP_E3[j];
        if (c5 != 0.0) {
#pragma unroll
            for (int d = 0; d < 6; ++d) e5[d] = e5[d] + c5 * k[j][d];
        }
        if (c3 != 0.0) {
#pragma unroll
            for (int d = 0; d < 6; ++d) e3[d] = e3[d] + c3 * k[j][d];
        }
    }
    double s5 = 0.0, s3 = 0.0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const double sc = atol + rtol * np_max(__builtin_fabs(y[d]), __builtin_fabs(yh[d]));
        const double q5 = (e5[d] * h) / sc;
        const double q3 = (e3[d] * h) / sc;
        s5 += q5 * q5;
        s3 += q3 * q3;
    }
    s5 = p.reduce(s5);
    s3 = p.reduce(s3);
    if (s5 == 0.0 && s3 == 0.0) return 0.0;
    const double denom = s5 + 0.01 * s3;
    return (__builtin_fabs(h) * s5) / __builtin_sqrt(denom * (double)P::N);
}

/* The three extra stages of the dense-output extension (rk.py:1836-1857). */
template <class P>
__device__ __forceinline__ void dop853_ext_stages(const P& p, const double (&y_old)[6],
                                                  const double (&k)[DOP_K][6], double hs,
                                                  double (&kx)[DOP_SX - DOP_K][6]) {
#pragma unroll
    for (int sr = DOP_K; sr < DOP_SX; ++sr) {
        double ys[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < sr; ++r) {
                const double a = HH_DOP_A[sr][r];
                if (a != 0.0) acc = acc + a * (r < DOP_K ? k[r][d] : kx[r - DOP_K][d]);
            }
            ys[d] = y_old[d] + hs * acc;
        }
        p.rhs(ys, kx[sr - DOP_K]);
    }
}

/* Interpolant row i (0..6) of component d (rk.py:1859-1874). */
__device__ __forceinline__ double dense_row(int i, int d, const double (&y_old)[6],
                                            const double (&y_new)[6], const double (&k)[DOP_K][6],
                                            const double (&kx)[DOP_SX - DOP_K][6], double hs) {
    const double dy = y_new[d] - y_old[d];
    if (i == 0) return dy;
    if (i == 1) return hs * k[0][d] - dy;
    if (i == 2) return 2.0 * dy - hs * (k[DOP_S][d] + k[0][d]);
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < DOP_SX; ++r) {
        const double c = HH_DOP_D[i - 3][r];
        if (c != 0.0) acc = acc + c * (r < DOP_K ? k[r][d] : kx[r - DOP_K][d]);
    }
    return hs * acc;
}

/* Dense output at abscissa x (rk.py:1989-2003), each row formed as the
 * Horner recurrence consumes it, so the 7x6 interpolant is never held.    */
__device__ __forceinline__ void dense_value(const double (&y_old)[6], const double (&y_new)[6],
                                            const double (&k)[DOP_K][6],
                                            const double (&kx)[DOP_SX - DOP_K][6], double hs,
                                            double x, double (&out)[6]) {
    const double omx = 1.0 - x;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        double v = 0.0;
#pragma unroll
        for (int i = DOP_P - 1; i >= 0; --i) {
            v = v + dense_row(i, d, y_old, y_new, k, kx, hs);
            v = v * (((DOP_P - 1 - i) % 2 == 0) ? x : omx);
        }
        out[d] = v + y_old[d];
    }
}

/* rk.py:2005-2102 with the event g = y[ev_index] - ev_offset.  Only the event
 * component is interpolated inside the bisection; the full state is formed
 * once at the returned abscissa (each component of the reference's y_mid is
 * computed independently, so the result is the same).                     */
__device__ __forceinline__ int crossed_direction(double gl, double gm, int dir) {
    if (dir == 0) return (gl < 0.0 && gm > 0.0) || (gl > 0.0 && gm < 0.0);
    if (dir > 0) return gl < 0.0 && gm > 0.0;
    return gl > 0.0 && gm < 0.0;
}

__device__ __forceinline__ int event_crossed(double gp, double gn, int dir) {
    if (dir == 0) return (gp < 0.0 && gn > 0.0) || (gp > 0.0 && gn < 0.0) || (gn == 0.0);
    if (dir > 0) return (gp < 0.0 && gn > 0.0) || (gn == 0.0);
    return (gp > 0.0 && gn < 0.0) || (gn == 0.0);
}

__device__ __forceinline__ double pick6(const double (&v)[6], int i) {
    double r = v[0];
#pragma unroll
    for (int d = 1; d < 6; ++d) r = (i == d) ? v[d] : r;
    return r;
}

__device__ __forceinline__ double refine_bisect(double y_ev0, const double (&fc)[DOP_P],
                                                double ev_off, int dir, double h, double xtol,
                                                double gtol) {
    /* returns the abscissa x in [0, 1] of the hit (rk.py:2079-2102) */
    double a = 0.0, b = 1.0;
    double gl = y_ev0 - ev_off;
    for (int it = 0; it < 128; ++it) {
        const double mid = 0.5 * (a + b);
        double v = 0.0;
        const double omx = 1.0 - mid;
#pragma unroll
        for (int i = DOP_P - 1; i >= 0; --i) {
            v = v + fc[i];
            v = v * (((DOP_P - 1 - i) % 2 == 0) ? mid : omx);
        }
        const double gm = (v + y_ev0) - ev_off;
        if (__builtin_fabs(gm) <= gtol) return mid;
        if (crossed_direction(gl, gm, dir)) {
            b = mid;
        } else {
            a = mid;
            gl = gm;
        }
        if ((b - a) * __builtin_fabs(h) <= xtol) break;
    }
    return b;
}

/* ---------------------------------------------------------------------- */
/* fixed-step explicit RK (rk.py:154-215) over the tableaux of rk.py:755-783 */
/* ---------------------------------------------------------------------- */

/* Tableau tags: stages and coefficients (integrators/coefficients/rk4.py,
 * rk6.py, rk8.py).  After unrolling every coefficient is a constant, so the
 * zero tests below vanish at compile time.                                 */
struct RK4 {
    static constexpr int S = HH_RK4_STAGES, ORDER = 4;
    static __device__ __forceinline__ double a(int i, int j) { return HH_RK4_A[i][j]; }
    static __device__ __forceinline__ double b(int j) { return HH_RK4_B[j]; }
};
struct RK6 {
    static constexpr int S = HH_RK6_STAGES, ORDER = 6;
    static __device__ __forceinline__ double a(int i, int j) { return HH_RK6_A[i][j]; }
    static __device__ __forceinline__ double b(int j) { return HH_RK6_B[j]; }
};
struct RK8 {
    static constexpr int S = HH_RK8_STAGES, ORDER = 8;
    static __device__ __forceinline__ double a(int i, int j) { return HH_RK8_A[i][j]; }
    static __device__ __forceinline__ double b(int j) { return HH_RK8_B[j]; }
};

/* rk_step (rk.py:154-215); k0 receives the first stage, f(t, y) at the
 * step's start -- the derivative _integrate_fixed_rk stores for that state
 * (rk.py:579-587)                                                         */
template <class M, class P>
__device__ __forceinline__ void rk_step(const P& p, double (&y)[6], double h, double (&k0)[6]) {
    double k[M::S][6];
    p.rhs(y, k[0]);
#pragma unroll
    for (int d = 0; d < 6; ++d) k0[d] = k[0][d];
#pragma unroll
    for (int i = 1; i < M::S; ++i) {
        double ys[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) ys[d] = y[d];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double a = M::a(i, j);
            if (a != 0.0) {
                const double ha = h * a;
#pragma unroll
                for (int d = 0; d < 6; ++d) ys[d] = ys[d] + ha * k[j][d];
            }
        }
        p.rhs(ys, k[i]);
    }
#pragma unroll
    for (int j = 0; j < M::S; ++j) {
        const double b = M::b(j);
        if (b != 0.0) {
            const double hb = h * b;
#pragma unroll
            for (int d = 0; d < 6; ++d) y[d] = y[d] + hb * k[j][d];
        }
    }
}

template <class M, class P>
__device__ __forceinline__ void rk_step(const P& p, double (&y)[6], double h) {
    double k0[6];
    rk_step<M>(p, y, h, k0);
}

}  // namespace hh

#endif /* HH_DEVICE_H */

/* hh_wave.h -- one 42-d variational trajectory per wave (64 lanes).
 *
 * Lane L = 8*b + d holds entry d (0..5) of block b (0..6) of the 42-d state,
 * the blocks being those of Lane42 (b = 0: the state x, b = 1..6: STM column
 * b-1); lanes with d > 5 or b = 7 are idle.  Every per-component expression
 * is Lane42's (and so the reference's, dynamics/rtbp.py:115-141, 219-242 and
 * integrators/rk.py:1636-1699, 1836-1874, 2457-2484), evaluated one component
 * per lane instead of six, so results are bit-identical to Lane42's.  What
 * the wave shares is formed once, in parallel lanes:
 *   - the twelve quotients of the variational field, one per lane 0..11,
 *     then read back as wave-uniform values;
 *   - the three pow calls of the step-size controller, in lanes 0..2;
 *   - the error-norm block sums, each formed in Lane42's in-lane order and
 *     combined by the same block tree.
 * The per-attempt instruction stream is about a third of Lane42's, which is
 * what bounds a trajectory that runs alone on its SIMD.
 */
#ifndef HH_WAVE_H
#define HH_WAVE_H

#include "hh_device.h"

namespace hh {

/* value of a double in lane l (l wave-uniform) */
__device__ __forceinline__ double rlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double xor_sign(double v, unsigned long long s) {
    return __longlong_as_double(__double_as_longlong(v) ^ (long long)s);
}

struct Wave42 {
    double mu, mu2;
    int lane;
    int d;                    /* entry within the block                          */
    bool st;                  /* block 0                                         */
    bool act;                 /* holds a component (b < 7, d < 6)               */
    bool odd;                 /* quotient lane on the second primary's side      */
    int kind;                 /* quotient kind of lanes 0..11                    */
    int pv_addr;              /* ds_bpermute byte address of the lane holding the
                                 velocity-part entry this lane's row needs       */
    double cl;                /* mu (even lanes) or -mu2 (odd lanes): x + cl is
                                 x + mu or x - mu2 exactly                       */
    unsigned long long sign;  /* bit 63: this lane's derivative is negated       */
    const PowTab* pt;         /* hh_pow's tables (an LDS copy)                   */

    __device__ __forceinline__ void setup(int lane_, double mu_, int fwd,
                                          unsigned long long flip, const PowTab* pt_) {
        lane = lane_;
        pt = pt_;
        const int b = lane >> 3;
        d = lane & 7;
        st = (b == 0);
        act = (b < 7) && (d < 6);
        odd = (lane & 1) != 0;
        kind = (lane >> 1) % 6;
        mu = mu_;
        mu2 = 1.0 - mu_;
        const int c = act ? comp42(b, d) : 0;
        sign = (act && fwd == -1 && ((flip >> c) & 1ULL)) ? (1ULL << 63) : 0ULL;
        /* rows 0..2 need entry d+3; row 3 needs entry 4 and row 4 entry 3
         * (Coriolis terms of the column rows and of the state rows)          */
        const int src = (d < 3) ? d + 3 : ((d == 3) ? 4 : 3);
        pv_addr = ((lane & ~7) | src) << 2;
        cl = odd ? -mu2 : mu_;
    }

    /* accessors shared with the pair layout */
    __device__ __forceinline__ bool holds() const { return act; }
    __device__ __forceinline__ bool leader() const { return lane == 0; }
    __device__ __forceinline__ int comp() const { return comp42(lane >> 3, d); }

    __device__ __forceinline__ double gather_pv(double v) const {
        const int lo = __builtin_amdgcn_ds_bpermute(pv_addr, __double2loint(v));
        const int hi = __builtin_amdgcn_ds_bpermute(pv_addr, __double2hiint(v));
        return __hiloint2double(hi, lo);
    }

    /* derivative of this lane's component; v = this lane's component of the
     * stage state (idle lanes: any finite value)                            */
    __device__ __forceinline__ double rhs(double v) const {
        const double p0 = group_get<0>(v);
        const double p1 = group_get<1>(v);
        const double p2 = group_get<2>(v);
        const double pv = gather_pv(v);
        const double x = rlane(v, 0);
        const double y = rlane(v, 1);
        const double z = rlane(v, 2);
        const double xm = x + mu;
        const double xr = x - mu2;
        /* quotient lane k = kind*2 + odd (Lane42's q1..q6 of both lane kinds):
         *  k: 0 xm/r3  1 xr/R3  2 y/r3  3 y/R3  4 z/r3  5 z/R3
         *     6 mu2/r5 7 mu/R5  8 mu2/r3 9 mu/R3 10 (mu2 xm)/r5 11 (mu xr)/R5 */
        const double xo = x + cl;
        const double P = (xo * xo + y * y) + z * z;
        double P3, P5;
        hh_dev_pow15_25(P, &P3, &P5);
        const double m = odd ? mu : mu2;
        double num = xo;
        num = (kind == 1) ? y : num;
        num = (kind == 2) ? z : num;
        num = (kind == 3 || kind == 4) ? m : num;
        num = (kind == 5) ? m * xo : num;
        const double q = num / ((kind == 3 || kind == 5) ? P5 : P3);
        const double Q0 = rlane(q, 0), Q1 = rlane(q, 1), Q2 = rlane(q, 2), Q3 = rlane(q, 3);
        const double Q4 = rlane(q, 4), Q5 = rlane(q, 5), Q6 = rlane(q, 6), Q7 = rlane(q, 7);
        const double Q8 = rlane(q, 8), Q9 = rlane(q, 9), Q10 = rlane(q, 10), Q11 = rlane(q, 11);
        /* state rows (rtbp.py:233-242); pv is vy on lane 3, vx on lane 4 */
        const double s3 = ((x - mu2 * Q0) - mu * Q1) + 2.0 * pv;
        const double s4 = ((y - mu2 * Q2) - mu * Q3) - 2.0 * pv;
        const double s5 = (-mu2 * Q4) - mu * Q5;
        /* Hessian entries (rtbp.py:124-141) */
        const double a53 = Q6 * 3.0;
        const double b53 = Q7 * 3.0;
        const double c3 = Q8 + Q9;
        const double gx = Q10 + Q11;
        const double uxx = ((1.0 + a53 * (xm * xm)) + b53 * (xr * xr)) - c3;
        const double uyy = ((1.0 + a53 * (y * y)) + b53 * (y * y)) - c3;
        const double uzz = ((0.0 + a53 * (z * z)) + b53 * (z * z)) - c3;
        const double uxy = (3.0 * y) * gx;
        const double uxz = (3.0 * z) * gx;
        const double uyz = ((3.0 * y) * z) * (Q6 + Q7);
        /* row d of F @ Phi[:, j] (column_rhs) */
        const double cA = (d == 3) ? uxx : ((d == 4) ? uxy : uxz);
        const double cB = (d == 3) ? uxy : ((d == 4) ? uyy : uyz);
        const double cC = (d == 3) ? uxz : ((d == 4) ? uyz : uzz);
        const double acc = ((0.0 + cA * p0) + cB * p1) + cC * p2;
        const double E = (d == 3) ? 2.0 : -2.0;
        const double row = (d == 5) ? acc : acc + E * pv;
        const double sv = (d == 3) ? s3 : ((d == 4) ? s4 : s5);
        const double lo3 = st ? pv : (0.0 + pv);
        const double hi3 = st ? sv : row;
        return xor_sign((d < 3) ? lo3 : hi3, sign);
    }

    /* Lane42::reduce of in-lane sums over a block: sq = this lane's term */
    __device__ __forceinline__ double block_tree(double sq) const {
        const double g0 = group_get<0>(sq);
        const double g1 = group_get<1>(sq);
        const double g2 = group_get<2>(sq);
        const double g3 = group_get<3>(sq);
        const double g4 = group_get<4>(sq);
        const double g5 = group_get<5>(sq);
        const double blk = (((((0.0 + g0) + g1) + g2) + g3) + g4) + g5;
        const double B0 = rlane(blk, 0), B1 = rlane(blk, 8), B2 = rlane(blk, 16);
        const double B3 = rlane(blk, 24), B4 = rlane(blk, 32), B5 = rlane(blk, 40);
        const double B6 = rlane(blk, 48);
        return ((B0 + B1) + (B2 + B3)) + ((B4 + B5) + (B6 + 0.0));
    }

    /* dop853_attempt, one component per lane */
    __device__ __forceinline__ double attempt(double y, double (&k)[DOP_K], double h, double rtol,
                                              double atol, double& yh) const {
#pragma unroll
        for (int i = 1; i < DOP_S; ++i) {
            double ys = y;
#pragma unroll
            for (int j = 0; j < i; ++j) {
                const double a = HH_DOP_A[i][j];
                if (a != 0.0) {
                    const double ha = h * a;
                    ys = ys + ha * k[j];
                }
            }
            k[i] = rhs(ys);
        }
        yh = y;
#pragma unroll
        for (int j = 0; j < DOP_S; ++j) {
            const double bb = HH_DOP_B[j];
            if (bb != 0.0) {
                const double hb = h * bb;
                yh = yh + hb * k[j];
            }
        }
        k[DOP_S] = rhs(yh);
        double e5 = 0.0, e3 = 0.0;
#pragma unroll
        for (int j = 0; j < DOP_K; ++j) {
            const double c5 = HH_DOP_E5[j];
            const double c3 = HH_DOP_E3[j];
            if (c5 != 0.0) e5 = e5 + c5 * k[j];
            if (c3 != 0.0) e3 = e3 + c3 * k[j];
        }
        const double sc = atol + rtol * np_max(__builtin_fabs(y), __builtin_fabs(yh));
        const double q5 = (e5 * h) / sc;
        const double q3 = (e3 * h) / sc;
        const double s5 = block_tree(q5 * q5);
        const double s3 = block_tree(q3 * q3);
        if (s5 == 0.0 && s3 == 0.0) return 0.0;
        const double denom = s5 + 0.01 * s3;
        return (__builtin_fabs(h) * s5) / __builtin_sqrt(denom * 42.0);
    }

    /* pi_accept_factor(err, err_prev) and pi_reject_factor(err) with their
     * three pow calls in lanes 0..2                                         */
    __device__ __forceinline__ void step_factors(double err, double err_prev, double& f_acc,
                                                 double& f_rej) const {
        const double beta = 1.0 / (8.0 + 1.0);
        const double alpha = 0.4 * beta;
        const double err_exp = 1.0 / 8.0;
        const double px = (lane == 1) ? err_prev : err;
        const double pe = (lane == 0) ? -beta : ((lane == 1) ? alpha : -err_exp);
        const double pw = pow_tab(pt, px, pe);
        const double pb = rlane(pw, 0), pa = rlane(pw, 1), pr = rlane(pw, 2);
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

    /* dense-output value at tf of the step [t, t + hs] (FinalOut::accepted) */
    __device__ __forceinline__ double dense_at(double y_old, double y_new, const double (&k)[DOP_K],
                                               double hs, double xq) const {
        double kx[DOP_SX - DOP_K];
#pragma unroll
        for (int sr = DOP_K; sr < DOP_SX; ++sr) {
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < sr; ++r) {
                const double a = HH_DOP_A[sr][r];
                if (a != 0.0) acc = acc + a * (r < DOP_K ? k[r] : kx[r - DOP_K]);
            }
            kx[sr - DOP_K] = rhs(y_old + hs * acc);
        }
        const double omx = 1.0 - xq;
        const double dy = y_new - y_old;
        double v = 0.0;
#pragma unroll
        for (int i = DOP_P - 1; i >= 0; --i) {
            double row;
            if (i == 0) {
                row = dy;
            } else if (i == 1) {
                row = hs * k[0] - dy;
            } else if (i == 2) {
                row = 2.0 * dy - hs * (k[DOP_S] + k[0]);
            } else {
                double acc = 0.0;
#pragma unroll
                for (int r = 0; r < DOP_SX; ++r) {
                    const double c = HH_DOP_D[i - 3][r];
                    if (c != 0.0) acc = acc + c * (r < DOP_K ? k[r] : kx[r - DOP_K]);
                }
                row = hs * acc;
            }
            v = v + row;
            v = v * (((DOP_P - 1 - i) % 2 == 0) ? xq : omx);
        }
        return v + y_old;
    }
};

}  // namespace hh

#endif /* HH_WAVE_H */

/* hh_group6.h -- one 6-d CR3BP trajectory per aligned 8-lane group, one
 * component per lane (lane d = 0..5 of the group holds y[d]; lanes 6 and 7
 * help with the shared work).  Every per-component expression is Lane6's
 * (rhs6, dop853_attempt, dense output: dynamics/rtbp.py:30-74,
 * integrators/rk.py:1636-1699, 1836-1874, 2457-2484), so results are
 * bit-identical; the group shares out what Lane6 does serially:
 *   - the two distances: even lanes form r1^3, odd lanes r2^3, swapped in
 *     pairs;
 *   - the six quotients of the field: lanes 3..5 form the two of their own
 *     component (lanes 0..2 take a velocity component);
 *   - the error-norm divisions (one pair per lane) and the three pow calls
 *     of the step controller (lanes 0..2).
 * Sums over the six components are formed in Lane6's order (from 0.0, d =
 * 0..5).  A step attempt is about 40% of Lane6's instruction stream: the
 * latency layout for small batches of event searches.
 */
#ifndef HH_GROUP6_H
#define HH_GROUP6_H

#include "hh_device.h"

namespace hh {

__device__ __forceinline__ double bperm(double v, int addr) {
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

struct Group6 {
    double mu, one_mu;
    int lane, d;
    bool act, odd;
    int base;                 /* lane & ~7, as a bpermute byte address       */
    int pv_addr;              /* the velocity entry this lane's row needs     */
    double cl;                /* mu (even lanes) or -one_mu (odd lanes)       */
    double c1;                /* one_mu, or -one_mu on lane 5                 */
    double c2;                /* 2.0 on lane 3, -2.0 otherwise                */
    unsigned long long sign;  /* bit 63: this lane's derivative is negated    */
    const PowTab* pt;         /* hh_pow's tables (an LDS copy)                */

    __device__ __forceinline__ void setup(int lane_, double mu_, int fwd,
                                          unsigned long long flip, const PowTab* pt_) {
        lane = lane_;
        pt = pt_;
        d = lane & 7;
        act = d < 6;
        odd = (d & 1) != 0;
        base = (lane & ~7) << 2;
        mu = mu_;
        one_mu = 1.0 - mu_;
        const int src = (d < 3) ? d + 3 : ((d == 3) ? 4 : 3);
        pv_addr = base + (src << 2);
        cl = odd ? -one_mu : mu_;
        c1 = (d == 5) ? -one_mu : one_mu;
        c2 = (d == 3) ? 2.0 : -2.0;
        sign = (act && fwd == -1 && ((flip >> d) & 1ULL)) ? (1ULL << 63) : 0ULL;
    }

    /* value of entry k (0..5, any lane-uniform index) of this lane's group */
    __device__ __forceinline__ double get(double v, int k) const { return bperm(v, base + (k << 2)); }

    /* rhs6 for this lane's component; v = this lane's stage component */
    __device__ __forceinline__ double rhs(double v) const {
        const double x = group_get<0>(v);
        const double y = group_get<1>(v);
        const double z = group_get<2>(v);
        const double pv = bperm(v, pv_addr);
        const double xm = x + mu;
        /* even lanes: r1 = |(x + mu, y, z)|, odd lanes: r2 = |(x - one_mu, y, z)| */
        const double xo = x + cl;
        const double rs = hh_dev_sqrt((xo * xo + y * y) + z * z);
        const double rc = rs * (rs * rs);
        const double rco = group_xor<1>(rc);
        const double r1c = odd ? rco : rc;
        const double r2c = odd ? rc : rco;
        const double xm1 = (x - 1.0) + mu;
        /* rows 3..5: (one_mu xm)/r1c and (mu xm1)/r2c; (one_mu y)/r1c and
         * (mu y)/r2c; ((-one_mu) z)/r1c and (mu z)/r2c                      */
        const double u = (d == 3) ? xm : ((d == 4) ? y : z);
        const double w = (d == 3) ? xm1 : ((d == 4) ? y : z);
        const double q1 = (c1 * u) / r1c;
        const double q2 = (mu * w) / r2c;
        const double head = c2 * pv + ((d == 3) ? x : y);   /* 2 vy + x, -2 vx + y */
        const double o45 = (head - q1) - q2;
        const double o6 = q1 - q2;
        const double o = (d < 3) ? pv : ((d == 5) ? o6 : o45);
        return __longlong_as_double(__double_as_longlong(o) ^ (long long)sign);
    }

    /* sum over the group's six components of this lane's term, from 0.0 in
     * component order (all lanes of the group get the sum)                  */
    __device__ __forceinline__ double sum6(double term) const {
        const double g0 = group_get<0>(term);
        const double g1 = group_get<1>(term);
        const double g2 = group_get<2>(term);
        const double g3 = group_get<3>(term);
        const double g4 = group_get<4>(term);
        const double g5 = group_get<5>(term);
        return (((((0.0 + g0) + g1) + g2) + g3) + g4) + g5;
    }

    /* initial_step (rk.py:2443-2448) */
    __device__ __forceinline__ double initial_step(double y, double f0, double rtol, double atol,
                                                   double min_step, double max_step) const {
        const double sc = atol + rtol * __builtin_fabs(y);
        const double a = y / sc;
        const double b = f0 / sc;
        const double s0 = sum6(a * a);
        const double s1 = sum6(b * b);
        const double sn = __builtin_sqrt(6.0);
        const double d0 = __builtin_sqrt(s0) / sn;
        const double d1 = __builtin_sqrt(s1) / sn;
        return select_initial_step(d0, d1, min_step, max_step);
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
        const double s5 = sum6(q5 * q5);
        const double s3 = sum6(q3 * q3);
        if (s5 == 0.0 && s3 == 0.0) return 0.0;
        const double denom = s5 + 0.01 * s3;
        return (__builtin_fabs(h) * s5) / __builtin_sqrt(denom * 6.0);
    }

    /* pi_accept_factor / pi_reject_factor with their pow calls in lanes 0..2 */
    __device__ __forceinline__ void step_factors(double err, double err_prev, double& f_acc,
                                                 double& f_rej) const {
        const double beta = 1.0 / (8.0 + 1.0);
        const double alpha = 0.4 * beta;
        const double err_exp = 1.0 / 8.0;
        const double px = (d == 1) ? err_prev : err;
        const double pe = (d == 0) ? -beta : ((d == 1) ? alpha : -err_exp);
        const double pw = pow_tab(pt, px, pe);
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

    /* the three extension stages and the seven interpolant rows of this
     * lane's component (rk.py:1836-1874)                                    */
    __device__ __forceinline__ void dense_rows(double y_old, double y_new, const double (&k)[DOP_K],
                                               double hs, double (&row)[DOP_P]) const {
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
        const double dy = y_new - y_old;
        row[0] = dy;
        row[1] = hs * k[0] - dy;
        row[2] = 2.0 * dy - hs * (k[DOP_S] + k[0]);
#pragma unroll
        for (int i = 3; i < DOP_P; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < DOP_SX; ++r) {
                const double c = HH_DOP_D[i - 3][r];
                if (c != 0.0) acc = acc + c * (r < DOP_K ? k[r] : kx[r - DOP_K]);
            }
            row[i] = hs * acc;
        }
    }
};

/* rk_step (rk.py:154-215) for this lane's component; k0 receives the first
 * stage (the derivative at the step's start)                              */
template <class M>
__device__ __forceinline__ void rk_step_group(const Group6& p, double& y, double h, double& k0) {
    double k[M::S];
    k[0] = p.rhs(y);
    k0 = k[0];
#pragma unroll
    for (int i = 1; i < M::S; ++i) {
        double ys = y;
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double a = M::a(i, j);
            if (a != 0.0) {
                const double ha = h * a;
                ys = ys + ha * k[j];
            }
        }
        k[i] = p.rhs(ys);
    }
#pragma unroll
    for (int j = 0; j < M::S; ++j) {
        const double b = M::b(j);
        if (b != 0.0) {
            const double hb = h * b;
            y = y + hb * k[j];
        }
    }
}

/* dense_value of one component from its rows (rk.py:1989-2003) */
__device__ __forceinline__ double dense_from_rows(const double (&row)[DOP_P], double y_old,
                                                  double x) {
    const double omx = 1.0 - x;
    double v = 0.0;
#pragma unroll
    for (int i = DOP_P - 1; i >= 0; --i) {
        v = v + row[i];
        v = v * (((DOP_P - 1 - i) % 2 == 0) ? x : omx);
    }
    return v + y_old;
}

}  // namespace hh

#endif /* HH_GROUP6_H */

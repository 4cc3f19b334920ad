/* hh_pair.h -- one 42-d variational trajectory per pair of waves.
 *
 * The two waves of a 128-thread workgroup share one trajectory, in the lane
 * layout of Wave42 (lane L = 8*b + d holds entry d of block b):
 *   - the state wave (wave 0) holds block 0, the state x, and forms per
 *     stage the twelve quotients of the field (lanes 0..11) and the state
 *     rows of the field (dynamics/rtbp.py:228-242);
 *   - the STM wave (wave 1) holds blocks 1..6, the six STM columns, and forms
 *     the Hessian of the effective potential (rtbp.py:115-141) from the
 *     quotients of the same stage and then the rows of F @ Phi (rtbp.py:
 *     219-225).
 * The state of a stage does not depend on Phi, so the two instruction
 * streams run side by side on two SIMDs: per right-hand side the state wave
 * writes the stage's quotients and x, y, z to LDS and both waves meet at
 * one barrier; the STM wave then forms its rows from them while the state
 * wave goes on to its state rows and the next stage.  Norms over all 42
 * components take one more barrier.  Each wave evaluates exactly Wave42's
 * expressions for its lanes, so results are bit-identical to Wave42's (and
 * so to Lane42's and the reference's).  Error-norm and initial-step sums
 * exchange their block sums through LDS and are combined in Wave42's
 * block-tree order by both waves, which then take the same step-size
 * decisions.  The critical path is the state wave's chain (field quotients,
 * powers, divisions, step control): about 6.2 us per attempt against 7.2 us
 * for Wave42 on one trajectory alone.
 */
#ifndef HH_PAIR_H
#define HH_PAIR_H

#include "hh_wave.h"

namespace hh {

/* LDS of one pair: stage records in two buffers used alternately (the
 * record of a right-hand side is read by the STM wave before the barrier of
 * the next one, so the buffer written next is free) and the block sums.    */
struct PairLds {
    double stage[2][16];   /* [buffer][0..11 quotients, 12..14 x, y, z] */
    double bsum[2][8];     /* block sums of two norms, blocks 0..6      */
};

struct Pair42 {
    Wave42 w;               /* lane layout, constants and shared helpers  */
    int role;               /* 0: state wave, 1: STM wave                  */
    PairLds* lds;
    mutable int seq;        /* right-hand sides evaluated so far            */

    __device__ __forceinline__ void setup(int lane, int role_, double mu, int fwd,
                                          unsigned long long flip, PairLds* l,
                                          const PowTab* pt) {
        w.setup(lane, mu, fwd, flip, pt);
        role = role_;
        lds = l;
        seq = 0;
        /* the state wave holds block 0 only, the STM wave blocks 1..6 */
        w.act = w.act && ((role == 0) == w.st);
    }

    __device__ __forceinline__ bool holds() const { return w.act; }
    __device__ __forceinline__ bool leader() const { return role == 0 && w.lane == 0; }
    __device__ __forceinline__ int comp() const { return w.comp(); }

    __device__ __forceinline__ void step_factors(double err, double err_prev, double& f_acc,
                                                 double& f_rej) const {
        w.step_factors(err, err_prev, f_acc, f_rej);
    }

    /* derivative of this lane's component; v = this lane's component of the
     * stage state.  One barrier per call, in both waves: the state wave
     * writes the stage record before it, the STM wave reads it after.  (A
     * one-way handoff through an LDS sequence number, letting the state wave
     * run ahead, measured slower on gfx950.)                                */
    __device__ __forceinline__ double rhs(double v) const {
        double* s = lds->stage[seq & 1];
        seq += 1;
        const double mu = w.mu, mu2 = w.mu2;
        const int d = w.d;
        const double pv = w.gather_pv(v);
        if (role == 0) {
            const double x = rlane(v, 0);
            const double y = rlane(v, 1);
            const double z = rlane(v, 2);
            const double xo = x + w.cl;
            const double P = (xo * xo + y * y) + z * z;
            double P3, P5;
            hh_dev_pow15_25(P, &P3, &P5);
            const double m = w.odd ? mu : mu2;
            const int kind = w.kind;
            double num = xo;
            num = (kind == 1) ? y : num;
            num = (kind == 2) ? z : num;
            num = (kind == 3 || kind == 4) ? m : num;
            num = (kind == 5) ? m * xo : num;
            const double q = num / ((kind == 3 || kind == 5) ? P5 : P3);
            if (w.lane < 12) s[w.lane] = q;
            if (w.lane < 3) s[12 + w.lane] = v;
            __syncthreads();
            const double Q0 = rlane(q, 0), Q1 = rlane(q, 1), Q2 = rlane(q, 2);
            const double Q3 = rlane(q, 3), Q4 = rlane(q, 4), Q5 = rlane(q, 5);
            /* state rows (rtbp.py:233-242); pv is vy on lane 3, vx on lane 4 */
            const double s3 = ((x - mu2 * Q0) - mu * Q1) + 2.0 * pv;
            const double s4 = ((y - mu2 * Q2) - mu * Q3) - 2.0 * pv;
            const double s5 = (-mu2 * Q4) - mu * Q5;
            const double sv = (d == 3) ? s3 : ((d == 4) ? s4 : s5);
            return xor_sign((d < 3) ? pv : sv, w.sign);
        }
        const double p0 = group_get<0>(v);
        const double p1 = group_get<1>(v);
        const double p2 = group_get<2>(v);
        __syncthreads();
        const double Q6 = s[6], Q7 = s[7], Q8 = s[8], Q9 = s[9], Q10 = s[10], Q11 = s[11];
        const double x = s[12], y = s[13], z = s[14];
        const double xm = x + mu;
        const double xr = x - mu2;
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
        return xor_sign((d < 3) ? (0.0 + pv) : row, w.sign);
    }

    /* Wave42::block_tree of two sums at once over both waves: each wave
     * contributes the block sums of its own blocks.  One barrier.          */
    __device__ __forceinline__ void pair_tree(double sa, double sb, double& ra, double& rb) const {
        const double ga = sum_block(sa);
        const double gb = sum_block(sb);
        const int b = w.lane >> 3;
        const bool mine = (w.lane & 7) == 0 && b < 7 && ((role == 0) == (b == 0));
        if (mine) {
            lds->bsum[0][b] = ga;
            lds->bsum[1][b] = gb;
        }
        __syncthreads();
        const double* A = lds->bsum[0];
        const double* B = lds->bsum[1];
        ra = ((A[0] + A[1]) + (A[2] + A[3])) + ((A[4] + A[5]) + (A[6] + 0.0));
        rb = ((B[0] + B[1]) + (B[2] + B[3])) + ((B[4] + B[5]) + (B[6] + 0.0));
    }

    /* Lane42's in-lane sum over a block (entries 0..5 from 0.0) */
    __device__ __forceinline__ static double sum_block(double sq) {
        const double g0 = group_get<0>(sq);
        const double g1 = group_get<1>(sq);
        const double g2 = group_get<2>(sq);
        const double g3 = group_get<3>(sq);
        const double g4 = group_get<4>(sq);
        const double g5 = group_get<5>(sq);
        return (((((0.0 + g0) + g1) + g2) + g3) + g4) + g5;
    }

    /* Wave42::attempt with the pair's right-hand side and norm */
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
        double s5, s3;
        pair_tree(w.act ? q5 * q5 : 0.0, w.act ? q3 * q3 : 0.0, s5, s3);
        if (s5 == 0.0 && s3 == 0.0) return 0.0;
        const double denom = s5 + 0.01 * s3;
        return (__builtin_fabs(h) * s5) / __builtin_sqrt(denom * 42.0);
    }

    /* Wave42::dense_at with the pair's right-hand side */
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

#endif /* HH_PAIR_H */

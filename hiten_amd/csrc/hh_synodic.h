/* hh_synodic.h -- section crossings of stored trajectories (synodic maps).
 *
 * Restates _SynodicDetectionBackend.detect_on_trajectory
 * (src/hiten/algorithms/poincare/synodic/backend.py:644-784) for a batch of
 * trajectories sampled on a common number of points: one wave per
 * trajectory, one segment [k, k+1] per lane.  Per segment, in the
 * reference's order:
 *   - the plane value g = n . x - c at samples k-1 .. k+2;
 *   - segment_refine == 0: the on-surface test of sample k (backend.py:
 *     134-180), else a sign change (183-227) refined linearly (230-265) or
 *     by Newton on the cubic Hermite of g with central-difference slopes
 *     (268-366);
 *   - segment_refine > 0: the on-surface test at s = 0, then r + 1 equal
 *     sub-intervals, each tested on the (linear or Hermite) interpolant and
 *     refined by a linear step and Newton inside the sub-interval
 *     (443-621);
 *   - the hit state by cubic Hermite (central-difference velocities) or
 *     linearly.
 * Lanes count their candidates, a wave scan orders them by segment (the
 * reference's stable sort), and one uniform loop deduplicates them against
 * the last kept hit (369-440).  The Hermite basis follows
 * poincare/utils.py:54-142 (x ** 2 as x * x, as numba lowers it); the
 * state interpolation's (1 - s) ** 2 is a CPython float power there (the C
 * library's pow), which rounds a square differently from x * x in ~0.1% of
 * cases: hit states then differ by an ulp, hit times do not.
 */
#ifndef HH_SYNODIC_H
#define HH_SYNODIC_H

#include "hh_device.h"
#include "hh_wave.h"

namespace hh {

struct SynArgs {
    int n, m;
    const double* times;     /* [m], or [n][m] with per_traj_times          */
    int per_traj_times;
    const int* lengths;      /* [n] samples used per trajectory, or NULL     */
    const double* states;    /* [n][m][6]                                    */
    double nrm[6];
    double offset;
    int direction;           /* 0: both, 1: g rising, -1: g falling          */
    int pi, pj;              /* state components of the 2-d section point    */
    int cubic;
    int refine;              /* segment_refine                               */
    int newton_max_iter;
    int max_hits;            /* 0: no limit                                  */
    int cap;                 /* hit slots per trajectory in the outputs      */
    double tol_on, dedup_t, dedup_p;
    int* count;              /* [n] hits kept (may exceed cap)               */
    double* hit_t;           /* [n][cap]                                     */
    double* hit_x;           /* [n][cap][6]                                  */
    double* hit_p;           /* [n][cap][2]                                  */
};

__device__ __forceinline__ double syn_hermite(double s, double y0, double y1, double dy0,
                                              double dy1, double dt) {
    const double oms = 1.0 - s;
    const double h00 = (1.0 + 2.0 * s) * (oms * oms);
    const double h10 = s * (oms * oms);
    const double h01 = (s * s) * (3.0 - 2.0 * s);
    const double h11 = (s * s) * (s - 1.0);
    return ((h00 * y0 + (h10 * dy0) * dt) + h01 * y1) + (h11 * dy1) * dt;
}

__device__ __forceinline__ double syn_hermite_der(double s, double y0, double y1, double dy0,
                                                  double dy1, double dt) {
    const double oms = 1.0 - s;
    const double dh00 = ((6.0 * s) * (s - 1.0) + (oms * oms) * 2.0) - (2.0 * oms) * (1.0 + 2.0 * s);
    const double dh10 = (oms * oms) + s * (2.0 * (s - 1.0));
    const double dh01 = (6.0 * s) * oms - (2.0 * s) * (3.0 - 2.0 * s);
    const double dh11 = (2.0 * s) * (s - 1.0) + (s * s);
    return ((dh00 * y0 + (dh10 * dy0) * dt) + dh01 * y1) + (dh11 * dy1) * dt;
}

/* One trajectory's samples */
struct SynTraj {
    const SynArgs* a;
    size_t base;   /* first sample of the trajectory */
    size_t tbase;  /* first time of the trajectory   */
    int N;         /* samples                        */

    __device__ __forceinline__ double T(int j) const { return a->times[tbase + j]; }
    __device__ __forceinline__ double X(int j, int d) const {
        return a->states[(base + (size_t)j) * 6 + d];
    }
    /* g = states @ normal - offset, summed left to right */
    __device__ __forceinline__ double G(int j) const {
        double acc = X(j, 0) * a->nrm[0];
#pragma unroll
        for (int d = 1; d < 6; ++d) acc = acc + X(j, d) * a->nrm[d];
        return acc - a->offset;
    }
    /* sample j as three 16-byte loads (a sample is 48 bytes, 16-aligned) */
    __device__ __forceinline__ void load(int j, double (&x)[6], double& t) const {
        const double2* p = reinterpret_cast<const double2*>(a->states + (base + (size_t)j) * 6);
        const double2 u = p[0], v = p[1], w = p[2];
        x[0] = u.x; x[1] = u.y; x[2] = v.x; x[3] = v.y; x[4] = w.x; x[5] = w.y;
        t = T(j);
    }
    __device__ __forceinline__ double G(const double (&x)[6]) const {
        double acc = x[0] * a->nrm[0];
#pragma unroll
        for (int d = 1; d < 6; ++d) acc = acc + x[d] * a->nrm[d];
        return acc - a->offset;
    }
};

/* What one segment needs, loaded once */
struct SynSeg {
    int k;
    double t0, t1, dt, gk, gk1, gkm1, d0, d1;
    bool has_m1;
};

/* segment k from the plane values and times of samples k-1 .. k+2 (those
 * outside the trajectory are not read)                                    */
__device__ __forceinline__ SynSeg syn_segment(const SynArgs& a, int N, int k, double gkm1,
                                              double gk, double gk1, double gk2, double tkm1,
                                              double tk, double tk1, double tk2) {
    SynSeg g;
    g.k = k;
    g.t0 = tk;
    g.t1 = tk1;
    g.dt = g.t1 - g.t0;
    g.gk = gk;
    g.gk1 = gk1;
    g.has_m1 = k >= 1;
    g.gkm1 = g.has_m1 ? gkm1 : 0.0;
    g.d0 = 0.0;
    g.d1 = 0.0;
    if (a.cubic && g.dt > 0.0) {
        /* backend.py:316-323 / 537-545 */
        g.d0 = g.has_m1 ? (g.gk1 - g.gkm1) / (g.t1 - tkm1) : (g.gk1 - g.gk) / (g.t1 - g.t0);
        g.d1 = (k + 2 < N) ? (gk2 - g.gk) / (tk2 - g.t0) : (g.gk1 - g.gk) / (g.t1 - g.t0);
    }
    return g;
}

/* True when no sub-interval of the segment can show a sign change: the
 * linear interpolant keeps the sign of both ends; the Hermite cubic
 * differs from a convex combination of gk, gk1 by at most 4/27 dt (|d0| +
 * |d1|) (|h10|, |h11| <= 4/27 on [0, 1], h00 + h01 = 1), and a margin of
 * 1e-9 of the magnitudes is far above the rounding of the evaluated
 * values.  NaN fails every test.                                           */
__device__ __forceinline__ bool syn_no_sign_change(const SynArgs& a, const SynSeg& g) {
    const double lo = fmin(g.gk, g.gk1), hi = fmax(g.gk, g.gk1);
    if (!(a.cubic && g.dt > 0.0)) return lo > 0.0 || hi < 0.0;
    const double B = (4.0 / 27.0) * g.dt * (__builtin_fabs(g.d0) + __builtin_fabs(g.d1));
    const double M = 1e-9 * ((__builtin_fabs(lo) + __builtin_fabs(hi)) + B);
    return (lo - B > M) || (hi + B < -M);
}

/* on-surface test of sample k with the direction filter (backend.py:163-180) */
__device__ __forceinline__ bool syn_on_surface(const SynArgs& a, const SynSeg& g) {
    if (!(__builtin_fabs(g.gk) < a.tol_on)) return false;
    if (a.direction == 0) return true;
    if (a.direction == 1) return g.gk1 >= 0.0 || (g.has_m1 && g.gkm1 <= 0.0);
    return g.gk1 <= 0.0 || (g.has_m1 && g.gkm1 >= 0.0);
}

__device__ __forceinline__ bool syn_crosses(double lo, double hi, int direction) {
    if (direction == 0) return (lo * hi <= 0.0) && (lo != hi);
    if (direction == 1) return (lo < 0.0) && (hi >= 0.0);
    return (lo > 0.0) && (hi <= 0.0);
}

__device__ __forceinline__ double syn_interp(const SynArgs& a, const SynSeg& g, double s) {
    if (a.cubic && g.dt > 0.0) return syn_hermite(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
    return (1.0 - s) * g.gk + s * g.gk1;
}

/* Walk the candidates of a segment in the reference's order.  Returns the
 * number of candidates when want < 0; else the parameter s of candidate
 * `want` (or -1.0 for the on-surface sample at s = 0, whose state is
 * sample k itself).                                                       */
__device__ __forceinline__ double syn_walk(const SynArgs& a, const SynSeg& g, int want,
                                           int* count) {
    int c = 0;
    if (a.refine <= 0) {
        if (syn_on_surface(a, g)) {
            *count = 1;
            return (want == 0) ? -1.0 : 0.0;
        }
        if (!syn_crosses(g.gk, g.gk1, a.direction)) {
            *count = 0;
            return 0.0;
        }
        *count = 1;
        if (want != 0) return 0.0;
        /* numpy.minimum(1, numpy.maximum(0, alpha)): NaN propagates */
        double al = g.gk / (g.gk - g.gk1);
        if (al == al) {
            al = (al >= 0.0) ? al : 0.0;
            al = (al <= 1.0) ? al : 1.0;
        }
        double s = al;
        if (a.cubic && g.dt > 0.0) {
            for (int it = 0; it < a.newton_max_iter; ++it) {
                const double f = syn_hermite(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
                const double df = syn_hermite_der(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
                if (df == 0.0) break;
                s = s - f / df;
                if (s < 0.0) { s = 0.0; break; }
                if (s > 1.0) { s = 1.0; break; }
            }
        }
        return s;
    }
    const bool left = syn_on_surface(a, g);
    if (left) {
        if (want == 0) {
            *count = 1;
            return -1.0;
        }
        ++c;
    }
    if (want < 0 && syn_no_sign_change(a, g)) {
        *count = c;
        return 0.0;
    }
    const int r = a.refine;
    const double step = 1.0 / (double)(r + 1);
    double g_prev = 0.0;
    bool have_prev = false;
    for (int m = 0; m <= r; ++m) {
        const double s_lo = (double)m * step;
        const double s_hi = (double)(m + 1) * step;
        if (s_hi > 1.0 + 1e-15) break;
        if (left && m == 0) continue;
        /* the value at s_lo is the previous sub-interval's value at s_hi:
         * the same expression of the same s, so the same bits            */
        const double g_lo = have_prev ? g_prev : syn_interp(a, g, s_lo);
        const double g_hi = syn_interp(a, g, s_hi);
        g_prev = g_hi;
        have_prev = true;
        if (!syn_crosses(g_lo, g_hi, a.direction)) continue;
        if (c == want) {
            double s;
            if (g_lo == g_hi) {
                s = 0.5 * (s_lo + s_hi);
            } else {
                /* Python's min(1.0, max(0.0, alpha)): a NaN gives 0.0 */
                double al = g_lo / (g_lo - g_hi);
                al = (0.0 < al) ? al : 0.0;
                al = (al < 1.0) ? al : 1.0;
                s = s_lo + al * (s_hi - s_lo);
            }
            if (a.cubic && g.dt > 0.0) {
                for (int it = 0; it < a.newton_max_iter; ++it) {
                    const double f = syn_hermite(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
                    const double df = syn_hermite_der(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
                    if (df == 0.0) break;
                    s = s - f / df;
                    if (s < s_lo) { s = s_lo; break; }
                    if (s > s_hi) { s = s_hi; break; }
                }
            }
            *count = c + 1;
            return s;
        }
        ++c;
    }
    *count = c;
    return 0.0;
}

/* syn_walk for one candidate (want >= 0) with the sub-interval located
 * first and refined after the walk, the same operations in the same order:
 * lanes of a wave refining candidates of different segments then run one
 * refinement loop together instead of one per sub-interval.               */
__device__ __forceinline__ double syn_walk_deferred(const SynArgs& a, const SynSeg& g, int want,
                                                    int* count) {
    int c = 0;
    if (a.refine <= 0) return syn_walk(a, g, want, count);
    const bool left = syn_on_surface(a, g);
    if (left) {
        if (want == 0) {
            *count = 1;
            return -1.0;
        }
        ++c;
    }
    const int r = a.refine;
    const double step = 1.0 / (double)(r + 1);
    double g_prev = 0.0;
    bool have_prev = false;
    bool found = false;
    double f_slo = 0.0, f_shi = 0.0, f_glo = 0.0, f_ghi = 0.0;
    for (int m = 0; m <= r; ++m) {
        const double s_lo = (double)m * step;
        const double s_hi = (double)(m + 1) * step;
        if (s_hi > 1.0 + 1e-15) break;
        if (left && m == 0) continue;
        /* the value at s_lo is the previous sub-interval's value at s_hi:
         * the same expression of the same s, so the same bits            */
        const double g_lo = have_prev ? g_prev : syn_interp(a, g, s_lo);
        const double g_hi = syn_interp(a, g, s_hi);
        g_prev = g_hi;
        have_prev = true;
        if (!syn_crosses(g_lo, g_hi, a.direction)) continue;
        if (c == want) {
            found = true;
            f_slo = s_lo;
            f_shi = s_hi;
            f_glo = g_lo;
            f_ghi = g_hi;
            break;
        }
        ++c;
    }
    if (!found) {
        *count = c;
        return 0.0;
    }
    double s;
    if (f_glo == f_ghi) {
        s = 0.5 * (f_slo + f_shi);
    } else {
        /* Python's min(1.0, max(0.0, alpha)): a NaN gives 0.0 */
        double al = f_glo / (f_glo - f_ghi);
        al = (0.0 < al) ? al : 0.0;
        al = (al < 1.0) ? al : 1.0;
        s = f_slo + al * (f_shi - f_slo);
    }
    if (a.cubic && g.dt > 0.0) {
        for (int it = 0; it < a.newton_max_iter; ++it) {
            const double f = syn_hermite(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
            const double df = syn_hermite_der(s, g.gk, g.gk1, g.d0, g.d1, g.dt);
            if (df == 0.0) break;
            s = s - f / df;
            if (s < f_slo) { s = f_slo; break; }
            if (s > f_shi) { s = f_shi; break; }
        }
    }
    *count = c + 1;
    return s;
}

/* hit state at parameter s of segment k (backend.py:345-360 / 591-606) */
__device__ __forceinline__ void syn_state(const SynTraj& tr, const SynSeg& g, double s,
                                          double (&x)[6]) {
    const int k = g.k;
    if (tr.a->cubic && g.dt > 0.0 && k >= 1 && k + 2 < tr.N) {
        const double tm1 = tr.T(k - 1), tp2 = tr.T(k + 2);
        const double oms = 1.0 - s;
        const double h00 = (1.0 + 2.0 * s) * (oms * oms);
        const double h10 = s * (oms * oms);
        const double h01 = (s * s) * (3.0 - 2.0 * s);
        const double h11 = (s * s) * (s - 1.0);
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const double xk = tr.X(k, d), xk1 = tr.X(k + 1, d);
            const double dx0 = (xk1 - tr.X(k - 1, d)) / (g.t1 - tm1);
            const double dx1 = (tr.X(k + 2, d) - xk) / (tp2 - g.t0);
            x[d] = ((h00 * xk + (h10 * dx0) * g.dt) + h01 * xk1) + (h11 * dx1) * g.dt;
        }
    } else {
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const double xk = tr.X(k, d);
            x[d] = xk + s * (tr.X(k + 1, d) - xk);
        }
    }
}

/* Trajectory i, whole, by one wave: its segments 64 at a time (the next
 * chunk's samples in flight while this one is processed), candidates
 * ordered by a wave scan and deduplicated by one uniform loop.             */
__device__ __forceinline__ void syn_detect_wave(SynArgs a, int i, int lane) {
    SynTraj tr;
    tr.a = &a;
    tr.base = (size_t)i * a.m;
    tr.tbase = a.per_traj_times ? (size_t)i * a.m : 0;
    tr.N = a.lengths ? min(a.lengths[i], a.m) : a.m;
    const int nseg = tr.N - 1;
    int kept = 0;
    bool done = false;
    double last_t = 0.0, last_u = 0.0, last_v = 0.0;
    /* sample k0 + lane of each chunk, and the extra sample of lanes 0, 62,
     * 63 (k0 - 1, k0 + 64, k0 + 65), loaded one chunk ahead              */
    auto extra = [&](int k0) {
        return (lane == 0) ? k0 - 1 : ((lane == 62) ? k0 + 64 : ((lane == 63) ? k0 + 65 : -1));
    };
    double xo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, to = 0.0;
    double xe[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, te = 0.0;
    if (lane < tr.N) tr.load(lane, xo, to);
    {
        const int jx = extra(0);
        if (jx >= 0 && jx < tr.N) tr.load(jx, xe, te);
    }
    for (int k0 = 0; k0 < nseg && !done; k0 += 64) {
        const int k = k0 + lane;
        /* plane value and time of sample k0 + lane; lane 0 also takes
         * sample k0 - 1, lanes 62 and 63 samples k0 + 64 and k0 + 65      */
        const double gj = (k < tr.N) ? tr.G(xo) : 0.0;
        const double tj = (k < tr.N) ? to : 0.0;
        const int jx = extra(k0);
        const bool ex = jx >= 0 && jx < tr.N;
        const double gx = ex ? tr.G(xe) : 0.0;
        const double tx = ex ? te : 0.0;
        /* next chunk's loads, in flight while this chunk is processed */
        if (k + 64 < tr.N) tr.load(k + 64, xo, to);
        {
            const int jn = extra(k0 + 64);
            if (jn >= 0 && jn < tr.N) tr.load(jn, xe, te);
        }
        const double g_up1 = __shfl(gj, (lane + 1) & 63), t_up1 = __shfl(tj, (lane + 1) & 63);
        const double g_up2 = __shfl(gj, (lane + 2) & 63), t_up2 = __shfl(tj, (lane + 2) & 63);
        const double g_dn1 = __shfl(gj, (lane + 63) & 63), t_dn1 = __shfl(tj, (lane + 63) & 63);
        const double g_x62 = __shfl(gx, 62), t_x62 = __shfl(tx, 62);
        const double g_x63 = __shfl(gx, 63), t_x63 = __shfl(tx, 63);
        const double gk1 = (lane < 63) ? g_up1 : g_x62, tk1 = (lane < 63) ? t_up1 : t_x62;
        const double gk2 = (lane < 62) ? g_up2 : ((lane == 62) ? g_x62 : g_x63);
        const double tk2 = (lane < 62) ? t_up2 : ((lane == 62) ? t_x62 : t_x63);
        const double gkm1 = (lane > 0) ? g_dn1 : gx, tkm1 = (lane > 0) ? t_dn1 : tx;
        SynSeg g;
        int cnt = 0;
        if (k < nseg) {
            g = syn_segment(a, tr.N, k, gkm1, gj, gk1, gk2, tkm1, tj, tk1, tk2);
            syn_walk(a, g, -1, &cnt);
        }
        /* inclusive scan of the counts over the wave */
        int inc = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(inc, o);
            if (lane >= o) inc += v;
        }
        const int exc = inc - cnt;
        const int total = __shfl(inc, 63);
        for (int j = 0; j < total && !done; ++j) {
            const unsigned long long own = __ballot(exc <= j && j < inc);
            const int src = __ffsll((long long)own) - 1;
            double th = 0.0, x[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (lane == src) {
                int c2 = 0;
                const double s = syn_walk(a, g, j - exc, &c2);
                if (s < 0.0) {
                    th = g.t0;
#pragma unroll
                    for (int d = 0; d < 6; ++d) x[d] = tr.X(g.k, d);
                } else {
                    th = (1.0 - s) * g.t0 + s * g.t1;
                    syn_state(tr, g, s, x);
                }
            }
            th = __shfl(th, src);
#pragma unroll
            for (int d = 0; d < 6; ++d) x[d] = __shfl(x[d], src);
            double u = x[0], v = x[0];
#pragma unroll
            for (int d = 1; d < 6; ++d) {
                u = (a.pi == d) ? x[d] : u;
                v = (a.pj == d) ? x[d] : v;
            }
            if (kept > 0) {
                if (__builtin_fabs(th - last_t) <= a.dedup_t) continue;
                const double du = u - last_u, dv = v - last_v;
                if ((du * du + dv * dv) <= (a.dedup_p * a.dedup_p)) continue;
            }
            if (lane == 0 && kept < a.cap) {
                const size_t o = (size_t)i * a.cap + kept;
                a.hit_t[o] = th;
#pragma unroll
                for (int d = 0; d < 6; ++d) a.hit_x[o * 6 + d] = x[d];
                a.hit_p[o * 2] = u;
                a.hit_p[o * 2 + 1] = v;
            }
            ++kept;
            last_t = th;
            last_u = u;
            last_v = v;
            if (a.max_hits > 0 && kept >= a.max_hits) done = true;
        }
    }
    if (lane == 0) a.count[i] = kept;
}

/* Four waves per SIMD: 4,096 trajectories are one wave each and all of them
 * are resident at once (1,024 SIMDs), which hides the chunk loads better
 * than two waves per SIMD without spills (measured 0.60 against 0.75 ms on
 * 4,096 x 4,713 samples, same bits), though the 128-VGPR budget spills 66
 * registers (268 B per lane of scratch).                                   */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_synodic_detect(SynArgs a) {
    const int i = (int)((blockIdx.x * (unsigned)BLOCK + threadIdx.x) >> 6);
    if (i >= a.n) return;
    syn_detect_wave(a, i, threadIdx.x & 63);
}

/* time and state of candidate c of segment g (backend.py:345-360 / 591-606)
 * from the segment's samples k-1 .. k+2 already in registers (xm1, x2 and
 * tm1, tp2 are read only where syn_state reads them): syn_state's
 * expressions on the same values                                          */
__device__ __forceinline__ void syn_candidate(const SynArgs& a, const SynSeg& g, int N, int c,
                                              const double (&xm1)[6], const double (&x0)[6],
                                              const double (&x1)[6], const double (&x2)[6],
                                              double tm1, double tp2, double& th,
                                              double (&x)[6]) {
    int c2 = 0;
    const double s = syn_walk_deferred(a, g, c, &c2);
    const int k = g.k;
    if (s < 0.0) {
        th = g.t0;
#pragma unroll
        for (int d = 0; d < 6; ++d) x[d] = x0[d];
    } else {
        th = (1.0 - s) * g.t0 + s * g.t1;
        if (a.cubic && g.dt > 0.0 && k >= 1 && k + 2 < N) {
            const double oms = 1.0 - s;
            const double h00 = (1.0 + 2.0 * s) * (oms * oms);
            const double h10 = s * (oms * oms);
            const double h01 = (s * s) * (3.0 - 2.0 * s);
            const double h11 = (s * s) * (s - 1.0);
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                const double dx0 = (x1[d] - xm1[d]) / (g.t1 - tm1);
                const double dx1 = (x2[d] - x0[d]) / (tp2 - g.t0);
                x[d] = ((h00 * x0[d] + (h10 * dx0) * g.dt) + h01 * x1[d]) + (h11 * dx1) * g.dt;
            }
        } else {
#pragma unroll
            for (int d = 0; d < 6; ++d) x[d] = x0[d] + s * (x1[d] - x0[d]);
        }
    }
}

/* The multi-pass form.  Pass 1 scans every part of SYN_PART segments of
 * every trajectory with a wave of its own and appends its candidates, with
 * their places in segment order, to compact lists; pass 2 refines the
 * listed candidates, one per lane of a dense grid; pass 3 deduplicates
 * each trajectory's refined candidates in order.  Candidates come from the
 * same segment values, walk and refinement as in syn_detect_wave, so the
 * hits are the same bits.                                                 */
constexpr int SYN_PART = 512;     /* segments per part (8 chunks of 64)        */
constexpr int SYN_CAP = 32;       /* refined candidates recorded per part      */
constexpr int SYN_REC = 7;        /* doubles of a refined candidate: t, x[6]   */
constexpr int SYN_LISTS = 8;      /* candidate lists (part index % SYN_LISTS),
                                     so the appends of a pass meet on 8
                                     counters                                  */

struct SynWs {
    int parts;                    /* parts per trajectory                      */
    int* nlist;                   /* [SYN_LISTS * 32] entries of each list (one
                                     128-byte line per counter; zeroed per call) */
    int4* list;                   /* [SYN_LISTS][cap_list] (part index, segment,
                                     candidate, place) per candidate            */
    int cap_list;                 /* entries per list                          */
    int* ncand;                   /* [n][parts] candidates of the part (more
                                     than SYN_CAP: counted, not listed)        */
    double* rec;                  /* [n][parts][SYN_REC][SYN_CAP] refined
                                     candidates, component-major per part      */
};

/* Pass 1 for part p of trajectory i (one wave); `seg` is the wave's LDS
 * list of candidates (SYN_CAP entries).                                   */
__device__ __forceinline__ void syn_scan_part(SynArgs a, const SynWs& w, int i, int p,
                                              int lane, int4* seg) {
    SynTraj tr;
    tr.a = &a;
    tr.base = (size_t)i * a.m;
    tr.tbase = a.per_traj_times ? (size_t)i * a.m : 0;
    tr.N = a.lengths ? min(a.lengths[i], a.m) : a.m;
    const int nseg = tr.N - 1;
    const int s0 = p * SYN_PART;
    const int s1 = min(s0 + SYN_PART, nseg);
    const int unit = i * w.parts + p;
    int found = 0;        /* candidates so far */
    auto extra = [&](int k0) {
        return (lane == 0) ? k0 - 1 : ((lane == 62) ? k0 + 64 : ((lane == 63) ? k0 + 65 : -1));
    };
    double xo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, to = 0.0;
    double xe[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, te = 0.0;
    if (s0 < s1) {
        if (s0 + lane < tr.N) tr.load(s0 + lane, xo, to);
        const int jx = extra(s0);
        if (jx >= 0 && jx < tr.N) tr.load(jx, xe, te);
    }
    for (int k0 = s0; k0 < s1; k0 += 64) {
        const int k = k0 + lane;
        const double gj = (k < tr.N) ? tr.G(xo) : 0.0;
        const double tj = (k < tr.N) ? to : 0.0;
        const int jx = extra(k0);
        const bool ex = jx >= 0 && jx < tr.N;
        const double gx = ex ? tr.G(xe) : 0.0;
        const double tx = ex ? te : 0.0;
        if (k0 + 64 < s1) {
            if (k + 64 < tr.N) tr.load(k + 64, xo, to);
            const int jn = extra(k0 + 64);
            if (jn >= 0 && jn < tr.N) tr.load(jn, xe, te);
        }
        const double g_up1 = __shfl(gj, (lane + 1) & 63), t_up1 = __shfl(tj, (lane + 1) & 63);
        const double g_up2 = __shfl(gj, (lane + 2) & 63), t_up2 = __shfl(tj, (lane + 2) & 63);
        const double g_dn1 = __shfl(gj, (lane + 63) & 63), t_dn1 = __shfl(tj, (lane + 63) & 63);
        const double g_x62 = __shfl(gx, 62), t_x62 = __shfl(tx, 62);
        const double g_x63 = __shfl(gx, 63), t_x63 = __shfl(tx, 63);
        const double gk1 = (lane < 63) ? g_up1 : g_x62, tk1 = (lane < 63) ? t_up1 : t_x62;
        const double gk2 = (lane < 62) ? g_up2 : ((lane == 62) ? g_x62 : g_x63);
        const double tk2 = (lane < 62) ? t_up2 : ((lane == 62) ? t_x62 : t_x63);
        const double gkm1 = (lane > 0) ? g_dn1 : gx, tkm1 = (lane > 0) ? t_dn1 : tx;
        int cnt = 0;
        if (k < s1) {
            const SynSeg g = syn_segment(a, tr.N, k, gkm1, gj, gk1, gk2, tkm1, tj, tk1, tk2);
            syn_walk(a, g, -1, &cnt);
        }
        if (__ballot(cnt > 0)) {
            /* places in segment order: an exclusive scan of the counts */
            int inc = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(inc, o);
                if (lane >= o) inc += v;
            }
            const int first = found + inc - cnt;
            for (int c = 0; c < cnt && first + c < SYN_CAP; ++c)
                seg[first + c] = make_int4(unit, k, c, first + c);
            found += __shfl(inc, 63);
        }
    }
    if (lane == 0) w.ncand[unit] = found;
    if (found == 0 || found > SYN_CAP) return;   /* nothing, or detected whole by the merge */
    /* found <= SYN_CAP: every candidate is in seg[]; append them to list
     * unit % SYN_LISTS                                                    */
    __builtin_amdgcn_wave_barrier();
    const int li = unit % SYN_LISTS;
    int base = 0;
    if (lane == 0) base = atomicAdd(w.nlist + li * 32, found);
    base = __shfl(base, 0);
    if (lane < found) w.list[(size_t)li * w.cap_list + base + lane] = seg[lane];
}

/* Four waves per SIMD fit without spills; three, with a second chunk in
 * flight per wave, measured slower (288 against 247 us on 4,096 x 4,713
 * samples).                                                               */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_synodic_scan(SynArgs a, SynWs w) {
    __shared__ int4 seg[BLOCK / 64][SYN_CAP];
    const long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    if (wave >= (long long)a.n * w.parts) return;
    syn_scan_part(a, w, (int)(wave / w.parts), (int)(wave % w.parts), threadIdx.x & 63,
                  seg[threadIdx.x >> 6]);
}

/* Pass 2: every listed candidate in a lane of its own (a grid-stride loop
 * over the lists laid end to end), refined into its place in its part's
 * records.  The list lengths are loaded together up front.               */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_synodic_refine(SynArgs a, SynWs w) {
    const int stride = (int)(gridDim.x * BLOCK);
    int start[SYN_LISTS + 1];   /* entries of the lists before list li */
#pragma unroll
    for (int li = 0; li < SYN_LISTS; ++li) start[li + 1] = w.nlist[li * 32];
    start[0] = 0;
#pragma unroll
    for (int li = 0; li < SYN_LISTS; ++li) start[li + 1] += start[li];
    for (int q = (int)(blockIdx.x * BLOCK + threadIdx.x); q < start[SYN_LISTS]; q += stride) {
        int li = 0;
#pragma unroll
        for (int j = 1; j < SYN_LISTS; ++j) li = (q >= start[j]) ? j : li;
        const int e = q - start[li];
        {
            const int4 ent = w.list[(size_t)li * w.cap_list + e];
            const int i = ent.x / w.parts;
            SynTraj tr;
            tr.a = &a;
            tr.base = (size_t)i * a.m;
            tr.tbase = a.per_traj_times ? (size_t)i * a.m : 0;
            tr.N = a.lengths ? min(a.lengths[i], a.m) : a.m;
            const int k = ent.y;
            const bool m1 = k >= 1, p2 = k + 2 < tr.N;
            /* the four samples around the segment, loaded together up front */
            double x0[6], x1[6], t0, t1;
            double xm1[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, x2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            double tm1 = 0.0, t2 = 0.0;
            tr.load(k, x0, t0);
            tr.load(k + 1, x1, t1);
            if (m1) tr.load(k - 1, xm1, tm1);
            if (p2) tr.load(k + 2, x2, t2);
            const double gk = tr.G(x0), gk1 = tr.G(x1);
            const double gkm1 = m1 ? tr.G(xm1) : 0.0, gk2 = p2 ? tr.G(x2) : 0.0;
            const SynSeg g = syn_segment(a, tr.N, k, gkm1, gk, gk1, gk2, tm1, t0, t1, t2);
            double th, x[6];
            syn_candidate(a, g, tr.N, ent.z, xm1, x0, x1, x2, tm1, t2, th, x);
            double* rec = w.rec + (size_t)ent.x * (SYN_CAP * SYN_REC) + ent.w;
            rec[0] = th;
#pragma unroll
            for (int d = 0; d < 6; ++d) rec[(1 + d) * SYN_CAP] = x[d];
        }
    }
}

/* Pass 3: one wave per trajectory.  Lane l loads record r0 + l of the
 * trajectory's records (all parts, in order); the records are then
 * deduplicated in that order against the last kept hit (backend.py:369-440)
 * as in syn_detect_wave.  A trajectory with a part of more than SYN_CAP
 * candidates (none recorded for that part) is detected whole by the wave,
 * as syn_detect_wave does.                                                */
__device__ __forceinline__ void syn_merge_traj(SynArgs a, const SynWs& w, int i, int lane) {
    const int* nc = w.ncand + (size_t)i * w.parts;
    /* the parts' candidate counts, 64 at a time one per lane; with at most
     * 64 parts lane p keeps the number of records before part p           */
    int np = 0;
    bool over = false;
    int pre = 0;
    for (int p0 = 0; p0 < w.parts; p0 += 64) {
        const int c = (p0 + lane < w.parts) ? nc[p0 + lane] : 0;
        over = over || __any(c > SYN_CAP);
        int inc = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(inc, o);
            if (lane >= o) inc += v;
        }
        if (p0 == 0) pre = inc - c;
        np += __shfl(inc, 63);
    }
    if (over) {
        syn_detect_wave(a, i, lane);
        return;
    }
    int kept = 0;
    bool done = false;
    double last_t = 0.0, last_u = 0.0, last_v = 0.0;
    const double* rec_all = w.rec + (size_t)i * w.parts * (SYN_CAP * SYN_REC);
    for (int r0 = 0; r0 < np && !done; r0 += 64) {
        const int e = r0 + lane;
        double th = 0.0, x[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (e < np) {
            int part = 0, idx = e;
            if (w.parts <= 64) {
                /* the last part whose first record is at or before e (a
                 * part with no records shares its start with the next)    */
                int first = 0;
                for (int q = 1; q < w.parts; ++q) {
                    const int pq = __builtin_amdgcn_readlane(pre, q);
                    if (e >= pq) {
                        part = q;
                        first = pq;
                    }
                }
                idx = e - first;
            } else {
                while (idx >= nc[part]) idx -= nc[part++];
            }
            const double* r = rec_all + (size_t)part * (SYN_CAP * SYN_REC) + idx;
            th = r[0];
#pragma unroll
            for (int d = 0; d < 6; ++d) x[d] = r[(1 + d) * SYN_CAP];
        }
        const int total = min(64, np - r0);
        for (int j = 0; j < total && !done; ++j) {
            const double tj = rlane(th, j);
            double xj[6];
#pragma unroll
            for (int d = 0; d < 6; ++d) xj[d] = rlane(x[d], j);
            double u = xj[0], v = xj[0];
#pragma unroll
            for (int d = 1; d < 6; ++d) {
                u = (a.pi == d) ? xj[d] : u;
                v = (a.pj == d) ? xj[d] : v;
            }
            if (kept > 0) {
                if (__builtin_fabs(tj - last_t) <= a.dedup_t) continue;
                const double du = u - last_u, dv = v - last_v;
                if ((du * du + dv * dv) <= (a.dedup_p * a.dedup_p)) continue;
            }
            if (lane == 0 && kept < a.cap) {
                const size_t o = (size_t)i * a.cap + kept;
                a.hit_t[o] = tj;
#pragma unroll
                for (int d = 0; d < 6; ++d) a.hit_x[o * 6 + d] = xj[d];
                a.hit_p[o * 2] = u;
                a.hit_p[o * 2 + 1] = v;
            }
            ++kept;
            last_t = tj;
            last_u = u;
            last_v = v;
            if (a.max_hits > 0 && kept >= a.max_hits) done = true;
        }
    }
    if (lane == 0) a.count[i] = kept;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_synodic_merge(SynArgs a, SynWs w) {
    const int i = (int)((blockIdx.x * (unsigned)BLOCK + threadIdx.x) >> 6);
    if (i >= a.n) return;
    syn_merge_traj(a, w, i, threadIdx.x & 63);
}

}  // namespace hh

#endif /* HH_SYNODIC_H */

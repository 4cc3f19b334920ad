/* hh_kernels.hip -- HIP kernels and the C-ABI of the CR3BP propagation path.
 *
 * Kernels (gfx950, wave64):
 *   k_rk_fixed6<M>      fixed-step RK of tableau M (RK4, RK6, RK8), 6-d, one
 *                       trajectory per lane (integrators/rk.py:531-588 over
 *                       numpy.linspace grid); k_rk_fixed6_group per 8-lane
 *                       group, k_rk_fixed42 on the 42-d variational system
 *   k_dop853_final6     DOP853 to t_f, 6-d, one trajectory per lane; the
 *                       returned state is the dense-output value at t_f
 *                       (rk.py:2375-2549 with t_eval = [t0, tf])
 *   k_dop853_event6     DOP853 until the plane event y[i] = c, 6-d
 *                       (rk.py:2678-2803, 2005-2102)
 *   k_dop853_final42    DOP853 to t_f on the 42-d variational system, one
 *                       trajectory per 8-lane group; `persistent` template
 *                       flag selects the work-stealing form that refills a
 *                       group from a global counter as soon as its
 *                       trajectory ends
 * State arrays are structure-of-arrays: component c of trajectory i lives at
 * [c * n + i], c in the reference's order ([Phi row-major | x] for 42-d).
 */
#include "hh_device.h"
#include "hh_wave.h"
#include "hh_pair.h"
#include "hh_group6.h"
#include "hh_synodic.h"
#include "hh_newton.h"
#include "../../include/hiten_hip.h"

#include <float.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <type_traits>

using namespace hh;

namespace {

thread_local char g_err[512] = "";
/* persistent 42-d kernels: see hh_set_tuning */
hh_tuning_t g_tuning = {2000, 256, 2000, 4, 2, 0};

int fail(int code, const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return HH_ERR_HIP;
}

/* bits of the 6 components a lane owns that a backward _DirectedSystem
 * negates (dynamics/base.py:298-304)                                       */
__device__ __forceinline__ unsigned lane_flip_bits(int fwd, unsigned long long flip, int b,
                                                   bool is42) {
    unsigned m = 0;
    if (fwd != -1) return 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int c = is42 ? comp42(b < 7 ? b : 1, d) : d;
        m |= (unsigned)((flip >> c) & 1ULL) << d;
    }
    return m;
}

/* ------------------------------------------------------------------------ */
/* per-trajectory DOP853 state machine, final-value mode                     */
/* ------------------------------------------------------------------------ */
struct StepParams {
    double rtol, atol, max_step, min_step;
    long long max_steps;
};

template <class P>
struct Traj {
    double y[6];
    double k[DOP_K][6];
    double t, tf, h, err_prev;
    int nacc, nrej, status;
    bool done;
};

/* Start a trajectory: k0 = f(y0), initial step (rk.py:2441-2448).  A span
 * that numpy.isclose calls zero (dynamics/base.py:421) or a reversed span
 * returns y0 unchanged, as the reference does.                              */
template <class P>
__device__ __forceinline__ void traj_init(const P& p, Traj<P>& s, double t0, double tf,
                                          const StepParams& sp) {
    s.t = t0;
    s.tf = tf;
    s.err_prev = -1.0;
    s.nacc = 0;
    s.nrej = 0;
    s.status = ST_OK;
    s.done = isclose(t0, tf) || !((t0 - tf) * 1.0 < 0.0);
    p.rhs(s.y, s.k[0]);
    s.h = initial_step(p, s.y, s.k[0], sp.rtol, sp.atol, sp.min_step, sp.max_step);
}

/* Output policies of traj_iterate.
 *
 * FinalOut keeps the dense-output value at tf (rk.py:2503-2543 for
 * t_eval[-1]); GridOut streams the dense output at every point of
 * numpy.linspace(t0, tf, m) as soon as the accepted step that covers it is
 * known: point t_q belongs to the segment [ts[j], ts[j+1]) with
 * j = searchsorted(ts, t_q, 'right') - 1, clipped to the last segment
 * (rk.py:2503-2509), so it is emitted after the step ending at ts[j+1] when
 * t_q < ts[j+1], or after the final step.                                 */
struct FinalOut {
    double y[6];
    __device__ __forceinline__ void init(const double (&y0)[6]) {
#pragma unroll
        for (int d = 0; d < 6; ++d) y[d] = y0[d];
    }
    __device__ __forceinline__ void constant(const double (&y0)[6]) { init(y0); }
    __device__ __forceinline__ void stop(const double (&ylast)[6]) { init(ylast); }
    template <class P>
    __device__ __forceinline__ void accepted(const P& p, const double (&y)[6], const double (&yh)[6],
                                             const double (&k)[DOP_K][6], double t, double t_new,
                                             double tf, bool last) {
        if (!last) return;
        const double hs = t_new - t;
        if (hs == 0.0) {
            init(y);
            return;
        }
        double kx[DOP_SX - DOP_K][6];
        dop853_ext_stages(p, y, k, hs, kx);
        dense_value(y, yh, k, kx, hs, (tf - t) / hs, this->y);
    }
};

struct GridOut {
    double* out;        /* [n][m][dim] base of this trajectory */
    const double* teval;  /* explicit output times shared by all trajectories, or NULL */
    int m, q, dim, blk;
    double t0, tf, step, delta;
    bool zstep;

    __device__ __forceinline__ double tq(int i) const {
        if (teval) return teval[i];
        /* numpy.linspace(t0, tf, m) */
        if (i == m - 1) return tf;
        if (zstep) return ((double)i / (double)(m - 1)) * delta + t0;
        return (double)i * step + t0;
    }
    __device__ __forceinline__ void write(int i, const double (&v)[6]) const {
        if (dim == 6) {
#pragma unroll
            for (int d = 0; d < 6; ++d) out[(size_t)i * 6 + d] = v[d];
        } else if (blk < 7) {
#pragma unroll
            for (int d = 0; d < 6; ++d) out[(size_t)i * 42 + comp42(blk, d)] = v[d];
        }
    }
    __device__ __forceinline__ void init(const double (&)[6]) { q = 0; }
    __device__ __forceinline__ void constant(const double (&y0)[6]) {
        for (int i = 0; i < m; ++i) write(i, y0);
        q = m;
    }
    __device__ __forceinline__ void stop(const double (&ylast)[6]) {
        for (; q < m; ++q) write(q, ylast);
    }
    template <class P>
    __device__ __forceinline__ void accepted(const P& p, const double (&y)[6], const double (&yh)[6],
                                             const double (&k)[DOP_K][6], double t, double t_new,
                                             double, bool last) {
        if (q >= m || !(last || tq(q) < t_new)) return;
        const double hs = t_new - t;
        if (hs == 0.0) {
            for (; q < m; ++q) write(q, y);
            return;
        }
        double kx[DOP_SX - DOP_K][6];
        dop853_ext_stages(p, y, k, hs, kx);
        for (; q < m; ++q) {
            const double t_q = tq(q);
            if (!(last || t_q < t_new)) break;
            double v[6];
            dense_value(y, yh, k, kx, hs, (t_q - t) / hs, v);
            write(q, v);
        }
    }
};

/* One pass of the `while (t - tf) < 0` loop of rk.py:2452-2484.  The
 * output policy sees every accepted step; the step that reaches tf marks
 * the trajectory done.                                                      */
template <class P, class O>
__device__ __forceinline__ void traj_iterate(const P& p, Traj<P>& s, const StepParams& sp, O& o) {
    if (s.nacc + s.nrej >= sp.max_steps) {
        s.status = ST_MAXSTEPS;
        s.done = true;
        o.stop(s.y);
        return;
    }
    double h = clamp_step(s.h, sp.max_step, sp.min_step);
    h = adjust_step_to_endpoint(s.t, h, s.tf);
    double yh[6];
    const double err = dop853_attempt(p, s.y, s.k, h, sp.rtol, sp.atol, yh);
    if (err != err) {
        s.status = ST_NONFINITE;
        s.done = true;
        o.stop(s.y);
        return;
    }
    /* Lane42: both factors from one pow per lane of the group, formed
     * before the group-uniform branch; Lane6: the factor its branch needs */
    double f_acc = 0.0, f_rej = 0.0;
    if constexpr (P::N == 42) group_step_factors(p.blk, err, s.err_prev, f_acc, f_rej);
    if (err <= 1.0) {
        const double t_new = s.t + h;
        const bool last = !((t_new - s.tf) * 1.0 < 0.0);
        o.accepted(p, s.y, yh, s.k, s.t, t_new, s.tf, last);
        s.done = last;
        s.t = t_new;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            s.y[d] = yh[d];
            s.k[0][d] = s.k[DOP_S][d];
        }
        if constexpr (P::N != 42) f_acc = pi_accept_factor(err, s.err_prev);
        s.h = h * f_acc;
        s.err_prev = err;
        s.nacc += 1;
    } else {
        if constexpr (P::N != 42) f_rej = pi_reject_factor(err);
        s.h = clamp_step(h * f_rej, sp.max_step, sp.min_step);
        s.nrej += 1;
    }
}

/* ------------------------------------------------------------------------ */
/* kernels                                                                   */
/* ------------------------------------------------------------------------ */

/* time of grid point s: t_vals[s] when given, else numpy.linspace(t0, tf, steps) */
__device__ __forceinline__ double grid_time(const double* t_vals, int s, int steps, double t0,
                                            double tf, double step, double delta, bool zstep) {
    if (t_vals) return t_vals[s];
    if (s == steps - 1) return tf;
    if (zstep) return ((double)s / (double)(steps - 1)) * delta + t0;
    return (double)s * step + t0;
}

/* rk.py:531-588 with h = t[i+1] - t[i].  A span that numpy.isclose calls zero
 * keeps y0 at every point (dynamics/base.py:421-424).                       */
template <class M, class P>
__device__ __forceinline__ void rk_run(const P& p, double (&y)[6], const double* t_vals,
                                        int steps, double t0, double tf, double* traj_base,
                                        double* deriv_base, size_t traj_stride, int blk) {
    const double delta = tf - t0;
    const double step = delta / (double)(steps - 1);
    const bool zstep = (step == 0.0);
    const bool constant = t_vals ? isclose(t_vals[0], t_vals[steps - 1]) : isclose(t0, tf);
    auto store = [&](double* base, int s_idx, const double (&v)[6]) {
        if (!base) return;
        if (P::N == 6) {
#pragma unroll
            for (int d = 0; d < 6; ++d) base[((size_t)s_idx * 6 + d) * traj_stride] = v[d];
        } else if (blk < 7) {
#pragma unroll
            for (int d = 0; d < 6; ++d)
                base[((size_t)s_idx * 42 + comp42(blk, d)) * traj_stride] = v[d];
        }
    };
    store(traj_base, 0, y);
    double t_prev = grid_time(t_vals, 0, steps, t0, tf, step, delta, zstep);
    for (int s = 1; s < steps; ++s) {
        const double t_s = grid_time(t_vals, s, steps, t0, tf, step, delta, zstep);
        if (!constant) {
            double k0[6];
            rk_step<M>(p, y, t_s - t_prev, k0);
            store(deriv_base, s - 1, k0);   /* f at state s - 1 (rk.py:577, 587) */
        }
        t_prev = t_s;
        store(traj_base, s, y);
    }
    if (deriv_base) {
        /* the last state's derivative; on a constant grid every state is y0 */
        double f[6];
        p.rhs(y, f);
        if (constant) {
            for (int s = 0; s < steps; ++s) store(deriv_base, s, f);
        } else {
            store(deriv_base, steps - 1, f);
        }
    }
}

/* fixed-step RK of tableau M, 6-d, one trajectory per lane */
template <class M>
__global__ __launch_bounds__(256) void k_rk_fixed6(int n, const double* __restrict__ mu,
                                                   const double* __restrict__ t0a,
                                                   const double* __restrict__ tfa, int steps,
                                                   const double* __restrict__ t_vals, int fwd,
                                                   unsigned long long flip,
                                                   const double* __restrict__ y0,
                                                   double* __restrict__ yout,
                                                   double* __restrict__ traj,
                                                   double* __restrict__ deriv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lane6 p;
    p.mu = mu[i];
    p.flipbits = lane_flip_bits(fwd, flip, 0, false);
    p.flip_on = (fwd == -1) && flip != 0;
    double y[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) y[d] = y0[(size_t)d * n + i];
    rk_run<M>(p, y, t_vals, steps, t0a[i], tfa[i], traj ? traj + i : nullptr,
              deriv ? deriv + i : nullptr, (size_t)n, 0);
#pragma unroll
    for (int d = 0; d < 6; ++d) yout[(size_t)d * n + i] = y[d];
}

/* 42-d fixed-step RK, one trajectory per 8-lane group; traj: [steps][42][n] */
template <class M>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_rk_fixed42(int n, const double* __restrict__ mu, const double* __restrict__ t0a,
                   const double* __restrict__ tfa, int steps, const double* __restrict__ t_vals,
                   int fwd, unsigned long long flip, const double* __restrict__ y0,
                   double* __restrict__ yout, double* __restrict__ traj,
                   double* __restrict__ deriv) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gtid >> 3;
    const int b = gtid & 7;
    if (i >= n) return;
    Lane42 p;
    p.blk = b;
    p.mu = mu[i];
    p.mu2 = 1.0 - p.mu;
    p.flipbits = lane_flip_bits(fwd, flip, b, true);
    p.flip_on = (fwd == -1) && flip != 0;
    double y[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) y[d] = (b < 7) ? y0[(size_t)comp42(b, d) * n + i] : 0.0;
    rk_run<M>(p, y, t_vals, steps, t0a[i], tfa[i], traj ? traj + i : nullptr,
              deriv ? deriv + i : nullptr, (size_t)n, b);
    if (b < 7) {
#pragma unroll
        for (int d = 0; d < 6; ++d) yout[(size_t)comp42(b, d) * n + i] = y[d];
    }
}

/* _hermite_eval_dense (rk.py:286-327) of one component */
__device__ __forceinline__ double hermite1(double y0, double f0, double y1, double f1, double x,
                                           double h) {
    const double x2 = x * x;
    const double x3 = x2 * x;
    const double H00 = (2.0 * x3 - 3.0 * x2) + 1.0;
    const double H10 = (x3 - 2.0 * x2) + x;
    const double H01 = -2.0 * x3 + 3.0 * x2;
    const double H11 = x3 - x2;
    return ((H00 * y0 + H10 * (h * f0)) + H01 * y1) + H11 * (h * f1);
}

/* Fixed-step RK of tableau M until the plane event g = y[ev] - c, 6-d, one
 * trajectory per lane: _integrate_fixed_rk_until_event (rk.py:655-718) with
 * _hermite_refine_in_step (rk.py:330-391) over rk_run's grid.  Only the
 * event component is interpolated inside the bisection; the state is formed
 * at the returned abscissa (the reference's components are independent, so
 * the bits are the same).  Without a crossing t_hit is the last grid time
 * and y_hit = y_last = the last state.                                     */
template <class M>
__global__ __launch_bounds__(256) void k_rk_fixed_event6(
    int n, const double* __restrict__ mu, const double* __restrict__ t0a,
    const double* __restrict__ tfa, int steps, const double* __restrict__ t_vals, int fwd,
    unsigned long long flip, int ev_index, double ev_off, int dir, double xtol, double gtol,
    const double* __restrict__ y0, int* __restrict__ hit, double* __restrict__ t_hit,
    double* __restrict__ y_hit, double* __restrict__ y_last) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lane6 p;
    p.mu = mu[i];
    p.flipbits = lane_flip_bits(fwd, flip, 0, false);
    p.flip_on = (fwd == -1) && flip != 0;
    double y[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) y[d] = y0[(size_t)d * n + i];
    const double t0 = t_vals ? t_vals[0] : t0a[i];
    const double tf = t_vals ? t_vals[steps - 1] : tfa[i];
    const double delta = tf - t0;
    const double step = delta / (double)(steps - 1);
    const bool zstep = (step == 0.0);
    const bool constant = isclose(t0, tf);
    int found = 0;
    double th = grid_time(t_vals, steps - 1, steps, t0, tf, step, delta, zstep);
    double yh[6], yl[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) yh[d] = yl[d] = y[d];
    if (!constant) {
        double f[6];
        p.rhs(y, f);
        double gp = pick6(y, ev_index) - ev_off;
        double t_prev = grid_time(t_vals, 0, steps, t0, tf, step, delta, zstep);
        for (int s_idx = 1; s_idx < steps; ++s_idx) {
            const double t_s = grid_time(t_vals, s_idx, steps, t0, tf, step, delta, zstep);
            const double h = t_s - t_prev;
            double yn[6], fn[6];
#pragma unroll
            for (int d = 0; d < 6; ++d) yn[d] = y[d];
            rk_step<M>(p, yn, h);
            p.rhs(yn, fn);
            const double gn = pick6(yn, ev_index) - ev_off;
            if (event_crossed(gp, gn, dir)) {
                const double ye0 = pick6(y, ev_index), fe0 = pick6(f, ev_index);
                const double ye1 = pick6(yn, ev_index), fe1 = pick6(fn, ev_index);
                double a = 0.0, b = 1.0, gl = ye0 - ev_off, x = -1.0;
                for (int it = 0; it < 128; ++it) {
                    const double mid = 0.5 * (a + b);
                    const double gm = hermite1(ye0, fe0, ye1, fe1, mid, h) - ev_off;
                    if (__builtin_fabs(gm) <= gtol) {
                        x = mid;
                        break;
                    }
                    if (crossed_direction(gl, gm, dir)) {
                        b = mid;
                    } else {
                        a = mid;
                        gl = gm;
                    }
                    if ((b - a) * __builtin_fabs(h) <= xtol) break;
                }
                if (x < 0.0) x = b;
                found = 1;
                th = t_prev + x * h;
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    yh[d] = hermite1(y[d], f[d], yn[d], fn[d], x, h);
                    yl[d] = yn[d];
                }
                break;
            }
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                y[d] = yn[d];
                f[d] = fn[d];
                yh[d] = yl[d] = yn[d];
            }
            gp = gn;
            t_prev = t_s;
        }
    }
    hit[i] = found;
    t_hit[i] = th;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        y_hit[(size_t)d * n + i] = yh[d];
        y_last[(size_t)d * n + i] = yl[d];
    }
}

/* k_rk_fixed6 with one trajectory per aligned 8-lane group, one component
 * per lane (Group6): rk_run's grid, steps and stores, per component.        */
template <class M, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_rk_fixed6_group(
    int n, const double* __restrict__ mu, const double* __restrict__ t0a,
    const double* __restrict__ tfa, int steps, const double* __restrict__ t_vals, int fwd,
    unsigned long long flip, const double* __restrict__ y0, double* __restrict__ yout,
    double* __restrict__ traj, double* __restrict__ deriv) {
    const long long gtid = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const int i = (int)(gtid >> 3);
    if (i >= n) return;
    Group6 p;
    p.setup(threadIdx.x & 63, mu[i], fwd, flip, nullptr);   /* no step control */
    const int d = p.d;
    double y = p.act ? y0[(size_t)d * n + i] : 0.0;
    const double t0 = t0a[i], tf = tfa[i];
    const double delta = tf - t0;
    const double step = delta / (double)(steps - 1);
    const bool zstep = (step == 0.0);
    const bool constant = t_vals ? isclose(t_vals[0], t_vals[steps - 1]) : isclose(t0, tf);
    if (traj && p.act) traj[(size_t)d * n + i] = y;
    double t_prev = grid_time(t_vals, 0, steps, t0, tf, step, delta, zstep);
    for (int s_idx = 1; s_idx < steps; ++s_idx) {
        const double t_s = grid_time(t_vals, s_idx, steps, t0, tf, step, delta, zstep);
        if (!constant) {
            double k0;
            rk_step_group<M>(p, y, t_s - t_prev, k0);
            if (deriv && p.act) deriv[((size_t)(s_idx - 1) * 6 + d) * n + i] = k0;
        }
        t_prev = t_s;
        if (traj && p.act) traj[((size_t)s_idx * 6 + d) * n + i] = y;
    }
    if (deriv) {
        /* the last state's derivative; on a constant grid every state is y0 */
        const double f = p.rhs(y);
        if (p.act) {
            if (constant) {
                for (int s_idx = 0; s_idx < steps; ++s_idx)
                    deriv[((size_t)s_idx * 6 + d) * n + i] = f;
            } else {
                deriv[((size_t)(steps - 1) * 6 + d) * n + i] = f;
            }
        }
    }
    if (p.act) yout[(size_t)d * n + i] = y;
}

template <bool GRID>
__global__ __launch_bounds__(256) void k_dop853_run6(
    int n, const double* __restrict__ mu, const double* __restrict__ t0a,
    const double* __restrict__ tfa, StepParams sp, int fwd, unsigned long long flip,
    const double* __restrict__ y0, double* __restrict__ yout, int m,
    const double* __restrict__ teval, int* __restrict__ status, int* __restrict__ nacc,
    int* __restrict__ nrej) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lane6 p;
    p.mu = mu[i];
    p.flipbits = lane_flip_bits(fwd, flip, 0, false);
    p.flip_on = (fwd == -1) && flip != 0;
    Traj<Lane6> s;
#pragma unroll
    for (int d = 0; d < 6; ++d) s.y[d] = y0[(size_t)d * n + i];
    typename std::conditional<GRID, GridOut, FinalOut>::type o;
    if constexpr (GRID) {
        o.out = yout + (size_t)i * m * 6;
        o.teval = teval;
        o.m = m;
        o.dim = 6;
        o.blk = 0;
        o.t0 = t0a[i];
        o.tf = tfa[i];
        o.delta = o.tf - o.t0;
        o.step = o.delta / (double)(m - 1);
        o.zstep = (o.step == 0.0);
    }
    o.init(s.y);
    traj_init(p, s, t0a[i], tfa[i], sp);
    if (s.done) o.constant(s.y);
    while (!s.done) traj_iterate(p, s, sp, o);
    if constexpr (!GRID) {
#pragma unroll
        for (int d = 0; d < 6; ++d) yout[(size_t)d * n + i] = o.y[d];
    }
    if (status) status[i] = s.status;
    if (nacc) nacc[i] = s.nacc;
    if (nrej) nrej[i] = s.nrej;
}

/* k_dop853_run6 with one trajectory per aligned 8-lane group, one component
 * per lane (Group6).  traj_init / traj_iterate and the FinalOut / GridOut
 * policies, statement for statement, with each lane holding its component. */
/* A caller's request to stop a loop early: with CANCEL set, the loops below
 * read the flag (LDS) before each attempt and stop with ST_CANCELLED once it
 * is nonzero.  Without it they compile as before.                         */
template <bool CANCEL>
__device__ __forceinline__ bool cancelled(const int* flag) {
    if constexpr (CANCEL)
        return __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
    return false;
}

/* One trajectory of k_dop853_run6 in the Group6 layout, statement for
 * statement: the grid form writes this lane's component of every output
 * point to gout ([m][6]); the final form returns it in yfin.  Status and
 * counters are the same in all lanes of the group.                         */
template <bool GRID, bool CANCEL = false>
__device__ __forceinline__ void group6_run(const Group6& p, double y, double t0, double tf,
                                           const StepParams& sp, double* gout, int m,
                                           const double* teval, double& yfin, int& status,
                                           int& nacc, int& nrej, const int* cancel = nullptr) {
    const int d = p.d;
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    /* GridOut: output times numpy.linspace(t0, tf, m) or teval */
    const double gdelta = tf - t0;
    const double gstep = GRID ? gdelta / (double)(m - 1) : 0.0;
    const bool gzstep = (gstep == 0.0);
    auto tq = [&](int q) -> double {
        if (teval) return teval[q];
        if (q == m - 1) return tf;
        if (gzstep) return ((double)q / (double)(m - 1)) * gdelta + t0;
        return (double)q * gstep + t0;
    };
    auto put = [&](int q, double v) {
        if (p.act) gout[(size_t)q * 6 + d] = v;
    };
    int q = 0;
    yfin = y;   /* FinalOut */
    double t = t0, h, err_prev = -1.0;
    int na = 0, nr = 0, st = ST_OK;
    bool done = isclose(t0, tf) || !((t0 - tf) * 1.0 < 0.0);
    k[0] = p.rhs(y);
    h = p.initial_step(y, k[0], sp.rtol, sp.atol, sp.min_step, sp.max_step);
    if (done) {
        if constexpr (GRID) {
            for (int j = 0; j < m; ++j) put(j, y);
            q = m;
        }
    }
    while (!done) {
        if (na + nr >= sp.max_steps) {
            st = ST_MAXSTEPS;
            done = true;
            if constexpr (GRID) { for (; q < m; ++q) put(q, y); } else { yfin = y; }
            break;
        }
        if (cancelled<CANCEL>(cancel)) {
            st = ST_CANCELLED;
            yfin = y;
            break;
        }
        double hc = clamp_step(h, sp.max_step, sp.min_step);
        hc = adjust_step_to_endpoint(t, hc, tf);
        double yh;
        const double err = p.attempt(y, k, hc, sp.rtol, sp.atol, yh);
        if (err != err) {
            st = ST_NONFINITE;
            done = true;
            if constexpr (GRID) { for (; q < m; ++q) put(q, y); } else { yfin = y; }
            break;
        }
        double fa, fr;
        p.step_factors(err, err_prev, fa, fr);
        if (err <= 1.0) {
            const double t_new = t + hc;
            const bool last = !((t_new - tf) * 1.0 < 0.0);
            const double hs = t_new - t;
            if constexpr (GRID) {
                if (q < m && (last || tq(q) < t_new)) {
                    if (hs == 0.0) {
                        for (; q < m; ++q) put(q, y);
                    } else {
                        double row[DOP_P];
                        p.dense_rows(y, yh, k, hs, row);
                        for (; q < m; ++q) {
                            const double t_q = tq(q);
                            if (!(last || t_q < t_new)) break;
                            put(q, dense_from_rows(row, y, (t_q - t) / hs));
                        }
                    }
                }
            } else {
                if (last) {
                    if (hs == 0.0) {
                        yfin = y;
                    } else {
                        double row[DOP_P];
                        p.dense_rows(y, yh, k, hs, row);
                        yfin = dense_from_rows(row, y, (tf - t) / hs);
                    }
                }
            }
            done = last;
            t = t_new;
            y = yh;
            k[0] = k[DOP_S];
            h = hc * fa;
            err_prev = err;
            na += 1;
        } else {
            h = clamp_step(hc * fr, sp.max_step, sp.min_step);
            nr += 1;
        }
    }
    status = st;
    nacc = na;
    nrej = nr;
}

template <int BLOCK, bool GRID>
__global__ __launch_bounds__(BLOCK) void k_dop853_run6_group(
    int n, const double* __restrict__ mu, const double* __restrict__ t0a,
    const double* __restrict__ tfa, StepParams sp, int fwd, unsigned long long flip,
    const double* __restrict__ y0, double* __restrict__ yout, int m,
    const double* __restrict__ teval, int* __restrict__ status, int* __restrict__ nacc,
    int* __restrict__ nrej) {
    __shared__ PowTab powt;
    pow_tab_fill(&powt);
    __syncthreads();
    const long long gtid = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const int i = (int)(gtid >> 3);
    if (i >= n) return;
    Group6 p;
    p.setup(threadIdx.x & 63, mu[i], fwd, flip, &powt);
    const int d = p.d;
    const double y = p.act ? y0[(size_t)d * n + i] : 0.0;
    double yfin;
    int st, na, nr;
    group6_run<GRID>(p, y, t0a[i], tfa[i], sp, GRID ? yout + (size_t)i * m * 6 : nullptr, m,
                     teval, yfin, st, na, nr);
    if constexpr (!GRID) {
        if (p.act) yout[(size_t)d * n + i] = yfin;
    }
    if (d == 0) {
        if (status) status[i] = st;
        if (nacc) nacc[i] = na;
        if (nrej) nrej[i] = nr;
    }
}

struct EventParams {
    int index;
    double offset;
    int direction;
    double xtol, gtol;
};

__global__ __launch_bounds__(256) void k_dop853_event6(
    int n, const double* __restrict__ mu, const double* __restrict__ t0a,
    const double* __restrict__ tmaxa, StepParams sp, EventParams ev, int fwd,
    unsigned long long flip, const double* __restrict__ y0, int* __restrict__ hit,
    double* __restrict__ t_hit, double* __restrict__ y_hit, double* __restrict__ y_last,
    int* __restrict__ status, int* __restrict__ nacc, int* __restrict__ nrej) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lane6 p;
    p.mu = mu[i];
    p.flipbits = lane_flip_bits(fwd, flip, 0, false);
    p.flip_on = (fwd == -1) && flip != 0;
    double y[6], k[DOP_K][6];
#pragma unroll
    for (int d = 0; d < 6; ++d) y[d] = y0[(size_t)d * n + i];
    double t = t0a[i];
    const double tmax = tmaxa[i];
    p.rhs(y, k[0]);
    double gp = pick6(y, ev.index) - ev.offset;
    double h = initial_step(p, y, k[0], sp.rtol, sp.atol, sp.min_step, sp.max_step);
    double err_prev = -1.0;
    int na = 0, nr = 0, st = ST_OK, found = 0;
    double th = 0.0, yh_out[6], ylast[6];
    while ((t - tmax) * 1.0 < 0.0) {
        if (na + nr >= sp.max_steps) {
            st = ST_MAXSTEPS;
            break;
        }
        h = clamp_step(h, sp.max_step, sp.min_step);
        h = adjust_step_to_endpoint(t, h, tmax);
        double yh[6];
        const double err = dop853_attempt(p, y, k, h, sp.rtol, sp.atol, yh);
        if (err != err) {
            st = ST_NONFINITE;
            break;
        }
        if (err <= 1.0) {
            const double t_new = t + h;
            const double gn = pick6(yh, ev.index) - ev.offset;
            na += 1;
            if (event_crossed(gp, gn, ev.direction)) {
                double kx[DOP_SX - DOP_K][6];
                dop853_ext_stages(p, y, k, h, kx);
                double fc[DOP_P];
#pragma unroll
                for (int r = 0; r < DOP_P; ++r) {
                    double v = dense_row(r, 0, y, yh, k, kx, h);
#pragma unroll
                    for (int d = 1; d < 6; ++d)
                        v = (ev.index == d) ? dense_row(r, d, y, yh, k, kx, h) : v;
                    fc[r] = v;
                }
                const double x = refine_bisect(pick6(y, ev.index), fc, ev.offset, ev.direction,
                                               h, ev.xtol, ev.gtol);
                th = t + x * h;
                dense_value(y, yh, k, kx, h, x, yh_out);
#pragma unroll
                for (int d = 0; d < 6; ++d) ylast[d] = yh[d];
                found = 1;
                break;
            }
            t = t_new;
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                y[d] = yh[d];
                k[0][d] = k[DOP_S][d];
            }
            gp = gn;
            h = h * pi_accept_factor(err, err_prev);
            err_prev = err;
        } else {
            h = clamp_step(h * pi_reject_factor(err), sp.max_step, sp.min_step);
            nr += 1;
        }
    }
    if (!found) {
        th = t;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            yh_out[d] = y[d];
            ylast[d] = y[d];
        }
    }
    hit[i] = found;
    t_hit[i] = th;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        y_hit[(size_t)d * n + i] = yh_out[d];
        if (y_last) y_last[(size_t)d * n + i] = ylast[d];
    }
    if (status) status[i] = st;
    if (nacc) nacc[i] = na;
    if (nrej) nrej[i] = nr;
}

/* One event search of k_dop853_event6 in the Group6 layout, statement for
 * statement; the event component is read from its lane of the group.  Per
 * lane: this lane's components of y_hit and y_last; the flags, t_hit and
 * counters are the same in all lanes of the group.                         */
struct EventOut6 {
    int found, status, nacc, nrej;
    double t_hit, y_hit, y_last;
};

template <bool CANCEL = false>
__device__ __forceinline__ void group6_event(const Group6& p, double y, double t, double tmax,
                                             const StepParams& sp, const EventParams& ev,
                                             EventOut6& o, const int* cancel = nullptr) {
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    k[0] = p.rhs(y);
    const int ev_addr = p.base + (ev.index << 2);
    double gp = bperm(y, ev_addr) - ev.offset;
    double h = p.initial_step(y, k[0], sp.rtol, sp.atol, sp.min_step, sp.max_step);
    double err_prev = -1.0;
    int na = 0, nr = 0, st = ST_OK, found = 0;
    double th = 0.0, yh_out = 0.0, ylast = 0.0;
    while ((t - tmax) * 1.0 < 0.0) {
        if (na + nr >= sp.max_steps) {
            st = ST_MAXSTEPS;
            break;
        }
        if (cancelled<CANCEL>(cancel)) {
            st = ST_CANCELLED;
            break;
        }
        h = clamp_step(h, sp.max_step, sp.min_step);
        h = adjust_step_to_endpoint(t, h, tmax);
        double yh;
        const double err = p.attempt(y, k, h, sp.rtol, sp.atol, yh);
        if (err != err) {
            st = ST_NONFINITE;
            break;
        }
        double fa, fr;
        p.step_factors(err, err_prev, fa, fr);
        if (err <= 1.0) {
            const double t_new = t + h;
            const double gn = bperm(yh, ev_addr) - ev.offset;
            na += 1;
            if (event_crossed(gp, gn, ev.direction)) {
                double row[DOP_P], fc[DOP_P];
                p.dense_rows(y, yh, k, h, row);
#pragma unroll
                for (int r = 0; r < DOP_P; ++r) fc[r] = bperm(row[r], ev_addr);
                const double x = refine_bisect(bperm(y, ev_addr), fc, ev.offset, ev.direction, h,
                                               ev.xtol, ev.gtol);
                th = t + x * h;
                yh_out = dense_from_rows(row, y, x);
                ylast = yh;
                found = 1;
                break;
            }
            t = t_new;
            y = yh;
            k[0] = k[DOP_S];
            gp = gn;
            h = h * fa;
            err_prev = err;
        } else {
            h = clamp_step(h * fr, sp.max_step, sp.min_step);
            nr += 1;
        }
    }
    if (!found) {
        th = t;
        yh_out = y;
        ylast = y;
    }
    o.found = found;
    o.status = st;
    o.nacc = na;
    o.nrej = nr;
    o.t_hit = th;
    o.y_hit = yh_out;
    o.y_last = ylast;
}

/* k_dop853_event6 with one trajectory per aligned 8-lane group, one
 * component per lane (Group6): the latency form for small batches.          */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_dop853_event6_group(
    int n, const double* __restrict__ mu, const double* __restrict__ t0a,
    const double* __restrict__ tmaxa, StepParams sp, EventParams ev, int fwd,
    unsigned long long flip, const double* __restrict__ y0, int* __restrict__ hit,
    double* __restrict__ t_hit, double* __restrict__ y_hit, double* __restrict__ y_last,
    int* __restrict__ status, int* __restrict__ nacc, int* __restrict__ nrej) {
    __shared__ PowTab powt;
    pow_tab_fill(&powt);
    __syncthreads();
    const long long gtid = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const int i = (int)(gtid >> 3);
    if (i >= n) return;
    Group6 p;
    p.setup(threadIdx.x & 63, mu[i], fwd, flip, &powt);
    const int d = p.d;
    const double y = p.act ? y0[(size_t)d * n + i] : 0.0;
    EventOut6 o;
    group6_event(p, y, t0a[i], tmaxa[i], sp, ev, o);
    if (p.act) {
        y_hit[(size_t)d * n + i] = o.y_hit;
        if (y_last) y_last[(size_t)d * n + i] = o.y_last;
    }
    if (d == 0) {
        hit[i] = o.found;
        t_hit[i] = o.t_hit;
        if (status) status[i] = o.status;
        if (nacc) nacc[i] = o.nacc;
        if (nrej) nrej[i] = o.nrej;
    }
}

/* ---- 42-d, one trajectory per 8-lane group ------------------------------ */

struct Final42Args {
    int n;
    const double* mu;
    const double* t0;
    const double* tf;
    StepParams sp;
    int fwd;
    unsigned long long flip;
    const double* y0;
    double* yout;
    int* status;
    int* nacc;
    int* nrej;
    int* counter;
    const int* order;   /* work order of the persistent form, or NULL (index order) */
    int m;              /* grid points per trajectory (grid form)                  */
    const double* teval;  /* explicit output times (grid form), or NULL for linspace */
    int heavy;            /* persistent form: projected attempts above which a
                             trajectory is moved to a wave of its own, 0 = only
                             in the tail */
    int heavy_age;        /* attempts before the projection is trusted */
    int pair_heavy;       /* pair form: projected attempts from which a heavy
                             trajectory takes both waves of its workgroup
                             (0 = never; below it, one wave) */
    int refill_min;       /* pair form: idle groups a wave gathers before it
                             refills (1 = refill every idle group at once)   */
};

__device__ __forceinline__ void load42(const Final42Args& a, int traj, int b, Lane42& p,
                                       Traj<Lane42>& s) {
    p.mu = a.mu[traj];
    p.mu2 = 1.0 - p.mu;
    p.flipbits = lane_flip_bits(a.fwd, a.flip, b, true);
    p.flip_on = (a.fwd == -1) && a.flip != 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int c = comp42(b < 7 ? b : 1, d);
        s.y[d] = (b < 7) ? a.y0[(size_t)c * a.n + traj] : 0.0;
    }
    traj_init(p, s, a.t0[traj], a.tf[traj], a.sp);
}

__device__ __forceinline__ void store42(const Final42Args& a, int traj, int b,
                                        const Traj<Lane42>& s, const double (&yo)[6]) {
    if (b < 7) {
#pragma unroll
        for (int d = 0; d < 6; ++d) a.yout[(size_t)comp42(b, d) * a.n + traj] = yo[d];
    }
    if (b == 0) {
        if (a.status) a.status[traj] = s.status;
        if (a.nacc) a.nacc[traj] = s.nacc;
        if (a.nrej) a.nrej[traj] = s.nrej;
    }
}

template <int BLOCK, bool GRID>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_dop853_final42_plain(Final42Args a) {
    const int gtid = blockIdx.x * BLOCK + threadIdx.x;
    const int traj = gtid >> 3;
    const int b = gtid & 7;
    if (traj >= a.n) return;
    Lane42 p;
    p.blk = b;
    Traj<Lane42> s;
    load42(a, traj, b, p, s);
    typename std::conditional<GRID, GridOut, FinalOut>::type o;
    if constexpr (GRID) {
        o.out = a.yout + (size_t)traj * a.m * 42;
        o.teval = a.teval;
        o.m = a.m;
        o.dim = 42;
        o.blk = b;
        o.t0 = a.t0[traj];
        o.tf = a.tf[traj];
        o.delta = o.tf - o.t0;
        o.step = o.delta / (double)(a.m - 1);
        o.zstep = (o.step == 0.0);
    }
    o.init(s.y);
    if (s.done) o.constant(s.y);
    while (!s.done) traj_iterate(p, s, a.sp, o);
    if constexpr (GRID) {
        if (b == 0) {
            if (a.status) a.status[traj] = s.status;
            if (a.nacc) a.nacc[traj] = s.nacc;
            if (a.nrej) a.nrej[traj] = s.nrej;
        }
    } else {
        store42(a, traj, b, s, o.y);
    }
}

/* Per-attempt hook of wide_finish: scheduling only, never touches results */
struct NoHook {
    __device__ __forceinline__ void attempt(int, double) {}
    __device__ __forceinline__ void end() {}
};

/* traj_iterate + FinalOut for one trajectory in a wave-wide layout (Wave42
 * or Pair42), from the current (y, k[0], t, h, err_prev, counters) to its
 * end: this lane's component of the final state in `out`.                  */
template <class WL, class Hook = NoHook>
__device__ __forceinline__ void wide_loop(const StepParams& sp, const WL& W, double y,
                                          double (&k)[DOP_K], double t, double tf, double h,
                                          double err_prev, int& nacc, int& nrej, double& out,
                                          int& status, Hook hook = Hook()) {
    out = y;
    status = ST_OK;
    for (;;) {
        hook.attempt(nacc + nrej, t);
        if (nacc + nrej >= sp.max_steps) {
            status = ST_MAXSTEPS;
            out = y;
            break;
        }
        double hh = clamp_step(h, sp.max_step, sp.min_step);
        hh = adjust_step_to_endpoint(t, hh, tf);
        double yh;
        const double err = W.attempt(y, k, hh, sp.rtol, sp.atol, yh);
        if (err != err) {
            status = ST_NONFINITE;
            out = y;
            break;
        }
        double fa, fr;
        W.step_factors(err, err_prev, fa, fr);
        if (err <= 1.0) {
            const double t_new = t + hh;
            const bool last = !((t_new - tf) * 1.0 < 0.0);
            if (last) {
                const double hs = t_new - t;
                out = (hs == 0.0) ? y : W.dense_at(y, yh, k, hs, (tf - t) / hs);
            }
            t = t_new;
            y = yh;
            k[0] = k[DOP_S];
            h = hh * fa;
            err_prev = err;
            nacc += 1;
            if (last) break;
        } else {
            h = clamp_step(hh * fr, sp.max_step, sp.min_step);
            nrej += 1;
        }
    }
    hook.end();
}

/* wide_loop, then the stores of store42 */
template <class WL, class Hook = NoHook>
__device__ __forceinline__ void wide_finish(const Final42Args& a, const WL& W, int traj,
                                            double y, double (&k)[DOP_K], double t, double tf,
                                            double h, double err_prev, int nacc, int nrej,
                                            Hook hook = Hook()) {
    double out;
    int status;
    wide_loop(a.sp, W, y, k, t, tf, h, err_prev, nacc, nrej, out, status, hook);
    if (W.holds()) a.yout[(size_t)W.comp() * a.n + traj] = out;
    if (W.leader()) {
        if (a.status) a.status[traj] = status;
        if (a.nacc) a.nacc[traj] = nacc;
        if (a.nrej) a.nrej[traj] = nrej;
    }
}

/* Finish the trajectory of the group whose lane 0 is `src` with the
 * wave-wide layout: lane 8*b + d takes entry d of block b from lane src + b.
 * Raises the wave's issue priority meanwhile: such a trajectory is the
 * longest-running work on its SIMD.                                         */
template <class Hook = NoHook>
__device__ __forceinline__ void run_wide(const Final42Args& a, const PowTab* pt, int lane, int src,
                                         int traj_lane, const Lane42& p, const Traj<Lane42>& s,
                                         Hook hook = Hook()) {
    Wave42 W;
    W.setup(lane, rlane(p.mu, src), a.fwd, a.flip, pt);
    const int from = src + (lane >> 3);
    double y = 0.0;
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
#pragma unroll
    for (int dd = 0; dd < 6; ++dd) {
        const double vy = __shfl(s.y[dd], from);
        const double vk = __shfl(s.k[0][dd], from);
        if (W.act && W.d == dd) {
            y = vy;
            k[0] = vk;
        }
    }
    __builtin_amdgcn_s_setprio(3);
    wide_finish(a, W, __builtin_amdgcn_readlane(traj_lane, src), y, k, rlane(s.t, src),
                rlane(s.tf, src), rlane(s.h, src), rlane(s.err_prev, src),
                __builtin_amdgcn_readlane(s.nacc, src), __builtin_amdgcn_readlane(s.nrej, src),
                hook);
    __builtin_amdgcn_s_setprio(0);
}

/* One trajectory per wave from the start (latency form for small batches):
 * traj_init (rk.py:2441-2448) in the wave-wide layout, then wide_finish.    */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_dop853_final42_wide(Final42Args a) {
    __shared__ PowTab powt;
    pow_tab_fill(&powt);
    __syncthreads();
    const int traj = (int)((blockIdx.x * (unsigned)BLOCK + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (traj >= a.n) return;
    Wave42 W;
    W.setup(lane, a.mu[traj], a.fwd, a.flip, &powt);
    const double y0 = W.act ? a.y0[(size_t)comp42(lane >> 3, W.d) * a.n + traj] : 0.0;
    const double t0 = a.t0[traj], tf = a.tf[traj];
    if (isclose(t0, tf) || !((t0 - tf) * 1.0 < 0.0)) {
        /* zero or reversed span: y0 unchanged (dynamics/base.py:421-424) */
        if (W.act) a.yout[(size_t)comp42(lane >> 3, W.d) * a.n + traj] = y0;
        if (lane == 0) {
            if (a.status) a.status[traj] = ST_OK;
            if (a.nacc) a.nacc[traj] = 0;
            if (a.nrej) a.nrej[traj] = 0;
        }
        return;
    }
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    k[0] = W.rhs(y0);
    /* initial_step with the block-tree sums of Lane42 */
    const StepParams& sp = a.sp;
    const double sc = sp.atol + sp.rtol * __builtin_fabs(y0);
    const double ua = y0 / sc;
    const double ub = k[0] / sc;
    const double s0 = W.block_tree(W.act ? ua * ua : 0.0);
    const double s1 = W.block_tree(W.act ? ub * ub : 0.0);
    const double sn = __builtin_sqrt(42.0);
    const double d0 = __builtin_sqrt(s0) / sn;
    const double d1 = __builtin_sqrt(s1) / sn;
    const double h = select_initial_step(d0, d1, sp.min_step, sp.max_step);
    wide_finish(a, W, traj, y0, k, t0, tf, h, -1.0, 0, 0);
}

/* One trajectory per pair of waves (128-thread workgroup) from the start:
 * the state wave and the STM wave of Pair42 side by side on two SIMDs.
 * traj_init (rk.py:2441-2448) with the block-tree sums of Lane42, then
 * wide_loop; a zero or reversed span leaves y0 unchanged
 * (dynamics/base.py:421-424).  Both waves of the workgroup must call it.   */
__device__ __forceinline__ void pair_final(const Pair42& P, const StepParams& sp, double y0,
                                           double t0, double tf, double& out, int& status,
                                           int& nacc, int& nrej) {
    nacc = 0;
    nrej = 0;
    if (isclose(t0, tf) || !((t0 - tf) * 1.0 < 0.0)) {
        out = y0;
        status = ST_OK;
        return;
    }
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    k[0] = P.rhs(y0);
    const double sc = sp.atol + sp.rtol * __builtin_fabs(y0);
    const double ua = y0 / sc;
    const double ub = k[0] / sc;
    double s0, s1;
    P.pair_tree(P.holds() ? ua * ua : 0.0, P.holds() ? ub * ub : 0.0, s0, s1);
    const double sn = __builtin_sqrt(42.0);
    const double d0 = __builtin_sqrt(s0) / sn;
    const double d1 = __builtin_sqrt(s1) / sn;
    const double h = select_initial_step(d0, d1, sp.min_step, sp.max_step);
    wide_loop(sp, P, y0, k, t0, tf, h, -1.0, nacc, nrej, out, status);
}

__global__ __launch_bounds__(128) void k_dop853_final42_pair(Final42Args a) {
    __shared__ PairLds lds;
    __shared__ PowTab powt;
    pow_tab_fill(&powt);
    __syncthreads();
    const int traj = blockIdx.x;
    const int lane = threadIdx.x & 63;
    /* wave-uniform in a scalar register, so that every branch on the role is
     * a uniform one and each wave executes exactly its own barriers */
    const int role = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (traj >= a.n) return;
    Pair42 P;
    P.setup(lane, role, a.mu[traj], a.fwd, a.flip, &lds, &powt);
    const double y0 = P.holds() ? a.y0[(size_t)P.comp() * a.n + traj] : 0.0;
    double out;
    int status, nacc, nrej;
    pair_final(P, a.sp, y0, a.t0[traj], a.tf[traj], out, status, nacc, nrej);
    if (P.holds()) a.yout[(size_t)P.comp() * a.n + traj] = out;
    if (P.leader()) {
        if (a.status) a.status[traj] = status;
        if (a.nacc) a.nacc[traj] = nacc;
        if (a.nrej) a.nrej[traj] = nrej;
    }
}

/* Work-stealing form.  Each wave keeps its 8 groups busy: after every step
 * attempt, groups whose trajectory ended take the next indices from a global
 * counter (one atomic per wave, ranks by ballot popcount).  A wave exits when
 * the counter has passed n and all of its groups are idle, so every wave
 * terminates once every trajectory has (each trajectory is bounded by
 * max_steps).                                                               */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_dop853_final42_persist(Final42Args a) {
    __shared__ PowTab powt;
    pow_tab_fill(&powt);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int b = lane & 7;
    const int grp = lane >> 3;
    Lane42 p;
    p.blk = b;
    Traj<Lane42> s;
    FinalOut o;
    int traj = -1;
    bool exhausted = false;
    s.done = true;
    for (;;) {
        /* A trajectory that has taken `heavy` attempts stops the refill of its
         * wave, so that it is soon alone there and moves to the wave-wide form */
        bool heavy = false;
        if (a.heavy > 0 && traj >= 0) {
            /* projected attempts = age * span / elapsed time */
            const int age = s.nacc + s.nrej;
            if (age >= a.heavy_age) {
                const double t0 = a.t0[traj];
                heavy = (double)age * (s.tf - t0) >= (double)a.heavy * (s.t - t0);
            }
        }
        const bool hold = __any(heavy);
        const bool need = (traj < 0) && !exhausted && !hold;
        const unsigned long long m = __ballot(need && b == 0);
        if (m) {
            const int cnt = __popcll(m);
            const int leader = __ffsll((long long)m) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(a.counter, cnt);
            base = __shfl(base, leader);
            if (need) {
                const unsigned long long below = m & ((1ULL << (grp * 8)) - 1ULL);
                const int idx = base + __popcll(below);
                if (idx < a.n) {
                    traj = a.order ? a.order[idx] : idx;
                    load42(a, traj, b, p, s);
                    o.init(s.y);
                } else {
                    exhausted = true;
                }
            }
        }
        if (traj >= 0 && s.done) {
            /* zero-span trajectories finish at load time */
            store42(a, traj, b, s, o.y);
            traj = -1;
            continue;
        }
        if (!__any(traj >= 0)) {
            if (__all(exhausted || traj >= 0)) break;
            continue;
        }
        const unsigned long long live = __ballot(traj >= 0 && b == 0);
        if (__popcll(live) == 1 && (hold || __any(exhausted))) {
            /* one trajectory left in the wave: finish it one component per lane */
            run_wide(a, &powt, lane, __ffsll((long long)live) - 1, traj, p, s);
            traj = -1;
            continue;
        }
        if (traj >= 0) {
            traj_iterate(p, s, a.sp, o);
            if (s.done) {
                store42(a, traj, b, s, o.y);
                traj = -1;
            }
        }
    }
}

/* ---- persistent form with a resume queue and pair tails --------------- */

/* Workspace of the pair-tail persistent form (caller memory, reset by the
 * library on the stream before each launch):
 *   ctr[0] fresh-work counter        ctr[1] resume slots claimed (tail)
 *   ctr[2] resume slots taken (head) ctr[3] trajectories not yet finished
 *   ctr[4] spills that found no slot (an error that the capacity rules
 *          out; such a trajectory would end with status HH_STATUS_DROPPED)
 *   ctr[5] slots taken by waves that hand their work away to sleep beside
 *          a reserved SIMD (at most SLEEP_BUDGET(n); past it they keep
 *          working instead)
 *   Capacity: every pairing or wave-wide run finishes one trajectory and
 *   spills at most 15 (the other 7 groups of its wave, 8 of the partner
 *   wave), so those spills take at most 15 n slots; the sleepers' spills are
 *   bounded by their own budget; cap = 15 n + 64 + SLEEP_BUDGET(n).
 *   rq[cap]        trajectory ids of the resume queue, -1 until written
 *   spill[89][n]   resumable state of a spilled trajectory: y (42, entry
 *                  6*b + d of Lane42 block b), k0 (42), t, h, err_prev, nacc,
 *                  nrej.  (t, h, err_prev, counters and the FSAL stage are
 *                  the whole state of the step loop, so a trajectory resumed
 *                  elsewhere continues bit for bit.)                      */
constexpr int RQ_CTR = 64;          /* ints reserved for the counters      */
constexpr int SPILL_W = 89;

struct ResumeArgs {
    int* ctr;
    int* rq;
    double* spill;
    int cap;
    int* simd_tab;   /* SIMD_TAB ints: nonzero = that SIMD is reserved         */
    int exclusive;   /* projected attempts from which a wave-wide or pair run
                        reserves its SIMD (0 = never)                          */
};

/* This wave's SIMD as an index < SIMD_TAB, from the hardware id registers
 * (XCC, shader engine, array, CU, SIMD).  Only used to find the other waves
 * that share a SIMD; a wrong layout would cost time, never results.        */
constexpr int SIMD_TAB = 65536;
__device__ __forceinline__ int simd_key() {
    const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);    /* HW_ID, 32 bits  */
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20);   /* XCC_ID, 4 bits  */
    const unsigned simd = (hw >> 4) & 3u, cu = (hw >> 8) & 15u, sh = (hw >> 12) & 1u,
                   se = (hw >> 13) & 7u;
    return (int)(((xcc & 15u) << 12) | (se << 9) | (sh << 8) | (cu << 4) | simd);
}

/* Reserves the wave's SIMD for a long run once the trajectory's projected
 * attempts, age * span / elapsed, reach r.exclusive (checked every 32
 * attempts): the other waves of the persistent kernel on that SIMD spill
 * their work and sleep until end() releases it.                            */
struct SimdReserve {
    int* tab;
    int key, threshold, lane;
    double t0, span;
    bool held;
    __device__ __forceinline__ void attempt(int age, double t) {
        if (held || !tab || threshold <= 0 || (age & 31) != 0 || age == 0) return;
        if ((double)age * span >= (double)threshold * (t - t0)) {
            int old = 1;
            if (lane == 0) old = atomicCAS(&tab[key], 0, 1);
            held = __builtin_amdgcn_readfirstlane(__shfl(old, 0)) == 0;
        }
    }
    __device__ __forceinline__ void end() {
        if (held && lane == 0) atomicExch(&tab[key], 0);
        held = false;
    }
};

__host__ __device__ constexpr long long sleep_budget(long long n) { return 16 * n + 64; }
__host__ __device__ constexpr long long resume_cap(long long n) {
    return 15 * n + 64 + sleep_budget(n);
}
__host__ __device__ constexpr long long ws_align(long long bytes) { return (bytes + 255) / 256 * 256; }
/* Workspace bytes of the pair-tail persistent form for n trajectories */
__host__ __device__ constexpr long long pair_workspace(long long n) {
    return ws_align((RQ_CTR + resume_cap(n)) * 4) + ws_align((long long)SPILL_W * n * 8) +
           (long long)SIMD_TAB * 4;
}

/* The trajectory a pair takes over, in Wave42's block/entry order */
struct PairHandoff {
    double y[8][8];
    double k[8][8];
    double t, tf, h, err_prev, mu;
    int nacc, nrej, traj;
};

/* Write the resumable state of every group with `me` set to the spill area
 * and publish its id in the resume queue (one slot claim per wave).       */
__device__ __forceinline__ void spill42(const Final42Args& a, const ResumeArgs& r, bool me,
                                        int traj, int b, int grp, int lane,
                                        const Traj<Lane42>& s) {
    const unsigned long long m = __ballot(me && b == 0);
    if (!m) return;
    int base = 0;
    if (lane == 0) base = atomicAdd(&r.ctr[1], __popcll(m));
    base = __shfl(base, 0);
    const int slot = base + __popcll(m & ((1ULL << (grp * 8)) - 1ULL));
    const size_t n = (size_t)a.n;
    if (me) {
        if (b < 7) {
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                r.spill[(size_t)(b * 6 + d) * n + traj] = s.y[d];
                r.spill[(size_t)(42 + b * 6 + d) * n + traj] = s.k[0][d];
            }
        }
        if (b == 0) {
            r.spill[84 * n + traj] = s.t;
            r.spill[85 * n + traj] = s.h;
            r.spill[86 * n + traj] = s.err_prev;
            r.spill[87 * n + traj] = (double)s.nacc;
            r.spill[88 * n + traj] = (double)s.nrej;
        }
    }
    __threadfence();
    if (me && b == 0) {
        if (slot < r.cap) {
            __hip_atomic_store(&r.rq[slot], traj, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            /* no slot (ruled out by the capacity above): the trajectory ends
             * here and says so instead of leaving its outputs unwritten */
            atomicAdd(&r.ctr[4], 1);
            if (a.status) a.status[traj] = ST_DROPPED;
            if (a.nacc) a.nacc[traj] = s.nacc;
            if (a.nrej) a.nrej[traj] = s.nrej;
            atomicSub(&r.ctr[3], 1);
        }
    }
}

/* Continue a spilled trajectory in this lane's group (inverse of spill42) */
__device__ __forceinline__ void resume42(const Final42Args& a, const ResumeArgs& r, int traj,
                                         int b, Lane42& p, Traj<Lane42>& s) {
    const size_t n = (size_t)a.n;
    p.mu = a.mu[traj];
    p.mu2 = 1.0 - p.mu;
    p.flipbits = lane_flip_bits(a.fwd, a.flip, b, true);
    p.flip_on = (a.fwd == -1) && a.flip != 0;
    const int bb = (b < 7) ? b : 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const double vy = r.spill[(size_t)(bb * 6 + d) * n + traj];
        const double vk = r.spill[(size_t)(42 + bb * 6 + d) * n + traj];
        s.y[d] = (b < 7) ? vy : 0.0;
        s.k[0][d] = (b < 7) ? vk : 0.0;
    }
    s.t = r.spill[84 * n + traj];
    s.h = r.spill[85 * n + traj];
    s.err_prev = r.spill[86 * n + traj];
    s.nacc = (int)r.spill[87 * n + traj];
    s.nrej = (int)r.spill[88 * n + traj];
    s.tf = a.tf[traj];
    s.status = ST_OK;
    s.done = false;
}

/* Finish the handed-over trajectory with both waves of the workgroup */
__device__ __forceinline__ void run_pair(const Final42Args& a, const ResumeArgs& r,
                                         const PairHandoff& ho, PairLds* plds, const PowTab* pt,
                                         int wid, int lane, int key) {
    Pair42 P;
    P.setup(lane, wid, ho.mu, a.fwd, a.flip, plds, pt);
    const int bb = lane >> 3, dd = lane & 7;
    const double y = P.holds() ? ho.y[bb][dd] : 0.0;
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    k[0] = P.holds() ? ho.k[bb][dd] : 0.0;
    SimdReserve hook{r.simd_tab, key, r.exclusive, lane, a.t0[ho.traj],
                     ho.tf - a.t0[ho.traj], false};
    __builtin_amdgcn_s_setprio(3);
    wide_finish(a, P, ho.traj, y, k, ho.t, ho.tf, ho.h, ho.err_prev, ho.nacc, ho.nrej, hook);
    __builtin_amdgcn_s_setprio(0);
    if (P.leader()) atomicSub(&r.ctr[3], 1);
}

/* Work-stealing form with two waves per workgroup.  Each wave keeps its 8
 * Lane42 groups busy from the resume queue first, then from the fresh
 * counter.  A trajectory flagged heavy (projected attempts, as in the form
 * above) -- or the last one of a wave once no work is left -- moves on at
 * once: its wave spills every other trajectory it holds to the resume queue,
 * where any wave picks them up, and finishes it in the Wave42 layout.  Past
 * the second threshold it is finished by both waves of its workgroup in the
 * Pair42 layout instead: its wave claims the workgroup's pair (LDS
 * compare-and-swap), both waves spill what they hold and meet at a barrier;
 * a wave that finds its partner's claim joins it at its next iteration.
 * Waves leave once every trajectory has finished; each trajectory is bounded
 * by max_steps and every pairing or wave-wide run finishes one, so every
 * wave leaves.  The only waits are the barriers of a pair (both waves reach
 * them) and the wait for a claimed resume slot, whose producer writes it
 * right after claiming it.                                                 */
template <int BLOCK, int WPE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void k_dop853_final42_resume(Final42Args a, ResumeArgs r) {
    static_assert(BLOCK == 128, "the pair form takes two waves per workgroup");
    __shared__ PairLds plds;
    __shared__ PairHandoff ho;
    __shared__ int pair_state;
    __shared__ PowTab powt;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = lane & 7;
    const int grp = lane >> 3;
    if (threadIdx.x == 0) pair_state = -1;
    pow_tab_fill(&powt);
    __syncthreads();
    const int key = simd_key();
    Lane42 p;
    p.blk = b;
    Traj<Lane42> s;
    FinalOut o;
    int traj = -1;
    bool fresh_done = false;  /* the fresh counter has passed n                */
    bool queue_dry = false;   /* the last look at the resume queue found it empty */
    s.done = true;
    for (;;) {
        /* heavy trajectories (projected attempts = age * span / elapsed time):
         * past a.pair_heavy they take both waves of the workgroup, past
         * a.heavy -- or as the last one of a wave once no work is left -- the
         * wave they are in                                                   */
        bool heavy = false, huge = false;
        if (traj >= 0) {
            const int age = s.nacc + s.nrej;
            if (age >= a.heavy_age) {
                const double t0 = a.t0[traj];
                const double lhs = (double)age * (s.tf - t0);
                heavy = a.heavy > 0 && lhs >= (double)a.heavy * (s.t - t0);
                huge = a.pair_heavy > 0 && lhs >= (double)a.pair_heavy * (s.t - t0);
            }
        }
        const unsigned long long live = __ballot(traj >= 0 && b == 0);
        if (__popcll(live) == 1 && fresh_done && queue_dry) heavy = heavy || (traj >= 0);
        const unsigned long long hm = __ballot(huge && b == 0);
        /* pair decisions are kept in scalar registers: every branch around a
         * barrier is then a uniform one, taken or skipped by the whole wave */
        int owner_src = -1;
        int join = 0;
        if (hm) {
            int old = 0;
            if (lane == 0) old = atomicCAS(&pair_state, -1, wid);
            old = __builtin_amdgcn_readfirstlane(__shfl(old, 0));
            if (old == -1) owner_src = __ffsll((long long)hm) - 1;
            else join = 1;
        } else {
            join = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&pair_state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= 0
                    ? 1 : 0);
        }
        owner_src = __builtin_amdgcn_readfirstlane(owner_src);
        if (owner_src >= 0 || join) {
            const bool moved = (owner_src >= 0) && (grp == (owner_src >> 3));
            spill42(a, r, traj >= 0 && !moved, traj, b, grp, lane, s);
            if (moved) {
                if (b < 7) {
#pragma unroll
                    for (int d = 0; d < 6; ++d) {
                        ho.y[b][d] = s.y[d];
                        ho.k[b][d] = s.k[0][d];
                    }
                }
                if (b == 0) {
                    ho.t = s.t;
                    ho.tf = s.tf;
                    ho.h = s.h;
                    ho.err_prev = s.err_prev;
                    ho.mu = p.mu;
                    ho.nacc = s.nacc;
                    ho.nrej = s.nrej;
                    ho.traj = traj;
                }
            }
            traj = -1;
            __syncthreads();
            run_pair(a, r, ho, &plds, &powt, wid, lane, key);
            __syncthreads();
            if (threadIdx.x == 0) pair_state = -1;
            __syncthreads();
            continue;
        }
        const unsigned long long wm = __ballot(heavy && b == 0);
        if (wm) {
            /* one wave: the other trajectories of the wave go to the resume
             * queue and this one is finished one component per lane       */
            const int src = __ffsll((long long)wm) - 1;
            const bool moved = grp == (src >> 3);
            spill42(a, r, traj >= 0 && !moved, traj, b, grp, lane, s);
            if (!moved) traj = -1;
            const int tr = __builtin_amdgcn_readlane(traj, src);
            const double t0 = a.t0[tr];
            run_wide(a, &powt, lane, src, traj, p, s,
                     SimdReserve{r.simd_tab, key, r.exclusive, lane, t0, a.tf[tr] - t0, false});
            if (lane == 0) atomicSub(&r.ctr[3], 1);
            traj = -1;
            continue;
        }
        /* a long run holds this SIMD: hand this wave's work to the others and
         * sleep until it is released (or a pair claim or the end wakes it) */
        bool sleep = r.exclusive > 0 &&
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(&r.simd_tab[key], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) != 0;
        if (sleep && live) {
            /* the slots this wave's spill takes come out of the sleepers'
             * budget; with the budget spent the wave keeps its work */
            const int cnt = __popcll(live);
            int ok = 1;
            if (lane == 0) {
                const int old = atomicAdd(&r.ctr[5], cnt);
                if ((long long)old + cnt > sleep_budget(a.n)) {
                    atomicSub(&r.ctr[5], cnt);
                    ok = 0;
                }
            }
            sleep = __builtin_amdgcn_readfirstlane(__shfl(ok, 0)) != 0;
        }
        if (sleep) {
            spill42(a, r, traj >= 0, traj, b, grp, lane, s);
            traj = -1;
            while (__builtin_amdgcn_readfirstlane(
                       __hip_atomic_load(&r.simd_tab[key], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT)) != 0 &&
                   __hip_atomic_load(&r.ctr[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0 &&
                   __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                       &pair_state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) < 0)
                __builtin_amdgcn_s_sleep(64);
            continue;
        }
        /* refill: the resume queue first, then fresh work; a wave that still
         * has live groups waits until refill_min of its groups are idle    */
        int finished = 0;
        const unsigned long long need = __ballot(traj < 0 && b == 0);
        if (need && (__popcll(need) >= a.refill_min || __popcll(live) == 0)) {
            int got = 0, rbase = 0, dry = 0;
            if (lane == 0) {
                const int tail = min(__hip_atomic_load(&r.ctr[1], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT), r.cap);
                const int head = __hip_atomic_load(&r.ctr[2], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                dry = tail <= head;
                if (!dry) {
                    const int take = min(tail - head, __popcll(need));
                    if (atomicCAS(&r.ctr[2], head, head + take) == head) {
                        got = take;
                        rbase = head;
                    }
                }
            }
            got = __shfl(got, 0);
            rbase = __shfl(rbase, 0);
            queue_dry = __shfl(dry, 0) != 0;
            const int rank = __popcll(need & ((1ULL << (grp * 8)) - 1ULL));
            if (traj < 0 && rank < got) {
                int id = -1;
                if (b == 0) {
                    /* the producer claimed this slot and writes it next */
                    while ((id = __hip_atomic_load(&r.rq[rbase + rank], __ATOMIC_ACQUIRE,
                                                   __HIP_MEMORY_SCOPE_AGENT)) < 0)
                        __builtin_amdgcn_s_sleep(2);
                }
                traj = __shfl(id, grp * 8);
                resume42(a, r, traj, b, p, s);
                o.init(s.y);
            }
            const bool want = (traj < 0) && !fresh_done;
            const unsigned long long m = __ballot(want && b == 0);
            if (m) {
                const int cnt = __popcll(m);
                const int leader = __ffsll((long long)m) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&r.ctr[0], cnt);
                base = __shfl(base, leader);
                bool past = false;
                if (want) {
                    const int idx = base + __popcll(m & ((1ULL << (grp * 8)) - 1ULL));
                    if (idx < a.n) {
                        traj = a.order ? a.order[idx] : idx;
                        load42(a, traj, b, p, s);
                        o.init(s.y);
                        if (s.done) {
                            /* zero or reversed span: finished at load time */
                            store42(a, traj, b, s, o.y);
                            traj = -1;
                            finished = 1;
                        }
                    } else {
                        past = true;
                    }
                }
                fresh_done = __any(past);
            }
        }
        if (traj >= 0) {
            traj_iterate(p, s, a.sp, o);
            if (s.done) {
                store42(a, traj, b, s, o.y);
                traj = -1;
                finished = 1;
            }
        }
        const unsigned long long fm = __ballot(finished && b == 0);
        if (fm && lane == 0) atomicSub(&r.ctr[3], __popcll(fm));
        if (!__any(traj >= 0)) {
            if (__hip_atomic_load(&r.ctr[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 0)
                break;
            __builtin_amdgcn_s_sleep(8);
        }
    }
}

__global__ void k_fill_int(int* p, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* ---- trajectory filters of the manifold computation ---------------------- */

/* Jacobi constant as common/energy.py:51-54 writes it (x**2 -> x*x, **0.5 ->
 * the correctly rounded square root) */
__device__ __forceinline__ double jacobi_c(double x, double y, double z, double vx, double vy,
                                           double vz, double mu1, double mu2) {
    const double r1 = __builtin_sqrt(((x + mu2) * (x + mu2) + y * y) + z * z);
    const double r2 = __builtin_sqrt(((x - mu1) * (x - mu1) + y * y) + z * z);
    return ((x * x + y * y) + 2.0 * (mu1 / r1 + mu2 / r2)) - ((vx * vx + vy * vy) + vz * vz);
}

/* NaN-propagating minimum (numpy.min) */
__device__ __forceinline__ double nan_min(double a, double b) {
    if (a != a || b != b) return a + b;
    return b < a ? b : a;
}

/* One wave per trajectory of an [n][m][6] state grid; lanes stride over the
 * samples.  min_r1/min_r2: numpy min of the primary distances of
 * services/manifold.py:416-420; max_err: _max_rel_energy_error
 * (common/energy.py:26-74), whose running maximum skips NaN terms.  Minimum
 * and maximum are exact, so the lane order does not change the result.    */
__global__ __launch_bounds__(256) void k_traj_filters(int n, int m, const double* __restrict__ mu,
                                                      const double* __restrict__ states,
                                                      double* __restrict__ min_r1,
                                                      double* __restrict__ min_r2,
                                                      double* __restrict__ max_err) {
    const int traj = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (traj >= n) return;
    const double mu2 = mu[traj];
    const double mu1 = 1.0 - mu2;
    const double* S = states + (size_t)traj * m * 6;
    const double C0 = jacobi_c(S[0], S[1], S[2], S[3], S[4], S[5], mu1, mu2);
    const double absC0 = __builtin_fabs(C0);
    double r1m = __builtin_inf(), r2m = __builtin_inf(), em = 0.0;
    for (int i = lane; i < m; i += 64) {
        const double* q = S + (size_t)i * 6;
        const double x = q[0], y = q[1], z = q[2];
        const double xa = x + mu2;
        const double xb = (x - 1.0) + mu2;
        r1m = nan_min(r1m, __builtin_sqrt((xa * xa + y * y) + z * z));
        r2m = nan_min(r2m, __builtin_sqrt((xb * xb + y * y) + z * z));
        if (i >= 1) {
            const double Ci = jacobi_c(x, y, z, q[3], q[4], q[5], mu1, mu2);
            const double rel = (absC0 > 1e-14) ? __builtin_fabs(Ci - C0) / absC0
                                               : __builtin_fabs(Ci - C0);
            if (rel > em) em = rel;
        }
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) {
        r1m = nan_min(r1m, __shfl_xor(r1m, sh));
        r2m = nan_min(r2m, __shfl_xor(r2m, sh));
        const double o = __shfl_xor(em, sh);
        em = (o > em) ? o : em;
    }
    if (lane == 0) {
        if (min_r1) min_r1[traj] = r1m;
        if (min_r2) min_r2[traj] = r2m;
        if (max_err) max_err[traj] = em;
    }
}

/* ---- unit probes: RHS and pow on given inputs --------------------------- */
__global__ void k_eval_rhs(int n, int dim, const double* __restrict__ mu,
                           const double* __restrict__ y, double* __restrict__ out) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    if (dim == 6) {
        const int i = gtid;
        if (i >= n) return;
        Lane6 p;
        p.mu = mu[i];
        p.flipbits = 0;
        p.flip_on = false;
        double s[6], o[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) s[d] = y[(size_t)d * n + i];
        p.rhs(s, o);
#pragma unroll
        for (int d = 0; d < 6; ++d) out[(size_t)d * n + i] = o[d];
    } else {
        const int i = gtid >> 3;
        const int b = gtid & 7;
        if (i >= n) return;
        Lane42 p;
        p.blk = b;
        p.mu = mu[i];
        p.mu2 = 1.0 - p.mu;
        p.flipbits = 0;
        p.flip_on = false;
        double s[6], o[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) s[d] = (b < 7) ? y[(size_t)comp42(b, d) * n + i] : 0.0;
        p.rhs(s, o);
        if (b < 7) {
#pragma unroll
            for (int d = 0; d < 6; ++d) out[(size_t)comp42(b, d) * n + i] = o[d];
        }
    }
}

__global__ void k_eval_math(int n, const double* __restrict__ x, const double* __restrict__ y,
                            double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], b = y[i];
    out[(size_t)0 * n + i] = __builtin_sqrt(a);
    out[(size_t)1 * n + i] = a / b;
    out[(size_t)2 * n + i] = hh_pow15(a);
    out[(size_t)3 * n + i] = hh_pow25(a);
    out[(size_t)4 * n + i] = hh_pow(a, b);
    out[(size_t)5 * n + i] = hh_dev_sqrt(a);
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return HH_OK;
}

int blocks_for(long long threads, int block) { return (int)((threads + block - 1) / block); }

StepParams make_sp(double rtol, double atol, double max_step, double min_step,
                   long long max_steps) {
    StepParams sp;
    sp.rtol = rtol;
    sp.atol = atol;
    sp.max_step = max_step;
    sp.min_step = min_step;
    sp.max_steps = max_steps;
    return sp;
}

/* the device-resident corrector, on the loops above */
#include "hh_correct.h"

template <class M>
void launch_rk_fixed(int n, int dim, const double* mu, const double* t0p,
                            const double* tfp, int steps, const double* t_vals, int fwd,
                            unsigned long long flip_mask, const double* y0, double* y_out,
                            double* traj_out, double* deriv_out, int variant, int block,
                            hipStream_t st) {
    if (dim == 6 && variant == HH_VARIANT_WIDE) {
        const int g = blocks_for((long long)n * 8, block);
        switch (block) {
            case 64: hipLaunchKernelGGL((k_rk_fixed6_group<M, 64>), dim3(g), dim3(64), 0, st, n, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out, traj_out, deriv_out); break;
            case 128: hipLaunchKernelGGL((k_rk_fixed6_group<M, 128>), dim3(g), dim3(128), 0, st, n, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out, traj_out, deriv_out); break;
            case 192: hipLaunchKernelGGL((k_rk_fixed6_group<M, 192>), dim3(g), dim3(192), 0, st, n, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out, traj_out, deriv_out); break;
            default: hipLaunchKernelGGL((k_rk_fixed6_group<M, 256>), dim3(g), dim3(256), 0, st, n, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out, traj_out, deriv_out); break;
        }
    } else if (dim == 6) {
        hipLaunchKernelGGL(k_rk_fixed6<M>, dim3(blocks_for(n, block)), dim3(block), 0, st, n, mu,
                           t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out, traj_out,
                           deriv_out);
    } else {
        hipLaunchKernelGGL(k_rk_fixed42<M>, dim3(blocks_for((long long)n * 8, block)),
                           dim3(block), 0, st, n, mu, t0p, tfp, steps, t_vals, fwd, flip_mask,
                           y0, y_out, traj_out, deriv_out);
    }
}

}  // namespace

/* ======================================================================== */
/* C-ABI                                                                     */
/* ======================================================================== */
template <bool GRID>
static int launch_run6_group(int n, const double* mu, const double* t0, const double* tf,
                             StepParams sp, int fwd, unsigned long long flip, const double* y0,
                             double* out, int m, const double* teval, int* status, int* nacc,
                             int* nrej, int block, hipStream_t st, const char* what) {
    const int g = blocks_for((long long)n * 8, block);
    switch (block) {
        case 64: hipLaunchKernelGGL((k_dop853_run6_group<64, GRID>), dim3(g), dim3(64), 0, st, n, mu, t0, tf, sp, fwd, flip, y0, out, m, teval, status, nacc, nrej); break;
        case 128: hipLaunchKernelGGL((k_dop853_run6_group<128, GRID>), dim3(g), dim3(128), 0, st, n, mu, t0, tf, sp, fwd, flip, y0, out, m, teval, status, nacc, nrej); break;
        case 192: hipLaunchKernelGGL((k_dop853_run6_group<192, GRID>), dim3(g), dim3(192), 0, st, n, mu, t0, tf, sp, fwd, flip, y0, out, m, teval, status, nacc, nrej); break;
        default: hipLaunchKernelGGL((k_dop853_run6_group<256, GRID>), dim3(g), dim3(256), 0, st, n, mu, t0, tf, sp, fwd, flip, y0, out, m, teval, status, nacc, nrej); break;
    }
    return check_launch(what);
}

extern "C" {

const char* hh_last_error(void) { return g_err; }

int hh_abi_version(void) { return HH_ABI_VERSION; }

int hh_set_tuning(const hh_tuning_t* t) {
    if (!t) return fail(HH_ERR_ARGS, "hh_set_tuning: null pointer");
    if (t->heavy_projected < 0 || t->min_age < 1 || t->pair_projected < 0 ||
        t->refill_min < 1 || t->refill_min > 8 || t->waves_per_simd < 1 ||
        t->waves_per_simd > 2 || t->exclusive_projected < 0)
        return fail(HH_ERR_ARGS, "hh_set_tuning: heavy_projected >= 0, min_age >= 1, "
                                 "pair_projected >= 0, refill_min in 1..8, waves_per_simd in "
                                 "1..2, exclusive_projected >= 0");
    g_tuning = *t;
    return HH_OK;
}

int hh_get_tuning(hh_tuning_t* t) {
    if (!t) return fail(HH_ERR_ARGS, "hh_get_tuning: null pointer");
    *t = g_tuning;
    return HH_OK;
}

int hh_rk_fixed(int order, int n, int dim, const double* mu, const double* t0, const double* tf,
                int steps, const double* t_vals, int fwd, unsigned long long flip_mask,
                const double* y0, double* y_out, double* traj_out, double* deriv_out,
                int variant, int block, void* stream) {
    if (order != 4 && order != 6 && order != 8)
        return fail(HH_ERR_ARGS, "hh_rk_fixed: order must be 4, 6 or 8");
    if (n < 0) return fail(HH_ERR_ARGS, "hh_rk_fixed: n < 0");
    if (n == 0) return HH_OK;
    if (dim != 6 && dim != 42) return fail(HH_ERR_ARGS, "hh_rk_fixed: dim must be 6 or 42");
    if (steps < 2) return fail(HH_ERR_ARGS, "hh_rk_fixed: steps must be >= 2");
    if (!mu || !y0 || !y_out || (!t_vals && (!t0 || !tf)))
        return fail(HH_ERR_ARGS, "hh_rk_fixed: null pointer");
    if (variant != HH_VARIANT_PLAIN && variant != HH_VARIANT_WIDE)
        return fail(HH_ERR_ARGS, "hh_rk_fixed: variant must be PLAIN or WIDE");
    if (block <= 0) block = dim == 6 ? 128 : 256;
    if (block > 256 || block % 64)
        return fail(HH_ERR_ARGS, "hh_rk_fixed: block must be 64..256, multiple of 64");
    const double* t0p = t0 ? t0 : mu;   /* unused when t_vals is given */
    const double* tfp = tf ? tf : mu;
    hipStream_t st = (hipStream_t)stream;
    if (order == 4)
        launch_rk_fixed<RK4>(n, dim, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out,
                             traj_out, deriv_out, variant, block, st);
    else if (order == 6)
        launch_rk_fixed<RK6>(n, dim, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out,
                             traj_out, deriv_out, variant, block, st);
    else
        launch_rk_fixed<RK8>(n, dim, mu, t0p, tfp, steps, t_vals, fwd, flip_mask, y0, y_out,
                             traj_out, deriv_out, variant, block, st);
    return check_launch("hh_rk_fixed launch");
}

int hh_rk_fixed_event(int order, int n, const double* mu, const double* t0, const double* tf,
                      int steps, const double* t_vals, int fwd, unsigned long long flip_mask,
                      int ev_index, double ev_offset, int direction, double xtol, double gtol,
                      const double* y0, int* hit, double* t_hit, double* y_hit, double* y_last,
                      int block, void* stream) {
    if (order != 4 && order != 6 && order != 8)
        return fail(HH_ERR_ARGS, "hh_rk_fixed_event: order must be 4, 6 or 8");
    if (n < 0) return fail(HH_ERR_ARGS, "hh_rk_fixed_event: n < 0");
    if (n == 0) return HH_OK;
    if (steps < 2) return fail(HH_ERR_ARGS, "hh_rk_fixed_event: steps must be >= 2");
    if (ev_index < 0 || ev_index > 5)
        return fail(HH_ERR_ARGS, "hh_rk_fixed_event: ev_index must be in 0..5");
    if (!mu || !y0 || !hit || !t_hit || !y_hit || !y_last || (!t_vals && (!t0 || !tf)))
        return fail(HH_ERR_ARGS, "hh_rk_fixed_event: null pointer");
    if (block <= 0) block = 128;
    if (block > 256 || block % 64)
        return fail(HH_ERR_ARGS, "hh_rk_fixed_event: block must be 64..256, multiple of 64");
    const double* t0p = t0 ? t0 : mu;
    const double* tfp = tf ? tf : mu;
    const int dir = direction > 0 ? 1 : (direction < 0 ? -1 : 0);
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid(blocks_for(n, block)), blk(block);
    if (order == 4)
        hipLaunchKernelGGL(k_rk_fixed_event6<RK4>, grid, blk, 0, st, n, mu, t0p, tfp, steps,
                           t_vals, fwd, flip_mask, ev_index, ev_offset, dir, xtol, gtol, y0, hit,
                           t_hit, y_hit, y_last);
    else if (order == 6)
        hipLaunchKernelGGL(k_rk_fixed_event6<RK6>, grid, blk, 0, st, n, mu, t0p, tfp, steps,
                           t_vals, fwd, flip_mask, ev_index, ev_offset, dir, xtol, gtol, y0, hit,
                           t_hit, y_hit, y_last);
    else
        hipLaunchKernelGGL(k_rk_fixed_event6<RK8>, grid, blk, 0, st, n, mu, t0p, tfp, steps,
                           t_vals, fwd, flip_mask, ev_index, ev_offset, dir, xtol, gtol, y0, hit,
                           t_hit, y_hit, y_last);
    return check_launch("hh_rk_fixed_event launch");
}

int hh_rk8_fixed(int n, int dim, const double* mu, const double* t0, const double* tf, int steps,
                 const double* t_vals, int fwd, unsigned long long flip_mask, const double* y0,
                 double* y_out, double* traj_out, int variant, int block, void* stream) {
    return hh_rk_fixed(8, n, dim, mu, t0, tf, steps, t_vals, fwd, flip_mask, y0, y_out,
                       traj_out, nullptr, variant, block, stream);
}

/* Resident workgroups of a kernel on the current device (CU count times the
 * occupancy per CU, capped at cap_per_cu when cap_per_cu > 0), queried once
 * per device and kernel: later calls -- e.g. inside a stream capture -- make
 * no HIP query.                                                            */
static int resident_grid(const void* kernel, int block, int cap_per_cu, int* grid_out) {
    struct Entry { int dev; const void* fn; int block; int cus; int per_cu; };
    static Entry table[64];
    static int used = 0;
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < used; ++i) {
        if (table[i].dev == dev && table[i].fn == kernel && table[i].block == block) {
            int per = table[i].per_cu > 0 ? table[i].per_cu : 1;
            if (cap_per_cu > 0 && per > cap_per_cu) per = cap_per_cu;
            *grid_out = table[i].cus * per;
            return HH_OK;
        }
    }
    int cus = 0, per_cu = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (used < 64) table[used++] = Entry{dev, kernel, block, cus, per_cu};
    int per = per_cu > 0 ? per_cu : 1;
    if (cap_per_cu > 0 && per > cap_per_cu) per = cap_per_cu;
    *grid_out = cus * per;
    return HH_OK;
}

long long hh_dop853_final_workspace(int n, int dim, int variant) {
    if (n <= 0 || dim != 42) return 0;
    if (variant == HH_VARIANT_PERSISTENT) return 256;
    if (variant == HH_VARIANT_PERSISTENT_PAIR) return pair_workspace(n);
    return 0;
}

int hh_dop853_final(int n, int dim, const double* mu, const double* t0, const double* tf,
                    double rtol, double atol, double max_step, double min_step,
                    long long max_steps, int fwd, unsigned long long flip_mask, const double* y0,
                    double* y_out, int* status, int* n_acc, int* n_rej, int variant, int block,
                    int grid, void* workspace, long long workspace_bytes, const int* order,
                    void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_dop853_final: n < 0");
    if (n == 0) return HH_OK;
    if (dim != 6 && dim != 42) return fail(HH_ERR_ARGS, "hh_dop853_final: dim must be 6 or 42");
    if (!mu || !t0 || !tf || !y0 || !y_out) return fail(HH_ERR_ARGS, "hh_dop853_final: null pointer");
    if (!(rtol >= 0.0) || !(atol > 0.0) || !(min_step > 0.0) || !(max_step >= min_step) || max_steps < 1)
        return fail(HH_ERR_ARGS, "hh_dop853_final: invalid tolerances/step bounds");
    StepParams sp = make_sp(rtol, atol, max_step, min_step, max_steps);
    hipStream_t st = (hipStream_t)stream;
    if (dim == 6) {
        if (variant != HH_VARIANT_PLAIN && variant != HH_VARIANT_WIDE)
            return fail(HH_ERR_ARGS, "hh_dop853_final: dim 6 supports the plain and wide variants");
        if (block <= 0) block = 128;
        if (block > 256 || block % 64) return fail(HH_ERR_ARGS, "hh_dop853_final: block must be 64..256, multiple of 64");
        if (variant == HH_VARIANT_WIDE)
            return launch_run6_group<false>(n, mu, t0, tf, sp, fwd, flip_mask, y0, y_out, 0,
                                            nullptr, status, n_acc, n_rej, block, st,
                                            "hh_dop853_final(6, wide) launch");
        hipLaunchKernelGGL(k_dop853_run6<false>, dim3(blocks_for(n, block)), dim3(block), 0, st, n,
                           mu, t0, tf, sp, fwd, flip_mask, y0, y_out, 0, (const double*)nullptr,
                           status, n_acc, n_rej);
        return check_launch("hh_dop853_final(6) launch");
    }
    Final42Args a;
    a.n = n; a.mu = mu; a.t0 = t0; a.tf = tf; a.sp = sp; a.fwd = fwd; a.flip = flip_mask;
    a.y0 = y0; a.yout = y_out; a.status = status; a.nacc = n_acc; a.nrej = n_rej;
    a.counter = (int*)workspace;
    a.order = order;
    a.m = 0;
    a.teval = nullptr;
    a.heavy = g_tuning.heavy_projected;
    a.heavy_age = g_tuning.min_age;
    a.pair_heavy = g_tuning.pair_projected;
    a.refill_min = g_tuning.refill_min;
    if (block <= 0) block = 256;
    if (block != 64 && block != 128 && block != 256 && block != 512)
        return fail(HH_ERR_ARGS, "hh_dop853_final: block must be 64, 128, 256 or 512");
    if (variant == HH_VARIANT_PLAIN) {
        const int g = blocks_for((long long)n * 8, block);
        switch (block) {
            case 64: hipLaunchKernelGGL((k_dop853_final42_plain<64, false>), dim3(g), dim3(64), 0, st, a); break;
            case 128: hipLaunchKernelGGL((k_dop853_final42_plain<128, false>), dim3(g), dim3(128), 0, st, a); break;
            case 256: hipLaunchKernelGGL((k_dop853_final42_plain<256, false>), dim3(g), dim3(256), 0, st, a); break;
            default: hipLaunchKernelGGL((k_dop853_final42_plain<512, false>), dim3(g), dim3(512), 0, st, a); break;
        }
        return check_launch("hh_dop853_final(42, plain) launch");
    }
    if (variant == HH_VARIANT_PAIR) {
        hipLaunchKernelGGL(k_dop853_final42_pair, dim3(n), dim3(128), 0, st, a);
        return check_launch("hh_dop853_final(42, pair) launch");
    }
    if (variant == HH_VARIANT_WIDE) {
        const int g = blocks_for((long long)n * 64, block);
        switch (block) {
            case 64: hipLaunchKernelGGL(k_dop853_final42_wide<64>, dim3(g), dim3(64), 0, st, a); break;
            case 128: hipLaunchKernelGGL(k_dop853_final42_wide<128>, dim3(g), dim3(128), 0, st, a); break;
            case 256: hipLaunchKernelGGL(k_dop853_final42_wide<256>, dim3(g), dim3(256), 0, st, a); break;
            default: hipLaunchKernelGGL(k_dop853_final42_wide<512>, dim3(g), dim3(512), 0, st, a); break;
        }
        return check_launch("hh_dop853_final(42, wide) launch");
    }
    if (variant != HH_VARIANT_PERSISTENT && variant != HH_VARIANT_PERSISTENT_PAIR)
        return fail(HH_ERR_ARGS, "hh_dop853_final: unknown variant");
    const long long need = hh_dop853_final_workspace(n, dim, variant);
    if (!workspace || workspace_bytes < need)
        return fail(HH_ERR_ARGS, "hh_dop853_final: persistent variants need a workspace of "
                                 "hh_dop853_final_workspace() bytes");
    if (variant == HH_VARIANT_PERSISTENT_PAIR) {
        ResumeArgs r;
        r.ctr = (int*)workspace;
        r.rq = r.ctr + RQ_CTR;
        r.cap = (int)resume_cap(n);
        r.spill = (double*)((char*)workspace + ws_align((long long)(RQ_CTR + r.cap) * 4));
        r.simd_tab = (int*)((char*)r.spill + ws_align((long long)SPILL_W * n * 8));
        r.exclusive = g_tuning.exclusive_projected;
        const int wpe = g_tuning.waves_per_simd;
        int g = grid;
        if (g <= 0) {
            /* 4 SIMDs per CU: wpe waves per SIMD are 2 * wpe two-wave workgroups */
            const int rc = resident_grid(
                wpe == 1 ? (const void*)k_dop853_final42_resume<128, 1>
                         : (const void*)k_dop853_final42_resume<128, 2>,
                128, 2 * wpe, &g);
            if (rc != HH_OK) return rc;
        }
        /* at most one workgroup per trajectory: each one may end up taking a
         * workgroup's pair of waves, and with fewer workgroups than heavy
         * trajectories those would take the pairs one after the other      */
        if (g > n) g = n;
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, r.ctr, 0, RQ_CTR);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, r.ctr + 3, n, 1);
        hipLaunchKernelGGL(k_fill_int, dim3((r.cap + 255) / 256), dim3(256), 0, st, r.rq, -1, r.cap);
        hipLaunchKernelGGL(k_fill_int, dim3(SIMD_TAB / 256), dim3(256), 0, st, r.simd_tab, 0, SIMD_TAB);
        if (wpe == 1)
            hipLaunchKernelGGL((k_dop853_final42_resume<128, 1>), dim3(g), dim3(128), 0, st, a, r);
        else
            hipLaunchKernelGGL((k_dop853_final42_resume<128, 2>), dim3(g), dim3(128), 0, st, a, r);
        return check_launch("hh_dop853_final(42, persistent pair) launch");
    }
    if (grid <= 0) {
        const void* fn = block == 64 ? (const void*)k_dop853_final42_persist<64>
                       : block == 128 ? (const void*)k_dop853_final42_persist<128>
                       : block == 256 ? (const void*)k_dop853_final42_persist<256>
                                      : (const void*)k_dop853_final42_persist<512>;
        const int rc = resident_grid(fn, block, 0, &grid);
        if (rc != HH_OK) return rc;
    }
    /* at most one wave per trajectory (each one may end up taking a wave) */
    const long long max_grid = ((long long)n * 64 + block - 1) / block;
    if (grid > max_grid) grid = (int)max_grid;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, a.counter, 0, 1);
    switch (block) {
        case 64: hipLaunchKernelGGL(k_dop853_final42_persist<64>, dim3(grid), dim3(64), 0, st, a); break;
        case 128: hipLaunchKernelGGL(k_dop853_final42_persist<128>, dim3(grid), dim3(128), 0, st, a); break;
        case 256: hipLaunchKernelGGL(k_dop853_final42_persist<256>, dim3(grid), dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL(k_dop853_final42_persist<512>, dim3(grid), dim3(512), 0, st, a); break;
    }
    return check_launch("hh_dop853_final(42, persistent) launch");
}

/* ---- hh_dop853_final without a caller workspace ------------------------- */

/* The workspace of the persistent forms, held by the library: the
 * persistent form's counter and the pair-tail form's counters, resume queue,
 * spill area and SIMD table for up to HH_STATIC_PAIR_MAX trajectories.
 * Each GPU the library runs on holds its own copy until the process ends.
 * In host code `scratch` names a host-side stand-in: the device address
 * comes from hipGetSymbolAddress (static_scratch below), once per device.  */
constexpr long long STATIC_SCRATCH_BYTES = pair_workspace(HH_STATIC_PAIR_MAX);
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

/* This device's address of `scratch` (looked up on the first call made with
 * the device current, then kept per device).                              */
static int static_scratch(void** out) {
    constexpr int MAX_DEVICES = 64;
    static void* ptr[MAX_DEVICES] = {};
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) return fail(HH_ERR_ARGS, "device index beyond 63");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(HH_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        ptr[dev] = p;
    }
    *out = ptr[dev];
    return HH_OK;
}

int hh_dop853_final_static(int n, int dim, const double* mu, const double* t0, const double* tf,
                           double rtol, double atol, double max_step, double min_step,
                           long long max_steps, int fwd, unsigned long long flip_mask,
                           const double* y0, double* y_out, int* status, int* n_acc, int* n_rej,
                           int variant, int block, int grid, const int* order, void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_dop853_final_static: n < 0");
    if (variant == HH_VARIANT_PERSISTENT_PAIR && n > HH_STATIC_PAIR_MAX)
        return fail(HH_ERR_ARGS, "hh_dop853_final_static: the persistent pair variant takes at "
                                 "most HH_STATIC_PAIR_MAX (16384) trajectories; use "
                                 "hh_dop853_final with a workspace for more");
    const long long need = hh_dop853_final_workspace(n, dim, variant);
    if (need > STATIC_SCRATCH_BYTES)
        return fail(HH_ERR_ARGS, "hh_dop853_final_static: workspace beyond the static scratch");
    void* ws = nullptr;
    if (need > 0) {
        const int rc = static_scratch(&ws);
        if (rc != HH_OK) return rc;
    }
    return hh_dop853_final(n, dim, mu, t0, tf, rtol, atol, max_step, min_step, max_steps, fwd,
                           flip_mask, y0, y_out, status, n_acc, n_rej, variant, block, grid, ws,
                           need > 0 ? STATIC_SCRATCH_BYTES : 0, order, stream);
}

int hh_dop853_grid(int n, int dim, const double* mu, const double* t0, const double* tf, int m,
                   const double* t_eval, double rtol, double atol, double max_step,
                   double min_step, long long max_steps, int fwd, unsigned long long flip_mask,
                   const double* y0, double* out, int* status, int* n_acc, int* n_rej,
                   int variant, int block, void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_dop853_grid: n < 0");
    if (n == 0) return HH_OK;
    if (dim != 6 && dim != 42) return fail(HH_ERR_ARGS, "hh_dop853_grid: dim must be 6 or 42");
    if (m < 2) return fail(HH_ERR_ARGS, "hh_dop853_grid: m must be >= 2");
    if (!mu || !t0 || !tf || !y0 || !out) return fail(HH_ERR_ARGS, "hh_dop853_grid: null pointer");
    if (!(rtol >= 0.0) || !(atol > 0.0) || !(min_step > 0.0) || !(max_step >= min_step) || max_steps < 1)
        return fail(HH_ERR_ARGS, "hh_dop853_grid: invalid tolerances/step bounds");
    StepParams sp = make_sp(rtol, atol, max_step, min_step, max_steps);
    hipStream_t st = (hipStream_t)stream;
    if (variant != HH_VARIANT_PLAIN && !(variant == HH_VARIANT_WIDE && dim == 6))
        return fail(HH_ERR_ARGS, "hh_dop853_grid: variant must be PLAIN, or WIDE for dim 6");
    if (dim == 6) {
        if (block <= 0) block = 128;
        if (block > 256 || block % 64) return fail(HH_ERR_ARGS, "hh_dop853_grid: block must be 64..256, multiple of 64");
        if (variant == HH_VARIANT_WIDE)
            return launch_run6_group<true>(n, mu, t0, tf, sp, fwd, flip_mask, y0, out, m, t_eval,
                                           status, n_acc, n_rej, block, st,
                                           "hh_dop853_grid(6, wide) launch");
        hipLaunchKernelGGL(k_dop853_run6<true>, dim3(blocks_for(n, block)), dim3(block), 0, st, n,
                           mu, t0, tf, sp, fwd, flip_mask, y0, out, m, t_eval, status, n_acc,
                           n_rej);
        return check_launch("hh_dop853_grid(6) launch");
    }
    Final42Args a;
    memset(&a, 0, sizeof a);
    a.n = n; a.mu = mu; a.t0 = t0; a.tf = tf; a.sp = sp; a.fwd = fwd; a.flip = flip_mask;
    a.y0 = y0; a.yout = out; a.status = status; a.nacc = n_acc; a.nrej = n_rej; a.m = m;
    a.teval = t_eval;
    if (block <= 0) block = 256;
    const int g = blocks_for((long long)n * 8, block);
    switch (block) {
        case 64: hipLaunchKernelGGL((k_dop853_final42_plain<64, true>), dim3(g), dim3(64), 0, st, a); break;
        case 128: hipLaunchKernelGGL((k_dop853_final42_plain<128, true>), dim3(g), dim3(128), 0, st, a); break;
        case 256: hipLaunchKernelGGL((k_dop853_final42_plain<256, true>), dim3(g), dim3(256), 0, st, a); break;
        case 512: hipLaunchKernelGGL((k_dop853_final42_plain<512, true>), dim3(g), dim3(512), 0, st, a); break;
        default: return fail(HH_ERR_ARGS, "hh_dop853_grid: block must be 64, 128, 256 or 512");
    }
    return check_launch("hh_dop853_grid(42) launch");
}

int hh_dop853_event(int n, const double* mu, const double* t0, const double* tmax, double rtol,
                    double atol, double max_step, double min_step, long long max_steps,
                    int ev_index, double ev_offset, int direction, double xtol, double gtol,
                    int fwd, unsigned long long flip_mask, const double* y0, int* hit,
                    double* t_hit, double* y_hit, double* y_last, int* status, int* n_acc,
                    int* n_rej, int variant, int block, void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_dop853_event: n < 0");
    if (n == 0) return HH_OK;
    if (!mu || !t0 || !tmax || !y0 || !hit || !t_hit || !y_hit)
        return fail(HH_ERR_ARGS, "hh_dop853_event: null pointer");
    if (ev_index < 0 || ev_index > 5) return fail(HH_ERR_ARGS, "hh_dop853_event: ev_index must be 0..5");
    if (direction < -1 || direction > 1) return fail(HH_ERR_ARGS, "hh_dop853_event: direction must be -1, 0 or 1");
    if (!(rtol >= 0.0) || !(atol > 0.0) || !(min_step > 0.0) || !(max_step >= min_step) || max_steps < 1)
        return fail(HH_ERR_ARGS, "hh_dop853_event: invalid tolerances/step bounds");
    if (block <= 0) block = 128;
    if (block > 256 || block % 64) return fail(HH_ERR_ARGS, "hh_dop853_event: block must be 64..256, multiple of 64");
    StepParams sp = make_sp(rtol, atol, max_step, min_step, max_steps);
    EventParams ev;
    ev.index = ev_index; ev.offset = ev_offset; ev.direction = direction; ev.xtol = xtol; ev.gtol = gtol;
    hipStream_t st = (hipStream_t)stream;
    if (variant == HH_VARIANT_WIDE) {
        const int g = blocks_for((long long)n * 8, block);
        switch (block) {
            case 64: hipLaunchKernelGGL(k_dop853_event6_group<64>, dim3(g), dim3(64), 0, st, n, mu, t0, tmax, sp, ev, fwd, flip_mask, y0, hit, t_hit, y_hit, y_last, status, n_acc, n_rej); break;
            case 128: hipLaunchKernelGGL(k_dop853_event6_group<128>, dim3(g), dim3(128), 0, st, n, mu, t0, tmax, sp, ev, fwd, flip_mask, y0, hit, t_hit, y_hit, y_last, status, n_acc, n_rej); break;
            case 192: hipLaunchKernelGGL(k_dop853_event6_group<192>, dim3(g), dim3(192), 0, st, n, mu, t0, tmax, sp, ev, fwd, flip_mask, y0, hit, t_hit, y_hit, y_last, status, n_acc, n_rej); break;
            default: hipLaunchKernelGGL(k_dop853_event6_group<256>, dim3(g), dim3(256), 0, st, n, mu, t0, tmax, sp, ev, fwd, flip_mask, y0, hit, t_hit, y_hit, y_last, status, n_acc, n_rej); break;
        }
        return check_launch("hh_dop853_event (group) launch");
    }
    if (variant != HH_VARIANT_PLAIN) return fail(HH_ERR_ARGS, "hh_dop853_event: variant must be PLAIN or WIDE");
    hipLaunchKernelGGL(k_dop853_event6, dim3(blocks_for(n, block)), dim3(block), 0, st, n, mu,
                       t0, tmax, sp, ev, fwd, flip_mask, y0, hit, t_hit, y_hit, y_last, status,
                       n_acc, n_rej);
    return check_launch("hh_dop853_event launch");
}

/* the windows and integration settings of _y_plane_crossing's search */
static int make_cross(CrossArgs& x, double window_start, double window_length,
                      double fallback_start, double fallback_length, int ev_index,
                      double ev_offset, int direction, double xtol, double gtol,
                      const char* who) {
    if (!(window_length >= 0.0) || !(fallback_length >= 0.0))
        return fail(HH_ERR_ARGS, who);
    x.ts[0] = window_start;
    x.ts[1] = fallback_start;
    x.span[0] = window_length;
    x.span[1] = fallback_length;
    for (int w = 0; w < 2; ++w) {
        /* numpy.isclose(0, v): alignment and search are skipped for ~0 */
        x.aligned[w] = !(fabs(0.0 - x.ts[w]) <= 1e-8 + 1e-5 * fabs(x.ts[w]));
        x.searched[w] = !(fabs(0.0 - x.span[w]) <= 1e-8 + 1e-5 * fabs(x.span[w]));
    }
    /* the reference's settings: DOP853 rtol = atol = 1e-12 everywhere;
     * max_step 1e4 for the alignment propagation (dynamics/base.py:447-450),
     * unbounded in the event search (singlehit/backend.py:202); min_step
     * 10 eps (integrators/rk.py:833-834)                                  */
    const double min_step = 10.0 * DBL_EPSILON;
    const long long max_steps = 1000000;
    x.sp_align = make_sp(1e-12, 1e-12, 1e4, min_step, max_steps);
    x.sp_event = make_sp(1e-12, 1e-12, INFINITY, min_step, max_steps);
    x.ev.index = ev_index;
    x.ev.offset = ev_offset;
    x.ev.direction = direction;
    x.ev.xtol = xtol;
    x.ev.gtol = gtol;
    x.flip_align = ~0ULL;                     /* backward alignment flips all six */
    return HH_OK;
}

int hh_correct(int n, const hh_correction_t* problems, hh_correction_result_t* results,
               double window_start, double window_length, double fallback_start,
               double fallback_length, void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_correct: n < 0");
    if (n == 0) return HH_OK;
    if (!problems || !results) return fail(HH_ERR_ARGS, "hh_correct: null pointer");
    CorrArgs a;
    a.n = n;
    a.prob = problems;
    a.res = results;
    /* the residual's search: y = 0, either direction, x/g tolerances 1e-12
     * (utils.py:13-123)                                                    */
    const int rc = make_cross(a.x, window_start, window_length, fallback_start, fallback_length,
                              1, 0.0, 0, 1e-12, 1e-12, "hh_correct: window lengths must be >= 0");
    if (rc != HH_OK) return rc;
    a.sp_stm = make_sp(1e-12, 1e-12, 1e4, 10.0 * DBL_EPSILON, 1000000);
    a.flip_stm = ((1ULL << 42) - 1) & ~((1ULL << 36) - 1);   /* the state block */
    hipLaunchKernelGGL(k_correct, dim3(n), dim3(192), 0, (hipStream_t)stream, a);
    return check_launch("hh_correct launch");
}

int hh_plane_crossing(int n, const double* mu, const double* y0, int forward, int ev_index,
                      double ev_offset, int direction, double xtol, double gtol,
                      double window_start, double window_length, double fallback_start,
                      double fallback_length, int* ok, double* t_hit, double* y_hit,
                      int* n_steps, void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_plane_crossing: n < 0");
    if (n == 0) return HH_OK;
    if (!mu || !y0 || !ok || !t_hit || !y_hit)
        return fail(HH_ERR_ARGS, "hh_plane_crossing: null pointer");
    if (ev_index < 0 || ev_index > 5)
        return fail(HH_ERR_ARGS, "hh_plane_crossing: ev_index must be 0..5");
    if (direction < -1 || direction > 1)
        return fail(HH_ERR_ARGS, "hh_plane_crossing: direction must be -1, 0 or 1");
    if (forward != 1 && forward != -1)
        return fail(HH_ERR_ARGS, "hh_plane_crossing: forward must be 1 or -1");
    CrossArgs x;
    const int rc = make_cross(x, window_start, window_length, fallback_start, fallback_length,
                              ev_index, ev_offset, direction, xtol, gtol,
                              "hh_plane_crossing: window lengths must be >= 0");
    if (rc != HH_OK) return rc;
    hipLaunchKernelGGL(k_plane_crossing<128>, dim3(blocks_for((long long)n * 8, 128)), dim3(128),
                       0, (hipStream_t)stream, n, mu, forward, y0, x, ok, t_hit, y_hit, n_steps);
    return check_launch("hh_plane_crossing launch");
}

int hh_correct_sizes(int* problem_bytes, int* result_bytes) {
    if (!problem_bytes || !result_bytes) return fail(HH_ERR_ARGS, "hh_correct_sizes: null pointer");
    *problem_bytes = (int)sizeof(hh_correction_t);
    *result_bytes = (int)sizeof(hh_correction_result_t);
    return HH_OK;
}

int hh_eval_solve2(int n, const double* J, const double* r, double* delta, double* cond,
                   void* stream) {
    if (n < 0 || (n > 0 && (!J || !r || !delta || !cond)))
        return fail(HH_ERR_ARGS, "hh_eval_solve2: invalid arguments");
    if (n == 0) return HH_OK;
    hipLaunchKernelGGL(k_eval_solve2, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       n, J, r, delta, cond);
    return check_launch("hh_eval_solve2 launch");
}

int hh_eval_rhs(int n, int dim, const double* mu, const double* y, double* out, void* stream) {
    if (n < 0 || (dim != 6 && dim != 42)) return fail(HH_ERR_ARGS, "hh_eval_rhs: bad n/dim");
    if (n == 0) return HH_OK;
    if (!mu || !y || !out) return fail(HH_ERR_ARGS, "hh_eval_rhs: null pointer");
    const long long threads = dim == 6 ? n : (long long)n * 8;
    hipLaunchKernelGGL(k_eval_rhs, dim3(blocks_for(threads, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, dim, mu, y, out);
    return check_launch("hh_eval_rhs launch");
}

int hh_eval_math(int n, const double* x, const double* y, double* out, void* stream) {
    if (n < 0) return fail(HH_ERR_ARGS, "hh_eval_math: n < 0");
    if (n == 0) return HH_OK;
    if (!x || !y || !out) return fail(HH_ERR_ARGS, "hh_eval_math: null pointer");
    hipLaunchKernelGGL(k_eval_math, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       n, x, y, out);
    return check_launch("hh_eval_math launch");
}

int hh_synodic_detect(int n, int m, const double* times, int per_traj_times,
                      const int* lengths, const double* states, const double* normal,
                      double offset, int direction, int plane_i, int plane_j, int cubic,
                      int segment_refine, double tol_on_surface, double dedup_time_tol,
                      double dedup_point_tol, int max_hits, int newton_max_iter, int cap,
                      int* count, double* hit_t, double* hit_x, double* hit_p, void* stream) {
    if (n < 0 || m < 0) return fail(HH_ERR_ARGS, "hh_synodic_detect: n < 0 or m < 0");
    if (n == 0) return HH_OK;
    if (!times || !states || !normal || !count)
        return fail(HH_ERR_ARGS, "hh_synodic_detect: null pointer");
    if (cap < 0 || (cap > 0 && (!hit_t || !hit_x || !hit_p)))
        return fail(HH_ERR_ARGS, "hh_synodic_detect: cap < 0 or null hit arrays");
    if (direction < -1 || direction > 1)
        return fail(HH_ERR_ARGS, "hh_synodic_detect: direction must be -1, 0 or 1");
    if (plane_i < 0 || plane_i > 5 || plane_j < 0 || plane_j > 5)
        return fail(HH_ERR_ARGS, "hh_synodic_detect: plane indices must be 0..5");
    if (segment_refine < 0 || newton_max_iter < 0 || max_hits < 0)
        return fail(HH_ERR_ARGS, "hh_synodic_detect: negative segment_refine, "
                                 "newton_max_iter or max_hits");
    SynArgs a;
    a.n = n;
    a.m = m;
    a.times = times;
    a.per_traj_times = per_traj_times ? 1 : 0;
    a.lengths = lengths;
    a.states = states;
    for (int d = 0; d < 6; ++d) a.nrm[d] = normal[d];
    a.offset = offset;
    a.direction = direction;
    a.pi = plane_i;
    a.pj = plane_j;
    a.cubic = cubic ? 1 : 0;
    a.refine = segment_refine;
    a.newton_max_iter = newton_max_iter;
    a.max_hits = max_hits;
    a.cap = cap;
    a.tol_on = tol_on_surface;
    a.dedup_t = dedup_time_tol;
    a.dedup_p = dedup_point_tol;
    a.count = count;
    a.hit_t = hit_t;
    a.hit_x = hit_x;
    a.hit_p = hit_p;
    hipLaunchKernelGGL(k_synodic_detect<256>, dim3(blocks_for((long long)n * 64, 256)), dim3(256),
                       0, (hipStream_t)stream, a);
    return check_launch("hh_synodic_detect launch");
}

static int syn_parts(int m) { return m >= 2 ? (m - 2) / SYN_PART + 1 : 1; }

long long hh_synodic_workspace(int n, int m) {
    if (n <= 0 || m < 0) return 0;
    const long long units = (long long)n * syn_parts(m);
    const long long cap_list = (units + SYN_LISTS - 1) / SYN_LISTS * SYN_CAP;
    return SYN_LISTS * 128 + ws_align(units * 4) + ws_align(SYN_LISTS * cap_list * 16) +
           units * SYN_CAP * SYN_REC * (long long)sizeof(double);
}

int hh_synodic_detect_ws(int n, int m, const double* times, int per_traj_times,
                         const int* lengths, const double* states, const double* normal,
                         double offset, int direction, int plane_i, int plane_j, int cubic,
                         int segment_refine, double tol_on_surface, double dedup_time_tol,
                         double dedup_point_tol, int max_hits, int newton_max_iter, int cap,
                         int* count, double* hit_t, double* hit_x, double* hit_p,
                         void* workspace, long long workspace_bytes, void* stream) {
    if (n < 0 || m < 0) return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: n < 0 or m < 0");
    if (n == 0) return HH_OK;
    if (!workspace || workspace_bytes < hh_synodic_workspace(n, m))
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: workspace missing or smaller than "
                                 "hh_synodic_workspace(n, m)");
    if ((reinterpret_cast<unsigned long long>(workspace) & 255) != 0)
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: workspace must be 256-byte aligned");
    /* the arguments are checked as hh_synodic_detect checks them */
    if (!times || !states || !normal || !count)
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: null pointer");
    if (cap < 0 || (cap > 0 && (!hit_t || !hit_x || !hit_p)))
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: cap < 0 or null hit arrays");
    if (direction < -1 || direction > 1)
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: direction must be -1, 0 or 1");
    if (plane_i < 0 || plane_i > 5 || plane_j < 0 || plane_j > 5)
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: plane indices must be 0..5");
    if (segment_refine < 0 || newton_max_iter < 0 || max_hits < 0)
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: negative segment_refine, "
                                 "newton_max_iter or max_hits");
    SynArgs a;
    a.n = n;
    a.m = m;
    a.times = times;
    a.per_traj_times = per_traj_times ? 1 : 0;
    a.lengths = lengths;
    a.states = states;
    for (int d = 0; d < 6; ++d) a.nrm[d] = normal[d];
    a.offset = offset;
    a.direction = direction;
    a.pi = plane_i;
    a.pj = plane_j;
    a.cubic = cubic ? 1 : 0;
    a.refine = segment_refine;
    a.newton_max_iter = newton_max_iter;
    a.max_hits = max_hits;
    a.cap = cap;
    a.tol_on = tol_on_surface;
    a.dedup_t = dedup_time_tol;
    a.dedup_p = dedup_point_tol;
    a.count = count;
    a.hit_t = hit_t;
    a.hit_x = hit_x;
    a.hit_p = hit_p;
    SynWs w;
    w.parts = syn_parts(m);
    const long long units = (long long)n * w.parts;
    if (units * SYN_CAP > 0x7fffffffLL)
        return fail(HH_ERR_ARGS, "hh_synodic_detect_ws: n * parts too large");
    const long long cap_list = (units + SYN_LISTS - 1) / SYN_LISTS * SYN_CAP;
    w.cap_list = (int)cap_list;
    char* base = static_cast<char*>(workspace);
    w.nlist = reinterpret_cast<int*>(base);
    w.ncand = reinterpret_cast<int*>(base + SYN_LISTS * 128);
    w.list = reinterpret_cast<int4*>(base + SYN_LISTS * 128 + ws_align(units * 4));
    w.rec = reinterpret_cast<double*>(base + SYN_LISTS * 128 + ws_align(units * 4) +
                                      ws_align(SYN_LISTS * cap_list * 16));
    hipStream_t st = (hipStream_t)stream;
    /* the list counters are cleared by a kernel, so the sequence can be
     * captured in a graph */
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(SYN_LISTS * 32), 0, st, w.nlist, 0,
                       SYN_LISTS * 32);
    hipLaunchKernelGGL(k_synodic_scan<256>, dim3(blocks_for(units * 64, 256)), dim3(256), 0, st,
                       a, w);
    hipLaunchKernelGGL(k_synodic_refine<256>,
                       dim3(min(blocks_for(units * SYN_CAP, 256), 256)), dim3(256), 0, st,
                       a, w);
    hipLaunchKernelGGL(k_synodic_merge<256>, dim3(blocks_for((long long)n * 64, 256)), dim3(256),
                       0, st, a, w);
    return check_launch("hh_synodic_detect_ws launch");
}

int hh_traj_filters(int n, int m, const double* mu, const double* states, double* min_r1,
                    double* min_r2, double* max_energy_err, void* stream) {
    if (n < 0 || m < 1) return fail(HH_ERR_ARGS, "hh_traj_filters: n < 0 or m < 1");
    if (n == 0) return HH_OK;
    if (!mu || !states) return fail(HH_ERR_ARGS, "hh_traj_filters: null pointer");
    const long long threads = (long long)n * 64;
    const int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(k_traj_filters, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n, m, mu,
                       states, min_r1, min_r2, max_energy_err);
    return check_launch("hh_traj_filters launch");
}

/* ---- stream-ordered copies and events for host loops ---------------------- */

int hh_memcpy_async(void* dst, const void* src, unsigned long long bytes, int to_device,
                    void* stream) {
    if (bytes == 0) return HH_OK;
    if (!dst || !src) return fail(HH_ERR_ARGS, "hh_memcpy_async: null pointer");
    const hipError_t e = hipMemcpyAsync(dst, src, (size_t)bytes,
                                        to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                                        (hipStream_t)stream);
    return e == hipSuccess ? HH_OK : hip_fail(e, "hipMemcpyAsync");
}

/* A stream with a hardware queue of its own: a stream created with a CU
 * mask (here every CU) is not multiplexed onto the runtime's shared queues,
 * so a long-running launch on it holds up no other stream's work.          */
void* hh_stream_create_dedicated(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        return nullptr;
    const int words = (cus + 31) / 32;
    uint32_t mask[64];
    if (words > 64) return nullptr;
    for (int w = 0; w < words; ++w) mask[w] = 0xFFFFFFFFu;
    hipStream_t s = nullptr;
    if (hipExtStreamCreateWithCUMask(&s, (uint32_t)words, mask) != hipSuccess) return nullptr;
    return (void*)s;
}

int hh_stream_destroy(void* stream) {
    if (!stream) return fail(HH_ERR_ARGS, "hh_stream_destroy: null stream");
    const hipError_t e = hipStreamDestroy((hipStream_t)stream);
    return e == hipSuccess ? HH_OK : hip_fail(e, "hipStreamDestroy");
}

/* Pinned, device-mapped, coherent host memory: device writes reach it
 * without a cache flush, so a host loop can poll what a running kernel
 * writes there.                                                            */
int hh_host_alloc(unsigned long long bytes, void** ptr) {
    if (!ptr || bytes == 0) return fail(HH_ERR_ARGS, "hh_host_alloc: invalid arguments");
    *ptr = nullptr;
    const hipError_t e = hipHostMalloc(ptr, (size_t)bytes,
                                       hipHostMallocMapped | hipHostMallocCoherent);
    return e == hipSuccess ? HH_OK : hip_fail(e, "hipHostMalloc");
}

int hh_host_free(void* ptr) {
    if (!ptr) return HH_OK;
    const hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? HH_OK : hip_fail(e, "hipHostFree");
}

void* hh_event_create(void) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return nullptr;
    return (void*)ev;
}

void hh_event_destroy(void* ev) {
    if (ev) (void)hipEventDestroy((hipEvent_t)ev);
}

int hh_event_record(void* ev, void* stream) {
    const hipError_t e = hipEventRecord((hipEvent_t)ev, (hipStream_t)stream);
    return e == hipSuccess ? HH_OK : hip_fail(e, "hipEventRecord");
}

int hh_event_query(void* ev) {
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    hip_fail(e, "hipEventQuery");
    return -HH_ERR_HIP;
}

}  /* extern "C" */

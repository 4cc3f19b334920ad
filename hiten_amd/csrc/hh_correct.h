/* hh_correct.h -- the device-resident single-shooting corrector.
 *
 * One workgroup of three waves runs one correction's whole Newton-Armijo
 * loop (the row functions of hh_newton.h: corrector/backends/newton.py:
 * 39-143, stepping/armijo.py:58-164) without returning to the host:
 *
 *   residual   _y_plane_crossing, or _z_plane_crossing / _x_plane_crossing
 *              as the problem's event plane asks (poincare/singlehit/
 *              backend.py:237-280, 315-371): the first window, then the fallback window (alignment
 *              propagation + event search), in the Group6 layout -- the
 *              sixteen 8-lane groups of waves 0 and 1 evaluate the residual,
 *              or all line-search trials of a step at once (in chunks of
 *              sixteen); as soon as the trials finished so far fix the
 *              line search's outcome (a trial accepted with all trials
 *              before it known), the groups still on later trials stop;
 *   Jacobian   Phi(t_event) from [I | x] (corrector/operators.py:298-307,
 *              437-452) in the Wave42 layout on wave 2 -- for the round's
 *              state 0 speculatively, while the groups search for its
 *              crossing (corr_spec_stm) -- then the Halo term
 *              (services/orbits.py:917-947);
 *   step       the 2x2 dense solve with its regularisation
 *              (corrector/backends/base.py:187-245), as NumPy's LAPACK
 *              forms it (corr_solve2), capped and spread over the trials.
 *
 * The integrations are the library's kernels' loops (group6_run,
 * group6_event, wide_loop), so each evaluated state's bits are those of
 * the batched path.  Two differences from the host-driven bank remain: the
 * Halo term's powers are correctly rounded here (the C library's pow is
 * within 0.52 ulp), and an exactly singular system takes a closed-form 2x2
 * minimum-norm solution where NumPy calls LAPACK's dgelsd.  Either can move
 * the iterates by an ulp, so results agree with the bank to the
 * corrector's tolerance, not bit for bit.
 *
 * Included by hh_kernels.hip after the integrator loops it uses.
 */
#ifndef HH_CORRECT_H
#define HH_CORRECT_H

#include <float.h>

#include "hh_newton.h"

constexpr int CORR_GROUPS = 16;   /* 8-lane groups of waves 0 and 1 */

/* The search of _y_plane_crossing without a time guess: two windows
 * (start, length), each an optional alignment propagation to its start and
 * an event search over its length.                                        */
struct CrossArgs {
    double ts[2], span[2];        /* the two search windows (start, length) */
    int aligned[2], searched[2];  /* not numpy.isclose(0, start / length)    */
    StepParams sp_align, sp_event;
    EventParams ev;
    unsigned long long flip_align;
};

struct CorrArgs {
    int n;
    const hh_correction_t* prob;
    hh_correction_result_t* res;
    CrossArgs x;
    StepParams sp_stm;
    unsigned long long flip_stm;
};

struct CorrLds {
    hh::NewtonRow row;
    double t[hh::NEWTON_JMAX];
    double xe[hh::NEWTON_JMAX][6];
    int ok[hh::NEWTON_JMAX];
    double phi[42];
    double spec_phi[42];  /* Phi of the round's state 0, integrated speculatively  */
    double spec_tf;       /* its event time, once known                         */
    int spec_known;       /* set when state 0's residual is in                  */
    int spec_ok;          /* state 0 crossed the section                        */
    int spec_valid;       /* spec_phi holds Phi(spec_tf) of state 0             */
    int use_spec;         /* this round's STM is state 0's: take spec_phi       */
    int n_eval;   /* states to evaluate this round: 1 (residual) or nal (trials) */
    int decided;  /* the line search's outcome is fixed: the other trials stop   */
    int known[hh::NEWTON_JMAX];   /* trial j evaluated (its results are in)     */
    int steps;    /* accepted integration steps so far                           */
    int more;     /* 1 while the line search is not yet decided                  */
    int fail;     /* the problem is not one this kernel takes                    */
    int ev_index; /* coordinate of the section plane: 0 x, 1 y, 2 z              */
};

/* The Halo term's powers, correctly rounded: squares of coordinates of
 * either sign, and s^1.5 of a sum of squares (hh_pow takes positive bases). */
struct DevPow {
    __device__ double operator()(double a, double b) const {
        if (b == 2.0) return a * a;
        if (b == 1.5) return hh_pow15(a);
        return hh_pow(a, b);
    }
};

/* numpy.linalg.cond(J) (2-norm) of a 2x2 matrix: sigma_max^2 / |det|, with
 * sigma_max = (|(a + d, c - b)| + |(a - d, c + b)|) / 2 and the determinant
 * formed with one rounding error (fma), so the ratio is accurate to a few
 * ulps at any condition (NumPy's, from dgesdd, to about eps * cond: near the
 * 1e8 threshold the two differ by ~2e-8 relative); NaN for NaN entries, inf
 * for a singular matrix.                                                  */
__device__ __forceinline__ double corr_cond2(const double (&J)[4]) {
    const double a = J[0], b = J[1], c = J[2], d = J[3];
    if (a != a || b != b || c != c || d != d) return NAN;
    const double p = a + d, q = c - b, u = a - d, v = c + b;
    const double smax = 0.5 * (__builtin_sqrt(p * p + q * q) + __builtin_sqrt(u * u + v * v));
    const double w = b * c;
    const double det = __builtin_fma(a, d, -w) + __builtin_fma(-b, c, w);
    if (smax == 0.0 || det == 0.0) return INFINITY;
    return smax / (__builtin_fabs(det) / smax);
}

/* numpy.linalg.solve(A, b) for 2x2: LAPACK dgesv as OpenBLAS forms it --
 * getf2 (partial pivoting on the first largest magnitude, the multiplier
 * scaled by the pivot's reciprocal, the Schur complement a11 - l a01), then
 * the unit-lower and upper triangular solves (fused multiply-adds in the
 * axpy updates, divisions by the diagonal).  Pinned against NumPy on random
 * systems.  Returns false for an exactly singular U (LinAlgError).        */
__device__ __forceinline__ bool corr_lu2(const double (&A)[4], const double (&b)[2],
                                         double (&x)[2]) {
    double a00 = A[0], a01 = A[1], a10 = A[2], a11 = A[3], b0 = b[0], b1 = b[1];
    if (__builtin_fabs(a10) > __builtin_fabs(a00)) {
        double s = a00; a00 = a10; a10 = s;
        s = a01; a01 = a11; a11 = s;
        s = b0; b0 = b1; b1 = s;
    }
    if (a00 == 0.0) return false;
    const double l = (__builtin_fabs(a00) >= DBL_MIN) ? a10 * (1.0 / a00) : a10 / a00;
    const double u11 = a11 - l * a01;
    if (u11 == 0.0) return false;
    const double y1 = __builtin_fma(-l, b0, b1);
    const double x1 = y1 / u11;
    x[1] = x1;
    x[0] = __builtin_fma(-a01, x1, b0) / a00;
    return true;
}

/* numpy.linalg.lstsq(A, b, rcond=None) for 2x2: the minimum-norm solution
 * over the singular values above eps * 2 * sigma_max, from a closed-form SVD
 * (one Jacobi rotation of A^T A).  LAPACK's dgelsd forms it differently:
 * agreement to rounding, not bits.                                         */
__device__ __forceinline__ void corr_lstsq2(const double (&A)[4], const double (&b)[2],
                                            double (&x)[2]) {
    const double a = A[0], bb = A[1], c = A[2], d = A[3];
    const double p = a * a + c * c, q = a * bb + c * d, r = bb * bb + d * d;
    const double th = 0.5 * atan2(2.0 * q, p - r);
    const double cs = cos(th), sn = sin(th);
    const double v[2][2] = {{cs, sn}, {-sn, cs}};
    double s[2], av[2][2];
    for (int i = 0; i < 2; ++i) {
        av[i][0] = a * v[i][0] + bb * v[i][1];
        av[i][1] = c * v[i][0] + d * v[i][1];
        s[i] = __builtin_sqrt(av[i][0] * av[i][0] + av[i][1] * av[i][1]);
    }
    const double smax = s[0] > s[1] ? s[0] : s[1];
    const double tol = (2.0 * DBL_EPSILON) * smax;
    x[0] = 0.0;
    x[1] = 0.0;
    for (int i = 0; i < 2; ++i) {
        if (!(s[i] > tol)) continue;
        const double g = (av[i][0] * b[0] + av[i][1] * b[1]) / (s[i] * s[i]);
        x[0] += v[i][0] * g;
        x[1] += v[i][1] * g;
    }
}

/* _solve_delta_dense (backends/base.py:187-245) for a square 2x2 system:
 * Tikhonov 1e-12 on the diagonal when cond(J) is NaN or above 1e8, then
 * solve, or least squares when the regularised system is singular.       */
__device__ __forceinline__ void corr_solve2(const double (&J)[4], const double (&r)[2],
                                            double (&d)[2]) {
    const double cj = corr_cond2(J);
    const bool reg = (cj != cj) || cj > 1e8;
    /* J + numpy.eye(2) * 1e-12: the off-diagonal entries gain +0.0 */
    const double A[4] = {reg ? J[0] + 1e-12 : J[0], reg ? J[1] + 0.0 : J[1],
                         reg ? J[2] + 0.0 : J[2], reg ? J[3] + 1e-12 : J[3]};
    const double b[2] = {-r[0], -r[1]};
    if (!corr_lu2(A, b, d)) corr_lstsq2(A, b, d);
}

/* corr_solve2 and corr_cond2 per thread, for tests against NumPy */
__global__ void k_eval_solve2(int n, const double* __restrict__ J, const double* __restrict__ r,
                              double* __restrict__ delta, double* __restrict__ cond) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Ji[4] = {J[4 * i], J[4 * i + 1], J[4 * i + 2], J[4 * i + 3]};
    const double ri[2] = {r[2 * i], r[2 * i + 1]};
    double d[2];
    corr_solve2(Ji, ri, d);
    delta[2 * i] = d[0];
    delta[2 * i + 1] = d[1];
    cond[i] = corr_cond2(Ji);
}

/* _y_plane_crossing (or the crossing of plane ev_index) of one state in
 * this lane's 8-lane group (y: this lane's component): the first window,
 * then the fallback one, each an optional alignment propagation to the
 * window's start and an event search over its length (singlehit/
 * backend.py:202-280).  ok, t and (this lane's component of) x_event as the
 * batched path forms them.                                                 */
template <bool CANCEL>
__device__ __forceinline__ bool corr_residual(const Group6& ga, const Group6& ge,
                                              const CrossArgs& a, int ev_index, double y,
                                              int& ok, double& t, double& xe, int& steps,
                                              const int* cancel) {
    EventParams ev = a.ev;
    ev.index = ev_index;
    ok = 0;
    t = NAN;
    xe = NAN;
    steps = 0;
    for (int w = 0; w < 2; ++w) {
        double ya = y;
        if (a.aligned[w]) {
            int st, na, nr;
            group6_run<false, CANCEL>(ga, y, 0.0, a.ts[w], a.sp_align, nullptr, 0, nullptr, ya,
                                      st, na, nr, cancel);
            steps += na;
            if (st == ST_CANCELLED) return false;
        }
        if (!a.searched[w]) continue;
        EventOut6 o;
        group6_event<CANCEL>(ge, ya, 0.0, a.span[w], a.sp_event, ev, o, cancel);
        steps += o.nacc;
        if (o.status == ST_CANCELLED) return false;
        const double t_rel = o.found ? o.t_hit : a.span[w];
        if (o.found && t_rel < a.span[w] && t_rel >= 0.0) {
            ok = 1;
            t = a.ts[w] + t_rel;
            xe = o.y_hit;
            return true;
        }
    }
    return true;
}

/* After trial j is in: whether the trials known so far, in order, fix the
 * line search (armijo.py:111-164 is sequential: the first accepted trial
 * ends it, and trials after it are never looked at); if so, set `decided`
 * so that the groups still integrating later trials stop.  Each group
 * publishes its trial before scanning, so of two groups that finish
 * together at least one sees both; if neither did, the round would only
 * end later, when every group has finished -- the decision itself is taken
 * by thread 0 afterwards from the results, never from this flag.        */
__device__ __forceinline__ void corr_try_decide(CorrLds& L, int n_eval) {
    const hh::NewtonRow& w = L.row;
    for (int j = 0; j < n_eval; ++j) {
        if (!__hip_atomic_load(&L.known[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
        if (!L.ok[j]) continue;
        double r[2];
        for (int q = 0; q < 2; ++q) r[q] = L.xe[j][w.res[q]] - w.tgt[q];
        if (hh::newton_inf_norm(r, 2) <= (1.0 - w.c * w.alpha[j]) * w.rn) break;
    }
    int expect = 0;
    __hip_atomic_compare_exchange_strong(&L.decided, &expect, 1, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

/* Phi(t_event) of the round's state 0 in the Wave42 layout, begun before
 * t_event is known.  DOP853's steps do not depend on the end time until a
 * step would pass it (integrators/utils.py:205-207 clips that step; the
 * initial step, utils.py:149-157, ignores it), so the integration runs
 * unclipped until state 0's event search publishes t_event, then returns
 * to the last state from which no attempt reached t_event and finishes as
 * the sequential integration does from there: same attempts, same bits.
 * It keeps the starting state and the last two accepted states with the
 * furthest time their attempts reached; if an older state's attempts
 * reached t_event it starts over.  Returns 1 with this lane's component of
 * the final state in `out`, 0 when state 0 has no crossing (no STM is
 * needed), -1 when it gives up (a non-finite error norm before t_event is
 * known, or the attempt cap): the caller then integrates as usual.       */
__device__ __forceinline__ int corr_spec_stm(const Wave42& W, const StepParams& sp, double y0,
                                             CorrLds& L, double& out, int& nacc_out) {
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    k[0] = W.rhs(y0);
    const double k00 = k[0];
    const double sc = sp.atol + sp.rtol * __builtin_fabs(y0);
    const double ua = y0 / sc;
    const double ub = k[0] / sc;
    const double s0 = W.block_tree(W.act ? ua * ua : 0.0);
    const double s1 = W.block_tree(W.act ? ub * ub : 0.0);
    const double sn = __builtin_sqrt(42.0);
    const double h0 = select_initial_step(__builtin_sqrt(s0) / sn, __builtin_sqrt(s1) / sn,
                                          sp.min_step, sp.max_step);
    /* current state */
    double y = y0, t = 0.0, h = h0, ep = -1.0;
    int na = 0, nr = 0;
    /* B: the last accepted state (the start at first), A: the one before;
     * rB, rA: furthest t + h their attempts reached; older: the same over
     * every state before A                                                */
    double yB = y0, kB = k00, tB = 0.0, hB = h0, epB = -1.0, rB = -INFINITY;
    int naB = 0, nrB = 0;
    double yA = y0, kA = k00, tA = 0.0, hA = h0, epA = -1.0, rA = -INFINITY;
    int naA = 0, nrA = 0;
    double older = -INFINITY;
    const long long cap = sp.max_steps;
    for (;;) {
        if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(
                &L.spec_known, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)))
            break;
        if (na + nr >= cap) return -1;
        const double hh = clamp_step(h, sp.max_step, sp.min_step);
        rB = (t + hh > rB) ? t + hh : rB;
        double yh;
        const double err = W.attempt(y, k, hh, sp.rtol, sp.atol, yh);
        if (err != err) return -1;
        double fa, fr;
        W.step_factors(err, ep, fa, fr);
        if (err <= 1.0) {
            older = (rA > older) ? rA : older;
            yA = yB; kA = kB; tA = tB; hA = hB; epA = epB; naA = naB; nrA = nrB; rA = rB;
            t = t + hh;
            y = yh;
            k[0] = k[DOP_S];
            h = hh * fa;
            ep = err;
            na += 1;
            yB = y; kB = k[0]; tB = t; hB = h; epB = ep; naB = na; nrB = nr; rB = -INFINITY;
        } else {
            h = clamp_step(hh * fr, sp.max_step, sp.min_step);
            nr += 1;
        }
    }
    if (!L.spec_ok) return 0;
    const double tf = L.spec_tf;
    if (isclose(0.0, tf) || !((0.0 - tf) * 1.0 < 0.0)) {
        /* zero or reversed span: y0 unchanged (dynamics/base.py:421-424) */
        out = y0;
        nacc_out = 0;
        return 1;
    }
    if (older >= tf) {
        y = y0; k[0] = k00; t = 0.0; h = h0; ep = -1.0; na = 0; nr = 0;
    } else if (rA >= tf) {
        y = yA; k[0] = kA; t = tA; h = hA; ep = epA; na = naA; nr = nrA;
    } else if (rB >= tf) {
        y = yB; k[0] = kB; t = tB; h = hB; ep = epB; na = naB; nr = nrB;
    }
    int status;
    wide_loop(sp, W, y, k, t, tf, h, ep, na, nr, out, status);
    nacc_out = na;
    return 1;
}

/* Phi(tf) from y0 in the Wave42 layout, the end time known from the start
 * (k_dop853_final42_wide's loop).                                         */
__device__ __forceinline__ void corr_stm(const Wave42& W, const StepParams& sp, double y0,
                                         double tf, double& out, int& nacc) {
    nacc = 0;
    if (isclose(0.0, tf) || !((0.0 - tf) * 1.0 < 0.0)) {
        out = y0;
        return;
    }
    double k[DOP_K];
#pragma unroll
    for (int j = 0; j < DOP_K; ++j) k[j] = 0.0;
    k[0] = W.rhs(y0);
    const double sc = sp.atol + sp.rtol * __builtin_fabs(y0);
    const double ua = y0 / sc;
    const double ub = k[0] / sc;
    const double s0 = W.block_tree(W.act ? ua * ua : 0.0);
    const double s1 = W.block_tree(W.act ? ub * ub : 0.0);
    const double sn = __builtin_sqrt(42.0);
    const double h = select_initial_step(__builtin_sqrt(s0) / sn, __builtin_sqrt(s1) / sn,
                                         sp.min_step, sp.max_step);
    int nr = 0, status;
    wide_loop(sp, W, y0, k, 0.0, tf, h, -1.0, nacc, nr, out, status);
}

/* This lane's component of [I | state]: the identity block of Phi, then
 * the full 6-d state of the controls q.                                   */
__device__ __forceinline__ double corr_stm_start(const Wave42& W, const hh::NewtonRow& w,
                                                 const double* q) {
    if (!W.holds()) return 0.0;
    const int c = W.comp();
    if (c < 36) return (c / 6 == c % 6) ? 1.0 : 0.0;
    double v = w.base[c - 36];
    for (int j = 0; j < 2; ++j)
        if (w.ctrl[j] == c - 36) v = q[j];
    return v;
}

/* One correction per 192-thread workgroup: waves 0 and 1 are the sixteen
 * Group6 groups that evaluate residuals and line-search trials, wave 2
 * integrates the STMs in the Wave42 layout -- in each round speculatively,
 * for the round's state 0 (the residual's state, or the alpha = 1 trial,
 * which is the next iterate in most steps) while the groups search for
 * the crossings.  Thread 0 keeps the row in LDS and takes the decisions
 * between barriers; every branch the waves take on LDS values is read after
 * a barrier, so all threads meet every barrier.  Each loop ends: every
 * integration is bounded by max_steps and the Newton loop by max_attempts,
 * and the STM wave's speculation ends when state 0's search publishes its
 * event time, which the first group always does.                         */
__global__ __launch_bounds__(192) void k_correct(CorrArgs a) {
    __shared__ CorrLds L;
    __shared__ PowTab powt;
    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int grp = tid >> 3;
    const int wid = __builtin_amdgcn_readfirstlane((int)(tid >> 6));
    if (i >= a.n) return;
    pow_tab_fill(&powt);
    if (tid == 0) {
        const hh_correction_t& pb = a.prob[i];
        L.fail = 1;
        /* the section's coordinate: y (default), x or z */
        L.ev_index = pb.event_plane == HH_PLANE_X ? 0 : (pb.event_plane == HH_PLANE_Z ? 2 : 1);
        const bool plane_ok = pb.event_plane == HH_PLANE_Y || pb.event_plane == HH_PLANE_X ||
                              pb.event_plane == HH_PLANE_Z;
        if (pb.n_residuals == 2 && pb.n_controls == 2 && plane_ok)
            L.fail = hh::newton_init(L.row, 2, 2, pb.mu, pb.x0, pb.residual_indices,
                                     pb.control_indices, pb.target, pb.halo, pb.forward,
                                     pb.max_attempts, pb.tol, pb.max_delta, pb.alpha_reduction,
                                     pb.min_alpha, pb.armijo_c) != 0;
        L.n_eval = 1;
        L.steps = 0;
        L.decided = 0;
        L.known[0] = 0;
        L.spec_known = 0;
        L.spec_valid = 0;
    }
    __syncthreads();
    if (!L.fail) {
        const double mu = L.row.mu;
        Group6 ga, ge;
        ga.setup(lane, mu, L.row.fwd, a.x.flip_align, &powt);
        ge.setup(lane, mu, 1, 0ULL, &powt);
        Wave42 W;
        W.setup(lane, mu, 1, a.flip_stm, &powt);
        for (;;) {
            /* the residual, or the line-search trials of the step; the STM
             * wave meanwhile integrates state 0's STM                      */
            const int n_eval = L.n_eval;
            const bool trials = L.row.stage == hh::NS_ARMIJO;
            for (int j0 = 0; j0 < n_eval; j0 += CORR_GROUPS) {
                const int j = j0 + grp;
                if (wid < 2 && j < n_eval) {
                    const double* q = trials ? L.row.trial[j] : L.row.x;
                    double y = 0.0;
                    if (ge.act) {
                        y = L.row.base[ge.d];
                        for (int c = 0; c < 2; ++c)
                            if (L.row.ctrl[c] == ge.d) y = q[c];
                    }
                    int ok, steps;
                    double t, xe;
                    const bool in = corr_residual<true>(ga, ge, a.x, L.ev_index, y, ok, t,
                                                        xe, steps, &L.decided);
                    if (in && ge.act) L.xe[j][ge.d] = xe;
                    if (ge.d == 0) {
                        atomicAdd(&L.steps, steps);
                        if (in) {
                            L.t[j] = t;
                            L.ok[j] = ok;
                        }
                        if (j == 0) {
                            L.spec_tf = t;
                            L.spec_ok = in && ok;
                        }
                    }
                    /* this group's results, then its flags, then the scan */
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    if (ge.d == 0 && j == 0)
                        __hip_atomic_store(&L.spec_known, 1, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (in && ge.d == 0 && trials) {
                        __hip_atomic_store(&L.known[j], 1, __ATOMIC_SEQ_CST,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                        corr_try_decide(L, n_eval);
                    }
                } else if (wid == 2 && j0 == 0) {
                    const double y0 = corr_stm_start(W, L.row, trials ? L.row.trial[0] : L.row.x);
                    double out = 0.0;
                    int na = 0;
                    const int r = corr_spec_stm(W, a.sp_stm, y0, L, out, na);
                    if (r == 1 && W.holds()) L.spec_phi[W.comp()] = out;
                    if (lane == 0) {
                        L.spec_valid = (r == 1);
                        L.steps += na;
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    const int done_n = (j0 + CORR_GROUPS < n_eval) ? j0 + CORR_GROUPS : n_eval;
                    L.more = (trials && !hh::newton_armijo_decided(L.row, 2, done_n, &L.xe[0][0],
                                                                   L.ok)) ? 1 : 0;
                }
                __syncthreads();
                if (!L.more) break;
            }
            if (tid == 0) {
                int taken = 0;
                if (L.row.stage == hh::NS_RESID)
                    hh::newton_residual(L.row, 2, L.t[0], L.xe[0], L.ok[0]);
                else
                    taken = hh::newton_armijo(L.row, 2, 2, L.t, &L.xe[0][0], L.ok);
                /* the STM this step needs is state 0's when state 0 is the
                 * next iterate: its speculative integration ended at the
                 * same event time                                            */
                L.use_spec = (L.row.stage == hh::NS_JAC) && taken == 0 && L.spec_valid;
            }
            __syncthreads();
            if (L.row.stage == hh::NS_DONE) break;
            /* Phi(t_event) from [I | x] */
            if (wid == 2) {
                if (L.use_spec) {
                    if (W.holds()) L.phi[W.comp()] = L.spec_phi[W.comp()];
                } else {
                    double out;
                    int na;
                    corr_stm(W, a.sp_stm, corr_stm_start(W, L.row, L.row.x), L.row.t_ev, out, na);
                    if (W.holds()) L.phi[W.comp()] = out;
                    if (lane == 0) L.steps += na;
                }
            }
            __syncthreads();
            if (tid == 0) {
                double J[4], r[2], d[2];
                hh::newton_jacobian(L.row, 2, 2, L.phi, J, r, DevPow());
                corr_solve2(J, r, d);
                hh::newton_set_step(L.row, 2, d);
                L.n_eval = L.row.nal;
                L.decided = 0;
                for (int j = 0; j < L.row.nal; ++j) L.known[j] = 0;
                L.spec_known = 0;
                L.spec_valid = 0;
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        hh_correction_result_t& o = a.res[i];
        if (L.fail) {
            for (int q = 0; q < 6; ++q) o.x[q] = a.prob[i].x0[q];
            o.residual_norm = NAN;
            o.half_period = NAN;
            o.iterations = 0;
            o.n_event_searches = 0;
            o.n_stm = 0;
            o.converged = 0;
            o.error = HH_CORRECT_INVALID;
            o.n_steps = 0;
        } else {
            hh::newton_full_state(L.row, L.row.x, 2, o.x);
            o.residual_norm = L.row.rn;
            o.half_period = L.row.half_period;
            o.iterations = L.row.k;
            o.n_event_searches = L.row.n_ev;
            o.n_stm = L.row.n_stm;
            o.converged = L.row.ok;
            o.error = L.row.err;
            o.n_steps = L.steps;
        }
        /* the result before its flag, for a host that polls mapped memory */
        __threadfence_system();
        __hip_atomic_store(&o.done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}


/* _y_plane_crossing for a batch of states in one launch (the windows of
 * poincare/singlehit/backend.py:237-280 without a time guess): one state
 * per 8-lane group, the fallback window taken in the same group.  Outputs
 * as section_crossing_batch builds them: ok, t = start + t_rel, and the
 * crossing state ([6][n], structure of arrays).                           */
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_plane_crossing(int n, const double* __restrict__ mu,
                                                          int fwd, const double* __restrict__ y0,
                                                          CrossArgs x, int* __restrict__ ok,
                                                          double* __restrict__ t_hit,
                                                          double* __restrict__ y_hit,
                                                          int* __restrict__ n_steps) {
    __shared__ PowTab powt;
    pow_tab_fill(&powt);
    __syncthreads();
    const long long gtid = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const int i = (int)(gtid >> 3);
    if (i >= n) return;
    const int lane = threadIdx.x & 63;
    Group6 ga, ge;
    ga.setup(lane, mu[i], fwd, x.flip_align, &powt);
    ge.setup(lane, mu[i], 1, 0ULL, &powt);
    const int d = ge.d;
    const double y = ge.act ? y0[(size_t)d * n + i] : 0.0;
    int okv, steps;
    double tv, xv;
    corr_residual<false>(ga, ge, x, x.ev.index, y, okv, tv, xv, steps, nullptr);
    if (ge.act) y_hit[(size_t)d * n + i] = xv;
    if (d == 0) {
        ok[i] = okv;
        t_hit[i] = tv;
        if (n_steps) n_steps[i] = steps;
    }
}

#endif

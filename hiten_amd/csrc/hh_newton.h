/* hh_newton.h -- one correction of the single-shooting Newton loop with its
 * Armijo line search, as functions of one row: the reference's
 * _NewtonBackend.run (corrector/backends/newton.py:39-143) with
 * _ArmijoLineSearch (corrector/stepping/armijo.py:58-164), the infinity norm
 * (corrector/interfaces.py:44-46) and the Halo extra Jacobian term
 * (services/orbits.py:917-947).  Shared by the host-side bank of
 * corrections (hh_bank.cpp) and the device-resident corrector (k_correct),
 * so both take the same decisions in the same arithmetic.
 *
 * A row moves through stages:
 *   RESID   waits for the event search at x (the residual),
 *   JAC     waits for the STM at the event time (Jacobian),
 *   STEP    waits for the Newton step (the dense solve),
 *   ARMIJO  waits for the event searches of every trial step
 *           alpha = 1, r, r^2, ... >= min_alpha (evaluated speculatively;
 *           armijo.py:111-164 is replayed over them in order),
 *   DONE.
 * Every arithmetic step is the reference's scalar expression, operation for
 * operation (the library is compiled with -ffp-contract=off).  The powers
 * of the Halo term go through a caller-supplied pow: the C library's on the
 * host (what NumPy's scalar ** calls), the correctly rounded one on the
 * device.
 */
#ifndef HH_NEWTON_H
#define HH_NEWTON_H

#include <math.h>

#include "../../include/hiten_hip.h"

#define HH_HD __host__ __device__

namespace hh {

constexpr int NEWTON_NMAX = 6;    /* residuals / controls per correction */
constexpr int NEWTON_JMAX = 64;   /* line-search trials per step          */

enum NewtonStage { NS_RESID = 0, NS_JAC = 1, NS_STEP = 2, NS_ARMIJO = 3, NS_DONE = 4 };

struct NewtonRow {
    double mu, base[6], tgt[NEWTON_NMAX];
    int res[NEWTON_NMAX], ctrl[NEWTON_NMAX];
    int halo, fwd, max_att;
    double tol, max_delta, c;
    double x[NEWTON_NMAX];
    int k, n_ev, n_stm, stage;
    double r[NEWTON_NMAX], rn, t_ev, x_ev[6];
    int nal;
    double alpha[NEWTON_JMAX], trial[NEWTON_JMAX][NEWTON_NMAX];
    int ok, err;
    double half_period;
};

/* numpy.max(numpy.abs(v)): NaN-propagating */
HH_HD inline double newton_inf_norm(const double* v, int n) {
    double m = fabs(v[0]);
    for (int i = 1; i < n; ++i) {
        const double a = fabs(v[i]);
        if (a != a || m != m) m = NAN;
        else if (a > m) m = a;
    }
    return m;
}

/* the full 6-d state: the base state with the controls replaced */
HH_HD inline void newton_full_state(const NewtonRow& w, const double* params, int nc,
                                    double* out) {
    for (int i = 0; i < 6; ++i) out[i] = w.base[i];
    for (int i = 0; i < nc; ++i) out[w.ctrl[i]] = params[i];
}

HH_HD inline void newton_finish(NewtonRow& w, int ok, int err, double t) {
    w.ok = ok;
    w.err = err;
    w.half_period = t;
    w.stage = NS_DONE;
}

/* A new row in stage RESID; the line-search steps alpha = 1, r, r^2, ...
 * >= min_alpha (armijo.py:101-110).  Returns 0, or -1 for invalid indices
 * or more than NEWTON_JMAX steps.                                          */
HH_HD inline int newton_init(NewtonRow& w, int nr, int nc, double mu, const double* x0,
                             const int* res, const int* ctrl, const double* tgt, int halo,
                             int fwd, int max_att, double tol, double max_delta,
                             double alpha_reduction, double min_alpha, double armijo_c) {
    w.mu = mu;
    for (int i = 0; i < 6; ++i) w.base[i] = x0[i];
    for (int i = 0; i < nr; ++i) {
        if (res[i] < 0 || res[i] > 5) return -1;
        w.res[i] = res[i];
        w.tgt[i] = tgt[i];
    }
    for (int i = 0; i < nc; ++i) {
        if (ctrl[i] < 0 || ctrl[i] > 5) return -1;
        w.ctrl[i] = ctrl[i];
        w.x[i] = x0[ctrl[i]];
    }
    w.halo = halo;
    w.fwd = fwd;
    w.max_att = max_att;
    w.tol = tol;
    w.max_delta = max_delta;
    w.c = armijo_c;
    w.k = 0;
    w.n_ev = 0;
    w.n_stm = 0;
    w.ok = 0;
    w.err = 0;
    double a = 1.0;
    w.nal = 0;
    while (a >= min_alpha) {
        if (w.nal == NEWTON_JMAX) return -1;
        w.alpha[w.nal++] = a;
        a = a * alpha_reduction;
    }
    w.stage = NS_RESID;
    w.rn = NAN;
    w.t_ev = NAN;
    w.half_period = NAN;
    return 0;
}

/* top of the Newton loop with the residual at x known (newton.py:100-120) */
HH_HD inline void newton_residual(NewtonRow& w, int nr, double t, const double* xe, int ok) {
    w.n_ev += 1;
    if (!ok) {
        w.rn = NAN;
        newton_finish(w, 0, HH_BANK_NO_CROSSING, NAN);
        return;
    }
    double r[NEWTON_NMAX];
    for (int i = 0; i < nr; ++i) r[i] = xe[w.res[i]] - w.tgt[i];
    const double rn = newton_inf_norm(r, nr);
    w.rn = rn;
    if (rn < w.tol) {
        newton_finish(w, 1, HH_BANK_CONVERGED, t);
        return;
    }
    if (w.k >= w.max_att) {
        newton_finish(w, 0, HH_BANK_MAX_ITERATIONS, NAN);
        return;
    }
    for (int i = 0; i < nr; ++i) w.r[i] = r[i];
    w.t_ev = t;
    for (int i = 0; i < 6; ++i) w.x_ev[i] = xe[i];
    w.stage = NS_JAC;
}

/* Whether the trial results 0..n-1 (of w.nal) already fix the line search's
 * outcome: some trial in order is accepted.  Lets a caller that evaluates
 * the trials in chunks stop early (armijo.py:111-164 is sequential).       */
HH_HD inline bool newton_armijo_decided(const NewtonRow& w, int nr, int n, const double* xe,
                                        const int* ok) {
    for (int j = 0; j < n; ++j) {
        if (!ok[j]) continue;
        double r[NEWTON_NMAX];
        for (int q = 0; q < nr; ++q) r[q] = xe[6 * j + w.res[q]] - w.tgt[q];
        if (newton_inf_norm(r, nr) <= (1.0 - w.c * w.alpha[j]) * w.rn) return true;
    }
    return n >= w.nal;
}

/* armijo.py:111-164 over the speculatively evaluated trials: the first
 * accepted one, else the best one, else failure; the accepted trial's
 * event search is the residual at the next iterate.  Returns the trial
 * taken, or -1.                                                           */
HH_HD inline int newton_armijo(NewtonRow& w, int nr, int nc, const double* t, const double* xe,
                               const int* ok) {
    int chosen = -1, best = -1;
    double best_norm = w.rn;
    for (int j = 0; j < w.nal; ++j) {
        w.n_ev += 1;
        if (!ok[j]) continue;
        double r[NEWTON_NMAX];
        for (int q = 0; q < nr; ++q) r[q] = xe[6 * j + w.res[q]] - w.tgt[q];
        const double nt = newton_inf_norm(r, nr);
        if (nt <= (1.0 - w.c * w.alpha[j]) * w.rn) {
            chosen = j;
            break;
        }
        if (nt < best_norm) {
            best = j;
            best_norm = nt;
        }
    }
    if (chosen < 0) chosen = best;
    if (chosen < 0) {
        newton_finish(w, 0, HH_BANK_LINE_SEARCH_FAILED, NAN);
        return -1;
    }
    for (int q = 0; q < nc; ++q) w.x[q] = w.trial[chosen][q];
    w.k += 1;
    w.stage = NS_RESID;
    newton_residual(w, nr, t[chosen], xe + 6 * chosen, ok[chosen]);
    return chosen;
}

/* services/orbits.py:917-947 */
template <class Pow>
HH_HD inline void newton_halo_term(double mu, const double* X, const double* Phi, double H[2][2],
                                   Pow P_) {
    const double x = X[0], y = X[1], z = X[2], vy0 = X[4];
    const double mu2 = 1 - mu;
    const double rho_1 = 1 / P_(P_(x + mu, 2.0) + P_(y, 2.0) + P_(z, 2.0), 1.5);
    const double rho_2 = 1 / P_(P_(x - mu2, 2.0) + P_(y, 2.0) + P_(z, 2.0), 1.5);
    const double omega_x = -(mu2 * (x + mu) * rho_1) - (mu * (x - mu2) * rho_2) + x;
    const double DD[2] = {2 * vy0 + omega_x, -(mu2 * z * rho_1) - (mu * z * rho_2)};
    double vy = vy0;
    if (fabs(vy) < 1e-9) vy = (vy != 0) ? ((vy > 0) ? 1e-9 : -1e-9) : 1e-9;
    const double Pr[2] = {Phi[6 * 1 + 0], Phi[6 * 1 + 4]};
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) H[a][c] = (DD[a] * Pr[c]) / vy;
}

/* J = Phi(t_event)[res][:, ctrl] - extra term (operators.py:437-452) and the
 * residual it goes with; the row moves to STEP.                            */
template <class Pow>
HH_HD inline void newton_jacobian(NewtonRow& w, int nr, int nc, const double* P, double* J,
                                  double* r, Pow pw) {
    for (int a = 0; a < nr; ++a)
        for (int c = 0; c < nc; ++c) J[a * nc + c] = P[6 * w.res[a] + w.ctrl[c]];
    if (w.halo) {
        double H[2][2];
        newton_halo_term(w.mu, w.x_ev, P, H, pw);
        for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c) J[a * 2 + c] = J[a * 2 + c] - H[a][c];
    }
    for (int a = 0; a < nr; ++a) r[a] = w.r[a];
    w.n_stm += 1;
    w.stage = NS_STEP;
}

/* The Newton step delta, capped to max_delta in the infinity norm
 * (armijo.py:101-110), and the trial points x + alpha_j delta.             */
HH_HD inline void newton_set_step(NewtonRow& w, int nc, const double* delta) {
    double d[NEWTON_NMAX];
    for (int c = 0; c < nc; ++c) d[c] = delta[c];
    if (!isinf(w.max_delta)) {
        const double dn = newton_inf_norm(d, nc);
        if (dn > w.max_delta)
            for (int c = 0; c < nc; ++c) d[c] = d[c] * (w.max_delta / dn);
    }
    for (int j = 0; j < w.nal; ++j)
        for (int c = 0; c < nc; ++c) w.trial[j][c] = w.x[c] + w.alpha[j] * d[c];
    w.stage = NS_ARMIJO;
}

}  // namespace hh

#endif

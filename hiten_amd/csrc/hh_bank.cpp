/* hh_bank.cpp -- host-side state machine of batched single-shooting
 * corrections (the reference's _NewtonBackend.run with its Armijo line
 * search, corrector/backends/newton.py:39-143 and stepping/armijo.py:58-164),
 * kept in native code so that a serving loop spends microseconds, not
 * Python-interpreter time, between the GPU batches of its corrections.
 *
 * One row per correction.  Stages:
 *   RESID   waits for the event search at x (the residual),
 *   JAC     waits for the STM at the event time (Jacobian),
 *   STEP    waits for the Newton step (the dense solve stays with NumPy's
 *           LAPACK, whose bits it must reproduce),
 *   ARMIJO  waits for the event searches of every trial step
 *           alpha = 1, r, r^2, ... >= min_alpha (evaluated speculatively in one
 *           batch; armijo.py:111-164 is replayed over them in order),
 *   DONE.
 * Every arithmetic step is the reference's scalar expression, operation for
 * operation (compiled with -ffp-contract=off); powers in the Halo term use the
 * C library pow, which is what NumPy's scalar ** calls.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "../../include/hiten_hip.h"

namespace {

constexpr int NMAX = 6;
constexpr int JMAX = 64;

enum Stage { RESID = 0, JAC = 1, STEP = 2, ARMIJO = 3, DONE = 4 };

struct Row {
    double mu, base[6], tgt[NMAX];
    int res[NMAX], ctrl[NMAX];
    int halo, fwd, max_att;
    double tol, max_delta, c;
    double x[NMAX];
    int k, n_ev, n_stm, stage;
    double r[NMAX], rn, t_ev, x_ev[6];
    int nal;
    double alpha[JMAX], trial[JMAX][NMAX];
    int ok, err;
    double half_period;
};

struct Bank {
    int nr, nc;
    std::vector<Row> rows;
    std::vector<int> free_rows;   /* released rows, reused by hh_bank_add */
};

/* numpy.max(numpy.abs(v)): NaN-propagating */
double inf_norm(const double* v, int n) {
    double m = fabs(v[0]);
    for (int i = 1; i < n; ++i) {
        const double a = fabs(v[i]);
        if (a != a || m != m) m = NAN;
        else if (a > m) m = a;
    }
    return m;
}

void full_state(const Row& w, const double* params, int nc, double* out) {
    for (int i = 0; i < 6; ++i) out[i] = w.base[i];
    for (int i = 0; i < nc; ++i) out[w.ctrl[i]] = params[i];
}

void finish(Row& w, int ok, int err, double t) {
    w.ok = ok;
    w.err = err;
    w.half_period = t;
    w.stage = DONE;
}

/* top of the Newton loop with the residual at x known */
void residual(Bank& b, Row& w, double t, const double* xe, int ok) {
    w.n_ev += 1;
    if (!ok) {
        w.rn = NAN;
        finish(w, 0, HH_BANK_NO_CROSSING, NAN);
        return;
    }
    double r[NMAX];
    for (int i = 0; i < b.nr; ++i) r[i] = xe[w.res[i]] - w.tgt[i];
    const double rn = inf_norm(r, b.nr);
    w.rn = rn;
    if (rn < w.tol) {
        finish(w, 1, HH_BANK_CONVERGED, t);
        return;
    }
    if (w.k >= w.max_att) {
        finish(w, 0, HH_BANK_MAX_ITERATIONS, NAN);
        return;
    }
    for (int i = 0; i < b.nr; ++i) w.r[i] = r[i];
    w.t_ev = t;
    for (int i = 0; i < 6; ++i) w.x_ev[i] = xe[i];
    w.stage = JAC;
}

/* The C library pow called as a function (NumPy's scalar ** calls it): a
 * compiler would otherwise be free to turn pow(v, 2) into v * v, which the
 * library pow does not always equal.                                       */
double (*volatile libm_pow)(double, double) = pow;

/* services/orbits.py:917-947 */
void halo_term(double mu, const double* X, const double* Phi, double H[2][2]) {
    double (*const P_)(double, double) = libm_pow;
    const double x = X[0], y = X[1], z = X[2], vy0 = X[4];
    const double mu2 = 1 - mu;
    const double rho_1 = 1 / P_(P_(x + mu, 2.0) + P_(y, 2.0) + P_(z, 2.0), 1.5);
    const double rho_2 = 1 / P_(P_(x - mu2, 2.0) + P_(y, 2.0) + P_(z, 2.0), 1.5);
    const double omega_x = -(mu2 * (x + mu) * rho_1) - (mu * (x - mu2) * rho_2) + x;
    const double DD[2] = {2 * vy0 + omega_x, -(mu2 * z * rho_1) - (mu * z * rho_2)};
    double vy = vy0;
    if (fabs(vy) < 1e-9) vy = (vy != 0) ? ((vy > 0) ? 1e-9 : -1e-9) : 1e-9;
    const double P[2] = {Phi[6 * 1 + 0], Phi[6 * 1 + 4]};
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) H[a][c] = (DD[a] * P[c]) / vy;
}

Bank* as_bank(void* p) { return static_cast<Bank*>(p); }

bool valid_rows(Bank* b, int n, const int* rows) {
    if (!b || n < 0 || (n > 0 && !rows)) return false;
    for (int i = 0; i < n; ++i)
        if (rows[i] < 0 || rows[i] >= (int)b->rows.size()) return false;
    return true;
}

}  // namespace

extern "C" {

void* hh_bank_create(int nr, int nc) {
    if (nr < 1 || nr > NMAX || nc < 1 || nc > NMAX) return nullptr;
    Bank* b = new (std::nothrow) Bank;
    if (!b) return nullptr;
    b->nr = nr;
    b->nc = nc;
    return b;
}

void hh_bank_destroy(void* bank) { delete as_bank(bank); }

int hh_bank_size(void* bank) { return bank ? (int)as_bank(bank)->rows.size() : -1; }

int hh_bank_add(void* bank, double mu, const double* x0, const int* res, const int* ctrl,
                const double* tgt, int halo, int fwd, int max_attempts, double tol,
                double max_delta, double alpha_reduction, double min_alpha, double armijo_c) {
    Bank* b = as_bank(bank);
    if (!b || !x0 || !res || !ctrl || !tgt) return -HH_ERR_ARGS;
    if (halo && (b->nr != 2 || b->nc != 2)) return -HH_ERR_ARGS;
    Row w;
    memset(&w, 0, sizeof w);
    w.mu = mu;
    for (int i = 0; i < 6; ++i) w.base[i] = x0[i];
    for (int i = 0; i < b->nr; ++i) {
        if (res[i] < 0 || res[i] > 5) return -HH_ERR_ARGS;
        w.res[i] = res[i];
        w.tgt[i] = tgt[i];
    }
    for (int i = 0; i < b->nc; ++i) {
        if (ctrl[i] < 0 || ctrl[i] > 5) return -HH_ERR_ARGS;
        w.ctrl[i] = ctrl[i];
        w.x[i] = x0[ctrl[i]];
    }
    w.halo = halo;
    w.fwd = fwd;
    w.max_att = max_attempts;
    w.tol = tol;
    w.max_delta = max_delta;
    w.c = armijo_c;
    double a = 1.0;
    w.nal = 0;
    while (a >= min_alpha) {
        if (w.nal == JMAX) return -HH_ERR_ARGS;
        w.alpha[w.nal++] = a;
        a = a * alpha_reduction;
    }
    w.stage = RESID;
    w.rn = NAN;
    w.half_period = NAN;
    if (!b->free_rows.empty()) {
        const int row = b->free_rows.back();
        b->free_rows.pop_back();
        b->rows[row] = w;
        return row;
    }
    b->rows.push_back(w);
    return (int)b->rows.size() - 1;
}

int hh_bank_release(void* bank, int row) {
    Bank* b = as_bank(bank);
    if (!b || row < 0 || row >= (int)b->rows.size() || b->rows[row].stage != DONE)
        return HH_ERR_ARGS;
    b->rows[row].stage = -1;      /* free: no longer DONE, not in any stage */
    b->free_rows.push_back(row);
    return HH_OK;
}

int hh_bank_stages(void* bank, int n, const int* rows, int* stage) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !stage) return HH_ERR_ARGS;
    for (int i = 0; i < n; ++i) stage[i] = b->rows[rows[i]].stage;
    return HH_OK;
}

int hh_bank_event_inputs(void* bank, int n, const int* rows, int* counts, double* states,
                         int capacity) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !counts) return -HH_ERR_ARGS;
    int total = 0;
    for (int i = 0; i < n; ++i) {
        const Row& w = b->rows[rows[i]];
        if (w.stage == RESID) counts[i] = 1;
        else if (w.stage == ARMIJO) counts[i] = w.nal;
        else return -HH_ERR_ARGS;
        total += counts[i];
    }
    if (!states) return total;
    if (total > capacity) return -HH_ERR_ARGS;
    int o = 0;
    for (int i = 0; i < n; ++i) {
        const Row& w = b->rows[rows[i]];
        if (w.stage == RESID) {
            full_state(w, w.x, b->nc, states + 6 * o++);
        } else {
            for (int j = 0; j < w.nal; ++j) full_state(w, w.trial[j], b->nc, states + 6 * o++);
        }
    }
    return total;
}

int hh_bank_feed_events(void* bank, int n, const int* rows, const double* t, const double* xe,
                        const int* ok) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !t || !xe || !ok) return HH_ERR_ARGS;
    int o = 0;
    for (int i = 0; i < n; ++i) {
        Row& w = b->rows[rows[i]];
        if (w.stage == RESID) {
            residual(*b, w, t[o], xe + 6 * o, ok[o]);
            o += 1;
            continue;
        }
        if (w.stage != ARMIJO) return HH_ERR_ARGS;
        /* armijo.py:111-164 over the speculatively evaluated trials */
        int chosen = -1, best = -1;
        double best_norm = w.rn;
        for (int j = 0; j < w.nal; ++j) {
            w.n_ev += 1;
            if (!ok[o + j]) continue;
            double r[NMAX];
            for (int q = 0; q < b->nr; ++q) r[q] = xe[6 * (o + j) + w.res[q]] - w.tgt[q];
            const double nt = inf_norm(r, b->nr);
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
            finish(w, 0, HH_BANK_LINE_SEARCH_FAILED, NAN);
        } else {
            for (int q = 0; q < b->nc; ++q) w.x[q] = w.trial[chosen][q];
            w.k += 1;
            w.stage = RESID;
            residual(*b, w, t[o + chosen], xe + 6 * (o + chosen), ok[o + chosen]);
        }
        o += w.nal;
    }
    return HH_OK;
}

int hh_bank_stm_inputs(void* bank, int n, const int* rows, double* states, double* t_event) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !states || !t_event) return HH_ERR_ARGS;
    for (int i = 0; i < n; ++i) {
        const Row& w = b->rows[rows[i]];
        if (w.stage != JAC) return HH_ERR_ARGS;
        full_state(w, w.x, b->nc, states + 6 * i);
        t_event[i] = w.t_ev;
    }
    return HH_OK;
}

int hh_bank_jacobians(void* bank, int n, const int* rows, const double* phi, double* J,
                      double* r) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !phi || !J || !r) return HH_ERR_ARGS;
    const int nr = b->nr, nc = b->nc;
    for (int i = 0; i < n; ++i) {
        Row& w = b->rows[rows[i]];
        if (w.stage != JAC) return HH_ERR_ARGS;
        const double* P = phi + 36 * i;
        double* Ji = J + (size_t)nr * nc * i;
        for (int a = 0; a < nr; ++a)
            for (int c = 0; c < nc; ++c) Ji[a * nc + c] = P[6 * w.res[a] + w.ctrl[c]];
        if (w.halo) {
            double H[2][2];
            halo_term(w.mu, w.x_ev, P, H);
            for (int a = 0; a < 2; ++a)
                for (int c = 0; c < 2; ++c) Ji[a * 2 + c] = Ji[a * 2 + c] - H[a][c];
        }
        for (int a = 0; a < nr; ++a) r[(size_t)nr * i + a] = w.r[a];
        w.n_stm += 1;
        w.stage = STEP;
    }
    return HH_OK;
}

int hh_bank_set_steps(void* bank, int n, const int* rows, const double* delta) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !delta) return HH_ERR_ARGS;
    const int nc = b->nc;
    for (int i = 0; i < n; ++i) {
        Row& w = b->rows[rows[i]];
        if (w.stage != STEP) return HH_ERR_ARGS;
        double d[NMAX];
        for (int c = 0; c < nc; ++c) d[c] = delta[(size_t)nc * i + c];
        if (!isinf(w.max_delta)) {
            const double dn = inf_norm(d, nc);
            if (dn > w.max_delta)
                for (int c = 0; c < nc; ++c) d[c] = d[c] * (w.max_delta / dn);
        }
        for (int j = 0; j < w.nal; ++j)
            for (int c = 0; c < nc; ++c) w.trial[j][c] = w.x[c] + w.alpha[j] * d[c];
        w.stage = ARMIJO;
    }
    return HH_OK;
}

int hh_bank_result(void* bank, int row, double* x_full, double* residual_norm, int* iterations,
                   int* n_event_searches, int* n_stm, int* converged, int* error,
                   double* half_period) {
    Bank* b = as_bank(bank);
    if (!b || row < 0 || row >= (int)b->rows.size()) return HH_ERR_ARGS;
    const Row& w = b->rows[row];
    if (w.stage != DONE) return HH_ERR_ARGS;
    if (x_full) full_state(w, w.x, b->nc, x_full);
    if (residual_norm) *residual_norm = w.rn;
    if (iterations) *iterations = w.k;
    if (n_event_searches) *n_event_searches = w.n_ev;
    if (n_stm) *n_stm = w.n_stm;
    if (converged) *converged = w.ok;
    if (error) *error = w.err;
    if (half_period) *half_period = w.half_period;
    return HH_OK;
}

}  /* extern "C" */

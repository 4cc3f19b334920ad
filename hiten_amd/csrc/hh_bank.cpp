/* hh_bank.cpp -- host-side state machine of batched single-shooting
 * corrections (the reference's _NewtonBackend.run with its Armijo line
 * search, corrector/backends/newton.py:39-143 and stepping/armijo.py:58-164),
 * kept in native code so that a serving loop spends microseconds, not
 * Python-interpreter time, between the GPU batches of its corrections.
 *
 * One row per correction; the stages and the arithmetic of a row are the
 * shared functions of hh_newton.h.  Between the GPU batches the caller does
 * the dense solve of the STEP stage with NumPy's LAPACK, whose bits the
 * reference's results carry; the Halo term's powers use the C library pow,
 * which is what NumPy's scalar ** calls.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "hh_newton.h"

namespace {

using hh::NewtonRow;
using Row = hh::NewtonRow;

constexpr int NMAX = hh::NEWTON_NMAX;

struct Bank {
    int nr, nc;
    std::vector<Row> rows;
    std::vector<int> free_rows;   /* released rows, reused by hh_bank_add */
};

/* The C library pow called as a function (NumPy's scalar ** calls it): a
 * compiler would otherwise be free to turn pow(v, 2) into v * v, which the
 * library pow does not always equal.                                       */
double (*volatile libm_pow)(double, double) = pow;

struct LibmPow {
    double operator()(double a, double b) const { return libm_pow(a, b); }
};

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
    if (hh::newton_init(w, b->nr, b->nc, mu, x0, res, ctrl, tgt, halo, fwd, max_attempts, tol,
                        max_delta, alpha_reduction, min_alpha, armijo_c) != 0)
        return -HH_ERR_ARGS;
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
    if (!b || row < 0 || row >= (int)b->rows.size() || b->rows[row].stage != hh::NS_DONE)
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
        if (w.stage == hh::NS_RESID) counts[i] = 1;
        else if (w.stage == hh::NS_ARMIJO) counts[i] = w.nal;
        else return -HH_ERR_ARGS;
        total += counts[i];
    }
    if (!states) return total;
    if (total > capacity) return -HH_ERR_ARGS;
    int o = 0;
    for (int i = 0; i < n; ++i) {
        const Row& w = b->rows[rows[i]];
        if (w.stage == hh::NS_RESID) {
            hh::newton_full_state(w, w.x, b->nc, states + 6 * o++);
        } else {
            for (int j = 0; j < w.nal; ++j)
                hh::newton_full_state(w, w.trial[j], b->nc, states + 6 * o++);
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
        if (w.stage == hh::NS_RESID) {
            hh::newton_residual(w, b->nr, t[o], xe + 6 * o, ok[o]);
            o += 1;
            continue;
        }
        if (w.stage != hh::NS_ARMIJO) return HH_ERR_ARGS;
        const int nal = w.nal;
        hh::newton_armijo(w, b->nr, b->nc, t + o, xe + 6 * o, ok + o);
        o += nal;
    }
    return HH_OK;
}

int hh_bank_stm_inputs(void* bank, int n, const int* rows, double* states, double* t_event) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !states || !t_event) return HH_ERR_ARGS;
    for (int i = 0; i < n; ++i) {
        const Row& w = b->rows[rows[i]];
        if (w.stage != hh::NS_JAC) return HH_ERR_ARGS;
        hh::newton_full_state(w, w.x, b->nc, states + 6 * i);
        t_event[i] = w.t_ev;
    }
    return HH_OK;
}

int hh_bank_jacobians(void* bank, int n, const int* rows, const double* phi, double* J,
                      double* r) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !phi || !J || !r) return HH_ERR_ARGS;
    const int nr = b->nr, nc = b->nc;
    for (int i = 0; i < n; ++i)
        if (b->rows[rows[i]].stage != hh::NS_JAC) return HH_ERR_ARGS;
    for (int i = 0; i < n; ++i)
        hh::newton_jacobian(b->rows[rows[i]], nr, nc, phi + 36 * i, J + (size_t)nr * nc * i,
                            r + (size_t)nr * i, LibmPow());
    return HH_OK;
}

int hh_bank_set_steps(void* bank, int n, const int* rows, const double* delta) {
    Bank* b = as_bank(bank);
    if (!valid_rows(b, n, rows) || !delta) return HH_ERR_ARGS;
    const int nc = b->nc;
    for (int i = 0; i < n; ++i)
        if (b->rows[rows[i]].stage != hh::NS_STEP) return HH_ERR_ARGS;
    for (int i = 0; i < n; ++i) hh::newton_set_step(b->rows[rows[i]], nc, delta + (size_t)nc * i);
    return HH_OK;
}

int hh_bank_result(void* bank, int row, double* x_full, double* residual_norm, int* iterations,
                   int* n_event_searches, int* n_stm, int* converged, int* error,
                   double* half_period) {
    Bank* b = as_bank(bank);
    if (!b || row < 0 || row >= (int)b->rows.size()) return HH_ERR_ARGS;
    const Row& w = b->rows[row];
    if (w.stage != hh::NS_DONE) return HH_ERR_ARGS;
    if (x_full) hh::newton_full_state(w, w.x, b->nc, x_full);
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

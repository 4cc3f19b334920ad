/* hh_oracle.h -- CPU restatement of hiten's CR3BP propagation hot path.
 *
 * TEST INFRASTRUCTURE.  This library is the checker for the HIP product: it
 * may only be loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  Nothing in hiten_amd/ links, imports or calls it.
 *
 * It restates, operation for operation, the numba kernels of hiten 0.5.4:
 *   dynamics/rtbp.py:30-255       CR3BP field, Jacobian, variational field
 *   dynamics/base.py:272-310      _DirectedSystem sign flips
 *   integrators/rk.py:154-215     explicit RK step
 *   integrators/rk.py:531-588     fixed-step driver
 *   integrators/rk.py:655-718     fixed-step driver with events (+286-391)
 *   integrators/rk.py:1636-1708   DOP853 step
 *   integrators/rk.py:1790-2003   DOP853 dense output
 *   integrators/rk.py:2005-2102   in-step event refinement
 *   integrators/rk.py:2375-2549   adaptive driver with dense output
 *   integrators/rk.py:2678-2803   adaptive driver until event
 *   integrators/utils.py:13-296   crossing predicates, step control
 * Where the reference delegates to a third-party routine whose bit-level
 * behaviour is implementation-defined (BLAS ddot / dnrm2 inside np.dot and
 * np.linalg.norm), the summation order is fixed here and documented at
 * or_sumsq().  Parity status of each function is listed in the header of
 * hh_oracle.c.
 */
#ifndef HH_ORACLE_H
#define HH_ORACLE_H

#include <stdint.h>

#define OR_KIND_CR3BP6 0
#define OR_KIND_CR3BP42 1
#define OR_KIND_LINEAR 2

#define OR_OK 0
#define OR_ERR_MAXSTEPS 1
#define OR_ERR_NONFINITE 2
#define OR_ERR_ARGS 3
#define OR_ERR_NOMEM 4

typedef struct {
    int kind;             /* OR_KIND_*                                        */
    int dim;              /* 6, 42, or 1..6 for LINEAR                        */
    double mu;            /* CR3BP mass parameter                             */
    int fwd;              /* _DirectedSystem direction: +1 or -1              */
    uint64_t flip;        /* bit i: negate component i when fwd == -1         */
    int libm_pow;         /* 0: correctly rounded x**1.5, x**2.5 and controller
                           *    powers, x**2 -> x*x, x**3 -> x*(x*x) (numba);
                           * 1: rhs42's x**1.5, x**2.5 through the C library
                           *    pow (numba calls it for them);
                           * 2: every ** through the C library pow, as the
                           *    reference's source runs as plain Python on
                           *    NumPy scalars (fields and step controller)   */
    const double *lin_a;  /* LINEAR: dy = lin_a (dim x dim) @ y + lin_b       */
    const double *lin_b;
    long long nfev;       /* right-hand side evaluations (diagnostic)         */
    double *trace;        /* NULL, or trace_cap rows (t, h, err, accepted)    */
    long long trace_cap;  /* of the adaptive drivers' attempts, in order      */
    long long trace_len;  /* rows written (counts past trace_cap too)         */
} or_sys;

typedef struct {
    long long n_accepted;
    long long n_rejected;
    int status;
} or_stats;

/* vector fields */
void or_rhs(or_sys *s, double t, const double *y, double *dy);
void or_crtbp_accel(const double *state, double mu, double *out);
void or_jacobian_crtbp(double x, double y, double z, double mu, int libm_pow, double *F);
void or_var_equations(const double *phi_vec, double mu, int libm_pow, double *out);
double or_sumsq(const double *v, int dim);
double or_pow15(double x);
double or_pow25(double x);
double or_pow(double x, double y);

/* numpy.linspace(t0, tf, n) */
void or_linspace(double t0, double tf, int n, double *out);

/* fixed-step explicit RK8 over t_vals (rk.py:531-588) */
int or_rk_fixed(or_sys *s, const double *y0, const double *t_vals, int n_t,
                double *states, double *derivs);
/* the same for the tableau of order 4, 6 or 8 (rk.py:755-783) */
int or_rk_fixed_order(int order, or_sys *s, const double *y0, const double *t_vals, int n_t,
                      double *states, double *derivs);

/* fixed-step RK until event (rk.py:655-718) */
int or_rk_fixed_event(or_sys *s, const double *y0, const double *t_vals, int n_t,
                      int ev_index, double ev_offset, int direction, double xtol, double gtol,
                      int *hit, double *t_hit, double *y_hit, double *y_last);
int or_rk_fixed_event_order(int order, or_sys *s, const double *y0, const double *t_vals,
                            int n_t, int ev_index, double ev_offset, int direction, double xtol,
                            double gtol, int *hit, double *t_hit, double *y_hit,
                            double *y_last);

/* DOP853 with dense output at t_eval (rk.py:2375-2549) */
int or_dop853(or_sys *s, const double *y0, const double *t_eval, int m,
              double rtol, double atol, double max_step, double min_step,
              long long max_steps, double *y_out, double *derivs_out, or_stats *st);

/* DOP853 until the event g(t,y) = y[ev_index] - ev_offset (rk.py:2678-2803) */
int or_dop853_event(or_sys *s, const double *y0, double t0, double tmax,
                    double rtol, double atol, double max_step, double min_step,
                    long long max_steps, int ev_index, double ev_offset,
                    int direction, double xtol, double gtol,
                    int *hit, double *t_hit, double *y_hit, double *y_last, or_stats *st);

/* one DOP853 attempt; returns the scaled error norm.  k_out: 13 x dim or NULL */
double or_dop853_single_step(or_sys *s, double t, const double *y, double h, double rtol,
                             double atol, double *y_high, double *err5, double *err3,
                             double *k_out);

/* batched drivers for the CPU baseline (OpenMP, one trajectory per thread).
 * Arrays are trajectory-major: y0[i*dim + c].                              */
int or_batch_rk8_final(int n, const double *mu, const double *y0, double t0, double tf,
                       int steps, double *y_out, int n_threads);
int or_batch_rk_final(int order, int n, const double *mu, const double *y0, double t0, double tf,
                      int steps, double *y_out, int n_threads, int with_derivs);
int or_batch_dop853_final(int n, int dim, const double *mu, const double *y0,
                          const double *t0, const double *tf, double rtol, double atol,
                          double max_step, long long max_steps, double *y_out,
                          long long *n_acc, long long *n_rej, int *status, int n_threads);
int or_batch_dop853_event(int n, const double *mu, const double *y0, const double *tmax,
                          double rtol, double atol, double max_step, long long max_steps,
                          int ev_index, double ev_offset, int direction, double xtol,
                          double gtol, int *hit, double *t_hit, double *y_hit,
                          long long *n_acc, long long *n_rej, int n_threads);

/* Jacobi constant of one state as _max_rel_energy_error's inner _jacobi
 * writes it (common/energy.py:51-54) and that function over an m-sample
 * trajectory (common/energy.py:26-74), plus the primary-distance minima of
 * the manifold filter (services/manifold.py:416-420).                      */
double or_jacobi(const double *s, double mu);
double or_max_rel_energy_error(const double *states, int m, double mu);
void or_traj_filters(int n, int m, const double *mu, const double *states, double *min_r1,
                     double *min_r2, double *max_err);

#endif /* HH_ORACLE_H */

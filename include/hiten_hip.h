/* hiten_hip.h -- C-ABI of the MI355X (gfx950) CR3BP propagation library.
 *
 * Every entry point takes plain device pointers (FP64 unless stated) that the
 * caller owns, a trajectory count, scalars, and a hipStream_t passed as
 * void*.  Calls are asynchronous on that stream; the library never frees or
 * retains caller memory.  State arrays are structure-of-arrays: component c
 * of trajectory i is element [c * n + i], with components in the reference's
 * order ([x, y, z, vx, vy, vz] for 6-d; [Phi row-major (36) | x (6)] for
 * 42-d).  Return value: HH_OK or an HH_ERR_* code, with a message from
 * hh_last_error() (thread-local).
 *
 * Direction handling restates hiten's _DirectedSystem
 * (src/hiten/algorithms/dynamics/base.py:186-310): with fwd == -1 the
 * derivative of every component whose bit is set in flip_mask is negated.
 *
 * Step control restates integrators/utils.py:126-296 and rk.py:2443-2484
 * (initial step, PI controller, endpoint clamp); max_steps bounds the
 * attempts per trajectory (the reference has no bound, rk.py:2452) and is
 * reported through status = HH_STATUS_MAXSTEPS.
 */
#ifndef HITEN_HIP_H
#define HITEN_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define HH_ABI_VERSION 3

#define HH_OK 0
#define HH_ERR_ARGS 1
#define HH_ERR_HIP 2

#define HH_STATUS_OK 0
#define HH_STATUS_MAXSTEPS 1
#define HH_STATUS_NONFINITE 2
/* hh_dop853_final, variant HH_VARIANT_PERSISTENT_PAIR: a trajectory its
 * scheduler could not requeue (the workspace capacity rules this out; the
 * status exists so that such a loss is never silent).  y_out is undefined. */
#define HH_STATUS_DROPPED 4

#define HH_VARIANT_PLAIN 0
#define HH_VARIANT_PERSISTENT 1
#define HH_VARIANT_WIDE 2
#define HH_VARIANT_PAIR 3
#define HH_VARIANT_PERSISTENT_PAIR 4

/* Message describing the last error returned on this thread. */
const char* hh_last_error(void);

/* HH_ABI_VERSION of the built library. */
int hh_abi_version(void);

/* Tuning of the persistent 42-d kernels (process-wide; read at launch).
 * No counterpart in the reference: scheduling only, results are
 * bit-identical for every setting.
 *   heavy_projected : a trajectory that has taken at least min_age step
 *                     attempts and whose projected total, attempts *
 *                     (tf - t0) / (t - t0), reaches it is finished one
 *                     component per lane: HH_VARIANT_PERSISTENT stops the
 *                     refill of its wave and switches once it is alone
 *                     there; HH_VARIANT_PERSISTENT_PAIR moves the wave's
 *                     other trajectories to its resume queue and switches at
 *                     once.  0 = only when no work is left.  Default 2000.
 *   min_age         : default 256
 *   pair_projected  : HH_VARIANT_PERSISTENT_PAIR: from this projected total
 *                     both waves of the workgroup take the trajectory (pair
 *                     layout); 0 = never.  Default 2000.
 *   refill_min      : HH_VARIANT_PERSISTENT_PAIR: idle groups (1..8) a wave
 *                     with live groups gathers before it refills; default 4
 *   waves_per_simd  : HH_VARIANT_PERSISTENT_PAIR: 1 or 2 (default): the
 *                     register budget per lane (512 or 256 VGPRs) and the
 *                     resident grid
 *   exclusive_projected : HH_VARIANT_PERSISTENT_PAIR: from this projected
 *                     total a wave-wide or pair run reserves its SIMD; the
 *                     other waves on it hand their work to the resume queue
 *                     and sleep until it ends.  0 = never (default)       */
typedef struct {
    int heavy_projected;
    int min_age;
    int pair_projected;
    int refill_min;
    int waves_per_simd;
    int exclusive_projected;
} hh_tuning_t;
int hh_set_tuning(const hh_tuning_t* t);
int hh_get_tuning(hh_tuning_t* t);

/* Fixed-step 13-stage RK8 on the 6-d CR3BP field (dim 6) or the 42-d
 * variational system (dim 42) over the grid t_vals[0..steps-1] when t_vals
 * is given (shared by all trajectories), else numpy.linspace(t0[i], tf[i],
 * steps).  Step k uses h = t[k+1] - t[k].
 * Replaces: _FixedStepRK._integrate_fixed_rk (src/hiten/algorithms/
 * integrators/rk.py:531-588) reached through _propagate_dynsys(...,
 * method="fixed", order=8) (dynamics/base.py:430-434).
 *   mu, t0, tf : [n] (t0, tf may be NULL when t_vals is given)
 *   y0, y_out  : [dim][n]
 *   traj_out   : [steps][dim][n] every grid state, or NULL
 *   variant    : HH_VARIANT_PLAIN (one trajectory per lane, or per 8-lane
 *                group for dim 42) or HH_VARIANT_WIDE (dim 6: one
 *                trajectory per 8-lane group, one component per lane: the
 *                latency form for small batches); same results either way
 *   block      : threads per block (64..256, 0 = default)              */
int hh_rk8_fixed(int n, int dim, const double* mu, const double* t0, const double* tf, int steps,
                 const double* t_vals, int fwd, unsigned long long flip_mask, const double* y0,
                 double* y_out, double* traj_out, int variant, int block, void* stream);

/* hh_rk8_fixed for the tableau of the given order: 4 (classical RK4), 6
 * (the 7-stage tableau the reference names RK6) or 8.
 * Replaces: _RK4 / _RK6 / _RK8 (integrators/rk.py:755-783, tableaux
 * integrators/coefficients/rk4.py, rk6.py, rk8.py) through
 * _FixedStepRK._integrate_fixed_rk (rk.py:531-588), i.e.
 * _propagate_dynsys(..., method="fixed", order=4|6|8).
 *   deriv_out  : [steps][dim][n] the derivative f(t, y) at every grid state
 *                (the `derivs` _integrate_fixed_rk returns beside the
 *                states, rk.py:575-588: each step's first stage, and one
 *                more evaluation at the last state), or NULL
 * Other arguments as hh_rk8_fixed.                                         */
int hh_rk_fixed(int order, int n, int dim, const double* mu, const double* t0, const double* tf,
                int steps, const double* t_vals, int fwd, unsigned long long flip_mask,
                const double* y0, double* y_out, double* traj_out, double* deriv_out,
                int variant, int block, void* stream);

/* Fixed-step RK of the given order (4, 6 or 8) on the 6-d field until the
 * plane event g = y[ev_index] - ev_offset crosses zero in `direction`
 * (0 either, +1 rising, -1 falling; g = 0 at a step's end counts), then
 * bisection on the cubic Hermite interpolant of the step to xtol / gtol.
 * Replaces: _FixedStepRK._integrate_fixed_rk_until_event (integrators/
 * rk.py:655-718) with _hermite_refine_in_step (rk.py:330-391), reached by
 * _FixedStepRK.integrate(..., event_fn=...) (rk.py:474-513).
 *   grid        : as hh_rk_fixed (t_vals shared, else linspace(t0, tf, steps))
 *   hit, t_hit  : [n]; without a crossing hit = 0 and t_hit = the last grid time
 *   y0, y_hit, y_last : [6][n]; y_last is the state at the end of the step
 *                that crossed (the last grid state without a crossing)  */
int hh_rk_fixed_event(int order, int n, const double* mu, const double* t0, const double* tf,
                      int steps, const double* t_vals, int fwd, unsigned long long flip_mask,
                      int ev_index, double ev_offset, int direction, double xtol, double gtol,
                      const double* y0, int* hit, double* t_hit, double* y_hit, double* y_last,
                      int block, void* stream);

/* Adaptive DOP853 from t0[i] to tf[i]; y_out is the dense-output value at
 * tf[i], i.e. the last row of _propagate_dynsys(..., steps=2,
 * method="adaptive") and the final Phi of _compute_stm.
 * Replaces: _DOP853._integrate_dop853 (integrators/rk.py:2375-2549) with
 * the step kernel dop853_step_jit_kernel (rk.py:1636-1708) on
 * _crtbp_accel (dim 6, dynamics/rtbp.py:30-74) or _var_equations (dim 42,
 * dynamics/rtbp.py:167-255).
 *   dim        : 6 or 42      mu, t0, tf : [n]   y0, y_out : [dim][n]
 *   status, n_acc, n_rej : [n] int, each may be NULL
 *   variant    : HH_VARIANT_PLAIN (one launch, one trajectory per lane or
 *                8-lane group), HH_VARIANT_PERSISTENT (dim 42: work-
 *                stealing groups refilled from a counter; a trajectory
 *                left alone in its wave, or past the heavy threshold of
 *                hh_set_tuning, finishes one component per lane),
 *                HH_VARIANT_PERSISTENT_PAIR (dim 42: work stealing with two
 *                waves per workgroup and a resume queue; a trajectory past
 *                the heavy threshold, or the last one of a wave once no work
 *                is left, is finished by both waves of its workgroup in the
 *                pair layout while the other trajectories of the workgroup
 *                move to the resume queue), or
 *                HH_VARIANT_WIDE (dim 42: one trajectory per wave, one
 *                component per lane, from the start; dim 6: one trajectory
 *                per 8-lane group, one component per lane: the latency
 *                forms for batches smaller than the GPU's wave slots) or
 *                HH_VARIANT_PAIR (dim 42: one trajectory per 128-thread
 *                workgroup, the state and the STM integrated by two waves
 *                side by side; block is ignored)
 *   block      : threads per block (dim 6: 64..256; dim 42: 64/128/256/512)
 *   grid       : persistent blocks (0 = as many as fit at once)
 *   workspace  : device memory of at least hh_dop853_final_workspace(n,
 *                dim, variant) bytes (work counters, resume queue, spilled
 *                states), required by the persistent variants and reset by
 *                the library on the stream before the launch; may be NULL
 *                for the others.  Calls that share a workspace must not
 *                overlap.
 *   order      : [n] permutation of 0..n-1 giving the order in which the
 *                persistent variants hand out trajectories (e.g. longest
 *                t_f first), or NULL for index order; ignored by the others */
int hh_dop853_final(int n, int dim, const double* mu, const double* t0, const double* tf,
                    double rtol, double atol, double max_step, double min_step,
                    long long max_steps, int fwd, unsigned long long flip_mask,
                    const double* y0, double* y_out, int* status, int* n_acc, int* n_rej,
                    int variant, int block, int grid, void* workspace,
                    long long workspace_bytes, const int* order, void* stream);

/* Bytes of workspace hh_dop853_final needs for (n, dim, variant): 0 for the
 * variants that need none.  No counterpart in the reference.              */
long long hh_dop853_final_workspace(int n, int dim, int variant);

/* hh_dop853_final without a workspace argument: the persistent variants take
 * their counters, resume queue and spill area from a device array the
 * library holds on each GPU (13,959,936 bytes: the pair-tail form's
 * workspace for HH_STATIC_PAIR_MAX trajectories), and the call
 * allocates no memory -- for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Every other argument, the variants and
 * the results are hh_dop853_final's.  HH_VARIANT_PERSISTENT_PAIR takes at
 * most HH_STATIC_PAIR_MAX trajectories (more: HH_ERR_ARGS); the persistent
 * variant takes any n.  The array is reset on the stream before each
 * launch; calls that overlap on one GPU share it and race.              */
#define HH_STATIC_PAIR_MAX 16384
int hh_dop853_final_static(int n, int dim, const double* mu, const double* t0, const double* tf,
                           double rtol, double atol, double max_step, double min_step,
                           long long max_steps, int fwd, unsigned long long flip_mask,
                           const double* y0, double* y_out, int* status, int* n_acc, int* n_rej,
                           int variant, int block, int grid, const int* order, void* stream);

/* Adaptive DOP853 from t0[i] to tf[i] with dense output at the m points of
 * numpy.linspace(t0[i], tf[i], m), or at t_eval[0..m-1] when t_eval is given
 * (shared by all trajectories; then t0[i] = t_eval[0] and tf[i] =
 * t_eval[m-1] are expected): the states of _propagate_dynsys(..., steps=m,
 * method="adaptive"), the PHI of _compute_stm(..., steps=m), and the
 * states of _DOP853.integrate(system, y0, t_vals).
 * Replaces: _DOP853._integrate_dop853 with its dense-output phase
 * (integrators/rk.py:2375-2549) and _propagate_dynsys
 * (dynamics/base.py:346-458) for the CR3BP systems.  A span that
 * numpy.isclose calls zero repeats y0 (dynamics/base.py:421-424).
 *   out : [n][m][dim] (trajectory-major, the reference's (steps, dim)
 *         array per trajectory)
 *   variant : HH_VARIANT_PLAIN, or HH_VARIANT_WIDE for dim 6 (one
 *         trajectory per 8-lane group, one component per lane)           */
int hh_dop853_grid(int n, int dim, const double* mu, const double* t0, const double* tf, int m,
                   const double* t_eval, double rtol, double atol, double max_step,
                   double min_step, long long max_steps, int fwd, unsigned long long flip_mask,
                   const double* y0, double* out, int* status, int* n_acc, int* n_rej,
                   int variant, int block, void* stream);

/* Adaptive DOP853 (6-d) from t0[i] until the first crossing of the plane
 * y[ev_index] = ev_offset in `direction` (-1, 0, +1), refined by bisection
 * on the dense interpolant, or until tmax[i].
 * Replaces: _DOP853._integrate_dop853_until_event (integrators/rk.py:
 * 2678-2803) and _dop853_refine_in_step (rk.py:2005-2102), the engine of
 * _SingleHitBackend._cross_event_driven (poincare/singlehit/backend.py:
 * 164-235) with the event functions _g_x0/_g_y0/_g_z0 (backend.py:30-42).
 *   hit : [n] int            t_hit : [n]
 *   y_hit : [6][n] state at the hit (or at the end when no hit)
 *   y_last : [6][n] end state of the crossing step, or NULL              */
int hh_dop853_event(int n, const double* mu, const double* t0, const double* tmax, double rtol,
                    double atol, double max_step, double min_step, long long max_steps,
                    int ev_index, double ev_offset, int direction, double xtol, double gtol,
                    int fwd, unsigned long long flip_mask, const double* y0, int* hit,
                    double* t_hit, double* y_hit, double* y_last, int* status, int* n_acc,
                    int* n_rej, int variant, int block, void* stream);

/* Evaluate the CR3BP right-hand side at n states (dim 6: _crtbp_accel,
 * dynamics/rtbp.py:30-74; dim 42: _var_equations, rtbp.py:167-255).
 *   y, out : [dim][n]                                                     */
int hh_eval_rhs(int n, int dim, const double* mu, const double* y, double* out, void* stream);

/* Arithmetic probe used by the parity tests: out[0][i] = sqrt(x[i]),
 * out[1][i] = x[i] / y[i], out[2][i] = x[i]^1.5, out[3][i] = x[i]^2.5,
 * out[4][i] = x[i]^y[i] as the kernels compute them, out[5][i] = the
 * kernels' own square root of x[i] (the guarded lean form the stage loops
 * use).  out : [6][n].
 * No counterpart in the reference (it checks the arithmetic the reference
 * obtains from libm: np.sqrt, '/', and '**' in rtbp.py / utils.py).       */
int hh_eval_math(int n, const double* x, const double* y, double* out, void* stream);

/* Crossings of an affine section n . x = offset by stored trajectories
 * (synodic Poincare maps): trajectory i is samples 0 .. lengths[i]-1 (all m
 * when lengths is NULL) of states[i][*][6] at times[*] (shared) or
 * times[i][*] (per_traj_times).  Detection, refinement (linear or cubic
 * Hermite + Newton, optional segment_refine sub-intervals), direction filter
 * and deduplication as the reference; hits in time order.
 * Replaces: _SynodicDetectionBackend.detect_on_trajectory / run
 * (src/hiten/algorithms/poincare/synodic/backend.py:644-887).
 *   normal     : host pointer to 6 doubles
 *   direction  : 0 (both), 1 (n . x - offset rising), -1 (falling)
 *   plane_i/j  : state components of each hit's 2-d point
 *   max_hits   : hits kept per trajectory at most (0 = no limit)
 *   cap        : hit slots per trajectory in hit_t [n][cap], hit_x
 *                [n][cap][6], hit_p [n][cap][2]; count[i] is the number of
 *                hits kept, which may exceed cap (then only cap are stored) */
int hh_synodic_detect(int n, int m, const double* times, int per_traj_times,
                      const int* lengths, const double* states, const double* normal,
                      double offset, int direction, int plane_i, int plane_j, int cubic,
                      int segment_refine, double tol_on_surface, double dedup_time_tol,
                      double dedup_point_tol, int max_hits, int newton_max_iter, int cap,
                      int* count, double* hit_t, double* hit_x, double* hit_p, void* stream);

/* hh_synodic_detect in passes: every part of 512 segments of every
 * trajectory is scanned by a wave of its own, which lists the segments
 * holding candidates; the listed segments are refined one per lane; then
 * one wave per trajectory deduplicates the refined candidates in segment
 * order -- the same hits, bit for bit.  A trajectory with more than 32
 * candidates in a part is detected whole by its wave of the last pass, as
 * hh_synodic_detect does.  workspace: device memory of hh_synodic_workspace(n, m)
 * bytes, 256-byte aligned, owned by the caller (no memory is allocated).  */
long long hh_synodic_workspace(int n, int m);
int hh_synodic_detect_ws(int n, int m, const double* times, int per_traj_times,
                         const int* lengths, const double* states, const double* normal,
                         double offset, int direction, int plane_i, int plane_j, int cubic,
                         int segment_refine, double tol_on_surface, double dedup_time_tol,
                         double dedup_point_tol, int max_hits, int newton_max_iter, int cap,
                         int* count, double* hit_t, double* hit_x, double* hit_p,
                         void* workspace, long long workspace_bytes, void* stream);

/* Filters of a dense-output state grid states[n][m][6] (row-major per
 * trajectory), as the manifold computation applies them to every arc:
 * min_r1[i], min_r2[i] = minimum distance to the primaries, NaN if any
 * sample is NaN (services/manifold.py:416-420), and max_energy_err[i] =
 * _max_rel_energy_error of the samples (common/energy.py:26-74).
 * Any output pointer may be NULL.                                          */
int hh_traj_filters(int n, int m, const double* mu, const double* states, double* min_r1,
                    double* min_r2, double* max_energy_err, void* stream);

/* Stream-ordered plumbing for host loops that keep many small batches in
 * flight (the serving loop): an asynchronous copy between pinned host and
 * device memory, and completion events (created without timing).
 * hh_event_query returns 1 when the event has fired, 0 if not yet, < 0 on
 * error.  No counterpart in the reference.                                 */
int hh_memcpy_async(void* dst, const void* src, unsigned long long bytes, int to_device,
                    void* stream);
/* A stream with a hardware queue of its own (created with an all-CU mask):
 * long-running launches on it hold up no other stream.  NULL on error.     */
void* hh_stream_create_dedicated(void);
int hh_stream_destroy(void* stream);
/* Pinned, device-mapped, coherent host memory (the device pointer equals the
 * host pointer), for results a host loop polls while a kernel runs.        */
int hh_host_alloc(unsigned long long bytes, void** ptr);
int hh_host_free(void* ptr);
void* hh_event_create(void);
void hh_event_destroy(void* event);
int hh_event_record(void* event, void* stream);
int hh_event_query(void* event);

/* ---- correction bank (host code) ---------------------------------------
 * Batched single-shooting corrections as a native state machine: the
 * reference's _NewtonBackend.run (corrector/backends/newton.py:39-143) with
 * _ArmijoLineSearch (stepping/armijo.py:58-164), the residual/Jacobian
 * operators (corrector/operators.py:319-452) and the Halo extra term
 * (services/orbits.py:917-947).  The caller runs the event searches and STM
 * propagations on the GPU and the dense solve (corrector/backends/base.py:
 * 187-245) with NumPy, and feeds the results back.  Per row the stages are
 * event search (residual or all line-search trials) -> STM -> solve ->
 * event search ...  hh_bank_add returns the row, or -HH_ERR_ARGS.          */
#define HH_BANK_CONVERGED 0
#define HH_BANK_NO_CROSSING 1
#define HH_BANK_MAX_ITERATIONS 2
#define HH_BANK_LINE_SEARCH_FAILED 3
#define HH_BANK_RESID 0
#define HH_BANK_JAC 1
#define HH_BANK_STEP 2
#define HH_BANK_ARMIJO 3
#define HH_BANK_DONE 4
void* hh_bank_create(int n_residuals, int n_controls);
void hh_bank_destroy(void* bank);
int hh_bank_size(void* bank);
int hh_bank_add(void* bank, double mu, const double* x0, const int* residual_indices,
                const int* control_indices, const double* target, int halo_term, int forward,
                int max_attempts, double tol, double max_delta, double alpha_reduction,
                double min_alpha, double armijo_c);
int hh_bank_stages(void* bank, int n, const int* rows, int* stage);
/* states whose event searches the rows wait for, grouped by row ([total][6]);
 * with states == NULL only counts[n] and the total are returned            */
int hh_bank_event_inputs(void* bank, int n, const int* rows, int* counts, double* states,
                         int capacity);
/* t[total], xe[total][6], ok[total] of those searches                        */
int hh_bank_feed_events(void* bank, int n, const int* rows, const double* t, const double* xe,
                        const int* ok);
int hh_bank_stm_inputs(void* bank, int n, const int* rows, double* states, double* t_event);
/* phi[n][36] (row-major Phi at the event time) -> J[n][nr][nc], r[n][nr]    */
int hh_bank_jacobians(void* bank, int n, const int* rows, const double* phi, double* J,
                      double* r);
/* Newton steps delta[n][nc] -> capped, trial points, line-search stage      */
int hh_bank_set_steps(void* bank, int n, const int* rows, const double* delta);
int hh_bank_result(void* bank, int row, double* x_full, double* residual_norm, int* iterations,
                   int* n_event_searches, int* n_stm, int* converged, int* error,
                   double* half_period);
/* a finished row whose result has been read; hh_bank_add reuses it         */
int hh_bank_release(void* bank, int row);

/* ---- device-resident corrector -----------------------------------------
 * One correction per 192-thread workgroup, its whole Newton-Armijo loop on
 * the device: replaces the host-driven rounds of the bank above (the
 * reference's _NewtonBackend.run, corrector/backends/newton.py:39-143, with
 * _ArmijoLineSearch, stepping/armijo.py:58-164; residual and Jacobian
 * operators corrector/operators.py:319-452; _solve_delta_dense
 * corrector/backends/base.py:187-245; Halo term services/orbits.py:917-947)
 * for square 2x2 problems (the Halo and Lyapunov configurations).  The
 * event search of every residual is _y_plane_crossing's
 * (poincare/singlehit/backend.py:237-280) with the two windows given as
 * (start, length); the STM is _compute_stm's (corrector/operators.py:
 * 298-307).  Results agree with the bank's to the corrector's tolerance (the
 * Halo term's powers are correctly rounded on the device).
 *
 * problems[n] and results[n] are device (or host-mapped) memory.  Each
 * result is written when its correction ends, its `done` set to 1 last with
 * a system-scope release, so a host polling mapped memory can take it
 * before the launch completes (clear `done` before the launch).  error:
 * HH_BANK_* or HH_CORRECT_INVALID (not 2x2, an index outside 0..5, an
 * unknown event plane, or more than 64 line-search steps).
 * event_plane selects the section of the residual: HH_PLANE_Y (0, the
 * default of a zero-initialised record: _y_plane_crossing, Halo and
 * Lyapunov orbits, services/orbits.py:863, 1254), HH_PLANE_X (1) or
 * HH_PLANE_Z (2: _z_plane_crossing, vertical orbits, orbits.py:1445).     */
#define HH_CORRECT_INVALID 4
#define HH_PLANE_Y 0
#define HH_PLANE_X 1
#define HH_PLANE_Z 2
typedef struct {
    double mu;
    double x0[6];
    double target[6];        /* entries 0..n_residuals-1 */
    double tol, max_delta, alpha_reduction, min_alpha, armijo_c;
    int residual_indices[6]; /* entries 0..n_residuals-1 */
    int control_indices[6];  /* entries 0..n_controls-1  */
    int n_residuals, n_controls;
    int halo, forward, max_attempts, event_plane;
} hh_correction_t;

typedef struct {
    double x[6];             /* the full corrected state */
    double residual_norm;
    double half_period;      /* the event time of the converged state, or NaN */
    int iterations, n_event_searches, n_stm, converged, error;
    int n_steps;             /* accepted integration steps, all searches and STMs */
    int done, reserved;
} hh_correction_result_t;

int hh_correct(int n, const hh_correction_t* problems, hh_correction_result_t* results,
               double window_start, double window_length, double fallback_start,
               double fallback_length, void* stream);
/* _y_plane_crossing (poincare/singlehit/backend.py:237-280, no time guess)
 * for n states in one launch: the window (start, length), then the
 * fallback window for the states without a valid hit, each an alignment
 * propagation to its start (skipped where numpy.isclose(0, start); the
 * alignment honours `forward`, the event search does not) and an event
 * search for y[ev_index] = ev_offset in `direction`.  mu[n], y0[6][n] and
 * the outputs ok[n], t_hit[n] (start + t_rel), y_hit[6][n] and n_steps[n]
 * (accepted steps, may be NULL) are device arrays.                        */
int hh_plane_crossing(int n, const double* mu, const double* y0, int forward, int ev_index,
                      double ev_offset, int direction, double xtol, double gtol,
                      double window_start, double window_length, double fallback_start,
                      double fallback_length, int* ok, double* t_hit, double* y_hit,
                      int* n_steps, void* stream);
/* sizeof of the two records, for bindings that lay them out themselves     */
int hh_correct_sizes(int* problem_bytes, int* result_bytes);
/* Test entry: the device corrector's 2x2 step, delta[n][2] solving
 * J[n][2][2] delta = -r[n][2] as _solve_delta_dense does
 * (corrector/backends/base.py:187-245), and cond[n] = numpy.linalg.cond(J). */
int hh_eval_solve2(int n, const double* J, const double* r, double* delta, double* cond,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HITEN_HIP_H */

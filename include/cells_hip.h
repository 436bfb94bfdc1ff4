/* cells_hip: the C-ABI of the MI355X batched tissue step (libcells_hip.so).
 *
 * Plain C types only: pointers, sizes, int status codes. Device buffers are the caller's
 * (stream-ordered; no allocation on the stepping path). A status other than CELLS_OK leaves a
 * message for cells_last_error().
 *
 * Each entry point names the reference interface it replaces (paths under src/cells/ of the
 * reference, Eidosoma/Cells).
 */
#ifndef CELLS_HIP_H
#define CELLS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CELLS_OK 0
#define CELLS_EINVAL 1   /* a malformed argument or pattern (the reference's ValueError/ConfigError) */
#define CELLS_EHIP 2     /* a HIP runtime error */
#define CELLS_ENOMEM 3   /* device allocation failed at set-up */
#define CELLS_ENODEV 4   /* no GPU: the library never falls back to the CPU */

#define CELLS_METHOD_EULER 0 /* numerics/surface_imex.py:36 imex_euler_face_electromigration@1 */
#define CELLS_METHOD_ROS2 1  /* ROS2 (gamma = 1 + 1/sqrt 2), stage matrix C + gamma dt G */
/* numerics/surface_imex.py:38-45 imex_ars222_implicit_conductances@1 (N2): ARS(2,2,2) with the
 * routes' slope conductances, frozen at half the step, in the operator C + gamma dt G, the junction
 * gates by their exponential update (models/surface_pump_v3.py:3547-3955 imex_step) */
#define CELLS_METHOD_N2 2
/* numerics/surface_imex.py:53-57 imex_ars222_face_electromigration@1: on cells over a held bath
 * its implicit part is empty (no faces), a two-stage explicit tableau with weights
 * (delta, 1 - delta) */
#define CELLS_METHOD_ARS222 3
/* numerics/fixed_step.py:25-71 rk4_checked_step, the step of an empty tableau
 * (models/surface_pump_v3.py:3989-3998); worlds without water */
#define CELLS_METHOD_RK4 4
/* Or-ed into the method: factor each stage matrix with the reference's summation order (every
 * entry's sources folded in adjacent pairs, numerics/sparse_cholesky.py:474-490), bit for bit
 * the reference's factor, instead of the default multifrontal supernodal factor (same pattern and
 * positions, sums grouped by the elimination tree; equal to rounding). */
#define CELLS_FACTOR_PAIRWISE 0x100

/* The packed worlds: lattices of cells on one held bath each, back to back. The reference's
 * counterpart is the layout pytree of models/surface_pump_v3.py:140-173 (LAYOUT_TYPE) built by
 * initial_values (models/surface_tissue.py:106) for these worlds. Host pointers. */
typedef struct cells_world {
    int32_t S, W, water, calcium;
    int64_t ncell, nedge;
    const int64_t* cell_off; /* [W+1] */
    const int64_t* edge_off; /* [W+1] */
    double valence[4];
    double bath_c[4];
    double temperature, bath_V;
    double atp, adp, pi, km_na, km_k, km_atp, dg0;
    double ca_km_ca, ca_km_atp, ca_dg0;
    double gate_threshold, gate_minimum;
    double water_K, water_P0, water_vmin, water_vmax, bath_pressure;
    const double* vol0;
    const double* cap;
    const double* pam;     /* [S][ncell] */
    const double* nak_cap;
    const double* ca_cap;
    const double* n0;      /* [S][ncell] */
    const double* q0;
    const double* lpa;
    const double* imperm;
    const int64_t* snd;
    const int64_t* rcv;
    const double* paj;     /* [S][nedge] */
    const int64_t* inc_ptr;
    const int64_t* inc_edge;
    const int32_t* inc_out;
} cells_world;

/* Held availabilities (definitions/surface_pump_v2_spec.py:118-123, Control), one per kind. */
typedef struct cells_control {
    double membrane, junction, pump, calcium_pump, water;
} cells_control;

typedef struct cells_symbolic cells_symbolic;

/* Plan options read by cells_ensemble_create and cells_field_create (no counterpart in the
 * reference; the schedule knobs the autotuner of cells_amd/autotune.py sweeps): the lattice
 * plans' relaxed supernodes (largest joined front, 1..128, default 128; explicit zeros per
 * joined column, default 64) and the grid plans' (explicit zeros per joined column, default 3;
 * the depth of the tree's cut: the first holding at least grid_subtrees fronts, default 32);
 * widen_groups (0 or 1, default 1): a lattice level of few fronts takes wider lane groups;
 * largest_first (0 or 1, default 1): the per-world launches take the worlds largest first.
 * Process-wide; set them between creates, not during one. */
int cells_set_plan_options(int32_t relax_rows, int32_t relax_zeros, int32_t grid_relax_zeros,
                           int32_t grid_subtrees, int32_t widen_groups, int32_t largest_first);
int cells_get_plan_options(int32_t* relax_rows, int32_t* relax_zeros, int32_t* grid_relax_zeros,
                           int32_t* grid_subtrees, int32_t* widen_groups, int32_t* largest_first);
typedef struct cells_ensemble cells_ensemble;

const char* cells_last_error(void);

/* ---- host symbolic analysis: numerics/sparse_cholesky.py ---- */

/* minimum_degree (sparse_cholesky.py:137-162): order_out[n] */
int cells_minimum_degree(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols,
                         int32_t* order_out);
/* bind (sparse_cholesky.py:165-215); order may be NULL (minimum degree). */
int cells_bind(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols,
               const int32_t* order, cells_symbolic** out);
int cells_symbolic_sizes(const cells_symbolic* s, int64_t* nnz, int32_t* levels);
/* The Symbolic's integer arrays (sparse_cholesky.py:69-85): order, rank [n]; column_start [n+1];
 * entry rows and columns as nodes [nnz]; parent, level [n] by rank. Any pointer may be NULL. */
int cells_symbolic_export(const cells_symbolic* s, int32_t* order, int32_t* rank,
                          int64_t* column_start, int32_t* row_node, int32_t* column_node,
                          int32_t* parent, int32_t* level);
void cells_symbolic_free(cells_symbolic* s);
/* Relaxed supernodes of a bound pattern (no counterpart in the reference, whose factor is
 * column by column): a fundamental supernode whose parent supernode starts at its parent column
 * (its next column) and has no other child joins its parent's front when the front's order stays
 * within max_rows and the supernode's columns gain at most max_zeros explicit zeros; chains join
 * from the top down. The joined columns take the front's dense trapezoid of rows: pairs receives
 * those entries (node pairs i, j, flattened) and count their number; binding the pattern plus
 * them with s's order gives the same elimination tree with the joined fronts as its fundamental
 * supernodes (the multifrontal plan of cells_ensemble_create uses cells_set_plan_options'
 * relax_rows and relax_zeros).
 * pairs may be NULL to ask for the count; capacity is in pairs. Host only. */
int cells_relaxed_padding(const cells_symbolic* s, int32_t max_rows, int32_t max_zeros, int32_t* pairs,
                          int64_t capacity, int64_t* count);

/* bind's capacitance topology (numerics/surface_field.py:253-340, representative :238-250; the
 * owner checks of numerics/surface_transport.py:194-216 it relies on): nodes are the n cells then
 * the o outside nodes; each active membrane joins its cell and its outside node into one capacitor
 * component (union-find). A component holding a reservoir (a voltage clamp) is held; any other is
 * floating and fixes its representative: the gauge where it holds it, else its lowest outside
 * node. The gauge rule (no gauge with reservoirs, exactly one active gauge without), and every
 * floating component reaching a clamp or the gauge through conductive faces (active, with a
 * charged species of positive face factor) or return paths (outside-node pairs, the forms faces'),
 * are refused otherwise, CELLS_EINVAL with the reference's messages. Outputs (any may be NULL but
 * counts): component [n+o] (components numbered by their lowest active node, -1 inactive),
 * floating [n+o] (floating component index, -1 where held or inactive), fixed [n+o] (inactive,
 * reservoir or representative), offset_free [floating components] (its representative is not the
 * gauge; capacity n+o suffices), counts[0] components, counts[1] floating components. */
int cells_capacitor_topology(int32_t n, int32_t o, int64_t m, const int32_t* membrane_cell,
                             const int32_t* membrane_outside, const uint8_t* membrane_active, const uint8_t* cell_active,
                             const uint8_t* outside_active, const uint8_t* reservoir, const uint8_t* gauge, int64_t f,
                             const int32_t* face_sender, const int32_t* face_receiver, const uint8_t* face_conductive,
                             int64_t r, const int32_t* return_paths, int32_t* component, int32_t* floating,
                             uint8_t* fixed, uint8_t* offset_free, int32_t* counts);

/* product_plan (numerics/sparse_cholesky.py:444-464): the matrix-vector product's plan on the
 * symmetric pattern of n nodes with off-diagonal coordinates (rows, cols) in either or both
 * triangles and the whole diagonal. Row i's entries are its neighbours and itself in ascending
 * column order, rows in order (a value array on the pattern is row-major): row_start [n+1],
 * column [nnz], each row's diagonal entry [n]. The size call gives nnz. Malformed patterns are
 * CELLS_EINVAL, as _neighbours (:122-135) refuses them. */
int cells_product_plan_size(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, int64_t* nnz);
int cells_product_plan(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, int64_t* row_start,
                       int32_t* column, int64_t* diagonal);
/* product (numerics/sparse_cholesky.py:467-471): y = A x on the device for `count` systems on one
 * plan (system b: values + b values_stride, x and y + b vector_stride), each row's products formed
 * one by one and summed in adjacent pairs, step by step (the reference's fold, :311-322), so the
 * result is the reference's bit for bit. row_start, column, values, x and y are device pointers;
 * y must not overlap x or values; stream-ordered. */
int cells_product(void* stream, int32_t n, const int64_t* row_start, const int32_t* column, const double* values,
                  const double* x, double* y, int32_t count, int64_t values_stride, int64_t vector_stride);

/* ---- device ensemble: checked_step under the execution layer's loops ---- */

/* Binds the worlds once (models/surface_pump_v3.py:6851 model_for; numerics/surface_field.py:371
 * _bind_sparse): uploads parameters, runs the symbolic analysis once per lattice shape, and
 * allocates the stepping workspace on `device`. */
int cells_ensemble_create(const cells_world* w, int32_t method, int32_t device,
                          cells_ensemble** out);
void cells_ensemble_destroy(cells_ensemble* e);
/* Doubles in one state vector (cell, edge and world fields; models/surface_pump_v3.py:122-126). */
int64_t cells_state_size(const cells_ensemble* e);

/* nsteps checked steps of every world (models/surface_pump_v3.py:3989 checked_step, run as
 * execution/vector_windows.py:348-413 runs it under vmap, or as execution/batch.py:321-357's
 * lax.map when batched = 0: one world per launch). y_in, y_out: device [state size].
 * active, ok_out: device int32 [W]. A world stops at its first refused step and keeps its last
 * accepted state (refusals="world"); ok_out[w] = 0 then. batched = 1 replays the captured
 * step graph. Stream-ordered on `stream` (a hipStream_t, NULL for the null stream). The steps run
 * on y_out itself (with one ensemble-owned buffer beside it): y_out may equal y_in (in place, no
 * copy of the state) or must not overlap it (one copy in). A new y_out pointer captures the step
 * graph again. */
int cells_step(cells_ensemble* e, void* stream, const double* y_in, double* y_out, double dt,
               int32_t nsteps, const cells_control* control, const int32_t* active,
               int32_t* ok_out, int32_t batched);

/* Per-kernel timing of one step's launches (no counterpart in the reference, whose profile is
 * scripts/profile_step.py's per-piece timing): runs `reps` steps from y with a HIP event after
 * every launch on `stream`, and returns each launch's mean duration (ms) and kernel name (32
 * bytes each) in launch order. */
int cells_kernel_times(cells_ensemble* e, void* stream, const double* y, double dt,
                       const cells_control* control, int32_t reps, int32_t capacity,
                       double* ms_out, char* names_out, int32_t* count_out);

/* Explicit rates of every field at y (models/surface_pump_v2.py:370-450 explicit_rates at the
 * state's own potentials): rates_out has the state's layout (world fields: bath supply rates);
 * valid_out: device int32 [W]. */
int cells_rates(cells_ensemble* e, void* stream, const double* y, const cells_control* control,
                double* rates_out, int32_t* valid_out);

/* The readouts' energy rates per world at y (models/surface_pump_v3.py:2221-2262 readouts'
 * chemical_energy_rate_pJ_s, capacitor_energy_rate_pJ_s, dissipation_pJ_s, reservoir_power_pJ_s,
 * electrode_power_pJ_s, the metabolic work rate, energy_balance_pJ_s and electrode_current_A), for
 * worlds without water: out (device, double [W][CELLS_ENERGY_RATES]) in that order. */
#define CELLS_ENERGY_RATES 8
int cells_energy_rates(cells_ensemble* e, void* stream, const double* y, const cells_control* control,
                       double* out);

/* Device row checks of a state against its predecessor (models/surface_pump_checks.py:112
 * row_checks over models/surface_pump.py:685-696's LEDGERS with the water operator's,
 * operators/osmotic_water.py:240; bounds models/surface_pump.py:752-800 ledger_limits): the bits
 * of failed_out[w] name the failed checks (the reference's check names in parentheses), and
 * ratios_out[CELLS_LEDGER_RATIOS w + k] is the largest |residual| / bound of ledger k (CELLS_RATIO_*).
 * steps: accepted steps behind y (the rounding allowance 2.5 eps steps x magnitudes); the energy
 * ledger's relative bound is IMEX Euler's 2e-2 (surface_pump.py:63) for both methods. */
#define CELLS_CHECK_DOMAIN 1               /* domain */
#define CELLS_CHECK_LOCAL 2                /* local: material per cell and species */
#define CELLS_CHECK_TOTAL 4                /* total: cells against the bath supplies */
#define CELLS_CHECK_GLOBAL_CHARGE 8        /* global_charge: electrode and supply charge */
#define CELLS_CHECK_DISSIPATION 16         /* dissipation negative or decreasing */
#define CELLS_CHECK_WATER 32               /* water_local, water_outside, water_total */
#define CELLS_CHECK_OUTSIDE 64             /* outside: the held bath's material */
#define CELLS_CHECK_CAPACITOR 128          /* charge: each capacitor's node charge */
#define CELLS_CHECK_ELECTRODE 256          /* electrode: electrode charge against plate charge */
#define CELLS_CHECK_ENERGY 512             /* energy */
#define CELLS_CHECK_RESERVOIR_WORK 1024    /* reservoir_work */
#define CELLS_CHECK_ELECTRODE_WORK 2048    /* electrode_work */
#define CELLS_CHECK_METABOLIC_WORK 4096    /* metabolic_work (Na/K, then Ca) */
#define CELLS_CHECK_BATH_WATER_WORK 8192   /* bath_water_work */
#define CELLS_LEDGER_RATIOS 12
enum {
    CELLS_RATIO_LOCAL, CELLS_RATIO_TOTAL, CELLS_RATIO_GLOBAL_CHARGE, CELLS_RATIO_WATER,
    CELLS_RATIO_OUTSIDE, CELLS_RATIO_CAPACITOR, CELLS_RATIO_ELECTRODE, CELLS_RATIO_ENERGY,
    CELLS_RATIO_RESERVOIR_WORK, CELLS_RATIO_ELECTRODE_WORK, CELLS_RATIO_METABOLIC_WORK,
    CELLS_RATIO_BATH_WATER_WORK
};
int cells_device_checks(cells_ensemble* e, void* stream, const double* previous, const double* y,
                        double steps, double material_atol, int32_t* failed_out,
                        double* ratios_out);

/* Reverse of one ROS2 step at y for every world (no counterpart in the reference, which
 * differentiates only forward, analysis/stationary.py:505-517): from the cotangent ybar_out on the
 * step's output amounts and gates, writes ybar_in's amounts and gates (other fields untouched) and
 * adds the parameter cotangents pam_bar [S][ncell] (membrane area x permeability), cap_bar [ncell]
 * (Na/K capacity), paj_bar [S][nedge] (junction area x permeability). The stage solves reuse the
 * step's Cholesky factor (the stage matrix is symmetric). 3-ion worlds without water; ROS2 only. */
int cells_step_vjp(cells_ensemble* e, void* stream, const double* y, const double* ybar_out,
                   double dt, const cells_control* control, double* ybar_in, double* pam_bar,
                   double* cap_bar, double* paj_bar, int32_t* valid_out);

/* cells_step_vjp for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference. The
 * same kernels and results as cells_step_vjp, whose first call allocates the adjoint workspace;
 * this one takes it from a fixed 64 MiB device array of the library (per GPU, held until the
 * process ends) and calls no allocating routine. An ensemble whose workspace exceeds it
 * (4 + 3 S doubles per cell plus 3 + 3 S per junction, each buffer 256-byte aligned) is
 * CELLS_EINVAL. Every call on one GPU shares that array: calls that overlap on a GPU (other
 * streams, other threads) race; calls ordered on one stream do not. */
int cells_step_vjp_noalloc(cells_ensemble* e, void* stream, const double* y, const double* ybar_out,
                           double dt, const cells_control* control, double* ybar_in, double* pam_bar,
                           double* cap_bar, double* paj_bar, int32_t* valid_out);

/* The ledger checks of cells_device_checks as each step's epilogue (execution/vector_windows.py:
 * 374-410 with models/surface_pump_checks.py:112): for the cells_step calls that follow, every
 * live world's candidate state is checked against its current one before the commit, and a world
 * that fails refuses the step (keeps its state and stops, as a domain refusal does) with its
 * failure bits (the bits of cells_device_checks) or-ed for the call. steps_before is the ledgers'
 * step count of the next input state (the rounding allowance grows with it); it advances by
 * nsteps with each call. enabled = 0 turns the epilogue off. */
int cells_step_checks(cells_ensemble* e, int32_t enabled, double material_atol, int64_t steps_before);

/* The failure bits per world (int32, device) of the last cells_step call: the ledger bits of
 * cells_device_checks when the epilogue is on, and the step's own refusal reasons always:
 * CELLS_REFUSE_RESIDUAL (a stage solve's charge residual above its bound,
 * numerics/surface_imex.py:269-278 as models/surface_pump.py:325-330 checks it) and
 * CELLS_REFUSE_EXPLICIT_RATE (an explicit stage's membrane rate times dt above 1,
 * models/surface_pump_v3.py:3640-3642 with membrane_rate :4588). */
#define CELLS_REFUSE_RESIDUAL (1 << 16)
#define CELLS_REFUSE_EXPLICIT_RATE (1 << 17)
int cells_step_check_bits(cells_ensemble* e, void* stream, int32_t* bits_out);

/* The definition-time rate bounds (1/s) of the bound worlds at their declared state, which every
 * cells_step checks against its dt before it launches anything (the reference's step checks,
 * models/surface_pump.py:402-555 StepChecks with explicit_rate_bound and junction_gate_rate_bound,
 * surface_pump_v2.py:583-617 calcium_rate_bound, surface_pump_v3.py:4981-5017
 * implicit_conductance_rate_bound for the method whose route slopes are implicit): rates_out[0]
 * the explicit-rate bound (dt x it <= 1), [1] the junction gates' (dt x it <= 0.125; IMEX Euler
 * only, 0 otherwise), [2] the Ca-ATPase's (dt x it <= 1). A dt beyond any of them is
 * CELLS_EINVAL with the reference's ConfigError message. */
int cells_step_bounds(const cells_ensemble* e, double* rates_out);

/* The stage solves' residual bound (ROS2): a cell's |(C + gamma dt G) dphi - q| must stay within
 * atol + rtol x its row's magnitude, else the world refuses the step with
 * CELLS_REFUSE_RESIDUAL. Defaults: the reference's CHARGE_ATOL_C 3e-25 and CHARGE_RTOL 2e-12
 * (numerics/surface_field.py:47-48). */
int cells_set_stage_tolerance(cells_ensemble* e, double atol, double rtol);

/* Step tape for the adjoint (no counterpart in the reference, which has no adjoint). Reserves
 * `slots` step records in device memory (0 frees them); each record keeps what the reverse of
 * one ROS2 step reads: the input state, the stage-1 state, the factor values with the Jacobi
 * scales, the route slopes and both stage increments. bytes_out (may be NULL) receives the size.
 * Synchronises the device. */
int cells_tape_reserve(cells_ensemble* e, int32_t slots, int64_t* bytes_out);

/* One batched ROS2 step, y_in -> y_out exactly as cells_step with nsteps = 1 (same kernels, same
 * results), that also writes its record into tape slot `slot`. */
int cells_step_record(cells_ensemble* e, void* stream, const double* y_in, double* y_out, double dt,
                      const cells_control* control, int32_t slot, int32_t* ok_out);

/* cells_step_vjp from the record in tape slot `slot` (nothing recomputed forward: the recorded
 * factor serves both transposed stage solves). valid_out receives the recorded step's flags. */
int cells_step_vjp_recorded(cells_ensemble* e, void* stream, int32_t slot, const double* ybar_out,
                            double dt, const cells_control* control, double* ybar_in,
                            double* pam_bar, double* cap_bar, double* paj_bar, int32_t* valid_out);

/* Stationary analysis (analysis/stationary.py:473-517,919-1080; the reference solves one world
 * on the host with SciPy GMRES). Damped Newton for the steady state of every world's amounts and
 * junction gates, in the reference's coordinates (amounts as mM over the cell volume, gates as
 * fractions, stationary.py:117-123), stopping when a full correction is at most `tolerance` in
 * those units (the reference's 1e-10). Linear systems: restarted GMRES(30), right-preconditioned
 * by the per-cell S x S blocks of the Jacobian (and each gate's diagonal), Jacobian products from
 * forward tangents of the rate laws; with fp32_inner the Krylov basis is FP32 and each restart's
 * residual b - J x is FP64 (iterative refinement to linear_rtol). Converged worlds leave the
 * batch between Newton iterations. Worlds without water. y_out holds y0 with the steady amounts
 * and gates; per world converged_out (int), step_out (the last correction's max norm) and
 * iterations_out may be NULL; all are device pointers. */
typedef struct cells_stationary_options {
  double tolerance;    /* Newton: full correction max norm (1e-10) */
  int32_t max_newton;  /* Newton iterations (60) */
  int32_t max_cycles;  /* GMRES restarts per Newton step (20) */
  double linear_rtol;  /* |b - J x|_2 <= linear_rtol |b|_2 (1e-8) */
  int32_t fp32_inner;  /* 1: FP32 Krylov under FP64 refinement; 0: all FP64 */
} cells_stationary_options;

typedef struct cells_stationary_report {
  int32_t converged;          /* worlds converged */
  int32_t newton_iterations;  /* iterations of the longest-running world */
  int64_t krylov_steps;       /* Arnoldi steps over all worlds */
  double worst_step;          /* largest final correction */
} cells_stationary_report;

int cells_stationary(cells_ensemble* e, void* stream, const double* y0, double* y_out,
                     const cells_control* control, const cells_stationary_options* opt,
                     int32_t* converged_out, double* step_out, int32_t* iterations_out,
                     cells_stationary_report* report);

/* The stationary analysis' residual f(y) (rates of the coordinates, amounts then gates, S ncell +
 * nedge doubles) and its tangent product J(y) v, for tests. */
int cells_stationary_residual(cells_ensemble* e, void* stream, const double* y,
                              const cells_control* control, double* r_out, int32_t* valid_out);
int cells_stationary_jvp(cells_ensemble* e, void* stream, const double* y, const cells_control* control,
                         const double* v, double* out);

/* Vmem misfit of one sample row: loss_out[w] += weight sum_i (V_i - target_i)^2 over world w's
 * cells, and (when ybar is not NULL) its cotangent on the amounts added to ybar. */
int cells_vmem_loss(cells_ensemble* e, void* stream, const double* y, const double* target,
                    double weight, double* loss_out, double* ybar);

/* Replaces the per-cell and per-edge parameters the calibration learns (device or host
 * pointers; NULL keeps one): pam [S][ncell], nak_cap [ncell], paj [S][nedge]. */
int cells_set_parameters(cells_ensemble* e, void* stream, const double* pam, const double* nak_cap,
                         const double* paj);

/* The batched numeric factor and solve on the stage pattern (numerics/sparse_cholesky.py:474-514):
 * values: device [factor size] (each world's assembled entries at its factor positions), factored
 * in place; rhs/x: device [ncell] in node order, one system per world. The factor's layout is its
 * plan's: the bound pattern's columns (the reference's) in the reference-order mode; in the
 * multifrontal mode the relaxed pattern's, whose chains of supernodes carry explicit zeros
 * (cells_world_factor_positions locates any entry). */
int64_t cells_factor_size(const cells_ensemble* e);
int cells_factor(cells_ensemble* e, void* stream, double* values);
int cells_solve(cells_ensemble* e, void* stream, const double* values, const double* rhs,
                double* x);
/* Totals of the bound patterns (without the relaxed supernodes' zeros): factor entries over all
 * worlds, sum over worlds of sum_j c_j^2 (the reference's factorisation cost, c_j the entries of
 * column j), the most column levels of the plans, and the number of distinct lattice shapes. */
int cells_plan_stats(const cells_ensemble* e, int64_t* nnz_total, double* flops_total,
                     int32_t* max_levels, int32_t* shapes);

/* ---- spatial worlds: the cells' lattice in an extracellular finite-volume grid ----
 * The reference's counterpart is a resolved world with outside squares and faces
 * (numerics/surface_transport.py's face routes, numerics/surface_field.py:371 _bind_sparse's
 * operator pattern) stepped by numerics/surface_imex.py:36 imex_euler_face_electromigration@1
 * (models/surface_pump.py:292-362); the reference's sparse field stops at 65 536 nodes
 * (surface_field.py:84), below a 256 x 256 grid. Every world shares one grid of side x side nodes
 * (row-major, the outer ring held at the bath: cells_world's bath_c, potential 0) with faces
 * between 4-neighbours: horizontal faces r (side - 1) + c between (r, c) and (r, c + 1), then
 * vertical faces side (side - 1) + r side + c between (r, c) and (r + 1, c), the first node the
 * sender. Each cell's membrane faces one grid node instead of the bath. */
typedef struct cells_field_grid {
    int32_t side;
    int32_t pad;
    const int32_t* cell_node;  /* [ncell] world-local grid node of each cell's membrane */
    double node_volume;        /* pL, every node */
    double face_factor[4];     /* D_s A / L of every face, m^3/s */
    const double* node_n0;     /* [S][side*side] reference amounts, fmol (every world's) */
    const double* node_q0;     /* [W][side*side] fixed node charges, C */
} cells_field_grid;

typedef struct cells_field cells_field;

/* Binds spatial worlds (3 species, no water, no Ca-ATPase; one lattice shared by every world):
 * the cells as cells_ensemble_create binds them, and the operator's plan: the interior grid
 * nodes' 5-point pattern (each cell's capacitor eliminated into its node: the operator's node
 * rows (C + dt L) phi = q become dt L phi = q_node + q_cell), ordered by nested dissection,
 * fundamental supernodes, a multifrontal schedule. The operator (face weights, factor) is rebuilt
 * at a world's first step and then every refresh_every steps (the reference rebuilds it every
 * step, models/surface_pump.py:300-306). */
int cells_field_create(const cells_world* cells, const cells_field_grid* grid, int32_t refresh_every,
                       int32_t device, cells_field** out);
void cells_field_destroy(cells_field* f);
/* Doubles in one spatial state: cell fields (amount increments [S], membrane transfers [S],
 * membrane heat, Na/K cycles, pump heat, metabolic work) x ncell, edge fields (junction transfers
 * [S], junction heat, gate) x nedge, node fields (amount increments [S], potential) x W side^2,
 * face fields (transfers [S], heat) x W F, world fields (ring supplies [S], reservoir work) x W. */
int64_t cells_field_state_size(const cells_field* f);
/* nsteps IMEX Euler steps with implicit face electromigration of every world (models/
 * surface_pump.py:292-362): the step-start potentials (the field solve, cells_field_potentials,
 * or with the field solve off the state's node potentials: the last accepted step's operator
 * potentials); explicit membranes, junctions, pumps, gates and face diffusion at them; the stage
 * charge; the operator solve; the face drift with the operator's weights; face heat and
 * reservoir work; domain checks; where(valid, new, old). A world whose field solve misses its
 * target refuses the step. Buffers and refusals as cells_step. */
int cells_field_step(cells_field* f, void* stream, const double* y_in, double* y_out, double dt,
                     int32_t nsteps, const cells_control* control, const int32_t* active,
                     int32_t* ok_out);
/* The next step rebuilds the operator. */
int cells_field_refresh(cells_field* f);
/* The step-start potentials: field_solve = 1 (the default) solves the field at each step's start
 * as the reference's readouts do (numerics/surface_field.py:614-855 solve: each cell and its node
 * a floating capacitor component, q_i / C_i on top of one offset per interior node, the offsets
 * by current continuity through the faces and junctions to max |r| <= 0.01 (1e-23 + 2e-12 max
 * current)), by at most max_iterations (1..64, default 8) chord steps with the operator's factor;
 * field_solve = 0 takes the state's node potentials instead (the last operator solve's). */
int cells_field_set_start(cells_field* f, int32_t field_solve, int32_t max_iterations);
/* The field solve at state y (device), alone: phi_out [W][side^2] the node potentials (a cell's is
 * its node's plus q_i / C_i), iterations_out [W] the chord steps taken, ok_out [W] 1 where the
 * solve met its target and the faces' and nodes' domain held (each may be NULL but phi_out). The
 * reference's counterpart is models/surface_tissue.py:306 field_values (its potentials). Builds
 * the operator when the next step would; the refresh schedule does not advance. */
int cells_field_potentials(cells_field* f, void* stream, const double* y, double dt,
                           const cells_control* control, double* phi_out, int32_t* iterations_out,
                           int32_t* ok_out);
/* The field solve's energy terms at state y (device): the field solve as cells_field_potentials,
 * then with those potentials transport alone (membranes into their nodes, junctions, faces by
 * log-mean electrodiffusion; no pumps) priced as the reference's field readout does
 * (numerics/surface_field.py:792-835 energy terms). out [W][11], per world: chemical, capacitor,
 * dissipation, reservoir power, electrode power, energy balance, electrode current, capacitor
 * energy, membrane heat, junction heat, face heat. The balance (chemical + capacitor +
 * dissipation - reservoir - electrode) vanishes to rounding at the field solve's potentials.
 * ok_out [W] (may be NULL): 1 where the field solve met its target and the domain held. */
int cells_field_energy_rates(cells_field* f, void* stream, const double* y, double dt,
                             const cells_control* control, double* out, int32_t* ok_out);
/* The operator plan: factor entries per world and flops per factor (sum of column counts squared)
 * of the bound pattern, and the supernodes, largest front and supernodal tree levels of the plan
 * (relaxed supernodes: the chains joined with their explicit zeros). */
int cells_field_plan_stats(const cells_field* f, int64_t* nnz, double* flops, int32_t* fronts,
                           int32_t* max_front, int32_t* levels);
/* Per-kernel timing of `reps` steps from y (operator refreshed at the first): each kernel's total
 * per step over the steps that ran it (HIP events), in first-launch order. */
int cells_field_kernel_times(cells_field* f, void* stream, const double* y, double dt,
                             const cells_control* control, int32_t reps, int32_t capacity,
                             double* ms_out, char* names_out, int32_t* count_out);
/* Host copy of world w's plan positions: diag_pos [n] by node, edge_pos [E] by edge, and the
 * offset of its factor slice. */
int cells_world_positions(const cells_ensemble* e, int32_t w, int64_t* diag_pos,
                          int64_t* edge_pos, int64_t* factor_offset);
/* Positions in world w's factor slice of the entries (rows[x], cols[x]) (world-local nodes, either
 * triangle), -1 where the plan's pattern has none. */
int cells_world_factor_positions(const cells_ensemble* e, int32_t w, const int32_t* rows,
                                 const int32_t* cols, int64_t count, int64_t* out);

#ifdef __cplusplus
}
#endif
#endif

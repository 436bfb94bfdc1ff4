/* Oracle (test infrastructure only): FP64 CPU restatement of the reference's batched tissue step.
 *
 * Only tests, the smoke check and the benchmark's checker legs (its CPU baseline and its parity
 * checks, outside every timed region) may load this library.
 * The product (cells_amd/csrc, include/cells_hip.h) never links it.
 *
 * The world description below is the packed ensemble layout the product also takes
 * (include/cells_hip.h, struct cells_world); the two structs carry the same fields in the same
 * order so one Python description feeds both.
 */
#ifndef CELLS_REF_H
#define CELLS_REF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t S;        /* species (3: Na K Cl, or 4: Na K Cl Ca) */
    int32_t W;        /* worlds */
    int32_t water;    /* 1: Strang water half-steps around the ionic step */
    int32_t calcium;  /* 1: the Ca-ATPase on species 3 */
    int64_t ncell, nedge;
    const int64_t* cell_off; /* [W+1] */
    const int64_t* edge_off; /* [W+1] */
    double valence[4];
    double bath_c[4];        /* held bath, mM */
    double temperature, bath_V;
    /* Na/K-ATPase and Ca-ATPase constants (maintained pools) */
    double atp, adp, pi, km_na, km_k, km_atp, dg0;
    double ca_km_ca, ca_km_atp, ca_dg0;
    /* junction gate */
    double gate_threshold, gate_minimum;
    /* water: modulus, offset, volume bounds as fractions of the reference volume, bath pressure */
    double water_K, water_P0, water_vmin, water_vmax, bath_pressure;
    /* per cell [ncell] (species-major [S][ncell] where marked) */
    const double* vol0;   /* pL */
    const double* cap;    /* F */
    const double* pam;    /* [S][ncell] membrane area x permeability, m^3/s */
    const double* nak_cap;/* fmol cycles/s */
    const double* ca_cap; /* fmol cycles/s */
    const double* n0;     /* [S][ncell] reference amounts, fmol */
    const double* q0;     /* initial cell charge, C */
    const double* lpa;    /* water conductance, pL/(Pa s) */
    const double* imperm; /* impermeant osmolytes, fmol */
    /* per edge [nedge] (global cell indices, sender < receiver) */
    const int64_t* snd;
    const int64_t* rcv;
    const double* paj;    /* [S][nedge] junction area x permeability, m^3/s */
    /* per cell incidence: out-edges (cell is sender) ascending, then in-edges ascending */
    const int64_t* inc_ptr;  /* [ncell+1] */
    const int64_t* inc_edge; /* [2*nedge] */
    const int32_t* inc_out;  /* [ncell] number of out-edges at the head of each list */
} cells_world;

typedef struct {
    double membrane, junction, pump, calcium_pump, water;
} cells_control;

/* State layout: field-major SoA blocks, cells then edges then worlds. */
typedef struct {
    int32_t cell_fields, edge_fields, world_fields;
    /* cell field indices */
    int32_t dn, mtr, mdiss, pcyc, pdiss, pwork, ccyc, cdiss, cwork, dvol, wtr, wdiss;
    /* edge field indices */
    int32_t jtr, jdiss, gate;
    /* world field indices */
    int32_t sup, echg, rwork, ework, wsup, wwork;
    int64_t cell_base, edge_base, world_base, total;
} cells_state_layout;

void ref_state_layout(const cells_world* w, cells_state_layout* out);

/* Explicit rates of every state field at the given cell potentials (V), one world at a time.
 * rates has the state's layout. gm [S][ncell] and gj [S][nedge] receive the routes' slopes
 * dJ/d(dv) (may be NULL). Returns 1 when every rate is valid. */
int ref_rates(const cells_world* w, int32_t world, const double* y, const double* phi,
              const cells_control* c, double* rates, double* gm, double* gj);

/* Cell potentials from the state's charges: V_bath + q_i / C_i. */
void ref_potentials(const cells_world* w, int32_t world, const double* y, double* phi);

/* One checked step of every world (IMEX Euler of the reference, or the ROS2 W-method), with
 * water half-steps when w->water. out receives the new state, or the old one where a world
 * refuses; valid[W] receives each world's acceptance. */
int ref_step_euler(const cells_world* w, const double* y, double* out, double dt,
                   const cells_control* c, int32_t* valid);

/* An explicit stage's membrane charge relaxation rate (1/s) of cell i at state y; IMEX Euler
 * refuses a world where dt times it exceeds 1 at any cell. */
double ref_membrane_rate(const cells_world* w, const double* y, int64_t i, const cells_control* c);

/* The device row checks of world k's state y against its predecessor prev
 * (models/surface_pump_checks.py:112-192 row_checks; balances surface_pump_checks.py:34-109 with
 * surface_pump_v2.py:686-709's pump terms and operators/osmotic_water.py:254-295's water terms;
 * bounds surface_pump.py:752-800 ledger_limits and osmotic_water.py:341-344) for cells on one
 * held bath. Returns the failed checks' bits (the values of include/cells_hip.h's CELLS_CHECK_*);
 * ratios[12] receives each ledger's largest |residual| / bound in CELLS_RATIO_* order. energy_rtol:
 * the energy ledger's relative bound (2e-2, IMEX Euler's). terms[7] (may be NULL) receives the energy ledger's parts,
 * pJ: mixing free energy, capacitor energy, heat, work (metabolic, reservoir, electrode, bath
 * water), water's dilution and elastic energy, the residual (mixing + capacitor + heat - work +
 * water) and the flows its bound scales. */
int ref_ledgers(const cells_world* w, int32_t k, const double* prev, const double* y, double steps,
                double material_atol, double energy_rtol, double* ratios, double* terms);

/* Sparse symbolic structure of one world's stage matrix (the product of oracle/symbolic.py's
 * bind): column_start [n+1] and each entry's row rank [nnz]; order [n]; plus, per lattice entry,
 * its factor position: diag_pos [n] (node order), edge_pos [E] (world edge order). */
typedef struct {
    int64_t n, nnz;
    const int64_t* column_start;
    const int64_t* row_rank;
    const int64_t* order;
    const int64_t* diag_pos;
    const int64_t* edge_pos;
} cells_symbolic;

int ref_step_ros2(const cells_world* w, const cells_symbolic* sym /* [W] */, const double* y,
                  double* out, double dt, const cells_control* c, int32_t* valid);

/* One checked step of every world under N2, the reference's ARS(2,2,2) with implicit slope
 * conductances (numerics/surface_imex.py:38-79 cells.solver.imex_ars222_implicit_conductances@1, as
 * models/surface_pump_v3.py:3547-3955 imex_step runs it for cells on one held bath); sym as for
 * ref_step_ros2 (the operator C + gamma dt G has the same pattern). */
int ref_step_n2(const cells_world* w, const cells_symbolic* sym, const double* y, double* out, double dt,
                const cells_control* c, int32_t* valid_out);

/* The readouts' energy rates of world k at y (models/surface_pump_v3.py:2221-2262): out[8] =
 * chemical, capacitor, dissipation, reservoir power, electrode power, metabolic work rate, energy
 * balance, electrode current. Worlds without water. Returns the rates' validity. */
int ref_energy_rates(const cells_world* w, int32_t k, const double* y, const cells_control* c, double* out);

/* ARS(2,2,2) (numerics/surface_imex.py:53-57) for cells on a held bath, where its implicit part is
 * empty; and the classical RK4 of an empty tableau (numerics/fixed_step.py:25-71, worlds without
 * water), both as ref_step_euler is called. */
int ref_step_ars222(const cells_world* w, const double* y, double* out, double dt, const cells_control* c,
                    int32_t* valid_out);
int ref_step_rk4(const cells_world* w, const double* y, double* out, double dt, const cells_control* c,
                 int32_t* valid_out);

/* numerics/exponential_gates.py:37-51: the lag weight 1 - (1 - e^-z) / z and the gate advance from
 * the start and end kinetics (targets, time constants in s) over dt. */
double ref_lag_weight(double z);
double ref_gate_advance(double gate, double target_start, double tau_start, double target_end,
                        double tau_end, double dt);

/* Numeric factor in the reference's order (left-looking, pairwise folds in ascending column
 * order, below entries times the pivot's reciprocal) and the solve (forward then back, each pivot
 * applied as a multiplication by its reciprocal, the device path of the reference). */
void ref_factor(const cells_symbolic* s, double* values);
void ref_solve(const cells_symbolic* s, const double* factor, const double* rhs_node, double* x_node);

/* ---- spatial worlds: the cells' lattice embedded in an extracellular finite-volume grid ----
 * Every world shares one square grid of side x side nodes (row-major; the outer ring is held at
 * the bath: w->bath_c, potential 0) with faces between 4-neighbours: horizontal faces
 * r (side - 1) + c between (r, c) and (r, c + 1), then vertical faces side (side - 1) + r side + c
 * between (r, c) and (r + 1, c); the first node of a face is its sender. Each cell's membrane
 * faces one grid node instead of the bath. */
typedef struct {
    int32_t side;
    int32_t pad;
    const int32_t* cell_node;  /* [ncell] world-local grid node of each cell's membrane */
    double node_volume;        /* pL, every node */
    double face_factor[4];     /* D_s A / L of every face, m^3/s */
    const double* node_n0;     /* [S][side*side] reference amounts, fmol (shared by the worlds) */
    const double* node_q0;     /* [W][side*side] fixed node charges, C */
} cells_field_grid;

/* Field state layout: cells (amount increments [S], membrane transfers [S], membrane heat, Na/K
 * cycles, pump heat, metabolic work), edges (junction transfers [S], junction heat, gate), nodes
 * (amount increments [S], potential) over W side^2, faces (transfers [S], heat) over W F, worlds
 * (ring supplies [S], reservoir work). */
typedef struct {
    int32_t cell_fields, edge_fields, node_fields, face_fields, world_fields;
    int32_t dn, mtr, mdiss, pcyc, pdiss, pwork, jtr, jdiss, gate, dno, phi, ftr, fheat, sup, rwork;
    int64_t nodes, faces; /* per world */
    int64_t edge_base, node_base, face_base, world_base, total;
} cells_field_layout;

void ref_field_layout(const cells_world* w, const cells_field_grid* g, cells_field_layout* out);

/* The field solve of spatial world k at state y (numerics/surface_field.py:614-855): the
 * capacitor components' base and the damped Newton over their floating offsets for current
 * continuity. ks is world k's offset pattern over the interior nodes (row-major): diag_pos of each,
 * edge_pos each face's entry (-1 where it touches the ring) then each junction's (world-local
 * order). phi [ncell_k + side^2] receives the cells' then the nodes' potentials; iterations (may
 * be NULL) the Newton count. Returns the reference's `converged`. */
int ref_field_start(const cells_world* w, const cells_field_grid* g, const cells_symbolic* ks, int32_t k,
                    const double* y, const cells_control* c, double* phi, int32_t* iterations);

/* One IMEX Euler step with implicit face electromigration of every spatial world
 * (models/surface_pump.py:292-362 with the Euler tableau). The step-start potentials are
 * ref_field_start's where ksym [W] is given, else the state's node potentials (the carried form).
 * sym[k] is world k's operator system,
 * cells then interior nodes (row-major), whose diag_pos lists every node's diagonal entry and
 * whose edge_pos lists each cell's capacitor entry (cells in order) then each face's entry (-1
 * where a face touches the ring). op_weight [W][S][F] and op_factor (the worlds' factors back to
 * back, sym[k].nnz each) hold the frozen operator: refresh = 1 rebuilds them from y's
 * concentrations (step-start weights) before the solve. out receives the new state or the old
 * one where a world refuses; valid[W] the acceptance. */
int ref_field_step(const cells_world* w, const cells_field_grid* g, const cells_symbolic* sym,
                   const cells_symbolic* ksym, const double* y, double* out, double dt, const cells_control* c,
                   int32_t refresh, double* op_weight, double* op_scale, double* op_factor, int32_t* valid);

/* The field solve's energy rates (numerics/surface_field.py:792-835) of spatial world k at state y
 * with the field's potentials phi (its cells', then its nodes', as ref_field_start gives them):
 * out[11] = chemical, capacitor, dissipation, reservoir power, electrode power, energy balance,
 * electrode current, capacitor energy, membrane heat, junction heat, face heat. Returns whether
 * every value is finite. */
int ref_field_energy_rates(const cells_world* w, const cells_field_grid* g, int32_t k, const double* y,
                           const cells_control* c, const double* phi, double* out);

/* The reference's log-mean (numerics/extracellular_field.py:48-62): mean and log ratio. */
void ref_logmean(double left, double right, double* mean, double* log_ratio);

/* Point laws, exposed for closed-form tests. */
double ref_bernoulli(double x);
double ref_bernoulli_slope(double x);
double ref_bernoulli_derivative(double x);
double ref_selective_flux(double c_from, double c_to, double dv, double z, double T, double pa,
                          double avail);
double ref_route_slope(double c_from, double c_to, double dv, double z, double T, double pa,
                       double avail);
double ref_logaddexp(double a, double b);
void ref_nak_kinetics(const cells_world* w, double na_in, double k_out, double mu_na, double mu_k,
                      double cap, double avail, double* cycles, double* energy, double* drive);
int ref_nak_domain(const cells_world* w, double na_in, double k_out, double mu_na, double mu_k,
                   double cap, double avail);
void ref_ca_kinetics(const cells_world* w, double ca_in, double ca_out, double v, double cap,
                     double avail, double* cycles, double* energy, double* drive);
double ref_gate_rhs(double gate, double v, double threshold, double minimum);

#ifdef __cplusplus
}
#endif
#endif

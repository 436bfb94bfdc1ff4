// MI355X (gfx950) kernels and C-ABI of the batched tissue step.
//
// One step of every world (models/surface_pump_v3.py:3989 checked_step of the reference, for
// lattices of cells on one held bath) runs as a fixed sequence of launches over all worlds packed
// back to back:
//   k_edges        junction routes: GHK flux, heat, gate rate, slopes (thread per edge)
//   k_cells        membrane routes, Na/K- and Ca-ATPase, owner sums, stage-matrix diagonal
//                  (thread per cell; a wavefront reads one species of 64 consecutive cells)
//   k_factor       level-scheduled sparse Cholesky of C + gamma dt G (workgroup per world)
//   k_solve_stage  its two triangular solves per stage (workgroup per world, vector in LDS)
//                  (workgroup per world)
//   k_*_update     stage combinations and domain checks (thread per cell and per edge)
//   k_world_*      per-world ledgers, fixed-order reductions (workgroup per world)
//   k_water        Heun water half-steps (workgroup per world)
//   k_commit       per-world acceptance: the state buffer pair flips only where valid
// The state is field-major SoA (cells, then edges, then worlds). Two state buffers alternate per
// world (sel[w]), so a refused world keeps its old state without a copy.
//
// FP64 throughout; compiled with -ffp-contract=off so every product and sum rounds as written
// (the factor's sums follow the reference's pairwise order, numerics/sparse_cholesky.py:32-40).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cells_hip.h"
#include "symbolic.h"

// The workspace of cells_step_vjp_noalloc: a fixed device array in the code object (64 MiB, each
// GPU the library runs on holds its own copy until the process ends), so that entry point never
// allocates. Host code takes its address once per device with hipGetSymbolAddress and uses only
// that address; the name itself is a host-side stand-in, not a device address.
constexpr size_t kScratchDoubles = (size_t)8 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t err_ = (expr);                                                          \
    if (err_ != hipSuccess)                                                            \
      return fail(CELLS_EHIP, std::string(#expr) + ": " + hipGetErrorString(err_));    \
  } while (0)

constexpr double kAvogadro = 6.02214076e23;
constexpr double kF = kAvogadro * 1.602176634e-19;
constexpr double kR = kAvogadro * 1.380649e-23;
constexpr double kMolPerFmol = 1e-15;
constexpr double kChargePerFmol = kF * kMolPerFmol;
constexpr double kMaxAbsX = 50.0;
constexpr double kNakMaxAffinity = 80.0;
constexpr double kNakMaxMu = 1600.0;
constexpr double kCaMaxAffinity = 80.0;
constexpr double kCaMaxVoltage = 0.15;
constexpr double kGateMaxVoltage = 0.120;
constexpr double kWaterMaxRelaxation = 0.05;
constexpr double kExplicitStepBound = 1.0;   // models/surface_pump.py:116 EXPLICIT_STEP_BOUND
constexpr double kGateStepBound = 0.125;     // models/surface_pump.py:114 GATE_STEP_BOUND
constexpr double kChargeAtol = 3e-25;        // numerics/surface_field.py CHARGE_ATOL_C
constexpr double kChargeRtol = 2e-12;        // numerics/surface_field.py CHARGE_RTOL
// Refusal reasons or-ed into a world's failure bits by the step's own kernels (the ledger checks'
// bits are the low ones): a stage solve's charge residual above its bound, an explicit stage's
// membrane rate above the explicit-step bound.
constexpr int kRefuseResidual = 1 << 16;
constexpr int kRefuseExplicitRate = 1 << 17;
constexpr int kBlock = 256;

// ------------------------------------------------------------------------------------------
// Device-side view of the packed worlds (kernel argument, by value).

struct Params {
  int S, W, water, calcium;
  int ncell, nedge;
  double z[4], cb[4], log_cb[4], mu_b[4];
  double T, RT, Vb;
  double nak_drive, log_km_na, log_km_k, nak_atp_sat;  // nak_atp_sat = logaddexp(0, ln Km_atp - ln ATP)
  double ca_dk, ca_drive, log_ca_km, ca_atp_sat;
  double dg0_over_rt_nak_terms;  // delta_g0/RT + ln ADP + ln Pi - ln ATP - ln 1000 (Na/K domain)
  double ca_aff_base;            // delta_g0/RT + ln ADP + ln Pi - ln ATP - ln 1000 (Ca domain)
  double gate_thr, gate_min, log_0065;
  double wK, wP0, vmin, vmax, psib;
  // the energy ledger's relative bound of the step method (surface_pump.py:63-64 ENERGY_RTOL)
  double energy_rtol;
  const int* cell_off;
  const int* edge_off;
  const int* cell_world;
  const int* edge_world;
  const double* vol0;
  const double* cap;
  const double* pam;
  const double* nak_cap;
  const double* ca_cap;
  const double* n0;
  const double* q0;
  const double* lpa;
  const double* imperm;
  const int* snd;
  const int* rcv;
  const double* paj;
  const int* inc_ptr;
  const int* inc_edge;
  const int* inc_out;
  // worlds by cell count, largest first: the order in which whole-ensemble launches of the
  // per-world factor and solves take them (the largest start first, the smallest fill the end)
  const int* world_order;
  // per world: refusal reasons of the current cells_step call (kRefuse* bits), or-ed atomically
  int* refuse_bits;
  // per cell: its position in its world's solve vector (the world's first cell plus its rank)
  const int* cell_rank;
  // the rate kernels' cell potentials when set (N2's stage rates at the operator's potentials),
  // else each cell's from its state's charge
  const double* phi;
  // state layout
  int cf, ef, wf;
  int f_dn, f_mtr, f_mdiss, f_pcyc, f_pdiss, f_pwork, f_ccyc, f_cdiss, f_cwork, f_dvol, f_wtr, f_wdiss;
  int f_jtr, f_jdiss, f_gate;
  int f_sup, f_echg, f_rwork, f_ework, f_wsup, f_wwork;
  long edge_base, world_base, total;
};

struct Ctl {
  double membrane, junction, pump, calcium_pump, water;
};

struct ShapeDev {
  int n, E, levels;
  int nnz;
  long o_cs, o_entry, o_rowptr, o_rowent, o_lev, o_levcol, o_levent, o_node;
  long o_mask;  // offset in the mask pool (unsigned long long words)
  int lds_pt;   // first factor position held in LDS during the factor (nnz: none)
  // multifrontal schedule: supernode records (16 ints each: factor position of the first column,
  // width, front order, update offset, children begin/end (pool-absolute descriptor indices),
  // update-vector offset, first column, pool index of the last column's ranks below its diagonal),
  // per level 6 boundaries (thread, 8-lane, wavefront, workgroup-in-LDS and workgroup-in-HBM
  // fronts, end), children (4 ints each: rows below, rel list start, update offset, update-vector
  // offset), rel lists
  long o_sn, o_snlev, o_child, o_rel;
  int sn_levels;
  long usz;  // doubles of update matrices per world
  long vsz;  // doubles of the solve's update vectors per world
  int big_tri;  // doubles of the largest front held in HBM scratch (0: none)
};

// Multifrontal factor and supernodal solve: one workgroup of kMfThreads per world.
constexpr int kMfThreads = 256;
constexpr int kMfLds = 4096;                               // doubles of LDS per workgroup front (32 KiB, order <= 90)
constexpr int kMfOcta = 8;                                 // lanes of a small group
constexpr int kMfThreadTri = kMfLds / kMfThreads;          // a thread's front: order <= 5
constexpr int kMfOctaTri = kMfLds / (kMfThreads / kMfOcta);  // an 8-lane group's front: order <= 15
constexpr int kMfWaveTri = kMfLds / (kMfThreads / 64);     // a wavefront's front: order <= 44
constexpr int kMfMaxRows = 128;                            // fronts above kMfLds (order <= 128) use HBM
constexpr int kMfBlockLds = 5120;  // a lattice workgroup front's LDS (40 KiB, order <= 100): the factor's
                                   // 4 workgroups per CU (its 4 waves per SIMD) fill the CU's 160 KiB
constexpr int kMfBlockWaveTri = kMfBlockLds / (kMfThreads / 64);  // a lattice wavefront's front: order <= 50
constexpr int kRelaxRows = kMfMaxRows;                     // relaxed supernodes join up to this order,
constexpr int kRelaxZeros = 64;                            // with at most this many zeros per column

// Plan options read when an ensemble or a field is created (cells_set_plan_options; defaults from
// sweeps on the batched and spatial configurations). Process-wide; set them between creates.
struct PlanOptions {
  int relax_rows = kRelaxRows;   // lattice plans: largest joined front
  int relax_zeros = kRelaxZeros; // lattice plans: explicit zeros per joined column
  int grid_relax_zeros = 3;      // grid plans: explicit zeros per joined column
  int grid_subtrees = 32;        // grid plans: fronts at the depth where the tree is cut (at least)
  int widen_groups = 1;          // lattice plans: a level of few fronts takes wider lane groups
  int largest_first = 1;         // per-world launches take the worlds largest first (else in order)
};
PlanOptions g_plan;

// The largest front order whose packed triangle fits `tri` doubles.
constexpr int max_order(int tri) {
  int m = 0;
  while ((m + 1) * (m + 2) / 2 <= tri) ++m;
  return m;
}
constexpr int kMfBatch = 8;                                // global loads in flight per lane
constexpr int kAsmBatch = 4;                               // a front's assembled entries per lane per pass
constexpr int kRecInts = 16;                               // ints per front record in the plan pool
constexpr int kLevelBounds = 6;                            // grid plans, per level: first front of each kind, end

// The lattice plans' lane groups: a front is held by the smallest group whose share of the
// factor's LDS holds its packed triangle (1, 4, 8, 16, 32 or 64 lanes, then the workgroup with the
// front in LDS or in HBM scratch), so that a level's many small fronts spread over many groups
// instead of queueing on a few wide ones.
constexpr int kLatClasses = 8;
constexpr int kLatBounds = kLatClasses + 1;  // lattice plans, per level: first front of each class, end
constexpr int kLatGroupTri[6] = {16, 64, 128, 256, 512, kMfBlockWaveTri};  // factor LDS per group (doubles)
constexpr int kLatGroupLanes[6] = {1, 4, 8, 16, 32, 64};
// the groups of every class within one wavefront fit that wavefront's quarter of the LDS region, in which it
// takes its share of each class in turn
constexpr bool lattice_classes_fit() {
  for (int c = 0; c < 6; ++c)  // a wavefront's groups of each class within its quarter of the region
    if ((64 / kLatGroupLanes[c]) * kLatGroupTri[c] > kMfBlockLds / (kMfThreads / 64)) return false;
  return max_order(kMfBlockLds) <= kMfMaxRows && kMfBlockWaveTri == kMfBlockLds / (kMfThreads / 64);
}
static_assert(lattice_classes_fit(), "a lattice class's lane groups outgrow the factor's LDS region");
static_assert(kLatGroupTri[0] == kMfThreadTri && kLatGroupTri[2] == kMfOctaTri && kLatGroupLanes[2] == kMfOcta,
              "the factor's thread and 8-lane classes take the plan's shares");

// The world a workgroup of a per-world launch takes: by size, largest first, when the launch
// covers every world.
__device__ __forceinline__ int launch_world(const Params& P, int w0) {
  return (w0 == 0 && (int)gridDim.x == P.W) ? P.world_order[blockIdx.x] : w0 + (int)blockIdx.x;
}

__device__ __forceinline__ double& CF(const Params& P, double* y, int f, int i) {
  return y[(long)f * P.ncell + i];
}
__device__ __forceinline__ double CFc(const Params& P, const double* y, int f, int i) {
  return y[(long)f * P.ncell + i];
}
__device__ __forceinline__ double& EF(const Params& P, double* y, int f, int e) {
  return y[P.edge_base + (long)f * P.nedge + e];
}
__device__ __forceinline__ double EFc(const Params& P, const double* y, int f, int e) {
  return y[P.edge_base + (long)f * P.nedge + e];
}
__device__ __forceinline__ double& WF(const Params& P, double* y, int f, int w) {
  return y[P.world_base + (long)f * P.W + w];
}
__device__ __forceinline__ double WFc(const Params& P, const double* y, int f, int w) {
  return y[P.world_base + (long)f * P.W + w];
}

// ------------------------------------------------------------------------------------------
// Point laws (laws/electrochemistry.py:25-48, laws/junctions/selective.py:28-50,
// laws/routes/patch_electrodiffusion.py:43-54, laws/pumps/*, junction_voltage_relaxation.py).

__device__ __forceinline__ double bernoulli(double x) {
  if (fabs(x) < 1e-4) {
    double x2 = x * x;
    return 1.0 - x / 2.0 + x2 / 12.0 - (x2 * x2) / 720.0;
  }
  return x / expm1(x);
}

__device__ __forceinline__ double bernoulli_slope(double x) {
  if (fabs(x) < 1e-3) return -0.5 + x / 6.0 - (x * x * x) / 180.0;
  double em = exp(-fabs(x));
  if (x > 0) {
    double d = 1.0 - em;
    return (em * (1.0 - x) - em * em) / (d * d);
  }
  double d = em - 1.0;
  return (em - 1.0 - x * em) / (d * d);
}

__device__ __forceinline__ double logaddexp0(double t) {
  double m = t > 0.0 ? t : 0.0;
  return m + log1p(exp(-fabs(0.0 - t)));
}

// GHK route flux (sender -> receiver) and its slope dJ/d(dv).
__device__ __forceinline__ void route(double cf, double ct, double dv, double z, double T,
                                      double pa, double avail, bool slope, double& J, double& g) {
  double x = z * kF * dv / (kR * T);
  J = pa * avail / kMolPerFmol * (cf * bernoulli(-x) - ct * bernoulli(x));
  if (slope) {
    double k = z * kF / (kR * T);
    double xs = k * dv;
    double pre = pa * avail / kMolPerFmol;
    g = pre * k * (-cf * bernoulli_slope(-xs) - ct * bernoulli_slope(xs));
  }
}

// The Bernoulli values of a route at charge z and potential drop dv, as route() takes them:
// bf = B(-x), bt = B(x) with x = z F dv / (R T), and with slopes sf = B'(-xs), st = B'(xs) with
// xs = (z F / (R T)) dv. Every operation in x and xs is odd in z (a product or quotient whose
// rounding is symmetric in sign), so a species whose charge equals another's has bitwise the same
// values, and one whose charge is its negative has them swapped: the route's transcendentals are
// evaluated once per distinct |z| and the fluxes keep their bits.
struct RouteBasis {
  double bf, bt, sf, st;
};

__device__ __forceinline__ RouteBasis route_basis(double z, double dv, double T, bool slope) {
  RouteBasis r;
  const double x = z * kF * dv / (kR * T);
  r.bf = bernoulli(-x);
  r.bt = bernoulli(x);
  r.sf = 0.0;
  r.st = 0.0;
  if (slope) {
    const double xs = z * kF / (kR * T) * dv;
    r.sf = bernoulli_slope(-xs);
    r.st = bernoulli_slope(xs);
  }
  return r;
}

// Species s's basis from species 0's (b0), reused when |z_s| = |z_0| (uniform branches).
__device__ __forceinline__ RouteBasis species_basis(const double* z, int s, const RouteBasis& b0, double dv,
                                                    double T, bool slope) {
  if (s == 0 || z[s] == z[0]) return b0;
  if (z[s] == -z[0]) return RouteBasis{b0.bt, b0.bf, b0.st, b0.sf};
  return route_basis(z[s], dv, T, slope);
}

// route() on a precomputed basis: the same operations in the same order, so the same bits.
__device__ __forceinline__ void route_on(const RouteBasis& b, double cf, double ct, double z, double T,
                                         double pa, double avail, bool slope, double& J, double& g) {
  J = pa * avail / kMolPerFmol * (cf * b.bf - ct * b.bt);
  if (slope) {
    double k = z * kF / (kR * T);
    double pre = pa * avail / kMolPerFmol;
    g = pre * k * (-cf * b.sf - ct * b.st);
  }
}

__device__ __forceinline__ void cell_concentrations(const Params& P, const double* y,
                                                    const double* yv, int i, double* c,
                                                    double& vol) {
  vol = P.water ? P.vol0[i] + CFc(P, yv, P.f_dvol, i) : P.vol0[i];
#pragma unroll
  for (int s = 0; s < 4; ++s)
    if (s < P.S) c[s] = (P.n0[(long)s * P.ncell + i] + CFc(P, y, P.f_dn + s, i)) / vol;
}

__device__ __forceinline__ double charge_increment(const Params& P, const double* y, int i) {
  double t = P.z[0] * CFc(P, y, P.f_dn, i);
#pragma unroll
  for (int s = 1; s < 4; ++s)
    if (s < P.S) t = t + P.z[s] * CFc(P, y, P.f_dn + s, i);
  return kChargePerFmol * t;
}

__device__ __forceinline__ double potential(const Params& P, const double* y, int i) {
  double q = P.q0[i] + charge_increment(P, y, i);
  return P.Vb + q / P.cap[i];
}

__device__ __forceinline__ const double* current(const double* A, const double* B, int s) {
  return s ? B : A;
}
__device__ __forceinline__ double* next(double* A, double* B, int s) { return s ? A : B; }

// ------------------------------------------------------------------------------------------
// Junction routes (numerics/surface_transport.py:310-367, junction branch; gate rhs
// junction_voltage_relaxation.py:41-59; slopes models/surface_pump_v3.py:2747-2755).

// AHEAD (the ROS2 second stage, gdt = dt): the stage state's gate is formed here, y's gate + gdt
// K1's (the value the stage combination would have stored), and stored into the stage state Y1g as
// well; the junction heat and gate fields of the step's result are finished here.
template <bool SLOPES, bool AHEAD>
__device__ __forceinline__ void edges_body(const Params& P, const double* A, const double* B, const int* sel,
                                           const double* src_override, int vol_next, int e0, int e1, const Ctl& c,
                                           double* R, double* gj, double* Gj, int* valid, const double* K1g,
                                           double gdt, double* Y1g) {
  int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  int w = P.edge_world[e];
  int sw = sel[w];
  const double* y = src_override ? src_override : current(A, B, sw);
  const double* yv = vol_next ? next(const_cast<double*>(A), const_cast<double*>(B), sw) : y;
  int a = P.snd[e], b = P.rcv[e];
  double ca[4], cb[4], va, vb;
  cell_concentrations(P, y, yv, a, ca, va);
  cell_concentrations(P, y, yv, b, cb, vb);
  double dv = P.phi ? P.phi[a] - P.phi[b] : potential(P, y, a) - potential(P, y, b);
  double g;
  if constexpr (AHEAD) {
    g = EFc(P, current(A, B, sw), P.f_gate, e) + gdt * EFc(P, K1g, P.f_gate, e);
    EF(P, Y1g, P.f_gate, e) = g;
  } else {
    g = EFc(P, y, P.f_gate, e);
  }
  double avail = c.junction * g;
  bool ok = true;
  double diss = 0.0, G = 0.0;
  constexpr bool slopes = SLOPES;  // gj, Gj non-null exactly when SLOPES
  const RouteBasis b0 = route_basis(P.z[0], dv, P.T, slopes);
#pragma unroll
  for (int s = 0; s < 4; ++s) {  // species, unrolled (S <= 4)
    if (s >= P.S) break;
    double z = P.z[s];
    double pa = P.paj[(long)s * P.nedge + e];
    ok = ok && isfinite(ca[s]) && ca[s] > 0 && isfinite(cb[s]) && cb[s] > 0;
    double J, sl = 0.0;
    route_on(species_basis(P.z, s, b0, dv, P.T, slopes), ca[s], cb[s], z, P.T, pa, avail, slopes, J, sl);
    double mu = P.RT * (log(ca[s]) - log(cb[s])) + z * kF * dv;
    diss = s == 0 ? J * mu : diss + J * mu;
    ok = ok && fabs(z * kF * dv / P.RT) <= kMaxAbsX && isfinite(J);
    EF(P, R, P.f_jtr + s, e) = J;
    if (slopes) {
      gj[(long)s * P.nedge + e] = sl;
      G = s == 0 ? z * sl : G + z * sl;
    }
  }
  if (slopes) Gj[e] = kChargePerFmol * G;
  ok = ok && isfinite(1e-3 * diss);
  double shift = fabs(dv) - P.gate_thr;
  double alpha = 1.3 * exp(-77.0 * shift);
  double log_beta = P.log_0065 + 140.0 * shift;
  double beta = 20.0 * exp(-logaddexp0(-log_beta));
  const double grate = alpha * (1.0 - g) - beta * (g - P.gate_min);
  if constexpr (AHEAD) {
    // the junction heat and the gate take no stage-solve correction: (y + 1.5 dt k1) + 0.5 dt
    // (F1 - 2 k1) straight into the step's result, as the final combination would form them
    const double* yc = current(A, B, sw);
    double* o = next(const_cast<double*>(A), const_cast<double*>(B), sw);
    const double h1 = 1.5 * gdt, h2 = 0.5 * gdt;
    const double kd = EFc(P, K1g, P.f_jdiss, e), kg = EFc(P, K1g, P.f_gate, e);
    const double od = (EFc(P, yc, P.f_jdiss, e) + h1 * kd) + h2 * (1e-3 * diss + -2.0 * kd);
    const double og = (EFc(P, yc, P.f_gate, e) + h1 * kg) + h2 * (grate + -2.0 * kg);
    EF(P, o, P.f_jdiss, e) = od;
    EF(P, o, P.f_gate, e) = og;
    // the new state's domain on these fields (the final combination checks the transfers)
    ok = ok && isfinite(od) && isfinite(og) && og >= P.gate_min && og <= 1.0;
  } else {
    EF(P, R, P.f_jdiss, e) = 1e-3 * diss;
    EF(P, R, P.f_gate, e) = grate;
  }
  ok = ok && isfinite(dv) && fabs(dv) <= kGateMaxVoltage && isfinite(g) && g >= P.gate_min &&
       g <= 1.0;
  if (!ok) valid[w] = 0;
}

template <bool SLOPES>
__global__ __launch_bounds__(kBlock) void k_edges(Params P, const double* A, const double* B,
                                                  const int* sel, const double* src_override,
                                                  int vol_next, int e0, int e1, Ctl c, double* R,
                                                  double* gj, double* Gj, int* valid) {
  edges_body<SLOPES, false>(P, A, B, sel, src_override, vol_next, e0, e1, c, R, gj, Gj, valid, nullptr, 0.0,
                            nullptr);
}

// The ROS2 second stage's junction rates at Y1 (its amounts from the stage-1 combination, its gates
// formed here from y and K1 and stored into Y1).
__global__ __launch_bounds__(kBlock) void k_edges_ahead(Params P, const double* A, const double* B, const int* sel,
                                                        double* Y1, int e0, int e1, Ctl c, double* R,
                                                        const double* K1, double dt, int* valid) {
  edges_body<false, true>(P, A, B, sel, Y1, 0, e0, e1, c, R, nullptr, nullptr, valid, K1, dt, Y1);
}

// ------------------------------------------------------------------------------------------
// Membrane routes, pumps and owners (surface_transport.py:310-375, surface_pump_v2.py:370-450).
// With coef > 0 also the stage matrix's diagonal and its Jacobi scale
// (numerics/surface_imex.py:228-259 on C + coef G).

// LEDGERS (the ROS2 second stage, K1 set): the fields no stage solve corrects (membrane heat, the
// pumps' cycles, heat and work) go straight to the step's result, (y + 1.5 dt k1) + 0.5 dt (F1 - 2
// k1) with ldt = dt, as the final combination would form them, instead of into R.
template <bool SLOPES, bool LEDGERS>
__device__ __forceinline__ void cells_body(Params P, const double* A, const double* B, const int* sel,
                                           const double* src_override, int vol_next, int c0, int c1, Ctl c,
                                           double* R, double* contrib, double* gm, const double* Gj, double coef,
                                           double* mdiag, double* msc, double rate_dt, double* yrhs,
                                           const double* K1, int* valid, double ldt) {
  int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c1) return;
  int w = P.cell_world[i];
  int sw = sel[w];
  const double* y = src_override ? src_override : current(A, B, sw);
  const double* yv = vol_next ? next(const_cast<double*>(A), const_cast<double*>(B), sw) : y;
  double ci[4], vol;
  cell_concentrations(P, y, yv, i, ci, vol);
  double dv = (P.phi ? P.phi[i] : potential(P, y, i)) - P.Vb;
  bool ok = true;
  constexpr bool slopes = SLOPES;  // gm, Gj, mdiag, msc non-null exactly when SLOPES
  double Jm[4], mu[4];
  double diss = 0.0, Gm = 0.0;
  const RouteBasis b0 = route_basis(P.z[0], dv, P.T, slopes);
#pragma unroll
  for (int s = 0; s < 4; ++s) {  // species, unrolled (S <= 4)
    if (s >= P.S) break;
    double z = P.z[s];
    ok = ok && isfinite(ci[s]) && ci[s] > 0;
    double pa = P.pam[(long)s * P.ncell + i];
    double sl = 0.0;
    route_on(species_basis(P.z, s, b0, dv, P.T, slopes), ci[s], P.cb[s], z, P.T, pa, c.membrane, slopes, Jm[s],
             sl);
    mu[s] = P.RT * (log(ci[s]) - P.log_cb[s]) + z * kF * dv;
    diss = s == 0 ? Jm[s] * mu[s] : diss + Jm[s] * mu[s];
    ok = ok && fabs(z * kF * dv / P.RT) <= kMaxAbsX && isfinite(Jm[s]);
    if (slopes) {
      gm[(long)s * P.ncell + i] = sl;
      Gm = s == 0 ? z * sl : Gm + z * sl;
    }
  }
  // Na/K-ATPase (nak_atp_mass_action_positive.py:63-149)
  double cap = P.nak_cap[i];
  double energy = P.nak_drive + (-3.0) * mu[0] + 2.0 * mu[1];
  double a = energy / P.RT;
  double ls = -logaddexp0(3.0 * (P.log_km_na - log(ci[0]))) -
              logaddexp0(2.0 * (P.log_km_k - P.log_cb[1])) - P.nak_atp_sat;
  bool enabled = cap > 0 && c.pump > 0;
  double mag = exp(log(enabled ? cap : 1.0) + log(enabled ? c.pump : 1.0) + ls + (a > 0.0 ? a : 0.0));
  double dep = a >= 0 ? expm1(-a) : -expm1(a);
  double cyc = enabled ? mag * dep : 0.0;
  {
    double sn = mu[0] / P.RT, sk = mu[1] / P.RT;
    double aff = P.dg0_over_rt_nak_terms + (-3.0) * sn + 2.0 * sk;
    ok = ok && isfinite(ci[0]) && ci[0] > 0 && isfinite(cap) && cap >= 0 && isfinite(sn) &&
         fabs(sn) <= kNakMaxMu && isfinite(sk) && fabs(sk) <= kNakMaxMu && isfinite(aff) &&
         fabs(aff) <= kNakMaxAffinity;
  }
  double extra[4] = {3.0 * cyc, -2.0 * cyc, -0.0 * cyc, -0.0 * cyc};
  double ccyc = 0.0, cenergy = 0.0;
  if (P.calcium) {
    // ca_atp_mass_action_positive.py:70-150 via er_calcium.pump_rates(ci, co, -v)
    double ccap = P.ca_cap[i];
    double ce = P.ca_dk + P.RT * (P.log_cb[3] - log(ci[3])) + 2.0 * kF * (-dv);
    double aff = ce / P.RT;
    double cls = -logaddexp0(P.log_ca_km - log(ci[3]));
    cls -= P.ca_atp_sat;
    bool cen = ccap > 0 && c.calcium_pump > 0;
    double cmag = exp(log(cen ? ccap : 1.0) + log(cen ? c.calcium_pump : 1.0) + cls +
                      (aff > 0.0 ? aff : 0.0));
    double cdep = aff >= 0 ? expm1(-aff) : -expm1(aff);
    ccyc = cen ? cmag * cdep : 0.0;
    cenergy = ce;
    double daff = P.ca_aff_base + P.log_cb[3] - log(ci[3]) - 2.0 * kF * dv / P.RT;
    ok = ok && isfinite(ccap) && ccap >= 0 && isfinite(dv) && fabs(dv) <= kCaMaxVoltage &&
         isfinite(daff) && fabs(daff) <= kCaMaxAffinity;
    extra[3] = extra[3] - (-1.0) * ccyc;
  }
  int p0 = P.inc_ptr[i], p1 = P.inc_ptr[i + 1], pm = p0 + P.inc_out[i];
  // edges outer (each index loaded once), species inner; each species sums its outgoing, then
  // incoming, junction transfers in incidence order
  double jout[4] = {0.0, 0.0, 0.0, 0.0}, jin[4] = {0.0, 0.0, 0.0, 0.0};
  double g = kChargePerFmol * Gm;  // with slopes: the diagonal's conductances, in incidence order
  // two incident edges at a time, both edges' loads issued before either is used (a wavefront's
  // requests for the lines its cells share arrive closer together); the sums keep their order
  for (int pb = p0; pb < p1; pb += 2) {
    const int e0 = P.inc_edge[pb], e1 = pb + 1 < p1 ? P.inc_edge[pb + 1] : -1;
    double g0 = 0.0, g1 = 0.0, t0[4], t1[4];
    if (slopes) {
      g0 = Gj[e0];
      g1 = e1 >= 0 ? Gj[e1] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      t0[s] = s < P.S ? EFc(P, R, P.f_jtr + s, e0) : 0.0;
      t1[s] = (s < P.S && e1 >= 0) ? EFc(P, R, P.f_jtr + s, e1) : 0.0;
    }
    if (slopes) g += g0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      if (pb < pm) jout[s] += -t0[s];
      else jin[s] += t0[s];
    }
    if (e1 < 0) break;
    if (slopes) g += g1;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      if (pb + 1 < pm) jout[s] += -t1[s];
      else jin[s] += t1[s];
    }
  }
  double qrate = 0.0;  // the charge rate of the amounts' rates (less 2 K1's with K1: stage 2)
#pragma unroll
  for (int s = 0; s < 4; ++s) {  // species, unrolled (S <= 4)
    if (s >= P.S) break;
    double cell = -Jm[s];
    cell += jout[s];
    cell += jin[s];
    double r = cell - extra[s];
    CF(P, R, P.f_dn + s, i) = r;
    CF(P, R, P.f_mtr + s, i) = Jm[s];
    contrib[(long)s * P.ncell + i] = -(Jm[s] + extra[s]);
    ok = ok && isfinite(r);
    if (yrhs) {
      const double v = K1 ? r + -2.0 * CFc(P, K1, P.f_dn + s, i) : r;
      qrate = s == 0 ? P.z[0] * v : qrate + P.z[s] * v;
    }
  }
  double mdiss = 1e-3 * diss;
  if constexpr (LEDGERS) {
    const double* yc = current(A, B, sw);
    double* o = next(const_cast<double*>(A), const_cast<double*>(B), sw);
    const double h1 = 1.5 * ldt, h2 = 0.5 * ldt;
    const int nf = P.calcium ? 7 : 4;
    const double r[7] = {mdiss, cyc, -1e-3 * cyc * energy, -1e-3 * cyc * P.nak_drive, ccyc, -1e-3 * ccyc * cenergy,
                         -1e-3 * ccyc * P.ca_drive};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      if (k >= nf) break;
      const int f = k < 4 ? P.f_mdiss + k : P.f_ccyc + (k - 4);
      const double k1 = CFc(P, K1, f, i);
      const double k2 = r[k] + -2.0 * k1;
      const double v = (CFc(P, yc, f, i) + h1 * k1) + h2 * k2;
      CF(P, o, f, i) = v;
      ok = ok && isfinite(v);  // the new state's domain on these fields (the final combination has the rest)
    }
  } else {
    CF(P, R, P.f_mdiss, i) = mdiss;
    CF(P, R, P.f_pcyc, i) = cyc;
    CF(P, R, P.f_pdiss, i) = -1e-3 * cyc * energy;
    CF(P, R, P.f_pwork, i) = -1e-3 * cyc * P.nak_drive;
    if (P.calcium) {
      CF(P, R, P.f_ccyc, i) = ccyc;
      CF(P, R, P.f_cdiss, i) = -1e-3 * ccyc * cenergy;
      CF(P, R, P.f_cwork, i) = -1e-3 * ccyc * P.ca_drive;
    }
  }
  ok = ok && isfinite(mdiss) && isfinite(cyc) && isfinite(-1e-3 * cyc * energy);
  if (P.calcium) ok = ok && isfinite(ccyc) && isfinite(-1e-3 * ccyc * cenergy);
  if (P.water) {
    CF(P, R, P.f_dvol, i) = 0.0;
    CF(P, R, P.f_wtr, i) = 0.0;
    CF(P, R, P.f_wdiss, i) = 0.0;
  }
  double sc_i = 0.0;
  if (slopes) {
    double d = P.cap[i] + coef * g;
    ok = ok && d > 0;
    double sc = 1.0 / sqrt(d > 0 ? d : 1.0);
    mdiag[i] = (d * sc * sc + d * sc * sc) / 2;
    msc[i] = sc;
    sc_i = sc;
  }
  // the stage solve's right-hand side, scaled, at the cell's place in its world's solve vector
  // (the solve then starts at once instead of gathering it in rank order)
  // (yrhs set: msc is the stage's Jacobi scales, written here with slopes, read here without)
  if (yrhs) yrhs[P.cell_rank[i]] = (kChargePerFmol * qrate) * (slopes ? sc_i : msc[i]);
  if (rate_dt > 0.0) {
    // the explicit stage's membrane charge relaxation (models/surface_pump_v3.py:4588-4638
    // membrane_rate with one patch per cell and no kinds): F^2/RT sum_s z^2 (A P)_s max(c_in, c_out)
    // times the availability, over the cell's capacitance; dt x rate <= 1 (surface_pump.py:116)
    double t = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {  // species, unrolled (S <= 4)
      if (s >= P.S) break;
      const double z2 = P.z[s] * P.z[s];
      const double term = z2 * P.pam[(long)s * P.ncell + i] * fmax(ci[s], P.cb[s]);
      t = s == 0 ? term : t + term;
    }
    const double rate = kF * kF / P.RT * (t * c.membrane) / P.cap[i];
    if (!(rate_dt * rate <= kExplicitStepBound)) {
      ok = false;
      atomicOr(P.refuse_bits + w, kRefuseExplicitRate);
    }
  }
  if (!ok) valid[w] = 0;
}

template <bool SLOPES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SLOPES ? 4 : 5))) void k_cells(Params P, const double* A, const double* B,
                                                  const int* sel, const double* src_override,
                                                  int vol_next, int c0, int c1, Ctl c, double* R,
                                                  double* contrib, double* gm, const double* Gj,
                                                  double coef, double* mdiag, double* msc,
                                                  double rate_dt, double* yrhs, const double* K1,
                                                  int* valid) {
  cells_body<SLOPES, false>(P, A, B, sel, src_override, vol_next, c0, c1, c, R, contrib, gm, Gj, coef, mdiag, msc,
                            rate_dt, yrhs, K1, valid, 0.0);
}

// The ROS2 second stage's cell rates at Y1 (K1 set), their uncorrected ledger fields finished into
// the step's result (dt the step).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_cells_ahead(
    Params P, const double* A, const double* B, const int* sel, const double* Y1, int c0, int c1, Ctl c, double* R,
    double* contrib, const double* msc, double* yrhs, const double* K1, double dt, int* valid) {
  cells_body<false, true>(P, A, B, sel, Y1, 0, c0, c1, c, R, contrib, nullptr, nullptr, 0.0, nullptr,
                          const_cast<double*>(msc), 0.0, yrhs, K1, valid, dt);
}

// ------------------------------------------------------------------------------------------
// Deterministic block reductions: each thread folds its strided items in order, then a fixed
// LDS tree.

__device__ double block_sum(double v, double* sh) {
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] = sh[t] + sh[t + o];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ double block_max(double v, double* sh) {
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] = fmax(sh[t], sh[t + o]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------------------------------
// Domain checks of a new state (models/surface_tissue.py:391-420, surface_kinetic.py:422-451).

__device__ bool cell_domain(const Params& P, const double* y, const double* yv, int i) {
  bool ok = true;
  for (int f = 0; f < P.cf; ++f) ok = ok && isfinite(CFc(P, y, f, i));
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double n = P.n0[(long)s * P.ncell + i] + CFc(P, y, P.f_dn + s, i);
    ok = ok && isfinite(n) && n > 0;
  }
  if (P.water) {
    double v = P.vol0[i] + CFc(P, yv, P.f_dvol, i);
    ok = ok && v > P.vmin * P.vol0[i] && v < P.vmax * P.vol0[i];
  }
  return ok;
}

__device__ bool edge_domain(const Params& P, const double* y, int e) {
  bool ok = true;
  for (int f = 0; f < P.ef; ++f) ok = ok && isfinite(EFc(P, y, f, e));
  double g = EFc(P, y, P.f_gate, e);
  return ok && g >= P.gate_min && g <= 1.0;
}

__device__ __forceinline__ bool is_water_field(const Params& P, int f) {
  return P.water && (f == P.f_dvol || f == P.f_wtr || f == P.f_wdiss);
}

// cell_domain without the fields the ROS2 second stage's cell kernel finishes and checks itself
// (membrane heat, the pumps' cycles, heat and work): the amounts, transfers and water fields.
__device__ bool cell_domain_final(const Params& P, const double* y, const double* yv, int i) {
  bool ok = true;
  for (int f = 0; f < P.cf; ++f)
    if (f < P.f_mtr + P.S || is_water_field(P, f)) ok = ok && isfinite(CFc(P, y, f, i));
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double n = P.n0[(long)s * P.ncell + i] + CFc(P, y, P.f_dn + s, i);
    ok = ok && isfinite(n) && n > 0;
  }
  if (P.water) {
    double v = P.vol0[i] + CFc(P, yv, P.f_dvol, i);
    ok = ok && v > P.vmin * P.vol0[i] && v < P.vmax * P.vol0[i];
  }
  return ok;
}

// ------------------------------------------------------------------------------------------
// IMEX Euler (surface_pump.py:292-362 with no faces): y + dt r, the stage potentials' residual
// check, the domain. Threads [0, nc) take cells c0.., [nc, nc+ne) edges e0...

__global__ __launch_bounds__(kBlock) void k_euler_update(Params P, double* A, double* B,
                                                         const int* sel, int c0, int nc, int e0,
                                                         int ne, double dt, const double* R,
                                                         int* valid) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double h = 1.0 * dt;
  if (t < nc) {
    int i = c0 + t;
    int w = P.cell_world[i];
    const double* y = current(A, B, sel[w]);
    double* o = next(A, B, sel[w]);
    for (int f = 0; f < P.cf; ++f) {
      if (is_water_field(P, f)) continue;
      CF(P, o, f, i) = CFc(P, y, f, i) + h * CFc(P, R, f, i);
    }
    double q = P.q0[i] + charge_increment(P, o, i);
    double phi = P.Vb + q / P.cap[i];
    double cphi = P.cap[i] * (phi - P.Vb);
    double scale = fabs(q) > fabs(cphi) ? fabs(q) : fabs(cphi);
    bool ok = isfinite(phi) && fabs(cphi - q) <= 3e-25 + 2e-12 * scale;
    ok = ok && cell_domain(P, o, o, i);
    if (!ok) valid[w] = 0;
  } else if (t < nc + ne) {
    int e = e0 + (t - nc);
    int w = P.edge_world[e];
    const double* y = current(A, B, sel[w]);
    double* o = next(A, B, sel[w]);
    for (int f = 0; f < P.ef; ++f) EF(P, o, f, e) = EFc(P, y, f, e) + h * EFc(P, R, f, e);
    if (!edge_domain(P, o, e)) valid[w] = 0;
  }
}

// Per-world ledgers of an ionic step (surface_pump.py:346-359): bath supplies, reservoir work,
// electrode charge and work. sup_h1/sup_h2 weigh the two stage supply sums (Euler: dt, 0).
__global__ __launch_bounds__(kBlock) void k_world_ionic(Params P, double* A, double* B,
                                                        const int* sel, int w0,
                                                        const double* contrib1, double h1,
                                                        const double* contrib2, double h2,
                                                        int* valid) {
  __shared__ double sh[kBlock];
  int w = w0 + blockIdx.x;
  int sw = sel[w];
  const double* y = current(A, B, sw);
  double* o = next(A, B, sw);
  int c0 = P.cell_off[w], c1 = P.cell_off[w + 1];
  double applied[4];
  bool ok = true;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double a1 = 0.0, a2 = 0.0;
    for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
      a1 += contrib1[(long)s * P.ncell + i];
      if (contrib2) a2 += contrib2[(long)s * P.ncell + i];
    }
    double s1 = block_sum(a1, sh);
    double s2 = contrib2 ? block_sum(a2, sh) : 0.0;
    double sup = WFc(P, y, P.f_sup + s, w);
    double nv = contrib2 ? (sup + h1 * s1) + h2 * s2 : sup + h1 * s1;
    applied[s] = contrib2 ? h1 * s1 + h2 * s2 : h1 * s1;
    if (threadIdx.x == 0) WF(P, o, P.f_sup + s, w) = nv;
    ok = ok && isfinite(nv);
  }
  double qs = 0.0, qe = 0.0;
  for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
    qs += charge_increment(P, y, i);
    qe += charge_increment(P, o, i);
  }
  qs = block_sum(qs, sh);
  qe = block_sum(qe, sh);
  if (threadIdx.x == 0) {
    double electrode = -qe - -qs;
    double work = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < P.S) work = s == 0 ? P.mu_b[s] * applied[s] : work + P.mu_b[s] * applied[s];
    double ec = WFc(P, y, P.f_echg, w) + electrode;
    double rw = WFc(P, y, P.f_rwork, w) + 1e-3 * work;
    double ew = WFc(P, y, P.f_ework, w) + 1e12 * P.Vb * electrode;
    WF(P, o, P.f_echg, w) = ec;
    WF(P, o, P.f_rwork, w) = rw;
    WF(P, o, P.f_ework, w) = ew;
    ok = ok && isfinite(ec) && isfinite(rw) && isfinite(ew);
    if (!ok) valid[w] = 0;
  }
}

// ------------------------------------------------------------------------------------------
// Water half-step (operators/osmotic_water.py:127-221): Heun over h at frozen amounts, each cell's
// water route to its held bath. from_next = 0 reads the step's start state, 1 the new state.
// Writes the new state's water fields. Workgroup per world.

__device__ __forceinline__ double water_psi(const Params& P, const double* y, int i, double vol,
                                            double& total) {
  double t = 0.0;
#pragma unroll
  for (int s = 0; s < 4; ++s)
    if (s < P.S) t += P.n0[(long)s * P.ncell + i] + CFc(P, y, P.f_dn + s, i);
  t = t + P.imperm[i];
  total = t;
  double osm = P.RT * t / vol;
  double hydro = P.wK * (vol - P.vol0[i]) / P.vol0[i] + P.wP0;
  return hydro - osm;
}

__global__ __launch_bounds__(kBlock) void k_water(Params P, double* A, double* B, const int* sel,
                                                  int w0, int from_next, double h, Ctl c,
                                                  int* valid) {
  __shared__ double sh[kBlock];
  int w = w0 + blockIdx.x;
  int sw = sel[w];
  const double* y = from_next ? next(A, B, sw) : current(A, B, sw);
  double* o = next(A, B, sw);
  int c0 = P.cell_off[w], c1 = P.cell_off[w + 1];
  bool ok = isfinite(c.water) && c.water >= 0 && c.water <= 1;
  double maxrow = 0.0, sumrow = 0.0;
  for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
    double vol = P.vol0[i] + CFc(P, y, P.f_dvol, i);
    double total;
    (void)water_psi(P, y, i, vol, total);
    double slope = P.wK / P.vol0[i] + P.RT * total / (vol * vol);
    double edge = P.lpa[i] * c.water * (slope + 0.0);
    maxrow = fmax(maxrow, edge);
    sumrow += edge;
    ok = ok && vol > P.vmin * P.vol0[i] && vol < P.vmax * P.vol0[i];
    ok = ok && isfinite(CFc(P, y, P.f_wtr, i)) && isfinite(CFc(P, y, P.f_wdiss, i));
  }
  maxrow = block_max(maxrow, sh);
  sumrow = block_sum(sumrow, sh);
  double row = fmax(maxrow, sumrow);
  ok = ok && 2 * h * row <= kWaterMaxRelaxation;
  double f0sum = 0.0, f1sum = 0.0;
  for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
    double dvol = CFc(P, y, P.f_dvol, i), wtr = CFc(P, y, P.f_wtr, i), wd = CFc(P, y, P.f_wdiss, i);
    double total;
    double vol = P.vol0[i] + dvol;
    double drive0 = water_psi(P, y, i, vol, total) - P.psib;
    double flow0 = P.lpa[i] * c.water * drive0;
    double r0v = -flow0, r0t = flow0, r0h = 1e-3 * flow0 * drive0;
    double pv = dvol + h * r0v;
    double pvol = P.vol0[i] + pv;
    ok = ok && pvol > P.vmin * P.vol0[i] && pvol < P.vmax * P.vol0[i] && isfinite(wtr + h * r0t) &&
         isfinite(wd + h * r0h);
    double drive1 = water_psi(P, y, i, pvol, total) - P.psib;
    double flow1 = P.lpa[i] * c.water * drive1;
    double heat1 = 1e-3 * flow1 * drive1;
    double mv = 0.5 * (r0v + -flow1), mt = 0.5 * (r0t + flow1), mh = 0.5 * (r0h + heat1);
    double nv = dvol + h * mv, nt = wtr + h * mt, nh = wd + h * mh;
    CF(P, o, P.f_dvol, i) = nv;
    CF(P, o, P.f_wtr, i) = nt;
    CF(P, o, P.f_wdiss, i) = nh;
    double nvol = P.vol0[i] + nv;
    ok = ok && nvol > P.vmin * P.vol0[i] && nvol < P.vmax * P.vol0[i] && isfinite(nt) &&
         isfinite(nh);
    f0sum += flow0;
    f1sum += flow1;
  }
  f0sum = block_sum(f0sum, sh);
  f1sum = block_sum(f1sum, sh);
  if (threadIdx.x == 0) {
    double sup0 = -f0sum, sup1 = -f1sum;
    double supm = 0.5 * (sup0 + sup1);
    double supplied = h * supm;
    double ws = WFc(P, y, P.f_wsup, w), ww = WFc(P, y, P.f_wwork, w);
    WF(P, o, P.f_wsup, w) = ws + supplied;
    WF(P, o, P.f_wwork, w) = ww + 1e-3 * P.psib * supplied;
    ok = ok && isfinite(ws + supplied) && isfinite(ww + 1e-3 * P.psib * supplied);
  }
  if (!ok) valid[w] = 0;
}

// ------------------------------------------------------------------------------------------
// Sparse Cholesky on the stage pattern, workgroup per world (numerics/sparse_cholesky.py:474-514).
// Entry (r, k) = assembled - pairwise fold over m in rows(k) & rows(r) of L[r,m] L[k,m]; a pivot
// is its square root, a below entry times the reciprocal of its column's pivot. The fold is the
// reference's adjacent-pairs tree, evaluated as a binary-counter stack (the same additions).

struct PlanView {
  const int* cs;      // [n+1] column starts
  const int* rrank;   // [nnz] entry row ranks
  const int* crank;   // [nnz] entry column ranks
  const int* rowptr;  // [n+1]
  const int* rowcol;  // [nnz-n] columns of each row's strictly lower entries
  const int* rowpos;  // [nnz-n] their positions
  const int* levptr;  // [levels+1]
  const int* levcol;  // [n]
  const int* leventptr;  // [levels+1]
  const int* levent;     // [nnz]
  const int* order;      // [n] rank -> node
  const int* diagpos;    // [n] node -> diagonal position
  const int* edgepos;    // [E] world edge -> position
  const unsigned long long* mask;  // [2 (nnz-n)]: per (k, m) row-structure pair, column k's slots
                                   // that column m's tail reaches (bit j: the j-th entry of k)
};

// Pairwise sums in the reference's adjacent-pairs order (numerics/sparse_cholesky.py:244-322),
// held as a binary counter of complete blocks: s[j] is the sum of a block of 2^j consecutive terms
// when bit j of n is set. Adding a term merges it, oldest block on the left, into every block the
// increment of n carries through; the total folds the blocks from the newest up. The loops are
// unrolled over a fixed depth so every partial stays in registers.
template <int kDepth>
struct PairSumT {  // up to 2^kDepth - 1 terms
  double s[kDepth];
  unsigned n;
  __device__ __forceinline__ void reset() { n = 0u; }
  __device__ __forceinline__ void push(double t) {
    double carry = t;
    bool open = true;
#pragma unroll
    for (int j = 0; j < kDepth; ++j) {
      bool bit = (n >> j) & 1u;
      if (open && bit) carry = s[j] + carry;
      if (open && !bit) s[j] = carry;
      open = open && bit;
    }
    ++n;
  }
  __device__ __forceinline__ double sum() const {
    double r = 0.0;
    bool have = false;
#pragma unroll
    for (int j = 0; j < kDepth; ++j) {
      if ((n >> j) & 1u) {
        r = have ? s[j] + r : s[j];
        have = true;
      }
    }
    return r;
  }
};
using PairSum = PairSumT<13>;  // up to 8191 terms

// y = A x on a product plan (numerics/sparse_cholesky.py:467-471 `product`): thread per row, the
// row's products value x x[column] formed and rounded one by one in the plan's (ascending column)
// order and summed in adjacent pairs, step by step (the reference's fold), by the binary-counter
// stack; an empty row gives +0.0. blockIdx.y is the system: its values, x and y at the strides.
__global__ __launch_bounds__(256) void k_product(int n, const long long* row_start, const int* column,
                                                 const double* values, const double* x, double* y,
                                                 long long values_stride, long long vector_stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* v = values + (long long)blockIdx.y * values_stride;
  const double* xb = x + (long long)blockIdx.y * vector_stride;
  PairSumT<31> pw;
  pw.reset();
  for (long long p = row_start[i]; p < row_start[i + 1]; ++p) {
    const double t = v[p] * xb[column[p]];
    pw.push(t);
  }
  y[(long long)blockIdx.y * vector_stride + i] = pw.sum();
}

__device__ PlanView plan_view(const ShapeDev& s, const int* pool,
                              const unsigned long long* masks = nullptr) {
  PlanView v;
  v.mask = masks ? masks + s.o_mask : nullptr;
  v.cs = pool + s.o_cs;
  v.rrank = pool + s.o_entry;
  v.crank = pool + s.o_entry + s.nnz;
  v.rowptr = pool + s.o_rowptr;
  v.rowcol = pool + s.o_rowent;
  v.rowpos = pool + s.o_rowent + (s.nnz - s.n);
  v.levptr = pool + s.o_lev;
  v.leventptr = pool + s.o_lev + (s.levels + 1);
  v.levcol = pool + s.o_levcol;
  v.levent = pool + s.o_levent;
  v.order = pool + s.o_node;
  v.diagpos = pool + s.o_node + s.n;
  v.edgepos = pool + s.o_node + 2 * s.n;
  return v;
}

// The factor's values: the trailing columns (positions >= pt, where most update terms are read)
// live in LDS during the factor, the rest in HBM.
struct FactorStore {
  double* g;
  double* l;
  int pt;
  __device__ __forceinline__ double get(int p) const { return p >= pt ? l[p - pt] : g[p]; }
  __device__ __forceinline__ void set(int p, double x) {
    if (p >= pt) l[p - pt] = x;
    else g[p] = x;
  }
};

__device__ void factor_world(const ShapeDev& s, const PlanView& v, FactorStore fv) {
  for (int lev = 0; lev < s.levels; ++lev) {
    int a = v.leventptr[lev], b = v.leventptr[lev + 1];
    for (int q = a + threadIdx.x; q < b; q += blockDim.x) {
      int p = v.levent[q];
      int k = v.crank[p];
      int slot = p - v.cs[k];
      unsigned long long lo_mask = slot >= 64 ? ~0ull : ((1ull << slot) - 1ull);
      unsigned long long hi_mask = slot >= 64 ? ((1ull << (slot - 64)) - 1ull) : 0ull;
      int word = slot >> 6;
      unsigned long long bit = 1ull << (slot & 63);
      PairSum pw;
      pw.reset();
      int x = v.rowptr[k], xe = v.rowptr[k + 1];
      // four source columns at a time: their plan words first, then their factor values, so the
      // loads of a group are in flight together; the terms still enter the sum in column order
      for (; x + 4 <= xe; x += 4) {
        unsigned long long m0[4], m1[4];
        int base[4];
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          m0[u] = v.mask[2 * (long)(x + u)];
          m1[u] = v.mask[2 * (long)(x + u) + 1];
          base[u] = v.rowpos[x + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          unsigned long long mine = word ? m1[u] : m0[u];
          int idx = __popcll(m0[u] & lo_mask) + __popcll(m1[u] & hi_mask);
          bool has = (mine & bit) != 0ull;
          a[u] = has ? fv.get(base[u] + idx) : 0.0;
          b[u] = has ? fv.get(base[u]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          unsigned long long mine = word ? m1[u] : m0[u];
          if (mine & bit) pw.push(a[u] * b[u]);
        }
      }
      for (; x < xe; ++x) {
        unsigned long long m0 = v.mask[2 * (long)x], m1 = v.mask[2 * (long)x + 1];
        unsigned long long mine = word ? m1 : m0;
        if (mine & bit) {
          int base = v.rowpos[x];
          int idx = __popcll(m0 & lo_mask) + __popcll(m1 & hi_mask);
          pw.push(fv.get(base + idx) * fv.get(base));
        }
      }
      double acc = fv.get(p) - pw.sum();
      fv.set(p, slot == 0 ? sqrt(acc) : acc);
    }
    __syncthreads();
    for (int q = a + threadIdx.x; q < b; q += blockDim.x) {
      int p = v.levent[q];
      int k = v.crank[p];
      if (p != v.cs[k]) fv.set(p, fv.get(p) * (1.0 / fv.get(v.cs[k])));
    }
    __syncthreads();
  }
}

// y (LDS, by rank) <- L^-T L^-1 y, each pivot applied as a multiplication by its reciprocal.
// A wavefront's sum of len terms term(i), i ascending, in the adjacent-pairs order: each chunk of
// 64 terms is an adjacent-pairs tree across the lanes (absent lanes hold -0.0, the exact identity,
// which gives the odd-tail pass-through), full chunks enter a binary counter of 64-term blocks, and
// the last partial chunk is the newest partial of the fold. Uniform across the wave.
template <typename Term>
__device__ __forceinline__ double wave_pairwise(int len, Term term) {
  const int lane = threadIdx.x & 63;
  double blk[7];
  unsigned nb = 0u;
  double tail = 0.0;
  bool have_tail = false;
  for (int c0 = 0; c0 < len; c0 += 64) {
    int idx = c0 + lane;
    double t = idx < len ? term(idx) : -0.0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      double o = __shfl_down(t, off, 64);
      if ((lane & (2 * off - 1)) == 0) t = t + o;
    }
    double chunk = __shfl(t, 0, 64);
    if (len - c0 >= 64) {
      double carry = chunk;
      bool open = true;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        bool bit = (nb >> j) & 1u;
        if (open && bit) carry = blk[j] + carry;
        if (open && !bit) blk[j] = carry;
        open = open && bit;
      }
      ++nb;
    } else {
      tail = chunk;
      have_tail = true;
    }
  }
  double r = tail;
  bool have = have_tail;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    if ((nb >> j) & 1u) {
      r = have ? blk[j] + r : blk[j];
      have = true;
    }
  }
  return have ? r : 0.0;
}

// Columns with at most this many terms are summed by one thread; longer ones by a wavefront.
constexpr int kShortColumn = 16;

// y (by rank) <- L^-T L^-1 y, each pivot applied as a multiplication by its reciprocal
// (numerics/sparse_cholesky.py:493-514, the device form). Level by level; in each level a thread
// takes each short column and a wavefront each long one.
__device__ void solve_world(const ShapeDev& s, const PlanView& v, const double* fv, double* y) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (int lev = 0; lev < s.levels; ++lev) {
    int q0 = v.levptr[lev], q1 = v.levptr[lev + 1];
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      int k = v.levcol[q];
      int x0 = v.rowptr[k], x1 = v.rowptr[k + 1];
      if (x1 - x0 > kShortColumn) continue;
      PairSum pw;
      pw.reset();
      for (int x = x0; x < x1; ++x) pw.push(fv[v.rowpos[x]] * y[v.rowcol[x]]);
      double known = y[k] - pw.sum();
      y[k] = known * (1.0 / fv[v.cs[k]]);
    }
    for (int q = q0 + wave; q < q1; q += waves) {
      int k = v.levcol[q];
      int x0 = v.rowptr[k], x1 = v.rowptr[k + 1];
      if (x1 - x0 <= kShortColumn) continue;
      double sum = wave_pairwise(x1 - x0, [&](int i) { return fv[v.rowpos[x0 + i]] * y[v.rowcol[x0 + i]]; });
      if (lane == 0) {
        double known = y[k] - sum;
        y[k] = known * (1.0 / fv[v.cs[k]]);
      }
    }
    __syncthreads();
  }
  for (int lev = s.levels - 1; lev >= 0; --lev) {
    int q0 = v.levptr[lev], q1 = v.levptr[lev + 1];
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      int k = v.levcol[q];
      int p0 = v.cs[k] + 1, p1 = v.cs[k + 1];
      if (p1 - p0 > kShortColumn) continue;
      PairSum pw;
      pw.reset();
      for (int p = p0; p < p1; ++p) pw.push(fv[p] * y[v.rrank[p]]);
      double known = y[k] - pw.sum();
      y[k] = known * (1.0 / fv[v.cs[k]]);
    }
    for (int q = q0 + wave; q < q1; q += waves) {
      int k = v.levcol[q];
      int p0 = v.cs[k] + 1, p1 = v.cs[k + 1];
      if (p1 - p0 <= kShortColumn) continue;
      double sum = wave_pairwise(p1 - p0, [&](int i) { return fv[p0 + i] * y[v.rrank[p0 + i]]; });
      if (lane == 0) {
        double known = y[k] - sum;
        y[k] = known * (1.0 / fv[v.cs[k]]);
      }
    }
    __syncthreads();
  }
}

// Assemble (Jacobi-scaled, symmetrised; numerics/surface_imex.py:228-259) and factor
// C + coef G for each world (workgroup per world, no LDS: the factor lives in HBM/L2).
constexpr int kFactorBlock = 128;
constexpr int64_t kFactorLdsDoubles = 0;  // trailing factor values held in LDS (0: none; occupancy wins)
constexpr size_t kSolveLdsBytes = 0;  // the solve vector stays in HBM scratch: occupancy hides latency better

__global__ __launch_bounds__(kFactorBlock) void k_factor(Params P, const ShapeDev* shapes,
                                                         const int* shape_of, const int* pool,
                                                         const unsigned long long* masks,
                                                         const long* foff, int w0, double coef,
                                                         const double* mdiag, const double* msc,
                                                         const double* Gj, double* fvals,
                                                         int* valid) {
  extern __shared__ double lds[];
  int w = w0 + blockIdx.x;
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool, masks);
  double* fv = fvals + foff[w];
  int c0 = P.cell_off[w], e0 = P.edge_off[w];
  for (int p = threadIdx.x; p < s.nnz; p += blockDim.x) fv[p] = 0.0;
  __syncthreads();
  for (int i = threadIdx.x; i < s.n; i += blockDim.x) fv[v.diagpos[i]] = mdiag[c0 + i];
  for (int e = threadIdx.x; e < s.E; e += blockDim.x) {
    double val = -(coef * Gj[e0 + e]);
    double sa = msc[P.snd[e0 + e]], sb = msc[P.rcv[e0 + e]];
    fv[v.edgepos[e]] = (val * sa * sb + val * sb * sa) / 2;
  }
  __syncthreads();
  for (int p = s.lds_pt + threadIdx.x; p < s.nnz; p += blockDim.x) lds[p - s.lds_pt] = fv[p];
  __syncthreads();
  factor_world(s, v, FactorStore{fv, lds, s.lds_pt});
  for (int p = s.lds_pt + threadIdx.x; p < s.nnz; p += blockDim.x) fv[p] = lds[p - s.lds_pt];
  __syncthreads();
  bool ok = true;
  for (int p = threadIdx.x; p < s.nnz; p += blockDim.x) ok = ok && isfinite(fv[p]);
  if (!ok) valid[w] = 0;
}

// Which solve runs: the supernodal one (with the multifrontal factor) or the level-scheduled one
// in the reference's summation order.
struct SolveArgs {
  int sn;
  const long* voffw;  // per world: start of its update vectors in vbuf
  double* vbuf;
};

__device__ void sv_solve_world(const ShapeDev& s, const int* pool, const double* fv, double* y, double* vec,
                               double* lds);

template <bool SN>
__device__ __forceinline__ void solve_any(const ShapeDev& s, const PlanView& v, const int* pool,
                                          const double* fv, double* y, const SolveArgs& sa, int w,
                                          double* lds) {
  if constexpr (SN) sv_solve_world(s, pool, fv, y, sa.vbuf + sa.voffw[w], lds);
  else solve_world(s, v, fv, y);
}

// Solve (C + coef G) dphi = q(K) on each world's factor: rhs' = q s, dphi = x s.
template <bool SN>
__global__ __launch_bounds__(kMfThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_solve_stage(Params P, const ShapeDev* shapes,
                                                            const int* shape_of, const int* pool,
                                                            const long* foff, int w0,
                                                            const double* msc,
                                                            const double* fvals, const double* K,
                                                            const double* Kminus2, double* dphi,
                                                            double* ysc, int y_in_lds, SolveArgs sa,
                                                            int part, int rhs_ready) {
  extern __shared__ double lds[];
  int w = launch_world(P, w0);
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool);
  const double* fv = fvals + foff[w];
  int c0 = P.cell_off[w];
  double* y = y_in_lds ? lds : ysc + c0;
  if (part != 2 && !rhs_ready) {
    // the right-hand side by rank: kRhsBatch ranks per thread at once, so that their scattered
    // loads are in flight together (the stores come after every load of the batch)
    constexpr int kRhsBatch = 2;
    for (int k0 = threadIdx.x; k0 < s.n; k0 += kRhsBatch * (int)blockDim.x) {
      int idx[kRhsBatch];
#pragma unroll
      for (int u = 0; u < kRhsBatch; ++u) {
        const int k = k0 + u * (int)blockDim.x;
        idx[u] = k < s.n ? c0 + v.order[k] : -1;
      }
      double q[kRhsBatch], sc[kRhsBatch];
#pragma unroll
      for (int u = 0; u < kRhsBatch; ++u) {
        const int i = idx[u] < 0 ? c0 : idx[u];
        if (Kminus2) {
          // the charge of K - 2 K2' field by field (the stage-2 right-hand side F1 - 2 k1)
          double t2 = P.z[0] * (CFc(P, K, P.f_dn, i) + -2.0 * CFc(P, Kminus2, P.f_dn, i));
#pragma unroll
          for (int sp = 1; sp < 4; ++sp)
            if (sp < P.S)
            t2 = t2 + P.z[sp] * (CFc(P, K, P.f_dn + sp, i) + -2.0 * CFc(P, Kminus2, P.f_dn + sp, i));
          q[u] = kChargePerFmol * t2;
        } else {
          q[u] = charge_increment(P, K, i);
        }
        sc[u] = msc[i];
      }
#pragma unroll
      for (int u = 0; u < kRhsBatch; ++u)
        if (idx[u] >= 0) y[k0 + u * (int)blockDim.x] = q[u] * sc[u];
    }
  }
  if (part == 0) {
    __syncthreads();
    solve_any<SN>(s, v, pool, fv, y, sa, w, lds);
    __syncthreads();
  }
  if (part != 1) {
    for (int k = threadIdx.x; k < s.n; k += blockDim.x) {
      int i = c0 + v.order[k];
      dphi[i] = y[k] * msc[i];
    }
  }
}

// Test entry: factor given assembled values in place, per world.
__global__ __launch_bounds__(kFactorBlock) void k_factor_only(const ShapeDev* shapes,
                                                              const int* shape_of, const int* pool,
                                                              const unsigned long long* masks,
                                                              const long* foff, double* fvals) {
  int w = blockIdx.x;
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool, masks);
  extern __shared__ double lds[];
  double* fv = fvals + foff[w];
  for (int p = s.lds_pt + threadIdx.x; p < s.nnz; p += blockDim.x) lds[p - s.lds_pt] = fv[p];
  __syncthreads();
  factor_world(s, v, FactorStore{fv, lds, s.lds_pt});
  for (int p = s.lds_pt + threadIdx.x; p < s.nnz; p += blockDim.x) fv[p] = lds[p - s.lds_pt];
}

template <bool SN>
__global__ __launch_bounds__(kMfThreads) void k_solve_only(const ShapeDev* shapes, const int* shape_of,
                                                           const int* pool, const long* foff,
                                                           const int* cell_off, const double* fvals,
                                                           const double* rhs, double* x, double* ysc,
                                                           int y_in_lds, SolveArgs sa, int part) {
  extern __shared__ double shared_y[];
  int w = blockIdx.x;
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool);
  int c0 = cell_off[w];
  double* lds = y_in_lds ? shared_y : ysc + c0;
  if (part != 2)
    for (int k = threadIdx.x; k < s.n; k += blockDim.x) lds[k] = rhs[c0 + v.order[k]];
  if (part == 0) {
    __syncthreads();
    solve_any<SN>(s, v, pool, fvals + foff[w], lds, sa, w, shared_y);
    __syncthreads();
  }
  if (part != 1)
    for (int k = threadIdx.x; k < s.n; k += blockDim.x) x[c0 + v.order[k]] = lds[k];
}

// ------------------------------------------------------------------------------------------
// Multifrontal supernodal factor (the default): the same factor L on the same pattern and
// positions, with the sums grouped by the elimination tree instead of the reference's
// adjacent-pairs fold over each entry's sources. Supernode by supernode, leaves first: the front
// (dense, packed lower, column-major) is the supernode's assembled columns plus its children's
// update matrices (extend-add through the children's row positions), it is factored over the
// supernode's columns in LDS (pivot = sqrt, column times the pivot's reciprocal, right-looking
// rank-1 updates of the trailing triangle), the columns go back to the factor and the trailing
// triangle becomes the supernode's update matrix. One workgroup per world; a front is held by one
// thread, one wavefront or the whole workgroup according to its size.


__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Groups of up to 64 lanes lie in one wavefront and follow one front's control flow together.
template <int G>
__device__ __forceinline__ void group_sync() {
  if constexpr (G > 1 && G <= 64) wave_sync();
  if constexpr (G == kMfThreads) __syncthreads();
}

// Packed lower column-major triangle of order m: entry (i, j), i >= j, at colbase(j, m) + i.
__device__ __forceinline__ int colbase(int j, int m) { return j * m - (j * (j + 1)) / 2; }

// Entry e of a packed lower column-major triangle of order T: column b, row a (a >= b). Column b
// starts at b T - b (b - 1) / 2; the root of that quadratic, then one correction either way.
__device__ __forceinline__ void tri_locate(int T, int e, int& b, int& a) {
  const float h = (float)(2 * T + 1);
  b = (int)((h - sqrtf(h * h - 8.0f * (float)e)) * 0.5f);
  b = b < 0 ? 0 : (b > T - 1 ? T - 1 : b);
  if (b + 1 < T && (b + 1) * T - (b + 1) * b / 2 <= e) ++b;
  if (b * T - b * (b - 1) / 2 > e) --b;
  a = e - (b * T - b * (b - 1) / 2) + b;
}

// Lane layout of a group of G lanes over a front's rows and columns: RW lanes down a column (row
// lr + t RW, t < NT, covering order <= NT RW) and G / RW column groups.
template <int G>
struct FrontLanes {
  static constexpr int RW = G < 64 ? G : 64;
  static constexpr int CG = G / RW;
  static constexpr int rows = G == 1 ? max_order(kMfThreadTri) : G == 4 ? max_order(kLatGroupTri[1])
                              : G == kMfOcta ? max_order(kMfOctaTri) : G == 16 ? max_order(kLatGroupTri[3])
                              : G == 32 ? max_order(kLatGroupTri[4]) : G == 64 ? max_order(kMfWaveTri)
                              : kMfMaxRows;
  static constexpr int NT = (rows + RW - 1) / RW;  // rows per lane
};

constexpr int kMfKB = 8;  // columns per block of a wavefront's or workgroup's partial factor

typedef double mf_v4d __attribute__((ext_vector_type(4)));

// Front columns k >= je, rows i >= k: F(i, k) -= sum_{jb <= l < je} F(i, l) F(k, l), on 16 x 16
// tiles of v_mfma_f64_16x16x4f64 (lane l: A(l % 16, l / 16), B(l / 16, l % 16), result register r
// at (l / 16 + 4 r, l % 16)); the group's wavefronts share the tiles. F packed lower, stride 1.
template <int G>
__device__ void mf_trailing_update(double* F, int m, int jb, int je, int lane) {
  const int l = lane & 63, wv = lane >> 6;
  const int lrow = l & 15, lq = l >> 4;
  const int nt = (m - je + 15) >> 4;
  int tile = 0;
  for (int ti = 0; ti < nt; ++ti) {
    for (int tk = 0; tk <= ti; ++tk, ++tile) {
      if (tile % (G / 64) != wv) continue;
      const int i0 = je + 16 * ti, k0 = je + 16 * tk;
      mf_v4d c;
      const int kc = k0 + lrow;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + lq + 4 * r;
        c[r] = (i < m && kc < m && i >= kc) ? F[colbase(kc, m) + i] : 0.0;
      }
      for (int l0 = jb; l0 < je; l0 += 4) {
        const int col = l0 + lq;
        const int ia = i0 + lrow, kb = k0 + lrow;
        const double a = (col < je && ia < m) ? -F[colbase(col, m) + ia] : 0.0;
        const double b = (col < je && kb < m) ? F[colbase(col, m) + kb] : 0.0;
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + lq + 4 * r;
        if (i < m && kc < m && i >= kc) F[colbase(kc, m) + i] = c[r];
      }
    }
  }
}

// A double broadcast from lane `src` (uniform) of the wavefront.
__device__ __forceinline__ double readlane_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (long long)(unsigned)lo);
}

// Columns [jb, je) (at most kMfKB) of a packed lower front of order m <= 64 NT, factored by one
// wavefront in its registers: lane lr holds rows lr + 64 t of the block's columns. Column j: pivot
// d = sqrt(F(j, j)), the rows below times 1/d, then F(i, k) -= F(i, j) F(k, j) for the block's
// columns k > j, each F(k, j) broadcast from the lane that holds row k. These are the operations
// of the rank-1 steps in LDS, in the same order, so the block's values are the same bits.
template <int NT, int STRIDE>
__device__ __forceinline__ void mf_panel_registers(double* F, int m, int jb, int je, int lr) {
  double P[NT][kMfKB];
  const int kb = je - jb;
#pragma unroll
  for (int c = 0; c < kMfKB; ++c) {
    const int j = jb + c;
    const int cb = colbase(j, m);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int i = lr + 64 * t;
      P[t][c] = (c < kb && i >= j && i < m) ? F[(cb + i) * STRIDE] : 0.0;
    }
  }
#pragma unroll
  for (int c = 0; c < kMfKB; ++c) {
    if (c < kb) {
      const int j = jb + c;
      double pv = P[0][c];
#pragma unroll
      for (int t = 1; t < NT; ++t)
        if ((j >> 6) == t) pv = P[t][c];
      const double d = sqrt(readlane_d(pv, j & 63));
      const double rinv = 1.0 / d;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int i = lr + 64 * t;
        if (i > j && i < m) P[t][c] = P[t][c] * rinv;
        else if (i == j) P[t][c] = d;
      }
#pragma unroll
      for (int c2 = c + 1; c2 < kMfKB; ++c2) {
        if (c2 < kb) {
          const int k = jb + c2;
          double bv = P[0][c];
#pragma unroll
          for (int t = 1; t < NT; ++t)
            if ((k >> 6) == t) bv = P[t][c];
          const double bk = readlane_d(bv, k & 63);
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int i = lr + 64 * t;
            if (i >= k && i < m) P[t][c2] -= P[t][c] * bk;
          }
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kMfKB; ++c) {
    const int j = jb + c;
    const int cb = colbase(j, m);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int i = lr + 64 * t;
      if (c < kb && i >= j && i < m) F[(cb + i) * STRIDE] = P[t][c];
    }
  }
}

// One front, by G lanes (this lane = lane) with its packed triangle at F[e * STRIDE]; fronts of at
// most MR rows.
// The stage matrix's entries scattered straight into the fronts (the lattice factor): each front
// starts at zero, takes its columns' diagonal mdiag and its columns' junction entries
// -coef G_e Jacobi-scaled and symmetrised (the values the assembly would have stored), so the
// factor array is neither cleared nor read before it is written.
struct MfAssembly {
  const double* mdiag;  // per cell, the stage matrix's scaled diagonal
  const double* msc;    // per cell, the Jacobi scale
  const double* Gj;     // per junction, the conductance slope
  const int* snd;
  const int* rcv;
  const int* order;     // the world's rank -> node
  double coef;
  int c0, e0;           // the world's first cell and junction
};

// REG_PANEL: wavefront and workgroup fronts of at most 128 rows take each column block's rank-1
// steps in one wavefront's registers (mf_panel_registers) instead of in LDS.
template <int G, int STRIDE, int MR = FrontLanes<G>::rows, bool ASM = false, bool REG_PANEL = true>
__device__ void mf_front(const int* pool, int qa, double* fv, double* ub, double* F, int lane, int* valid_w,
                         const MfAssembly* am = nullptr) {
  using FL = FrontLanes<G>;
  constexpr int RW = FL::RW, CG = FL::CG, NT = (MR + RW - 1) / RW;
  const int lr = lane % RW, lc = lane / RW;
  const int4 ra4 = reinterpret_cast<const int4*>(pool)[4 * qa];
  const int4 rb4 = reinterpret_cast<const int4*>(pool)[4 * qa + 1];
  const int J = ra4.y, m = ra4.z, uoff = ra4.w, c0 = rb4.x, c1 = rb4.y;
  const long p0 = ra4.x;
  const int S = m - J, tri = m * (m + 1) / 2, Pn = colbase(J, m) + J;
  if constexpr (ASM) {
    const int2 ec = reinterpret_cast<const int2*>(pool)[(kRecInts * qa + 10) / 2];  // junction list
    const int2* el = reinterpret_cast<const int2*>(pool + ec.x);
    // kAsmBatch diagonal entries and kAsmBatch junction entries per lane at a time, each level of
    // their index chains (column -> cell -> diagonal; entry -> junction -> ends -> scales) loaded
    // for the whole batch before the next; the first batch's first level goes out before the
    // front is cleared
    for (int b0 = 0; b0 < J || b0 < ec.y; b0 += kAsmBatch * G) {
      int cell[kAsmBatch];
      int2 le[kAsmBatch];  // (front-local packed position, world-local junction)
#pragma unroll
      for (int u = 0; u < kAsmBatch; ++u) {
        const int jj = b0 + lane + u * G;
        cell[u] = jj < J ? am->order[rb4.w + jj] : -1;
        le[u] = jj < ec.y ? el[jj] : make_int2(-1, 0);
      }
      if (b0 == 0) {
        for (int e = lane; e < tri; e += G) F[e * STRIDE] = 0.0;
        group_sync<G>();
      }
      double dv[kAsmBatch], gv[kAsmBatch];
      int ea[kAsmBatch], eb[kAsmBatch];
#pragma unroll
      for (int u = 0; u < kAsmBatch; ++u) {
        dv[u] = cell[u] >= 0 ? am->mdiag[am->c0 + cell[u]] : 0.0;
        const int e = am->e0 + le[u].y;
        const bool j_in = le[u].x >= 0;
        gv[u] = j_in ? am->Gj[e] : 0.0;
        ea[u] = j_in ? am->snd[e] : 0;
        eb[u] = j_in ? am->rcv[e] : 0;
      }
      double sa[kAsmBatch], sb[kAsmBatch];
#pragma unroll
      for (int u = 0; u < kAsmBatch; ++u) {
        const bool j_in = le[u].x >= 0;
        sa[u] = j_in ? am->msc[ea[u]] : 0.0;
        sb[u] = j_in ? am->msc[eb[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kAsmBatch; ++u) {
        const int jj = b0 + lane + u * G;
        if (cell[u] >= 0) F[(colbase(jj, m) + jj) * STRIDE] = dv[u];
        if (le[u].x >= 0) {
          const double val = -(am->coef * gv[u]);
          F[le[u].x * STRIDE] = (val * sa[u] * sb[u] + val * sb[u] * sa[u]) / 2;
        }
      }
    }
    if (J == 0 && ec.y == 0) {
      for (int e = lane; e < tri; e += G) F[e * STRIDE] = 0.0;
    }
  } else {
    // the supernode's columns are contiguous in the factor and are the front's first J columns;
    // the update block starts at zero
    for (int e0 = lane; e0 < tri; e0 += kMfBatch * G) {
      double v[kMfBatch];
#pragma unroll
      for (int u = 0; u < kMfBatch; ++u) {
        const int e = e0 + u * G;
        v[u] = e < Pn ? fv[p0 + e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kMfBatch; ++u) {
        const int e = e0 + u * G;
        if (e < tri) F[e * STRIDE] = v[u];
      }
    }
  }
  group_sync<G>();
  // extend-add: child entry (a, b), a >= b, of the packed Sc x Sc update goes to front entry
  // (rel[a], rel[b]); lanes stride over the packed entries (coalesced), kMfBatch loads in flight
  for (int x = c0; x < c1; ++x) {
    const int4 kd = reinterpret_cast<const int4*>(pool)[x];  // (Sc, rel start, update offset, -)
    const int Sc = kd.x;
    const int* rl = pool + kd.y;
    const double* U = ub + kd.z;
    const int nU = Sc * (Sc + 1) / 2;
    for (int e0 = lane; e0 < nU; e0 += kMfBatch * G) {
      int ta[kMfBatch];
      double uv[kMfBatch];
#pragma unroll
      for (int u = 0; u < kMfBatch; ++u) {
        const int e = e0 + u * G;
        const bool in = e < nU;
        int b, a;
        tri_locate(Sc, in ? e : 0, b, a);
        const int ra = in ? rl[a] : 0, rb = in ? rl[b] : 0;
        uv[u] = in ? U[e] : 0.0;
        ta[u] = in ? colbase(rb, m) + ra : -1;
      }
#pragma unroll
      for (int u = 0; u < kMfBatch; ++u)
        if (ta[u] >= 0) F[ta[u] * STRIDE] += uv[u];
    }
    group_sync<G>();
  }
  // partial factor over the supernode's columns. Groups of whole wavefronts work in blocks of
  // kMfKB columns: the block's rank-1 steps in the registers of one wavefront (in a workgroup's
  // front the first; in LDS for fronts above 128 rows), then one rank-kMfKB update of everything
  // to its right on the matrix cores, so the group meets twice per block. Narrower groups take
  // rank-1 steps over the whole front in LDS (lanes own rows i = jj + 1 + lr + t RW).
  constexpr int KB = G >= 64 ? kMfKB : 1 << 20;
  for (int jb = 0; jb < J; jb += KB) {
    const int je = J < jb + KB ? J : jb + KB;
    const int kend = G >= 64 ? je : m;  // columns the rank-1 steps update
    if constexpr (REG_PANEL && G >= 64 && NT <= 2) {
      // the block's columns in one wavefront's registers (lane lr holds rows lr + 64 t), their
      // rank-1 steps with the pivots and the block's rows broadcast from the owning lanes: the
      // same operations in the same order as the steps in LDS below, with no barrier inside
      if (lc == 0) mf_panel_registers<NT, STRIDE>(F, m, jb, je, lr);
      group_sync<G>();  // the block's columns are factored
    } else {
      // rank-1 steps in LDS: over the whole front for narrower groups; for wavefront groups (fronts
      // above 128 rows) over the block, by the first wavefront alone in a workgroup's front
      constexpr bool WAVE_PANEL = G > 64;
      constexpr int PCG = WAVE_PANEL ? 1 : CG;  // column groups of the rank-1 steps
      if (!WAVE_PANEL || lc == 0) {
        for (int jj = jb; jj < je; ++jj) {
          const int cb = colbase(jj, m);
          const double d = sqrt(F[(cb + jj) * STRIDE]);
          const double rinv = 1.0 / d;
          double av[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int i = jj + 1 + lr + t * RW;
            av[t] = i < m ? F[(cb + i) * STRIDE] * rinv : 0.0;
          }
          if constexpr (WAVE_PANEL) wave_sync();
          else group_sync<G>();  // every lane has read the pivot and its rows
          if (lc == 0) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              const int i = jj + 1 + lr + t * RW;
              if (i < m) F[(cb + i) * STRIDE] = av[t];
            }
          }
          if (lane == 0) F[(cb + jj) * STRIDE] = d;
          if constexpr (WAVE_PANEL) wave_sync();
          else group_sync<G>();  // the scaled column is in place
          const int pc = WAVE_PANEL ? 0 : lc;
#pragma unroll 4
          for (int k = jj + 1 + pc; k < kend; k += PCG) {
            const double bk = F[(cb + k) * STRIDE];
            const int ck = colbase(k, m);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              const int i = jj + 1 + lr + t * RW;
              if (i >= k && i < m) F[(ck + i) * STRIDE] -= av[t] * bk;
            }
          }
          if constexpr (WAVE_PANEL) wave_sync();
          else group_sync<G>();
        }
      }
      if constexpr (WAVE_PANEL) group_sync<G>();  // the block's columns are factored
    }
    if constexpr (G >= 64) {
      if (je < m) {
        mf_trailing_update<G>(F, m, jb, je, lane);
        group_sync<G>();
      }
    }
  }
  bool ok = true;
  for (int e = lane; e < Pn; e += G) {
    const double v = F[e * STRIDE];
    ok = ok && isfinite(v);
    fv[p0 + e] = v;
  }
  if (!ok && valid_w) *valid_w = 0;  // a pivot at or below zero, or an overflow
  if (S > 0)
    for (int e = Pn + lane; e < tri; e += G) ub[uoff + e - Pn] = F[e * STRIDE];
  group_sync<G>();
}

// One class of a level's fronts in the factor: fronts [q0, q1) of the plan, a front per group of G
// lanes, each group's front at base + group * TRI (G = kMfThreads: one front at a time, at base).
// The thread index is made opaque here for the reason sv_class gives.
template <int G, int TRI, int MR, bool ASM>
__device__ __forceinline__ void mf_class(const int* pool, int qa0, int q0, int q1, double* fv, double* ub,
                                         double* base, int* valid_w, const MfAssembly* am) {
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  if constexpr (G == 1) {
    // a wavefront's slice of the region (base is the wave's), a thread's front interleaved across it
    const int lane = tid & 63;
    for (int q = q0 + tid; q < q1; q += kMfThreads)
      mf_front<1, 64, MR, ASM>(pool, qa0 + q, fv, ub, base + lane, 0, valid_w, am);
  } else if constexpr (G == kMfThreads) {
    for (int q = q0; q < q1; ++q) mf_front<kMfThreads, 1, MR, ASM>(pool, qa0 + q, fv, ub, base, tid, valid_w, am);
  } else {
    // groups of G lanes within one wavefront, each with its TRI doubles of the wave's slice
    const int lane = tid & 63;
    for (int q = q0 + tid / G; q < q1; q += kMfThreads / G)
      mf_front<G, 1, MR, ASM>(pool, qa0 + q, fv, ub, base + (lane / G) * TRI, tid % G, valid_w, am);
  }
}

// The factor of one world; fronts in HBM scratch hold at most BIGROWS rows.
template <int BIGROWS = kMfMaxRows, bool ASM = false>
__device__ __forceinline__ void mf_factor_world(const ShapeDev& s, const int* pool, double* fv, double* ub,
                                                double* lds, double* big, int* valid_w,
                                                const MfAssembly* am = nullptr) {
  const int* lev = pool + s.o_snlev;
  const int qa0 = (int)(s.o_sn / kRecInts);
  for (int L = 0; L < s.sn_levels; ++L) {
    // the fronts of level L by class: groups of 1, 4, 8, 16, 32 and 64 lanes, the workgroup with
    // its front in LDS, the workgroup in HBM scratch. Every group up to a wavefront lies within one
    // wavefront and works in that wavefront's slice of the LDS region (a quarter), so each wavefront
    // takes its share of every class in turn with no barrier between classes (its lanes reconverge
    // between them); the workgroup meets once, when the level's small fronts are done (their update
    // matrices handed on, the region free for the workgroup classes).
    const int* bd = lev + kLatBounds * L;
    if (bd[0] < bd[6]) {
      double* R = lds + (threadIdx.x >> 6) * kMfBlockWaveTri;
      if (bd[0] < bd[1]) {
        mf_class<1, kLatGroupTri[0], FrontLanes<1>::rows, ASM>(pool, qa0, bd[0], bd[1], fv, ub, R, valid_w, am);
        wave_sync();
      }
      if (bd[1] < bd[2]) {
        mf_class<4, kLatGroupTri[1], FrontLanes<4>::rows, ASM>(pool, qa0, bd[1], bd[2], fv, ub, R, valid_w, am);
        wave_sync();
      }
      if (bd[2] < bd[3]) {
        mf_class<kMfOcta, kLatGroupTri[2], FrontLanes<kMfOcta>::rows, ASM>(pool, qa0, bd[2], bd[3], fv, ub, R,
                                                                         valid_w, am);
        wave_sync();
      }
      if (bd[3] < bd[4]) {
        mf_class<16, kLatGroupTri[3], FrontLanes<16>::rows, ASM>(pool, qa0, bd[3], bd[4], fv, ub, R, valid_w, am);
        wave_sync();
      }
      if (bd[4] < bd[5]) {
        mf_class<32, kLatGroupTri[4], FrontLanes<32>::rows, ASM>(pool, qa0, bd[4], bd[5], fv, ub, R, valid_w, am);
        wave_sync();
      }
      if (bd[5] < bd[6])
        mf_class<64, kMfBlockWaveTri, max_order(kMfBlockWaveTri), ASM>(pool, qa0, bd[5], bd[6], fv, ub, R,
                                                                       valid_w, am);
      __syncthreads();
    }
    // the workgroup classes end each front at a barrier of their own
    mf_class<kMfThreads, 0, FrontLanes<kMfThreads>::rows, ASM>(pool, qa0, bd[6], bd[7], fv, ub, lds, valid_w, am);
    mf_class<kMfThreads, 0, BIGROWS, ASM>(pool, qa0, bd[7], bd[8], fv, ub, big, valid_w, am);
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Supernodal solve (the default, with the multifrontal factor): y (by rank) <- L^-T L^-1 y over
// the same fronts. Forward, leaves first: a front's vector is its columns' right-hand sides plus
// its children's update vectors (extend-add), its pivots run in registers (each pivot applied as
// a multiplication by its reciprocal, as the device form of the reference's solve does,
// numerics/sparse_cholesky.py:19-31), the solved values go to y and the rows below become the
// update vector. Backward, roots first: the values below a front are its ancestors' solution
// (gathered from y), one product with the front's rows below, then the pivots last to first.

// Rows per lane of a solve group: one thread, 8 lanes, or a wavefront (the factor's wavefront and
// workgroup fronts alike, order <= kMfMaxRows).
template <int G>
struct SolveLanes {
  static constexpr int rows = G == 1 ? max_order(kMfThreadTri) : G == 4 ? max_order(kLatGroupTri[1])
                              : G == kMfOcta ? max_order(kMfOctaTri) : G == 16 ? max_order(kLatGroupTri[3])
                              : G == 32 ? max_order(kLatGroupTri[4]) : kMfMaxRows;
  static constexpr int NT = (rows + G - 1) / G;
};
static_assert(FrontLanes<1>::NT * 1 >= max_order(kMfThreadTri) && FrontLanes<kMfThreads>::NT * 64 >= kMfMaxRows,
              "every front row has a lane");
// LDS of the solve's workgroup: separate regions for the thread, 8-lane and wavefront fronts (the
// three kinds of one level run without a barrier between them)
constexpr int kSvThreadLds = kMfThreads * SolveLanes<1>::NT;
constexpr int kSvOctaLds = (kMfThreads / kMfOcta) * kMfOcta * SolveLanes<kMfOcta>::NT;
constexpr int kSvWaveLds = (kMfThreads / 64) * 64 * SolveLanes<64>::NT;
constexpr int kSvLds = kSvThreadLds + kSvOctaLds + kSvWaveLds;
// The lattice solve: each wavefront runs its groups' fronts of a level class after class (the
// wavefront's lanes reconverge between classes), so every class of one wavefront shares one
// region, sized for the largest: 64 thread fronts, 16 four-lane, 8 eight-lane, 4 sixteen-lane,
// 2 thirty-two-lane fronts or one wavefront front.
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int kLatSvWave = cmax(cmax(64 * SolveLanes<1>::NT, 64 * SolveLanes<4>::NT),
                                cmax(cmax(64 * SolveLanes<kMfOcta>::NT, 64 * SolveLanes<16>::NT),
                                     cmax(64 * SolveLanes<32>::NT, 64 * SolveLanes<64>::NT)));
constexpr int kLatSvLds = (kMfThreads / 64) * kLatSvWave;

template <int RW>
__device__ __forceinline__ double group_shfl(double v, int src) {
  if constexpr (RW == 1) return v;
  else return __shfl(v, src, RW);
}

template <int NT>
__device__ __forceinline__ double pick(const double (&a)[NT], int idx) {
  double r = a[0];
#pragma unroll
  for (int t = 1; t < NT; ++t)
    if (idx == t) r = a[t];
  return r;
}

template <int G, int STRIDE>
__device__ void sv_forward_front(const int* pool, int qa, const double* __restrict__ fv, double* __restrict__ y,
                                 double* __restrict__ vec, double* Fv, int lane) {
  constexpr int NT = SolveLanes<G>::NT;
  const int4 ra4 = reinterpret_cast<const int4*>(pool)[4 * qa];
  const int4 rb4 = reinterpret_cast<const int4*>(pool)[4 * qa + 1];
  const long p0 = ra4.x;
  const int J = ra4.y, m = ra4.z, c0 = rb4.x, c1 = rb4.y, voff = rb4.z, j0 = rb4.w;
  for (int i = lane; i < m; i += G) Fv[i * STRIDE] = i < J ? y[j0 + i] : 0.0;
  group_sync<G>();
  for (int x = c0; x < c1; ++x) {
    const int4 kd = reinterpret_cast<const int4*>(pool)[x];  // (Sc, rel start, -, vector offset)
    const int* rl = pool + kd.y;
    const double* u = vec + kd.w;
    for (int a = lane; a < kd.x; a += G) Fv[rl[a] * STRIDE] += u[a];
    group_sync<G>();
  }
  double fr[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int i = lane + t * G;
    fr[t] = i < m ? Fv[i * STRIDE] : 0.0;
  }
  auto column = [&](int jj) {
    const long cbp = p0 + colbase(jj, m);
    const double d = fv[cbp + jj];
    double lv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int i = lane + t * G;
      lv[t] = (i > jj && i < m) ? fv[cbp + i] : 0.0;
    }
    const double xj = group_shfl<G>(pick(fr, jj / G), jj % G) * (1.0 / d);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int i = lane + t * G;
      if (i > jj && i < m) fr[t] -= lv[t] * xj;
    }
    if (lane == jj % G) y[j0 + jj] = xj;
  };
  if constexpr (G <= kMfOcta) {
    // a narrow group's front (J <= its rows): the columns unrolled, so its rows stay in registers
#pragma unroll
    for (int jj = 0; jj < SolveLanes<G>::rows; ++jj)
      if (jj < J) column(jj);
  } else {
#pragma unroll 4
    for (int jj = 0; jj < J; ++jj) column(jj);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int i = lane + t * G;
    if (i >= J && i < m) vec[voff + i - J] = fr[t];
  }
  group_sync<G>();
}

// UNROLL: loads in flight per lane in the products with the rows below.
template <int G, int STRIDE, int UNROLL = 1>
__device__ void sv_backward_front(const int* pool, int qa, const double* __restrict__ fv, double* __restrict__ y,
                                  double* Fv, int lane) {
  constexpr int NT = SolveLanes<G>::NT;
  const int4 ra4 = reinterpret_cast<const int4*>(pool)[4 * qa];
  const int4 rb4 = reinterpret_cast<const int4*>(pool)[4 * qa + 1];
  const int rows_below = pool[kRecInts * qa + 8];  // ranks of the last column's rows after its diagonal
  const long p0 = ra4.x;
  const int J = ra4.y, m = ra4.z, j0 = rb4.w;
  // the solution at the rows below
  for (int a = lane; a < m - J; a += G) Fv[(J + a) * STRIDE] = y[pool[rows_below + a]];
  group_sync<G>();
  double xr[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int k = lane + t * G;
    double acc = 0.0;
    if (k < J) {
      const long ck = p0 + colbase(k, m);
      acc = y[j0 + k];
#pragma unroll UNROLL
      for (int i = J; i < m; ++i) acc -= fv[ck + i] * Fv[i * STRIDE];
    }
    xr[t] = acc;
  }
  auto column = [&](int jj) {
    const double d = fv[p0 + colbase(jj, m) + jj];
    double lv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int k = lane + t * G;
      lv[t] = k < jj ? fv[p0 + colbase(k, m) + jj] : 0.0;
    }
    const double zj = group_shfl<G>(pick(xr, jj / G), jj % G) * (1.0 / d);
    if (lane == jj % G) y[j0 + jj] = zj;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int k = lane + t * G;
      if (k < jj) xr[t] -= lv[t] * zj;
    }
  };
  if constexpr (G <= kMfOcta) {
#pragma unroll
    for (int jj = SolveLanes<G>::rows - 1; jj >= 0; --jj)
      if (jj < J) column(jj);
  } else {
    for (int jj = J - 1; jj >= 0; --jj) column(jj);
  }
  group_sync<G>();
}

// One class of a level's fronts in the solve (fronts [q0, q1) of the plan, a front per group of G
// lanes; G = 64 also takes the workgroup classes' fronts). The thread index is made opaque here,
// so that no value of one class is computed once above the level loop and held in registers
// through every other class (which made the kernel spill).
template <int G, bool FORWARD, int UNROLL>
__device__ __forceinline__ void sv_class(const int* pool, int qa0, int q0, int q1, const double* __restrict__ fv,
                                         double* __restrict__ y, double* __restrict__ vec, double* R) {
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63;
  if constexpr (G == 1) {
    for (int q = q0 + tid; q < q1; q += kMfThreads) {
      if constexpr (FORWARD) sv_forward_front<1, 64>(pool, qa0 + q, fv, y, vec, R + lane, 0);
      else sv_backward_front<1, 64>(pool, qa0 + q, fv, y, R + lane, 0);
    }
    wave_sync();
  } else {
    double* F = R + (lane / G) * (G * SolveLanes<G>::NT);
    for (int q = q0 + tid / G; q < q1; q += kMfThreads / G) {
      if constexpr (FORWARD) sv_forward_front<G, 1>(pool, qa0 + q, fv, y, vec, F, tid % G);
      else sv_backward_front<G, 1, UNROLL>(pool, qa0 + q, fv, y, F, tid % G);
    }
  }
}

// The solve of one world's y (by rank) on its supernodal factor; blockDim.x == kMfThreads, LDS
// kLatSvLds doubles. A wavefront takes its share of each class of a level in turn (one front per
// group of 1, 4, 8, 16, 32 or 64 of its lanes; the workgroup classes' fronts go to wavefronts too,
// up to kMfMaxRows rows), in one region of its own.
__device__ void sv_solve_world(const ShapeDev& s, const int* pool, const double* fv, double* y, double* vec,
                               double* lds) {
  const int* lev = pool + s.o_snlev;
  const int qa0 = (int)(s.o_sn / kRecInts);
  double* R = lds + (threadIdx.x >> 6) * kLatSvWave;
  for (int L = 0; L < s.sn_levels; ++L) {
    const int* bd = lev + kLatBounds * L;
    sv_class<1, true, 1>(pool, qa0, bd[0], bd[1], fv, y, vec, R);
    sv_class<4, true, 1>(pool, qa0, bd[1], bd[2], fv, y, vec, R);
    sv_class<kMfOcta, true, 1>(pool, qa0, bd[2], bd[3], fv, y, vec, R);
    sv_class<16, true, 1>(pool, qa0, bd[3], bd[4], fv, y, vec, R);
    sv_class<32, true, 1>(pool, qa0, bd[4], bd[5], fv, y, vec, R);
    sv_class<64, true, 1>(pool, qa0, bd[5], bd[8], fv, y, vec, R);
    __syncthreads();
  }
  for (int L = s.sn_levels - 1; L >= 0; --L) {
    const int* bd = lev + kLatBounds * L;
    sv_class<1, false, 1>(pool, qa0, bd[0], bd[1], fv, y, vec, R);
    sv_class<4, false, 4>(pool, qa0, bd[1], bd[2], fv, y, vec, R);
    sv_class<kMfOcta, false, 4>(pool, qa0, bd[2], bd[3], fv, y, vec, R);
    sv_class<16, false, 8>(pool, qa0, bd[3], bd[4], fv, y, vec, R);
    sv_class<32, false, 8>(pool, qa0, bd[4], bd[5], fv, y, vec, R);
    sv_class<64, false, 8>(pool, qa0, bd[5], bd[8], fv, y, vec, R);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kMfThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_factor_mf(Params P, const ShapeDev* shapes,
                                                          const int* shape_of, const int* pool,
                                                          const long* foff, const long* uoffw, int w0,
                                                          double coef, const double* mdiag,
                                                          const double* msc, const double* Gj,
                                                          double* fvals, double* ubuf, double* bigbuf,
                                                          const long* boffw, int* valid) {
  extern __shared__ double lds[];
  int w = launch_world(P, w0);
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool);
  double* fv = fvals + foff[w];
  const MfAssembly am{mdiag, msc, Gj, P.snd, P.rcv, v.order, coef, P.cell_off[w], P.edge_off[w]};
  mf_factor_world<kMfMaxRows, true>(s, pool, fv, ubuf + uoffw[w], lds, bigbuf + boffw[w], valid + w, &am);
}

// Test entry: the multifrontal factor of assembled values in place, per world.
__global__ __launch_bounds__(kMfThreads) void k_factor_only_mf(const ShapeDev* shapes, const int* shape_of,
                                                               const int* pool, const long* foff,
                                                               const long* uoffw, double* fvals, double* ubuf,
                                                               double* bigbuf, const long* boffw) {
  extern __shared__ double lds[];
  int w = blockIdx.x;
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool);
  mf_factor_world(s, pool, fvals + foff[w], ubuf + uoffw[w], lds, bigbuf + boffw[w], nullptr);
}

// ------------------------------------------------------------------------------------------
// ROS2 stage combinations (the W-method's k = F + coef (df/dphi) dphi on the route fields).

// The W-method's correction coef (df/dphi) dphi of one cell: dn_c[s] on the amounts, dm_c[s] on
// the membrane transfers (the bath supply takes -dm_c). With Gj (RES) also the stage matrix's
// row times dphi, (C + coef G) dphi at the cell (its diagonal summed as k_cells sums it), and the
// row's magnitude (the sum of the absolute terms of that product).
// XB > 1: the incident edges are gathered XB at a time, every load of a batch issued before any is
// used, so a wavefront's requests for the lines its cells share arrive together (the lattice numbers
// a cell's out-edges and its left neighbour's next to each other); the sums keep their order.
template <bool RES, int XB = 1>
__device__ __forceinline__ void cell_correction(const Params& P, int i, double coef, const double* gm,
                                                const double* gj, const double* Gj, const double* dphi,
                                                double* dn_c, double* dm_c, double* mx = nullptr,
                                                double* mag = nullptr) {
  int p0 = P.inc_ptr[i], p1 = P.inc_ptr[i + 1], pm = p0 + P.inc_out[i];
  // edges outer, species inner: each edge's index and potential drop are loaded once; every
  // species still sums its outgoing, then incoming, edges in incidence order
  double out[4] = {0.0, 0.0, 0.0, 0.0}, in[4] = {0.0, 0.0, 0.0, 0.0};
  double gsum = 0.0, off = 0.0, offmag = 0.0;
  if constexpr (XB > 1) {
    for (int pb = p0; pb < p1; pb += XB) {
      int e[XB];
#pragma unroll
      for (int u = 0; u < XB; ++u) e[u] = pb + u < p1 ? P.inc_edge[pb + u] : -1;
      int sn[XB], rc[XB];
#pragma unroll
      for (int u = 0; u < XB; ++u) {
        sn[u] = e[u] >= 0 ? P.snd[e[u]] : i;
        rc[u] = e[u] >= 0 ? P.rcv[e[u]] : i;
      }
      double ds[XB], dr[XB], ge[XB], gv[XB][4];
#pragma unroll
      for (int u = 0; u < XB; ++u) {
        ds[u] = dphi[sn[u]];
        dr[u] = dphi[rc[u]];
        ge[u] = (RES && e[u] >= 0) ? Gj[e[u]] : 0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) gv[u][s] = (s < P.S && e[u] >= 0) ? gj[(long)s * P.nedge + e[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < XB; ++u) {
        if (e[u] < 0) break;
        const double d = ds[u] - dr[u];
        const bool outgoing = pb + u < pm;
        if constexpr (RES) {
          const double t = ge[u] * (outgoing ? dr[u] : ds[u]);
          gsum += ge[u];
          off += t;
          offmag += fabs(t);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (s >= P.S) break;
          const double t = coef * (gv[u][s] * d);
          if (outgoing) out[s] += -t;
          else in[s] += t;
        }
      }
    }
  }
  for (int p = XB > 1 ? p1 : p0; p < p1; ++p) {
    const int e = P.inc_edge[p];
    const double ds = dphi[P.snd[e]], dr = dphi[P.rcv[e]];
    const double d = ds - dr;
    const bool outgoing = p < pm;
    if constexpr (RES) {
      const double ge = Gj[e];
      const double t = ge * (outgoing ? dr : ds);
      gsum += ge;
      off += t;
      offmag += fabs(t);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      const double t = coef * (gj[(long)s * P.nedge + e] * d);
      if (outgoing) out[s] += -t;
      else in[s] += t;
    }
  }
  const double dp = dphi[i];
  double Gm = 0.0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    const double g = gm[(long)s * P.ncell + i];
    if constexpr (RES) Gm = s == 0 ? P.z[s] * g : Gm + P.z[s] * g;
    double dm = coef * (g * dp);
    double cell = -dm;
    cell += out[s];
    cell += in[s];
    dn_c[s] = cell;
    dm_c[s] = dm;
  }
  if constexpr (RES) {
    const double diag = P.cap[i] + coef * (kChargePerFmol * Gm + gsum);
    *mx = diag * dp - coef * off;
    *mag = fabs(diag * dp) + fabs(coef) * offmag;
  }
}

// The stage solve's charge residual at a cell (numerics/surface_imex.py:269-278 solve_operator's
// (C + a L) phi - q, checked as models/surface_pump.py:325-330 checks it): |M dphi - q| <=
// atol + rtol scale. The stage equation here is in rate form (q is the charge rate of a stage's
// rates, which can vanish at a cell whose neighbours' terms do not), so the scale is the largest
// of |q|, |M dphi| and the row's absolute terms.
__device__ __forceinline__ bool stage_residual_ok(double q, double mx, double mag, double atol, double rtol) {
  const double scale = fmax(fmax(fabs(q), fabs(mx)), mag);
  return isfinite(mx) && fabs(mx - q) <= atol + rtol * scale;
}

// Adds coef (df/dphi) dphi to K's transfers and amounts (and the cell's supply contribution);
// returns whether the stage solve's residual at the cell holds (q: the solve's right-hand side).
// MEMBRANE = false: the amounts alone (the membrane transfers and the supply contribution are
// corrected where they are next read).
template <bool MEMBRANE>
__device__ bool correct_cell(const Params& P, int i, double coef, const double* gm, const double* gj,
                             const double* Gj, const double* dphi, double q, double atol, double rtol,
                             double* K, double* contrib) {
  double dn_c[4], dm_c[4], mx, mag;
  cell_correction<true, 4>(P, i, coef, gm, gj, Gj, dphi, dn_c, dm_c, &mx, &mag);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    CF(P, K, P.f_dn + s, i) += dn_c[s];
    if constexpr (MEMBRANE) {
      CF(P, K, P.f_mtr + s, i) += dm_c[s];
      contrib[(long)s * P.ncell + i] = contrib[(long)s * P.ncell + i] - dm_c[s];
    }
  }
  return stage_residual_ok(q, mx, mag, atol, rtol);
}

__device__ void correct_edge(const Params& P, int e, double coef, const double* gj, const double* dphi,
                             double* K) {
  double d = dphi[P.snd[e]] - dphi[P.rcv[e]];
#pragma unroll
  for (int s = 0; s < 4; ++s)
    if (s < P.S) EF(P, K, P.f_jtr + s, e) += coef * (gj[(long)s * P.nedge + e] * d);
}

// Stage 1: the stage solve's residual (its right-hand side is K1's charge rate before the
// correction); K1 += correction(dphi1); Y1 = y + dt K1 on the fields the rates read (amounts,
// gates, and the water volume, which the stage carries from the water half-step).
// EDGES = false leaves all but the amounts to the kernels after it: the second stage's junction
// rates form Y1's gates (k_edges_ahead), and the final combination adds K1's junction and membrane
// transfer corrections and the supply contribution's itself.
template <bool EDGES>
__global__ __launch_bounds__(kBlock) void k_ros2_stage1(Params P, const double* A, const double* B,
                                                        const int* sel, int c0, int nc, int e0,
                                                        int ne, double dt, double coef,
                                                        const double* gm, const double* gj,
                                                        const double* Gj, const double* dphi,
                                                        double atol, double rtol, double* K1,
                                                        double* contrib1, double* Y1, int* valid) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nc) {
    int i = c0 + t;
    int w = P.cell_world[i];
    int sw = sel[w];
    const double* y = current(A, B, sw);
    const double q = charge_increment(P, K1, i);
    if (!correct_cell<EDGES>(P, i, coef, gm, gj, Gj, dphi, q, atol, rtol, K1, contrib1)) {
      valid[w] = 0;
      atomicOr(P.refuse_bits + w, kRefuseResidual);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < P.S)
      CF(P, Y1, P.f_dn + s, i) = CFc(P, y, P.f_dn + s, i) + dt * CFc(P, K1, P.f_dn + s, i);
    if (P.water) {
      const double* yv = next(const_cast<double*>(A), const_cast<double*>(B), sw);
      CF(P, Y1, P.f_dvol, i) = CFc(P, yv, P.f_dvol, i);
    }
  } else if (EDGES && t < nc + ne) {
    int e = e0 + (t - nc);
    const double* y = current(A, B, sel[P.edge_world[e]]);
    correct_edge(P, e, coef, gj, dphi, K1);
    EF(P, Y1, P.f_gate, e) = EFc(P, y, P.f_gate, e) + dt * EFc(P, K1, P.f_gate, e);
  }
}

// Final: k2 = (F1 - 2 k1) + correction(dphi2), held in registers; the stage-2 solve's residual
// (its right-hand side is the charge rate of F1 - 2 k1); y_new = (y + 1.5 dt k1) + 0.5 dt k2; the
// cell's bath-supply contribution of k2; domain. F1 is read from K2. With dphi1 (the stage-1
// combination left the junctions alone) k1's junction transfers take their correction here, as
// that combination would have stored it.
__global__ __launch_bounds__(kBlock) void k_ros2_final(Params P, double* A, double* B,
                                                       const int* sel, int c0, int nc, int e0,
                                                       int ne, double dt, double coef,
                                                       const double* gm, const double* gj,
                                                       const double* Gj, const double* dphi,
                                                       double atol, double rtol, const double* K1,
                                                       const double* K2, double* contrib1,
                                                       double* contrib2, const double* dphi1, int* valid) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double h1 = 1.5 * dt, h2 = 0.5 * dt;
  if (t < nc) {
    int i = c0 + t;
    int w = P.cell_world[i];
    const double* y = current(A, B, sel[w]);
    double* o = next(A, B, sel[w]);
    double dn_c[4], dm_c[4], mx, mag;
    cell_correction<true>(P, i, coef, gm, gj, Gj, dphi, dn_c, dm_c, &mx, &mag);
    // with dphi1, k1's membrane-transfer correction (cell_correction's, from the cell's own dphi1)
    double dm1[4] = {0.0, 0.0, 0.0, 0.0};
    if (dphi1) {
      const double dp1 = dphi1[i];
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < P.S) dm1[s] = coef * (gm[(long)s * P.ncell + i] * dp1);
    }
    double qt = 0.0;  // the charge of F1 - 2 k1, summed as the stage-2 solve sums it
    // the amounts and membrane transfers (the stage-2 rate kernels finished the ledger fields)
    for (int f = 0; f < P.f_mtr + P.S; ++f) {
      double k1 = CFc(P, K1, f, i);
      if (dphi1 && f >= P.f_mtr) k1 += dm1[f - P.f_mtr];
      double k2 = CFc(P, K2, f, i) + -2.0 * k1;
      if (f >= P.f_dn && f < P.f_dn + P.S) {
        const int sp = f - P.f_dn;
        qt = sp == 0 ? P.z[0] * k2 : qt + P.z[sp] * k2;
        k2 += dn_c[sp];
      }
      if (f >= P.f_mtr && f < P.f_mtr + P.S) k2 += dm_c[f - P.f_mtr];
      CF(P, o, f, i) = (CFc(P, y, f, i) + h1 * k1) + h2 * k2;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      long q = (long)s * P.ncell + i;
      double c1 = contrib1[q];
      if (dphi1) {
        c1 = c1 - dm1[s];
        contrib1[q] = c1;
      }
      contrib2[q] = (contrib2[q] + -2.0 * c1) - dm_c[s];
    }
    bool ok = cell_domain_final(P, o, o, i);
    if (!stage_residual_ok(kChargePerFmol * qt, mx, mag, atol, rtol)) {
      ok = false;
      atomicOr(P.refuse_bits + w, kRefuseResidual);
    }
    if (!ok) valid[w] = 0;
  } else if (t < nc + ne) {
    int e = e0 + (t - nc);
    int w = P.edge_world[e];
    const double* y = current(A, B, sel[w]);
    double* o = next(A, B, sel[w]);
    double d = dphi[P.snd[e]] - dphi[P.rcv[e]];
    const double d1 = dphi1 ? dphi1[P.snd[e]] - dphi1[P.rcv[e]] : 0.0;
    for (int f = 0; f < P.f_jtr + P.S; ++f) {  // the transfers (k_edges_ahead finished the rest)
      double k1 = EFc(P, K1, f, e);
      if (dphi1 && f >= P.f_jtr && f < P.f_jtr + P.S) k1 += coef * (gj[(long)(f - P.f_jtr) * P.nedge + e] * d1);
      double k2 = EFc(P, K2, f, e) + -2.0 * k1;
      if (f >= P.f_jtr && f < P.f_jtr + P.S) k2 += coef * (gj[(long)(f - P.f_jtr) * P.nedge + e] * d);
      EF(P, o, f, e) = (EFc(P, y, f, e) + h1 * k1) + h2 * k2;
    }
    // the transfers' domain (k_edges_ahead checked the junction heat and the gate)
    bool eok = true;
    for (int f = P.f_jtr; f < P.f_jtr + P.S; ++f) eok = eok && isfinite(EFc(P, o, f, e));
    if (!eok) valid[w] = 0;
  }
}

// ------------------------------------------------------------------------------------------
// Acceptance bookkeeping.

__global__ void k_init(int W, const int* active, int* sel, int* alive, int* valid) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  sel[w] = 0;
  alive[w] = active ? (active[w] != 0) : 1;
  valid[w] = 1;
}

__global__ void k_commit(int w0, int nw, int* sel, int* alive, int* valid) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nw) return;
  int w = w0 + t;
  if (alive[w]) {
    if (valid[w]) sel[w] ^= 1;
    else alive[w] = 0;
  }
  valid[w] = 1;
}

// out[field, item] = the selected buffer's value; threads over cells, edges, worlds. Items whose
// selected buffer is out itself are left as they are.
__global__ void k_gather(Params P, const double* A, const double* B, const int* sel, double* out,
                         const int* alive, int* ok_out) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < P.ncell) {
    int i = (int)t;
    const double* y = current(A, B, sel[P.cell_world[i]]);
    if (y != out)
      for (int f = 0; f < P.cf; ++f) CF(P, out, f, i) = CFc(P, y, f, i);
  } else if (t < (long)P.ncell + P.nedge) {
    int e = (int)(t - P.ncell);
    const double* y = current(A, B, sel[P.edge_world[e]]);
    if (y != out)
      for (int f = 0; f < P.ef; ++f) EF(P, out, f, e) = EFc(P, y, f, e);
  } else if (t < (long)P.ncell + P.nedge + P.W) {
    int w = (int)(t - P.ncell - P.nedge);
    const double* y = current(A, B, sel[w]);
    if (y != out)
      for (int f = 0; f < P.wf; ++f) WF(P, out, f, w) = WFc(P, y, f, w);
    if (ok_out) ok_out[w] = alive[w];
  }
}

// World supply rates for cells_rates.
__global__ __launch_bounds__(kBlock) void k_world_rates(Params P, const double* contrib, double* R) {
  __shared__ double sh[kBlock];
  int w = blockIdx.x;
  int c0 = P.cell_off[w], c1 = P.cell_off[w + 1];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double a = 0.0;
    for (int i = c0 + threadIdx.x; i < c1; i += kBlock) a += contrib[(long)s * P.ncell + i];
    a = block_sum(a, sh);
    if (threadIdx.x == 0) WF(P, R, P.f_sup + s, w) = a;
  }
  if (threadIdx.x == 0)
    for (int f = P.S; f < P.wf; ++f) WF(P, R, f, w) = 0.0;
}

// ------------------------------------------------------------------------------------------
// Device row checks (models/surface_pump_checks.py:112-192 row_checks; balances
// surface_pump_checks.py:34-109 tissue_balances with surface_pump_v2.py:686-709's pump terms and
// the water operator's, operators/osmotic_water.py:254-295; bounds surface_pump.py:752-800
// ledger_limits and osmotic_water.py:341-344): per world, the reference's ledgers for cells on one
// held bath. Bits of failed[w] (kCheck*) and ratios[w * kLedgerRatios + k] = the largest
// |residual| / bound of ledger k (kRatio*).
constexpr double kEps = 2.220446049250313e-16;
constexpr double kLedgerStepRounding = 2.5;   // surface_pump.py LEDGER_STEP_ROUNDING
constexpr double kComponentRounding = 2.5;    // surface_field.py COMPONENT_STEP_ROUNDING
constexpr double kDissipationAtol = 1e-12;
constexpr double kGlobalChargeAtol = 3e-25;   // surface_pump.py GLOBAL_CHARGE_ATOL_C
constexpr double kGlobalChargeRounding = 256.0 * kEps;
constexpr double kEnergyAtol = 2e-8;          // surface_tissue.py ENERGY_ATOL_PJ
constexpr double kWorkAtol = 3e-12;           // the work ledgers' floor (surface_pump.py:789-791)
constexpr double kVolumeAtol = 2e-12;         // osmotic_water.py VOLUME_ATOL_PL
constexpr double kBathWorkAtol = 3e-12;       // osmotic_water.py WORK_ATOL_PJ

constexpr int kCheckDomain = CELLS_CHECK_DOMAIN, kCheckLocal = CELLS_CHECK_LOCAL,
              kCheckTotal = CELLS_CHECK_TOTAL, kCheckGlobalCharge = CELLS_CHECK_GLOBAL_CHARGE,
              kCheckDissipation = CELLS_CHECK_DISSIPATION, kCheckWater = CELLS_CHECK_WATER,
              kCheckOutside = CELLS_CHECK_OUTSIDE, kCheckCapacitor = CELLS_CHECK_CAPACITOR,
              kCheckElectrode = CELLS_CHECK_ELECTRODE, kCheckEnergy = CELLS_CHECK_ENERGY,
              kCheckReservoirWork = CELLS_CHECK_RESERVOIR_WORK,
              kCheckElectrodeWork = CELLS_CHECK_ELECTRODE_WORK,
              kCheckMetabolic = CELLS_CHECK_METABOLIC_WORK,
              kCheckBathWaterWork = CELLS_CHECK_BATH_WATER_WORK;
constexpr int kRatioLocal = CELLS_RATIO_LOCAL, kRatioTotal = CELLS_RATIO_TOTAL,
              kRatioGlobalCharge = CELLS_RATIO_GLOBAL_CHARGE, kRatioWater = CELLS_RATIO_WATER,
              kRatioOutside = CELLS_RATIO_OUTSIDE, kRatioCapacitor = CELLS_RATIO_CAPACITOR,
              kRatioElectrode = CELLS_RATIO_ELECTRODE, kRatioEnergy = CELLS_RATIO_ENERGY,
              kRatioReservoirWork = CELLS_RATIO_RESERVOIR_WORK,
              kRatioElectrodeWork = CELLS_RATIO_ELECTRODE_WORK,
              kRatioMetabolic = CELLS_RATIO_METABOLIC_WORK,
              kRatioBathWaterWork = CELLS_RATIO_BATH_WATER_WORK, kLedgerRatios = CELLS_LEDGER_RATIOS;
static_assert(kRefuseResidual > kCheckBathWaterWork, "the step's refusal bits lie above the checks'");

__device__ __forceinline__ double ratio_of(double residual, double bound) {
  double m = fabs(residual);
  if (m == 0.0) return 0.0;
  double r = m / bound;
  return isnan(r) ? INFINITY : r;
}

// K sums (or maxima) over the workgroup at once: each wave folds by a butterfly (every lane ends
// with the same bits), then every thread adds the waves' partials in wave order.
template <int K, bool MAX>
__device__ void block_fold(double (&v)[K], double* sh) {
  constexpr int kWaves = kBlock / 64;
  static_assert(K * kWaves <= kBlock, "the fold's partials fit the shared row");
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int o = 32; o > 0; o >>= 1) {
      const double t = __shfl_xor(v[k], o, 64);
      v[k] = MAX ? fmax(v[k], t) : v[k] + t;
    }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k * kWaves + wave] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = sh[k * kWaves];
    for (int j = 1; j < kWaves; ++j) t = MAX ? fmax(t, sh[k * kWaves + j]) : t + sh[k * kWaves + j];
    v[k] = t;
  }
  __syncthreads();
}

// World w's ledger checks of the state y against its predecessor prev (whole workgroup); the
// failure bits and the kLedgerRatios bound ratios on thread 0.
__device__ int ledger_checks(const Params& P, const double* prev, const double* y, int w, double steps,
                           double material_atol, double* sh, double* rat) {
  int c0 = P.cell_off[w], c1 = P.cell_off[w + 1], e0 = P.edge_off[w], e1 = P.edge_off[w + 1];
  const double rounding = kLedgerStepRounding * kEps * steps;
  bool domain = true, dissip = false;
  // maxima: local material, water local, capacitor charge, metabolic work
  double mx[4] = {0.0, 0.0, 0.0, 0.0};
  // sums: per species the amount increments, their magnitudes, the membrane totals (transfers
  // and pump stoichiometry, to the bath) and their magnitudes; then the charge increment, its
  // magnitude, the capacitors' bath-side charge, the energy terms, the energy flows' magnitudes,
  // the volume increments, their magnitudes, the water transfers and their magnitudes
  constexpr int kSums = 25;
  double sm[kSums];
#pragma unroll
  for (int k = 0; k < kSums; ++k) sm[k] = 0.0;
  const double rt3 = 1e-3 * P.RT;
  for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
    domain = domain && cell_domain(P, y, y, i);
    const double cyc = CFc(P, y, P.f_pcyc, i);
    const double ccyc = P.calcium ? CFc(P, y, P.f_ccyc, i) : 0.0;
    int p0 = P.inc_ptr[i], p1 = P.inc_ptr[i + 1], pm = p0 + P.inc_out[i];
    // the junction net per species, edges outer (each index loaded once), each species summed in
    // incidence order
    double nets[4] = {0.0, 0.0, 0.0, 0.0}, netms[4] = {0.0, 0.0, 0.0, 0.0};
    for (int pb = p0; pb < p1; pb += 2) {  // two edges' loads issued first; the sums keep their order
      const int e0 = P.inc_edge[pb], e1 = pb + 1 < p1 ? P.inc_edge[pb + 1] : -1;
      double j0[4], j1[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        j0[s] = s < P.S ? EFc(P, y, P.f_jtr + s, e0) : 0.0;
        j1[s] = (s < P.S && e1 >= 0) ? EFc(P, y, P.f_jtr + s, e1) : 0.0;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (s >= P.S) break;
        if (pb < pm) nets[s] -= j0[s];
        else nets[s] += j0[s];
        netms[s] += fabs(j0[s]);
      }
      if (e1 < 0) break;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (s >= P.S) break;
        if (pb + 1 < pm) nets[s] -= j1[s];
        else nets[s] += j1[s];
        netms[s] += fabs(j1[s]);
      }
    }
    double zq = 0.0, zm = 0.0, chem = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      const double dn = CFc(P, y, P.f_dn + s, i), mtr = CFc(P, y, P.f_mtr + s, i);
      double pump = s == 0 ? 3.0 * cyc : s == 1 ? -2.0 * cyc : 0.0;
      if (s == 3 && P.calcium) pump += ccyc;
      const double hm = mtr + pump;  // the membrane total toward the bath
      const double hmm = fabs(mtr) + fabs(pump);
      const double res = dn + hm - nets[s];
      mx[0] = fmax(mx[0], ratio_of(res, material_atol + rounding * (fabs(dn) + hmm + netms[s])));
      sm[s] += dn;
      sm[4 + s] += fabs(dn);
      sm[8 + s] += hm;
      sm[12 + s] += hmm;
      zq += P.z[s] * dn;
      zm += fabs(P.z[s] * dn);
      // the free energy of mixing against the reference amounts at the reference volume
      const double a = P.n0[(long)s * P.ncell + i], vr = P.vol0[i];
      chem += dn * (log(a / vr) - 1.0) + (a + dn) * log1p(dn / a);
    }
    const double q = kChargePerFmol * zq, qm = kChargePerFmol * zm;
    sm[16] += q;
    sm[17] += qm;
    // the diagonal capacitor on the held bath: dp = q / C, its plate charge C dp (bath side -C dp)
    const double C = P.cap[i], q0 = P.q0[i];
    const double dp = q / C, plate = C * dp;
    const double base = P.Vb + q0 / C;
    const double image = C * (base - P.Vb);
    mx[2] = fmax(mx[2], ratio_of(image - q0 + plate - q, kChargeAtol + kChargeRtol * fabs(q0 + q)));
    sm[18] += plate;
    // energy: mixing, capacitor, heat less work (membrane, pumps; junctions below)
    const double v0 = q0 / C;
    const double mdiss = CFc(P, y, P.f_mdiss, i), pdiss = CFc(P, y, P.f_pdiss, i);
    const double pwork = CFc(P, y, P.f_pwork, i);
    double en = rt3 * chem + 0.5e12 * (C * dp * (2.0 * v0 + dp)) + mdiss + pdiss - pwork;
    double fl = fabs(mdiss) + fabs(pdiss) + fabs(pwork);
    {
      const double drive = 1e-3 * P.nak_drive * cyc;
      mx[3] = fmax(mx[3], ratio_of(pwork + drive, kWorkAtol + rounding * (fabs(pwork) + fabs(drive))));
    }
    if (P.calcium) {
      const double cdiss = CFc(P, y, P.f_cdiss, i), cwork = CFc(P, y, P.f_cwork, i);
      en += cdiss - cwork;
      fl += fabs(cdiss) + fabs(cwork);
      const double drive = 1e-3 * P.ca_drive * ccyc;
      mx[3] = fmax(mx[3], ratio_of(cwork + drive, kWorkAtol + rounding * (fabs(cwork) + fabs(drive))));
    }
    const int df[4] = {P.f_mdiss, P.f_pdiss, P.calcium ? P.f_cdiss : -1, P.water ? P.f_wdiss : -1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (df[k] < 0) continue;
      double v = CFc(P, y, df[k], i), u = CFc(P, prev, df[k], i);
      dissip = dissip || v < -kDissipationAtol || v - u < -kDissipationAtol;
    }
    if (P.water) {
      const double dv = CFc(P, y, P.f_dvol, i), tr = CFc(P, y, P.f_wtr, i);
      const double wdiss = CFc(P, y, P.f_wdiss, i);
      mx[1] = fmax(mx[1], ratio_of(dv + tr, kVolumeAtol + rounding * (fabs(dv) + fabs(tr))));
      sm[21] += dv;
      sm[22] += fabs(dv);
      sm[23] += tr;
      sm[24] += fabs(tr);
      // dilution of every osmolyte to the current volume and the wall's elastic energy
      double total = 0.0;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < P.S) total += P.n0[(long)s * P.ncell + i] + CFc(P, y, P.f_dn + s, i);
      total = total + P.imperm[i];
      const double vr = P.vol0[i], stiff = P.wK / vr;
      const double pressure = stiff * (vr - vr) + P.wP0;
      en += -rt3 * (total * log1p(dv / vr)) + 1e-3 * (pressure * dv + 0.5 * stiff * dv * dv) + wdiss;
      fl += fabs(wdiss);
    }
    sm[19] += en;
    sm[20] += fl;
  }
  for (int e = e0 + threadIdx.x; e < e1; e += kBlock) {
    domain = domain && edge_domain(P, y, e);
    double v = EFc(P, y, P.f_jdiss, e), u = EFc(P, prev, P.f_jdiss, e);
    dissip = dissip || v < -kDissipationAtol || v - u < -kDissipationAtol;
    sm[19] += v;
    sm[20] += fabs(v);
  }
  {
    double f[6] = {mx[0], mx[1], mx[2], mx[3], domain ? 0.0 : 1.0, dissip ? 1.0 : 0.0};
    block_fold<6, true>(f, sh);
    for (int k = 0; k < 4; ++k) mx[k] = f[k];
    domain = f[4] == 0.0;
    dissip = f[5] > 0.0;
  }
  block_fold<kSums, false>(sm, sh);
  if (threadIdx.x != 0) return 0;
  double rtot = 0.0, rout = 0.0, supq = 0.0, supm = 0.0, rwexp = 0.0, rwmag = 0.0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    const double sup = WFc(P, y, P.f_sup + s, w);
    rtot = fmax(rtot, ratio_of(sm[s] - sup, material_atol + rounding * (sm[4 + s] + fabs(sup))));
    // the held bath's amount stays put: what the membranes moved less what the bath supplied
    rout = fmax(rout, ratio_of(-sm[8 + s] - sup, material_atol + rounding * (sm[12 + s] + fabs(sup))));
    supq += P.z[s] * sup;
    supm += fabs(P.z[s] * sup);
    rwexp += P.mu_b[s] * sup;
    rwmag += fabs(P.mu_b[s]) * fabs(sup);
  }
  const double echg = WFc(P, y, P.f_echg, w), rwork = WFc(P, y, P.f_rwork, w);
  const double ework = WFc(P, y, P.f_ework, w);
  // global charge: the electrode's charge and the bath supplies' charge
  const double moved = kChargePerFmol * supm + fabs(echg);
  const double rq = ratio_of(echg + kChargePerFmol * supq,
                             kGlobalChargeAtol + kGlobalChargeRounding * moved +
                                 fmax(rounding * moved, kComponentRounding * kEps * steps * sm[17]));
  // the electrode's charge against the bath-side plate charge of the world's capacitors
  const double relec = ratio_of(echg + sm[18], kChargeAtol + kChargeRtol * fabs(echg));
  double energy = sm[19] - rwork - ework, flows = sm[20] + fabs(rwork) + fabs(ework);
  const double rres = ratio_of(rwork - 1e-3 * rwexp, kWorkAtol + rounding * (fabs(rwork) + 1e-3 * rwmag));
  const double eterm = 1e12 * P.Vb * echg;
  const double rew = ratio_of(ework - eterm, kWorkAtol + rounding * (fabs(ework) + fabs(eterm)));
  double rwat = mx[1], rbw = 0.0;
  if (P.water) {
    const double ws = WFc(P, y, P.f_wsup, w), ww = WFc(P, y, P.f_wwork, w);
    rwat = fmax(rwat, ratio_of(sm[21] - ws, kVolumeAtol + rounding * (sm[22] + fabs(ws))));
    rwat = fmax(rwat, ratio_of(ws + sm[23], kVolumeAtol + rounding * (sm[24] + fabs(ws))));
    const double bw = 1e-3 * P.psib * ws;
    rbw = ratio_of(ww - bw, kBathWorkAtol + rounding * (fabs(ww) + fabs(bw)));
    energy -= ww;
    flows += fabs(ww);
  }
  const double ren = ratio_of(energy, kEnergyAtol + P.energy_rtol * flows);
  int bits = 0;
  if (!domain) bits |= kCheckDomain;
  if (!(mx[0] <= 1.0)) bits |= kCheckLocal;
  if (!(rtot <= 1.0)) bits |= kCheckTotal;
  if (!(rq <= 1.0)) bits |= kCheckGlobalCharge;
  if (dissip) bits |= kCheckDissipation;
  if (!(rwat <= 1.0)) bits |= kCheckWater;
  if (!(rout <= 1.0)) bits |= kCheckOutside;
  if (!(mx[2] <= 1.0)) bits |= kCheckCapacitor;
  if (!(relec <= 1.0)) bits |= kCheckElectrode;
  if (!(ren <= 1.0)) bits |= kCheckEnergy;
  if (!(rres <= 1.0)) bits |= kCheckReservoirWork;
  if (!(rew <= 1.0)) bits |= kCheckElectrodeWork;
  if (!(mx[3] <= 1.0)) bits |= kCheckMetabolic;
  if (!(rbw <= 1.0)) bits |= kCheckBathWaterWork;
  rat[kRatioLocal] = mx[0];
  rat[kRatioTotal] = rtot;
  rat[kRatioGlobalCharge] = rq;
  rat[kRatioWater] = rwat;
  rat[kRatioOutside] = rout;
  rat[kRatioCapacitor] = mx[2];
  rat[kRatioElectrode] = relec;
  rat[kRatioEnergy] = ren;
  rat[kRatioReservoirWork] = rres;
  rat[kRatioElectrodeWork] = rew;
  rat[kRatioMetabolic] = mx[3];
  rat[kRatioBathWaterWork] = rbw;
  return bits;
}

// The readouts' energy rates of each world at its state (models/surface_pump_v3.py:2221-2262
// readouts, for cells on one held bath without water): from the rates R at the state's potentials,
// out[kEnergyRates w + k] = the chemical rate 1e-3 RT sum ln(c) dN/dt, the capacitor rate
// 1e12 sum C V dV/dt (dV = the capacitor solve of the charge rate), the dissipation (membrane,
// pump and junction heat), the reservoir power 1e-3 sum mu_s supply_s, the electrode power
// 1e12 V_bath I, the metabolic work rate, the energy balance chemical + capacitor + dissipation -
// reservoir - electrode - metabolic (zero in exact arithmetic), and the electrode current I
// (minus the cells' plate-charge rate).
constexpr int kEnergyRates = 8;

__global__ __launch_bounds__(kBlock) void k_energy_rates(Params P, const double* y, const double* R, double* out) {
  __shared__ double sh[kBlock];
  const int w = blockIdx.x;
  const int c0 = P.cell_off[w], c1 = P.cell_off[w + 1], e0 = P.edge_off[w], e1 = P.edge_off[w + 1];
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // chemical, capacitor, dissipation, metabolic, plate rate
  for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
    double ci[4], vol;
    cell_concentrations(P, y, y, i, ci, vol);
    double chem = 0.0, zq = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      const double r = CFc(P, R, P.f_dn + s, i);
      chem = s == 0 ? log(ci[s]) * r : chem + log(ci[s]) * r;
      zq = s == 0 ? P.z[s] * r : zq + P.z[s] * r;
    }
    const double qdot = kChargePerFmol * zq, C = P.cap[i];
    const double dvm = qdot / C;
    v[0] += chem;
    v[1] += C * (potential(P, y, i) - P.Vb) * dvm;
    double heat = CFc(P, R, P.f_mdiss, i) + CFc(P, R, P.f_pdiss, i);
    double work = CFc(P, R, P.f_pwork, i);
    if (P.calcium) {
      heat += CFc(P, R, P.f_cdiss, i);
      work += CFc(P, R, P.f_cwork, i);
    }
    v[2] += heat;
    v[3] += work;
    v[4] += C * dvm;
  }
  for (int e = e0 + threadIdx.x; e < e1; e += kBlock) v[2] += EFc(P, R, P.f_jdiss, e);
  block_fold<5, false>(v, sh);
  if (threadIdx.x != 0) return;
  double reservoir = 0.0;
#pragma unroll
  for (int s = 0; s < 4; ++s)
    if (s < P.S) reservoir += P.mu_b[s] * WFc(P, R, P.f_sup + s, w);
  reservoir *= 1e-3;
  const double current = -v[4];
  const double chemical = 1e-3 * P.RT * v[0], capacitor = 1e12 * v[1], electrode = 1e12 * P.Vb * current;
  double* o = out + (long)kEnergyRates * w;
  o[0] = chemical;
  o[1] = capacitor;
  o[2] = v[2];
  o[3] = reservoir;
  o[4] = electrode;
  o[5] = v[3];
  o[6] = chemical + capacitor + v[2] - (reservoir + electrode) - v[3];
  o[7] = current;
}

__global__ __launch_bounds__(kBlock) void k_device_checks(Params P, const double* prev, const double* y,
                                                          double steps, double material_atol,
                                                          int* failed, double* ratios) {
  __shared__ double sh[kBlock];
  const int w = blockIdx.x;
  double rat[kLedgerRatios];
  const int bits = ledger_checks(P, prev, y, w, steps, material_atol, sh, rat);
  if (threadIdx.x == 0) {
    failed[w] = bits;
    for (int k = 0; k < kLedgerRatios; ++k) ratios[(long)kLedgerRatios * w + k] = rat[k];
  }
}

// The checks as the step's epilogue: each live world's candidate state (the buffer the commit
// would select) against its current one; a failing world refuses the step (the commit keeps its
// state and stops it) and its bits are or-ed into bits_out. steps = the ledgers' step count
// after this step (the device counter).
__global__ __launch_bounds__(kBlock) void k_step_checks(Params P, const double* A, const double* B, const int* sel,
                                                        const int* alive, int* valid, const long long* counter,
                                                        double material_atol, int w0, int* bits_out) {
  __shared__ double sh[kBlock];
  const int w = w0 + blockIdx.x;
  if (!alive[w] || !valid[w]) return;
  const int sw = sel[w];
  const double* cur = current(A, B, sw);
  const double* cand = next(const_cast<double*>(A), const_cast<double*>(B), sw);
  double rat[kLedgerRatios];
  const int bits = ledger_checks(P, cur, cand, w, (double)counter[0], material_atol, sh, rat);
  if (threadIdx.x == 0 && bits) {
    valid[w] = 0;
    bits_out[w] |= bits;
  }
}

__global__ void k_tick(long long* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1;
}

// ------------------------------------------------------------------------------------------
// Adjoint of the ROS2 step (3-ion worlds, no water): reverse of the forward arithmetic restricted
// to the fields the dynamics read (amounts, gates). The reference differentiates only forward
// (analysis/stationary.py:505-517); this is the step's exact discrete adjoint, pinned by central
// differences of the forward step.

__device__ __forceinline__ double bernoulli_second(double x) {
  if (fabs(x) < 0.1) {
    double x2 = x * x;
    return 1.0 / 6.0 - x2 / 60.0 + x2 * x2 / 1008.0 - x2 * x2 * x2 / 21600.0 +
           x2 * x2 * x2 * x2 / 532224.0;
  }
  if (x > 0) {
    double em = exp(-x), d = 1.0 - em;
    return em * (x * (1.0 + em) - 2.0 * d) / (d * d * d);
  }
  double ex = exp(x), d = ex - 1.0;
  return ex * (x * (ex + 1.0) - 2.0 * d) / (d * d * d);
}

// The Bernoulli values the reverse route kernels take at x = (z F / RT) dv and -x: B, B' and (with
// SECOND) B''. As RouteBasis, x is odd in z operation by operation, so species of equal charge share
// them and species of opposite charge take them swapped.
struct VjpBasis {
  double Bm, Bp, Sm, Sp, Qm, Qp;
};

__device__ __forceinline__ VjpBasis vjp_basis(double x, bool second) {
  VjpBasis b;
  b.Bm = bernoulli(-x);
  b.Bp = bernoulli(x);
  b.Sm = bernoulli_slope(-x);
  b.Sp = bernoulli_slope(x);
  b.Qm = second ? bernoulli_second(-x) : 0.0;
  b.Qp = second ? bernoulli_second(x) : 0.0;
  return b;
}

__device__ __forceinline__ VjpBasis vjp_species_basis(const double* z, int s, const VjpBasis& b0, double x,
                                                      bool second) {
  if (s == 0 || z[s] == z[0]) return b0;
  if (z[s] == -z[0]) return VjpBasis{b0.Bp, b0.Bm, b0.Sp, b0.Sm, b0.Qp, b0.Qm};
  return vjp_basis(x, second);
}

// Per edge: pull the amount-rate cotangent fN [S][n] (J leaves the sender, enters the receiver),
// the gate-rate cotangent fg [E] and the slope cotangent gjbar [S][E] back to the edge's two
// concentrations (ca_bar, cb_bar [S][E]), its voltage (ev_bar [E]), its gate (added to gate_out
// [E] and, times h2, to gate_out2 [E]) and its parameters (paj_bar, added).
__global__ __launch_bounds__(kBlock) void k_edges_vjp(Params P, const double* y, Ctl c,
                                                      const double* fN, const double* fg,
                                                      const double* gjbar, double* ca_bar,
                                                      double* cb_bar, double* ev_bar,
                                                      double* gate_out, double* gate_out2, double h2,
                                                      double* paj_bar) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P.nedge) return;
  int a = P.snd[e], b = P.rcv[e];
  double ca[4], cb[4], va, vb;
  cell_concentrations(P, y, y, a, ca, va);
  cell_concentrations(P, y, y, b, cb, vb);
  double dv = potential(P, y, a) - potential(P, y, b);
  double g = EFc(P, y, P.f_gate, e);
  double unit = c.junction / kMolPerFmol;
  double vsum = 0.0, gbar = 0.0;
  const VjpBasis b0 = vjp_basis(P.z[0] * kF / P.RT * dv, gjbar != nullptr);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double z = P.z[s], kk = z * kF / P.RT, x = kk * dv;
    double pa = P.paj[(long)s * P.nedge + e];
    const VjpBasis bs = vjp_species_basis(P.z, s, b0, x, gjbar != nullptr);
    double Bm = bs.Bm, Bp = bs.Bp;
    double Sm = bs.Sm, Sp = bs.Sp;
    double A = pa * g * unit;
    double core = ca[s] * Bm - cb[s] * Bp;
    double score = -ca[s] * Sm - cb[s] * Sp;
    double gjv = A * kk * score;
    double Jbar = fN ? fN[(long)s * P.ncell + b] - fN[(long)s * P.ncell + a] : 0.0;
    double Gbar = gjbar ? gjbar[(long)s * P.nedge + e] : 0.0;
    double cab = Jbar * A * Bm, cbb = Jbar * (-A * Bp), dvb = Jbar * gjv;
    if (gjbar) {
      double Qm = bs.Qm, Qp = bs.Qp;
      cab += Gbar * (-A * kk * Sm);
      cbb += Gbar * (-A * kk * Sp);
      dvb += Gbar * (A * kk * kk * (ca[s] * Qm - cb[s] * Qp));
    }
    ca_bar[(long)s * P.nedge + e] = cab;
    cb_bar[(long)s * P.nedge + e] = cbb;
    vsum += dvb;
    gbar += Jbar * (pa * unit * core) + Gbar * (pa * unit * kk * score);
    paj_bar[(long)s * P.nedge + e] += Jbar * (g * unit * core) + Gbar * (g * unit * kk * score);
  }
  if (fg) {
    double shift = fabs(dv) - P.gate_thr;
    double alpha = 1.3 * exp(-77.0 * shift);
    double log_beta = P.log_0065 + 140.0 * shift;
    double beta = 20.0 * exp(-logaddexp0(-log_beta));
    double fb = fg[e];
    gbar += fb * -(alpha + beta);
    double sgn = dv > 0 ? 1.0 : (dv < 0 ? -1.0 : 0.0);
    double dal = -77.0 * alpha, dbe = 140.0 * beta * (1.0 - beta / 20.0);
    vsum += fb * (dal * (1.0 - g) - dbe * (g - P.gate_min)) * sgn;
  }
  ev_bar[e] = vsum;
  gate_out[e] += gbar;
  if (gate_out2) gate_out2[e] += h2 * gbar;
}

// Per cell: the membrane routes' and the Na/K-ATPase's pull-back plus the edges' partials, then
// to the amounts (c = N / V, phi = V_bath + q / C): adds to out (and h2 times to out2).
__global__ __launch_bounds__(kBlock) void k_cells_vjp(Params P, const double* y, Ctl c,
                                                      const double* fN, const double* gmbar,
                                                      const double* ca_bar, const double* cb_bar,
                                                      const double* ev_bar, double* out,
                                                      double* out2, double h2, double* pam_bar,
                                                      double* cap_bar) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.ncell) return;
  double ci[4], vol;
  cell_concentrations(P, y, y, i, ci, vol);
  double dv = potential(P, y, i) - P.Vb;
  double unit = c.membrane / kMolPerFmol;
  double cbar[4] = {0, 0, 0, 0};
  double vbar = 0.0;
  const VjpBasis b0 = vjp_basis(P.z[0] * kF / P.RT * dv, gmbar != nullptr);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double z = P.z[s], kk = z * kF / P.RT, x = kk * dv, co = P.cb[s];
    double pa = P.pam[(long)s * P.ncell + i];
    double A = pa * unit;
    const VjpBasis bs = vjp_species_basis(P.z, s, b0, x, gmbar != nullptr);
    double Bm = bs.Bm, Bp = bs.Bp;
    double Sm = bs.Sm, Sp = bs.Sp;
    double score = -ci[s] * Sm - co * Sp;
    double gmv = A * kk * score;
    double Jbar = fN ? -fN[(long)s * P.ncell + i] : 0.0;
    double Gbar = gmbar ? gmbar[(long)s * P.ncell + i] : 0.0;
    cbar[s] += Jbar * A * Bm;
    vbar += Jbar * gmv;
    if (gmbar) {
      double Qm = bs.Qm, Qp = bs.Qp;
      cbar[s] += Gbar * (-A * kk * Sm);
      vbar += Gbar * (A * kk * kk * (ci[s] * Qm - co * Qp));
    }
    pam_bar[(long)s * P.ncell + i] += Jbar * (unit * (ci[s] * Bm - co * Bp)) + Gbar * (unit * kk * score);
  }
  if (fN) {
    double cyc_bar = -3.0 * fN[i] + 2.0 * fN[P.ncell + i];
    double cap = P.nak_cap[i], av = c.pump;
    if (cap > 0 && av > 0 && cyc_bar != 0.0) {
      double mu_na = P.RT * (log(ci[0]) - P.log_cb[0]) + P.z[0] * kF * dv;
      double mu_k = P.RT * (log(ci[1]) - P.log_cb[1]) + P.z[1] * kF * dv;
      double a = (P.nak_drive + (-3.0) * mu_na + 2.0 * mu_k) / P.RT;
      double r = exp(3.0 * (P.log_km_na - log(ci[0])));
      double s1 = 1.0 / (1.0 + r);
      double s23 = exp(-logaddexp0(2.0 * (P.log_km_k - P.log_cb[1])) - P.nak_atp_sat);
      double sat = s1 * s23;
      double ea = exp(a);
      double abar = cyc_bar * (-cap * av * sat * ea);
      double mu_na_bar = abar * (-3.0 / P.RT), mu_k_bar = abar * (2.0 / P.RT);
      double ds1 = (3.0 * r / ci[0]) * s1 * s1;
      cbar[0] += cyc_bar * cap * av * ds1 * s23 * (1.0 - ea) + mu_na_bar * P.RT / ci[0];
      cbar[1] += mu_k_bar * P.RT / ci[1];
      vbar += (mu_na_bar * P.z[0] + mu_k_bar * P.z[1]) * kF;
      cap_bar[i] += cyc_bar * av * sat * (1.0 - ea);
    }
  }
  int p0 = P.inc_ptr[i], p1 = P.inc_ptr[i + 1], pm = p0 + P.inc_out[i];
  for (int p = p0; p < pm; ++p) {
    int e = P.inc_edge[p];
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < P.S) cbar[s] += ca_bar[(long)s * P.nedge + e];
    vbar += ev_bar[e];
  }
  for (int p = pm; p < p1; ++p) {
    int e = P.inc_edge[p];
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < P.S) cbar[s] += cb_bar[(long)s * P.nedge + e];
    vbar -= ev_bar[e];
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= P.S) break;
    double nb = cbar[s] / vol + vbar * kChargePerFmol * P.z[s] / P.cap[i];
    CF(P, out, P.f_dn + s, i) += nb;
    if (out2) out2[(long)s * P.ncell + i] += h2 * nb;
  }
}

// Seeds the reverse sweep from the output cotangent: ybar_in = ybar_out on amounts and gates;
// k1bar = 1.5 dt ybar_out, k2bar = 0.5 dt ybar_out (amounts [S][n] and gates [E]).
__global__ __launch_bounds__(kBlock) void k_adj_seed(Params P, const double* ybar, double dt,
                                                     double* ybar_in, double* k1bar, double* k2bar,
                                                     double* k1gbar, double* k2gbar) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < P.ncell) {
    int i = (int)t;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      double v = CFc(P, ybar, P.f_dn + s, i);
      CF(P, ybar_in, P.f_dn + s, i) = v;
      k1bar[(long)s * P.ncell + i] = 1.5 * dt * v;
      k2bar[(long)s * P.ncell + i] = 0.5 * dt * v;
    }
  } else if (t < (long)P.ncell + P.nedge) {
    int e = (int)(t - P.ncell);
    double v = EFc(P, ybar, P.f_gate, e);
    EF(P, ybar_in, P.f_gate, e) = v;
    k1gbar[e] = 1.5 * dt * v;
    k2gbar[e] = 0.5 * dt * v;
  }
}

// dbar = coef D^T u and the slopes' cotangent of D(g) delta against u (cells, then edges).
__global__ __launch_bounds__(kBlock) void k_adj_dtrans(Params P, double coef, const double* gm,
                                                       const double* gj, const double* u,
                                                       const double* delta, double* dbar,
                                                       double* gmbar, double* gjbar) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < P.ncell) {
    int i = (int)t;
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= P.S) break;
      double us = u[(long)s * P.ncell + i];
      acc += -gm[(long)s * P.ncell + i] * us;
      gmbar[(long)s * P.ncell + i] += coef * (us * -delta[i]);
    }
    int p0 = P.inc_ptr[i], p1 = P.inc_ptr[i + 1], pm = p0 + P.inc_out[i];
    for (int p = p0; p < p1; ++p) {
      int e = P.inc_edge[p];
      int a = P.snd[e], b = P.rcv[e];
      double tt = 0.0;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < P.S)
        tt += gj[(long)s * P.nedge + e] * (u[(long)s * P.ncell + b] - u[(long)s * P.ncell + a]);
      acc += p < pm ? tt : -tt;
    }
    dbar[i] = coef * acc;
  } else if (t < (long)P.ncell + P.nedge) {
    int e = (int)(t - P.ncell);
    int a = P.snd[e], b = P.rcv[e];
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s < P.S)
      gjbar[(long)s * P.nedge + e] +=
          coef * ((delta[a] - delta[b]) * (u[(long)s * P.ncell + b] - u[(long)s * P.ncell + a]));
  }
}

// w = M^-1 dbar on the step's factor (workgroup per world); then rate_bar += q^T w and the
// slopes' cotangent of M: -coef w delta^T through G = F' sum_s z_s slope_s.
template <bool SN>
__global__ __launch_bounds__(kMfThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_adj_solve(Params P, const ShapeDev* shapes,
                                                          const int* shape_of, const int* pool,
                                                          const long* foff, double coef,
                                                          const double* msc, const double* fvals,
                                                          const double* dbar, const double* delta,
                                                          double* wv, double* rate_bar, double* gmbar,
                                                          double* gjbar, double* ysc, int y_in_lds,
                                                          SolveArgs sa, int part) {
  extern __shared__ double shared_y[];
  int w = launch_world(P, 0);
  const ShapeDev s = shapes[shape_of[w]];
  const PlanView v = plan_view(s, pool);
  const double* fv = fvals + foff[w];
  int c0 = P.cell_off[w], e0 = P.edge_off[w];
  double* lds = y_in_lds ? shared_y : ysc + c0;
  if (part != 2)
    for (int k = threadIdx.x; k < s.n; k += blockDim.x) {
      int i = c0 + v.order[k];
      lds[k] = dbar[i] * msc[i];
    }
  if (part == 0) {
    __syncthreads();
    solve_any<SN>(s, v, pool, fv, lds, sa, w, shared_y);
    __syncthreads();
  }
  if (part == 1) return;
  for (int k = threadIdx.x; k < s.n; k += blockDim.x) {
    int i = c0 + v.order[k];
    wv[i] = lds[k] * msc[i];
  }
  __syncthreads();
  for (int i = c0 + threadIdx.x; i < c0 + s.n; i += blockDim.x)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q >= P.S) break;
      rate_bar[(long)q * P.ncell + i] += kChargePerFmol * P.z[q] * wv[i];
      gmbar[(long)q * P.ncell + i] += kChargePerFmol * P.z[q] * (-coef * wv[i] * delta[i]);
    }
  for (int e = e0 + threadIdx.x; e < e0 + s.E; e += blockDim.x) {
    int a = P.snd[e], b = P.rcv[e];
    double t = -coef * (wv[a] - wv[b]) * (delta[a] - delta[b]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < P.S) gjbar[(long)q * P.nedge + e] += kChargePerFmol * P.z[q] * t;
  }
}

// k1bar += -2 r2bar (amounts and gates).
__global__ __launch_bounds__(kBlock) void k_adj_axpy(long n1, double* k1bar, const double* r2bar,
                                                     long n2, double* k1gbar, const double* g2bar) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n1) k1bar[t] += -2.0 * r2bar[t];
  else if (t < n1 + n2) k1gbar[t - n1] += -2.0 * g2bar[t - n1];
}

__global__ __launch_bounds__(kBlock) void k_fill(long n, double* x, double v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

// The Vmem loss of a sample row: sum_i wgt (V_i - target_i)^2 per world into loss[w] (added),
// and its cotangent on the amounts added to ybar (dV_i/dN_si = F' z_s / C_i).
__global__ __launch_bounds__(kBlock) void k_vmem_loss(Params P, const double* y, const double* target,
                                                      double wgt, double* loss, double* ybar) {
  __shared__ double sh[kBlock];
  int w = blockIdx.x;
  int c0 = P.cell_off[w], c1 = P.cell_off[w + 1];
  double acc = 0.0;
  for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
    double v = potential(P, y, i) - P.Vb;
    double d = v - target[i];
    acc += wgt * d * d;
    if (ybar) {
      double g = 2.0 * wgt * d;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < P.S) CF(P, ybar, P.f_dn + s, i) += g * kChargePerFmol * P.z[s] / P.cap[i];
    }
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) loss[w] += acc;
}

int blocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

// ==========================================================================================
// Host side.

struct cells_symbolic {
  cells::Symbolic s;
};

#include "n2.inc"
#include "rk.inc"

struct cells_ensemble {
  int device = 0;
  int method = 0;
  Params P{};
  int S = 0, W = 0, ncell = 0, nedge = 0, max_n = 0;
  std::vector<int64_t> cell_off, edge_off;
  std::vector<int> shape_of_host;
  std::vector<cells::Symbolic> shapes_host;      // the plans' patterns (relaxed in the multifrontal mode)
  std::vector<int64_t> exact_nnz_host;           // per shape: nnz(L) of the bound pattern itself
  std::vector<double> exact_flops_host;          // per shape: its sum of column counts squared
  std::vector<long> foff_host;
  std::vector<ShapeDev> sd_host;
  std::vector<int> pool_host;
  long factor_total = 0;
  // device allocations
  std::vector<void*> allocations;
  ShapeDev* shapes = nullptr;
  int* shape_of = nullptr;
  int* pool = nullptr;
  unsigned long long* masks = nullptr;
  long* foff = nullptr;
  bool pairwise = false;  // factor in the reference's summation order (slot masks), else multifrontal
  long* uoffw = nullptr;  // per world: start of its update matrices in ubuf
  double* ubuf = nullptr;
  long* voffw = nullptr;  // per world: start of the solve's update vectors in vbuf
  double* vbuf = nullptr;
  long* boffw = nullptr;  // per world: start of its HBM front scratch in bigbuf
  double* bigbuf = nullptr;
  std::vector<cells::Supernodes> sn_host;
  size_t factor_lds = 0, solve_lds = 0;  // dynamic LDS bytes of the factor and solve kernels
  int y_in_lds = 1;
  double *A = nullptr, *B = nullptr, *Y1 = nullptr, *K1 = nullptr, *K2 = nullptr, *K3 = nullptr, *K4 = nullptr;
  double *contrib1 = nullptr, *contrib2 = nullptr, *gm = nullptr, *gj = nullptr, *Gj = nullptr;
  double *mdiag = nullptr, *msc = nullptr, *dphi = nullptr, *dphi2 = nullptr, *fvals = nullptr;
  double* ysc = nullptr;
  int *sel = nullptr, *alive = nullptr, *valid = nullptr;
  // the ledger checks as the step's epilogue (cells_step_checks)
  bool checks_on = false;
  double check_atol = 1e-12;
  long long check_steps = 0;  // the ledgers' step count of the next cells_step's input
  long long* step_counter = nullptr;
  int* check_bits = nullptr;  // per world: the ledger checks' bits and the step's refusal reasons
  // the stage solves' residual bound (numerics/surface_field.py:47-48 CHARGE_ATOL_C, CHARGE_RTOL)
  double res_atol = kChargeAtol, res_rtol = kChargeRtol;
  // rate bounds of the bound worlds at their declared state (1/s; cells_step_bounds): the explicit
  // rate, the junction gates' and the Ca-ATPase's
  double bound_explicit = 0.0, bound_gate = 0.0, bound_calcium = 0.0;
  // step records for the adjoint (cells_tape_reserve): per slot the step's input state, stage-1
  // state, factor values, Jacobi scales, route slopes and both stage increments, field-major
  int tape_slots = 0;
  std::vector<void*> tape_allocations;
  double *t_y = nullptr, *t_y1 = nullptr, *t_fv = nullptr, *t_msc = nullptr, *t_gm = nullptr;
  double *t_gj = nullptr, *t_d1 = nullptr, *t_d2 = nullptr;
  int* t_valid = nullptr;
  // the adjoint workspace (allocated at the first cells_step_vjp; laid out by adjoint_layout)
  double* adj_base = nullptr;
  // N2's stage buffers (CELLS_METHOD_N2)
  N2Buffers n2;
  hipStream_t capture_stream = nullptr;
  // what a captured step graph bakes in: one is replayed only while all of it is unchanged
  struct GraphKey {
    double dt = 0.0;
    Ctl ctl{};
    bool checks = false;
    double check_atol = 0.0, res_atol = 0.0, res_rtol = 0.0;
    const double* A = nullptr;  // the first state buffer (the caller's output buffer)
    bool operator==(const GraphKey& o) const {
      return dt == o.dt && A == o.A && ctl.membrane == o.ctl.membrane && ctl.junction == o.ctl.junction &&
             ctl.pump == o.ctl.pump && ctl.calcium_pump == o.ctl.calcium_pump && ctl.water == o.ctl.water &&
             checks == o.checks && check_atol == o.check_atol && res_atol == o.res_atol &&
             res_rtol == o.res_rtol;
    }
  };
  // the captured step graphs, most recently used first (a stepping loop whose output buffers
  // alternate between a few addresses replays one graph per address)
  static constexpr size_t kGraphs = 4;
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
  ~cells_ensemble() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    for (void* p : allocations) (void)hipFree(p);
    for (void* p : tape_allocations) (void)hipFree(p);
  }
};

// Drops the stationary analysis' workspace record of an ensemble (its buffers are the ensemble's
// allocations); defined with the stationary analysis.
void stationary_release(const cells_ensemble* e);

namespace {

template <typename T>
int upload(cells_ensemble* e, T** dst, const T* src, size_t count) {
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return fail(CELLS_ENOMEM, "hipMalloc failed at set-up");
  e->allocations.push_back(p);
  if (src && count) HIP_TRY(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
  *dst = static_cast<T*>(p);
  return CELLS_OK;
}

template <typename T>
int alloc(cells_ensemble* e, T** dst, size_t count) {
  return upload<T>(e, dst, nullptr, count);
}

std::vector<int> narrow(const int64_t* a, size_t n) {
  std::vector<int> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = static_cast<int>(a[i]);
  return out;
}

double host_logaddexp0(double t) {
  double m = t > 0.0 ? t : 0.0;
  return m + std::log1p(std::exp(-std::fabs(0.0 - t)));
}

// Which lanes hold a front of this order in the multifrontal factor: 0 one thread, 1 eight lanes,
// 2 one wavefront, 3 the workgroup in LDS, 4 the workgroup in HBM scratch, 5 none (the solve's
// wavefront holds at most kMfMaxRows rows).
int32_t front_class(int32_t rows) {
  int64_t tri = (int64_t)rows * (rows + 1) / 2;
  for (int c = 0; c < 6; ++c)
    if (tri <= kLatGroupTri[c]) return c;
  if (tri <= kMfBlockLds) return 6;
  if (rows <= kMfMaxRows) return 7;
  return 8;
}

// A level's fronts fill the workgroup's lanes: a lane-group front of a level with F fronts takes at
// least the widest group of which the workgroup holds F (so a level of few fronts runs them on
// wide groups, one front per group, and only crowded levels use the narrow ones).
int32_t widen_class(int32_t cls, int32_t level_fronts) {
  if (cls > 5 || level_fronts < 1 || !g_plan.widen_groups) return cls;
  static const int32_t lanes[6] = {1, 4, 8, 16, 32, 64};
  int32_t wide = 0;
  for (int c = 5; c >= 0; --c)
    if ((int64_t)level_fronts * lanes[c] <= kMfThreads) {
      wide = c;
      break;
    }
  return std::max(cls, wide);
}

int check_world(const cells_world* w) {
  if (!w) return fail(CELLS_EINVAL, "world is NULL");
  if (w->S < 3 || w->S > 4) return fail(CELLS_EINVAL, "species count must be 3 or 4");
  if (w->calcium && w->S != 4) return fail(CELLS_EINVAL, "the Ca-ATPase needs species 3 (Ca)");
  if (w->W < 1) return fail(CELLS_EINVAL, "need at least one world");
  if (w->ncell < 1 || w->ncell > (int64_t)1 << 30 || w->nedge < 0 || w->nedge > (int64_t)1 << 30)
    return fail(CELLS_EINVAL, "cell or edge count out of range");
  if (w->cell_off[0] != 0 || w->cell_off[w->W] != w->ncell || w->edge_off[0] != 0 ||
      w->edge_off[w->W] != w->nedge)
    return fail(CELLS_EINVAL, "world offsets disagree with the counts");
  for (int k = 0; k < w->W; ++k) {
    if (w->cell_off[k + 1] <= w->cell_off[k]) return fail(CELLS_EINVAL, "a world has no cells");
    if (w->edge_off[k + 1] < w->edge_off[k]) return fail(CELLS_EINVAL, "edge offsets decrease");
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
      int64_t a = w->snd[e], b = w->rcv[e];
      if (a < w->cell_off[k] || a >= w->cell_off[k + 1] || b < w->cell_off[k] ||
          b >= w->cell_off[k + 1] || a == b)
        return fail(CELLS_EINVAL, "an edge leaves its world or joins a cell to itself");
    }
  }
  {
    // The stage matrix takes one off-diagonal entry per pair of cells; the reference sums the
    // conductances of every junction between a pair into it (numerics/surface_field.py:413-418),
    // which the plans here do not: a repeated pair is refused rather than mis-assembled.
    std::vector<std::pair<int64_t, int64_t>> pairs;
    pairs.reserve((size_t)w->nedge);
    for (int64_t e = 0; e < w->nedge; ++e)
      pairs.emplace_back(std::min(w->snd[e], w->rcv[e]), std::max(w->snd[e], w->rcv[e]));
    std::sort(pairs.begin(), pairs.end());
    if (std::adjacent_find(pairs.begin(), pairs.end()) != pairs.end())
      return fail(CELLS_EINVAL, "two junctions join the same pair of cells (merge them into one)");
  }
  if (w->inc_ptr[0] != 0 || w->inc_ptr[w->ncell] != 2 * w->nedge)
    return fail(CELLS_EINVAL, "incidence lists disagree with the edges");
  for (int64_t i = 0; i < w->ncell; ++i) {
    int64_t p0 = w->inc_ptr[i], p1 = w->inc_ptr[i + 1];
    if (p1 < p0 || w->inc_out[i] < 0 || w->inc_out[i] > p1 - p0)
      return fail(CELLS_EINVAL, "malformed incidence list");
    for (int64_t p = p0; p < p1; ++p) {
      int64_t e = w->inc_edge[p];
      if (e < 0 || e >= w->nedge) return fail(CELLS_EINVAL, "incidence names an unknown edge");
      bool out = p < p0 + w->inc_out[i];
      if ((out && w->snd[e] != i) || (!out && w->rcv[e] != i))
        return fail(CELLS_EINVAL, "incidence list disagrees with the edge's ends");
    }
  }
  return CELLS_OK;
}

void fill_layout(Params& P, const cells_world* w) {
  int S = w->S, f = 0;
  P.f_dn = f; f += S;
  P.f_mtr = f; f += S;
  P.f_mdiss = f++;
  P.f_pcyc = f++;
  P.f_pdiss = f++;
  P.f_pwork = f++;
  P.f_ccyc = P.f_cdiss = P.f_cwork = -1;
  if (w->calcium) { P.f_ccyc = f++; P.f_cdiss = f++; P.f_cwork = f++; }
  P.f_dvol = P.f_wtr = P.f_wdiss = -1;
  if (w->water) { P.f_dvol = f++; P.f_wtr = f++; P.f_wdiss = f++; }
  P.cf = f;
  P.f_jtr = 0;
  P.f_jdiss = S;
  P.f_gate = S + 1;
  P.ef = S + 2;
  f = 0;
  P.f_sup = f; f += S;
  P.f_echg = f++;
  P.f_rwork = f++;
  P.f_ework = f++;
  P.f_wsup = P.f_wwork = -1;
  if (w->water) { P.f_wsup = f++; P.f_wwork = f++; }
  P.wf = f;
  P.edge_base = (long)P.cf * w->ncell;
  P.world_base = P.edge_base + (long)P.ef * w->nedge;
  P.total = P.world_base + (long)P.wf * w->W;
}

// Rate bounds (1/s) of the worlds at their declared state, from the definition alone as the
// reference's step checks take them (models/surface_pump.py:402-445 StepChecks; held bath, one
// patch per cell, no faces, no floating nodes):
// - explicit: IMEX Euler keeps the membranes and junctions explicit, explicit_rate_bound
//   (surface_pump.py:464-555): the larger of 2 max_i sum_j G_j / C_i over the junctions and
//   max_i G_m / C_i over the membranes, G = F^2/RT sum_s z_s^2 (A P)_s max(c_a, c_b), plus the
//   neutral species' material term (neutral_material_rate, :621); ROS2 puts the route slopes in
//   the stage matrix as N2 does, so what stays explicit is the material relaxation of
//   implicit_conductance_rate_bound (surface_pump_v3.py:4981-5017): 2 max over species and cells
//   of (A P_m B(-x) + sum_j A P_j) / V, x = z F V_0 / RT at the declared membrane voltage;
// - gate: junction_gate_rate_bound (surface_pump.py:447-461), 1.3 exp(77 threshold) + 20 where a
//   gate can close (minimum below 1), checked for IMEX Euler only (surface_pump_v3.py:5161-5164);
// - calcium: calcium_rate_bound (surface_pump_v2.py:583-588), 2 max_i cap_i / (Km V_i).
void rate_bounds(const cells_world* w, int method, double* bexp, double* bgate, double* bca) {
  const int S = w->S;
  const int64_t nc = w->ncell, ne = w->nedge;
  const double RT = kR * w->temperature, slope = kF * kF / RT;
  auto conc = [&](int s, int64_t i) { return w->n0[(int64_t)s * nc + i] / w->vol0[i]; };
  auto host_bernoulli = [](double x) {
    if (std::fabs(x) < 1e-4) {
      const double x2 = x * x;
      return 1.0 - x / 2.0 + x2 / 12.0 - (x2 * x2) / 720.0;
    }
    return x / std::expm1(x);
  };
  auto rows = [](double g, double c) { return c > 0 ? g / c : (g == 0 ? 0.0 : INFINITY); };
  *bexp = *bgate = *bca = 0.0;
  if (method == CELLS_METHOD_RK4) return;  // RK4 has none of the IMEX step checks (surface_pump.py:434-437)
  // the IMEX tableaux with explicit junctions and membranes (Euler, ARS(2,2,2)) against those whose
  // routes' slopes are implicit (N2, and this build's ROS2)
  const bool explicit_routes = method == CELLS_METHOD_EULER || method == CELLS_METHOD_ARS222;
  double best = 0.0;
  if (explicit_routes) {
    std::vector<double> jrow(nc, 0.0);
    for (int64_t e = 0; e < ne; ++e) {
      const int64_t a = w->snd[e], b = w->rcv[e];
      double t = 0.0;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < S)
        t += w->valence[s] * w->valence[s] * w->paj[(int64_t)s * ne + e] * std::max(conc(s, a), conc(s, b));
      const double g = slope * t;
      jrow[a] += g;
      jrow[b] += g;
    }
    double jmax = 0.0, mmax = 0.0;
    for (int64_t i = 0; i < nc; ++i) {
      double t = 0.0;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (s < S)
        t += w->valence[s] * w->valence[s] * w->pam[(int64_t)s * nc + i] * std::max(conc(s, i), w->bath_c[s]);
      jmax = std::max(jmax, rows(jrow[i], w->cap[i]));
      mmax = std::max(mmax, rows(slope * t, w->cap[i]));
    }
    best = std::max(2.0 * jmax, mmax);
  }
  // material relaxation through the routes: every species under ROS2, the neutral ones under Euler
  std::vector<double> mat((size_t)S * nc, 0.0);
  for (int64_t e = 0; e < ne; ++e)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s >= S) break;
      mat[(size_t)s * nc + w->snd[e]] += w->paj[(int64_t)s * ne + e];
      mat[(size_t)s * nc + w->rcv[e]] += w->paj[(int64_t)s * ne + e];
    }
  double material = 0.0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s >= S) break;
    if (explicit_routes && w->valence[s] != 0.0) continue;
    for (int64_t i = 0; i < nc; ++i) {
      const double v0 = w->q0[i] / w->cap[i];
      const double x = w->valence[s] * kF * v0 / RT;
      const double pm = w->pam[(int64_t)s * nc + i];
      const double route = explicit_routes ? pm : pm * host_bernoulli(-x);
      material = std::max(material, (route + mat[(size_t)s * nc + i]) / (w->vol0[i] * 1e-15));
    }
  }
  *bexp = std::max(best, 2.0 * material);
  *bgate = (explicit_routes && ne > 0 && w->gate_minimum < 1.0)
               ? 1.3 * std::exp(77.0 * w->gate_threshold) + 20.0 : 0.0;
  double ca = 0.0;
  if (w->calcium)
    for (int64_t i = 0; i < nc; ++i)
      if (w->ca_cap[i] > 0) ca = std::max(ca, w->ca_cap[i] / w->ca_km_ca / w->vol0[i]);
  *bca = 2.0 * ca;
}

// The step checks of one step size (models/surface_pump.py:431-444, surface_pump_v2.py:612-617,
// surface_pump_v3.py:5156-5164): ConfigError's messages as CELLS_EINVAL.
int check_step_size(const cells_ensemble* e, double dt) {
  if (dt * e->bound_explicit > kExplicitStepBound)
    return fail(CELLS_EINVAL, "IMEX step exceeds the explicit-rate stability bound");
  if (dt * e->bound_calcium > kExplicitStepBound)
    return fail(CELLS_EINVAL, "IMEX step exceeds the explicit calcium pump rate bound");
  if (dt * e->bound_gate > kGateStepBound)
    return fail(CELLS_EINVAL, "Resolved junction step exceeds conservative kinetic bound");
  return CELLS_OK;
}

// The lattice pattern of world k in local indices.
void world_pattern(const cells_world* w, int k, std::vector<int32_t>& rows, std::vector<int32_t>& cols) {
  int64_t c0 = w->cell_off[k];
  rows.clear();
  cols.clear();
  for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
    rows.push_back(static_cast<int32_t>(w->snd[e] - c0));
    cols.push_back(static_cast<int32_t>(w->rcv[e] - c0));
  }
}

}  // namespace

extern "C" {

const char* cells_last_error(void) { return g_error.c_str(); }

int cells_set_plan_options(int32_t relax_rows, int32_t relax_zeros, int32_t grid_relax_zeros,
                           int32_t grid_subtrees, int32_t widen_groups, int32_t largest_first) {
  if (relax_rows < 1 || relax_rows > kMfMaxRows || relax_zeros < 0 || grid_relax_zeros < 0 || grid_subtrees < 1 ||
      (widen_groups != 0 && widen_groups != 1) || (largest_first != 0 && largest_first != 1))
    return fail(CELLS_EINVAL, "plan options out of range");
  g_plan.relax_rows = relax_rows;
  g_plan.relax_zeros = relax_zeros;
  g_plan.grid_relax_zeros = grid_relax_zeros;
  g_plan.grid_subtrees = grid_subtrees;
  g_plan.widen_groups = widen_groups;
  g_plan.largest_first = largest_first;
  return CELLS_OK;
}

int cells_get_plan_options(int32_t* relax_rows, int32_t* relax_zeros, int32_t* grid_relax_zeros,
                           int32_t* grid_subtrees, int32_t* widen_groups, int32_t* largest_first) {
  if (relax_rows) *relax_rows = g_plan.relax_rows;
  if (relax_zeros) *relax_zeros = g_plan.relax_zeros;
  if (grid_relax_zeros) *grid_relax_zeros = g_plan.grid_relax_zeros;
  if (grid_subtrees) *grid_subtrees = g_plan.grid_subtrees;
  if (widen_groups) *widen_groups = g_plan.widen_groups;
  if (largest_first) *largest_first = g_plan.largest_first;
  return CELLS_OK;
}

int cells_capacitor_topology(int32_t n, int32_t o, int64_t m, const int32_t* membrane_cell,
                             const int32_t* membrane_outside, const uint8_t* membrane_active, const uint8_t* cell_active,
                             const uint8_t* outside_active, const uint8_t* reservoir, const uint8_t* gauge, int64_t f,
                             const int32_t* face_sender, const int32_t* face_receiver, const uint8_t* face_conductive,
                             int64_t r, const int32_t* return_paths, int32_t* component, int32_t* floating,
                             uint8_t* fixed, uint8_t* offset_free, int32_t* counts) {
  if (n < 0 || o < 1 || m < 0 || f < 0 || r < 0) return fail(CELLS_EINVAL, "negative layout size");
  if ((m > 0 && (!membrane_cell || !membrane_outside || !membrane_active)) || (n > 0 && !cell_active) ||
      !outside_active || !reservoir || !gauge || (f > 0 && (!face_sender || !face_receiver || !face_conductive)) ||
      (r > 0 && !return_paths) || !counts)
    return fail(CELLS_EINVAL, "NULL argument");
  try {
    const cells::CapacitorTopology t =
        cells::capacitor_topology(n, o, m, membrane_cell, membrane_outside, membrane_active, cell_active, outside_active,
                                  reservoir, gauge, f, face_sender, face_receiver, face_conductive, r, return_paths);
    counts[0] = t.components;
    counts[1] = (int32_t)t.offset_free.size();
    if (component) std::copy(t.component.begin(), t.component.end(), component);
    if (floating) std::copy(t.floating.begin(), t.floating.end(), floating);
    if (fixed) std::copy(t.fixed.begin(), t.fixed.end(), fixed);
    if (offset_free) std::copy(t.offset_free.begin(), t.offset_free.end(), offset_free);
    return CELLS_OK;
  } catch (const std::exception& ex) {
    return fail(CELLS_EINVAL, ex.what());
  }
}

int cells_product_plan_size(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, int64_t* nnz) {
  if (!nnz || (m > 0 && (!rows || !cols))) return fail(CELLS_EINVAL, "NULL argument");
  if (n < 1) return fail(CELLS_EINVAL, "a sparse pattern needs a positive size");
  try {
    *nnz = (int64_t)cells::product_plan(n, m, rows, cols).column.size();
    return CELLS_OK;
  } catch (const std::exception& ex) {
    return fail(CELLS_EINVAL, ex.what());
  }
}

int cells_product_plan(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, int64_t* row_start,
                       int32_t* column, int64_t* diagonal) {
  if (!row_start || !column || !diagonal || (m > 0 && (!rows || !cols))) return fail(CELLS_EINVAL, "NULL argument");
  if (n < 1) return fail(CELLS_EINVAL, "a sparse pattern needs a positive size");
  try {
    const cells::ProductPlan p = cells::product_plan(n, m, rows, cols);
    std::copy(p.row_start.begin(), p.row_start.end(), row_start);
    std::copy(p.column.begin(), p.column.end(), column);
    std::copy(p.diagonal.begin(), p.diagonal.end(), diagonal);
    return CELLS_OK;
  } catch (const std::exception& ex) {
    return fail(CELLS_EINVAL, ex.what());
  }
}

int cells_product(void* stream, int32_t n, const int64_t* row_start, const int32_t* column, const double* values,
                  const double* x, double* y, int32_t count, int64_t values_stride, int64_t vector_stride) {
  if (n < 0 || count < 0 || values_stride < 0 || vector_stride < 0) return fail(CELLS_EINVAL, "negative size or stride");
  if (n == 0 || count == 0) return CELLS_OK;
  if (!row_start || !column || !values || !x || !y) return fail(CELLS_EINVAL, "NULL argument");
  if (count > 65535) return fail(CELLS_EINVAL, "at most 65535 systems per call");
  if (x == y) return fail(CELLS_EINVAL, "y must not be x");
  hipLaunchKernelGGL(k_product, dim3((n + 255) / 256, count), dim3(256), 0, (hipStream_t)stream, n,
                     reinterpret_cast<const long long*>(row_start), column, values, x, y, (long long)values_stride,
                     (long long)vector_stride);
  if (hipGetLastError() != hipSuccess) return fail(CELLS_EHIP, "k_product launch failed");
  return CELLS_OK;
}

int cells_minimum_degree(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols,
                         int32_t* order_out) {
  try {
    if (n < 1) return fail(CELLS_EINVAL, "a sparse pattern needs a positive size");
    auto order = cells::minimum_degree(n, m, rows, cols);
    std::copy(order.begin(), order.end(), order_out);
    return CELLS_OK;
  } catch (const std::exception& ex) {
    return fail(CELLS_EINVAL, ex.what());
  }
}

int cells_bind(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols,
               const int32_t* order, cells_symbolic** out) {
  try {
    auto* s = new cells_symbolic;
    s->s = cells::bind(n, m, rows, cols, order);
    *out = s;
    return CELLS_OK;
  } catch (const std::exception& ex) {
    return fail(CELLS_EINVAL, ex.what());
  }
}

int cells_symbolic_sizes(const cells_symbolic* s, int64_t* nnz, int32_t* levels) {
  if (!s) return fail(CELLS_EINVAL, "symbolic is NULL");
  if (nnz) *nnz = s->s.nnz();
  if (levels) *levels = s->s.levels();
  return CELLS_OK;
}

int cells_symbolic_export(const cells_symbolic* sym, int32_t* order, int32_t* rank,
                          int64_t* column_start, int32_t* row_node, int32_t* column_node,
                          int32_t* parent, int32_t* level) {
  if (!sym) return fail(CELLS_EINVAL, "symbolic is NULL");
  const cells::Symbolic& s = sym->s;
  if (order) std::copy(s.order.begin(), s.order.end(), order);
  if (rank) std::copy(s.rank.begin(), s.rank.end(), rank);
  if (column_start) std::copy(s.column_start.begin(), s.column_start.end(), column_start);
  for (int64_t p = 0; p < s.nnz(); ++p) {
    if (row_node) row_node[p] = s.order[s.row_rank[p]];
    if (column_node) column_node[p] = s.order[s.col_rank[p]];
  }
  if (parent) std::copy(s.parent.begin(), s.parent.end(), parent);
  if (level) std::copy(s.level.begin(), s.level.end(), level);
  return CELLS_OK;
}

void cells_symbolic_free(cells_symbolic* s) { delete s; }

int cells_relaxed_padding(const cells_symbolic* s, int32_t max_rows, int32_t max_zeros, int32_t* pairs,
                          int64_t capacity, int64_t* count) {
  if (!s || !count || max_rows < 1 || max_zeros < 0) return fail(CELLS_EINVAL, "bad argument");
  const std::vector<int32_t> pad = cells::relaxed_padding(s->s, max_rows, max_zeros);
  *count = (int64_t)pad.size() / 2;
  if (pairs) {
    if ((int64_t)pad.size() > 2 * capacity) return fail(CELLS_EINVAL, "pairs buffer too small");
    std::copy(pad.begin(), pad.end(), pairs);
  }
  return CELLS_OK;
}

int cells_ensemble_create(const cells_world* w, int32_t method, int32_t device,
                          cells_ensemble** out) {
  int rc = check_world(w);
  if (rc) return rc;
  const bool pairwise = (method & CELLS_FACTOR_PAIRWISE) != 0;
  method &= ~CELLS_FACTOR_PAIRWISE;
  if (method != CELLS_METHOD_EULER && method != CELLS_METHOD_ROS2 && method != CELLS_METHOD_N2 &&
      method != CELLS_METHOD_ARS222 && method != CELLS_METHOD_RK4)
    return fail(CELLS_EINVAL, "unknown method");
  if (method == CELLS_METHOD_RK4 && w->water)
    return fail(CELLS_EINVAL, "RK4 covers worlds without water (the reference's water RK4 couples the water "
                              "rates into the right-hand side)");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(CELLS_ENODEV, "no GPU visible: the HIP step has no CPU fallback");
  if (device < 0 || device >= count) return fail(CELLS_EINVAL, "device out of range");
  HIP_TRY(hipSetDevice(device));
  std::unique_ptr<cells_ensemble> e(new cells_ensemble);
  e->device = device;
  e->method = method;
  e->pairwise = pairwise;
  e->S = w->S;
  e->W = w->W;
  e->ncell = (int)w->ncell;
  e->nedge = (int)w->nedge;
  e->cell_off.assign(w->cell_off, w->cell_off + w->W + 1);
  e->edge_off.assign(w->edge_off, w->edge_off + w->W + 1);
  Params& P = e->P;
  P.S = w->S;
  P.W = w->W;
  P.water = w->water;
  P.calcium = w->calcium;
  P.ncell = e->ncell;
  P.nedge = e->nedge;
  const double T = w->temperature, RT = kR * T;
  for (int s = 0; s < 4; ++s) {
    P.z[s] = s < w->S ? w->valence[s] : 0.0;
    P.cb[s] = s < w->S ? w->bath_c[s] : 1.0;
    P.log_cb[s] = std::log(P.cb[s]);
    P.mu_b[s] = RT * std::log(P.cb[s]) + P.z[s] * kF * w->bath_V;
  }
  P.T = T;
  P.RT = RT;
  P.Vb = w->bath_V;
  P.nak_drive = w->dg0 + kR * T * (std::log(w->adp) + std::log(w->pi) - std::log(w->atp) - std::log(1000.0));
  P.log_km_na = std::log(w->km_na);
  P.log_km_k = std::log(w->km_k);
  P.nak_atp_sat = host_logaddexp0(std::log(w->km_atp) - std::log(w->atp));
  P.dg0_over_rt_nak_terms = w->dg0 / RT + std::log(w->adp) + std::log(w->pi) - std::log(w->atp) - std::log(1000.0);
  double dg0k = w->ca_dg0 - kR * T * std::log(1000.0);
  P.ca_dk = dg0k + kR * T * (std::log(w->adp) + std::log(w->pi) - std::log(w->atp));
  P.ca_drive = w->ca_dg0 + kR * T * (std::log(w->adp) + std::log(w->pi) - std::log(w->atp) - std::log(1000.0));
  P.log_ca_km = std::log(w->ca_km_ca);
  P.ca_atp_sat = host_logaddexp0(std::log(w->ca_km_atp) - std::log(w->atp));
  P.ca_aff_base = w->ca_dg0 / RT + std::log(w->adp) + std::log(w->pi) - std::log(w->atp) - std::log(1000.0);
  P.gate_thr = w->gate_threshold;
  P.gate_min = w->gate_minimum;
  P.log_0065 = std::log(0.065);
  P.wK = w->water_K;
  P.wP0 = w->water_P0;
  P.vmin = w->water_vmin;
  P.vmax = w->water_vmax;
  {
    double total = 0.0;
    for (int s = 0; s < w->S; ++s) total += w->bath_c[s];
    P.psib = w->bath_pressure - kR * T * total;
  }
  // surface_pump.py:63-64 ENERGY_RTOL: IMEX Euler's 2e-2, N2's (ARS(2,2,2)'s) 2e-3. ROS2, which
  // the reference lacks, takes Euler's: its frozen-slope stages carry the stiff start's energy
  // error at first order (at 5 ms the first steps of the calcium worlds book 6e-3 of their flows)
  P.energy_rtol = method == CELLS_METHOD_N2 || method == CELLS_METHOD_ARS222 ? 2e-3
                  : method == CELLS_METHOD_RK4                             ? 0.0
                                                                           : 2e-2;
  fill_layout(P, w);
  if (!(T >= 200.0 && T <= 400.0)) return fail(CELLS_EINVAL, "temperature outside 200-400 K");
  if (!(w->gate_threshold >= 0 && w->gate_threshold <= kGateMaxVoltage && w->gate_minimum >= 0 &&
        w->gate_minimum <= 1))
    return fail(CELLS_EINVAL, "junction gate parameters outside their domain");
  // per-cell and per-edge arrays
  std::vector<int> cell_world(e->ncell), edge_world(e->nedge);
  for (int k = 0; k < e->W; ++k) {
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) cell_world[i] = k;
    for (int64_t q = w->edge_off[k]; q < w->edge_off[k + 1]; ++q) edge_world[q] = k;
    e->max_n = std::max<int>(e->max_n, (int)(w->cell_off[k + 1] - w->cell_off[k]));
  }
  auto co = narrow(w->cell_off, w->W + 1), eo = narrow(w->edge_off, w->W + 1);
  auto snd = narrow(w->snd, e->nedge), rcv = narrow(w->rcv, e->nedge);
  auto iptr = narrow(w->inc_ptr, e->ncell + 1), iedge = narrow(w->inc_edge, 2 * (size_t)e->nedge);
  int* d_i = nullptr;
  double* d_d = nullptr;
  size_t nc = e->ncell, ne = e->nedge, S = w->S;
#define UP_I(field, vec) if ((rc = upload(e.get(), &d_i, vec.data(), vec.size()))) return rc; P.field = d_i;
#define UP_D(field, src, count) if ((rc = upload(e.get(), &d_d, src, count))) return rc; P.field = d_d;
  UP_I(cell_off, co);
  UP_I(edge_off, eo);
  UP_I(cell_world, cell_world);
  UP_I(edge_world, edge_world);
  UP_I(snd, snd);
  UP_I(rcv, rcv);
  UP_I(inc_ptr, iptr);
  UP_I(inc_edge, iedge);
  if ((rc = upload(e.get(), &d_i, w->inc_out, nc))) return rc;
  P.inc_out = d_i;
  {
    std::vector<int> order(e->W);
    for (int k = 0; k < e->W; ++k) order[k] = k;
    if (g_plan.largest_first)
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return w->cell_off[a + 1] - w->cell_off[a] > w->cell_off[b + 1] - w->cell_off[b];
      });
    UP_I(world_order, order);
  }
  UP_D(vol0, w->vol0, nc);
  UP_D(cap, w->cap, nc);
  UP_D(pam, w->pam, S * nc);
  UP_D(nak_cap, w->nak_cap, nc);
  UP_D(ca_cap, w->ca_cap, nc);
  UP_D(n0, w->n0, S * nc);
  UP_D(q0, w->q0, nc);
  UP_D(lpa, w->lpa, nc);
  UP_D(imperm, w->imperm, nc);
  UP_D(paj, w->paj, S * ne);
#undef UP_I
#undef UP_D
  rate_bounds(w, method, &e->bound_explicit, &e->bound_gate, &e->bound_calcium);
  // workspace
  long total = P.total;
  if ((rc = alloc(e.get(), &e->A, total)) || (rc = alloc(e.get(), &e->B, total)) ||
      (rc = alloc(e.get(), &e->K1, total)) || (rc = alloc(e.get(), &e->contrib1, S * nc)) ||
      (rc = alloc(e.get(), &e->ysc, nc)) ||
      (rc = alloc(e.get(), &e->sel, e->W)) || (rc = alloc(e.get(), &e->alive, e->W)) ||
      (rc = alloc(e.get(), &e->valid, e->W)) || (rc = alloc(e.get(), &e->step_counter, 1)) ||
      (rc = alloc(e.get(), &e->check_bits, e->W)))
    return rc;
  P.refuse_bits = e->check_bits;
  if (method == CELLS_METHOD_N2) {
    N2Buffers& n = e->n2;
    if ((rc = alloc(e.get(), &n.E0, total)) || (rc = alloc(e.get(), &n.E1, total)) ||
        (rc = alloc(e.get(), &n.D1, total)) || (rc = alloc(e.get(), &n.D2, total)) ||
        (rc = alloc(e.get(), &n.known, total)) || (rc = alloc(e.get(), &n.stage, total)) ||
        (rc = alloc(e.get(), &n.diag, nc)) || (rc = alloc(e.get(), &n.q, nc)) ||
        (rc = alloc(e.get(), &n.cD1, S * nc)) || (rc = alloc(e.get(), &n.cD2, S * nc)) ||
        (rc = alloc(e.get(), &n.phi0, nc)) || (rc = alloc(e.get(), &n.target0, std::max<size_t>(ne, 1))) ||
        (rc = alloc(e.get(), &n.relax0, std::max<size_t>(ne, 1))))
      return rc;
  }
  if (method == CELLS_METHOD_RK4)
    if ((rc = alloc(e.get(), &e->K3, total)) || (rc = alloc(e.get(), &e->K4, total))) return rc;
  if (method == CELLS_METHOD_ARS222 || method == CELLS_METHOD_RK4)
    if ((rc = alloc(e.get(), &e->Y1, total)) || (rc = alloc(e.get(), &e->K2, total)) ||
        (rc = alloc(e.get(), &e->contrib2, S * nc)))
      return rc;
  if (method == CELLS_METHOD_ROS2 || method == CELLS_METHOD_N2) {
    if ((rc = alloc(e.get(), &e->Y1, total)) || (rc = alloc(e.get(), &e->K2, total)) ||
        (rc = alloc(e.get(), &e->contrib2, S * nc)) || (rc = alloc(e.get(), &e->gm, S * nc)) ||
        (rc = alloc(e.get(), &e->gj, std::max<size_t>(S * ne, 1))) ||
        (rc = alloc(e.get(), &e->Gj, std::max<size_t>(ne, 1))) || (rc = alloc(e.get(), &e->mdiag, nc)) ||
        (rc = alloc(e.get(), &e->msc, nc)) || (rc = alloc(e.get(), &e->dphi, nc)) ||
        (rc = alloc(e.get(), &e->dphi2, nc)))
      return rc;
  }
  // symbolic analysis once per lattice shape (worlds with the same cell count share it)
  {
    std::map<int, int> by_n;
    std::vector<int> rep;  // a representative world per shape
    e->shape_of_host.resize(e->W);
    for (int k = 0; k < e->W; ++k) {
      int n = (int)(w->cell_off[k + 1] - w->cell_off[k]);
      auto it = by_n.find(n);
      if (it == by_n.end()) {
        it = by_n.emplace(n, (int)rep.size()).first;
        rep.push_back(k);
      }
      e->shape_of_host[k] = it->second;
    }
    // the shape is keyed by n: verify every world's pattern equals its representative's
    for (int k = 0; k < e->W; ++k) {
      int r = rep[e->shape_of_host[k]];
      int64_t ek = w->edge_off[k + 1] - w->edge_off[k], er = w->edge_off[r + 1] - w->edge_off[r];
      if (ek != er) return fail(CELLS_EINVAL, "worlds of one cell count with different edges");
      for (int64_t q = 0; q < ek; ++q)
        if (w->snd[w->edge_off[k] + q] - w->cell_off[k] != w->snd[w->edge_off[r] + q] - w->cell_off[r] ||
            w->rcv[w->edge_off[k] + q] - w->cell_off[k] != w->rcv[w->edge_off[r] + q] - w->cell_off[r])
          return fail(CELLS_EINVAL, "worlds of one cell count with different lattices");
    }
    e->shapes_host.resize(rep.size());
    e->exact_nnz_host.resize(rep.size());
    e->exact_flops_host.resize(rep.size());
    e->sn_host.resize(rep.size());
    std::atomic<int> next_shape{0};
    std::vector<std::string> bad_msg(rep.size());  // per shape: the failure of its analysis, if any
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool_threads;
    const PlanOptions plan = g_plan;  // read once: the worker threads share it
    for (unsigned t = 0; t < hw; ++t)
      pool_threads.emplace_back([&]() {
        std::vector<int32_t> rows, cols;
        for (int q; (q = next_shape.fetch_add(1)) < (int)rep.size();) {
          try {
            world_pattern(w, rep[q], rows, cols);
            int n = (int)(w->cell_off[rep[q] + 1] - w->cell_off[rep[q]]);
            cells::Symbolic s = cells::bind(n, (int64_t)rows.size(), rows.data(), cols.data(), nullptr);
            e->exact_nnz_host[q] = s.nnz();
            double f = 0.0;
            for (int32_t k = 0; k < s.n; ++k) {
              const double c = (double)(s.column_start[k + 1] - s.column_start[k]);
              f += c * c;
            }
            e->exact_flops_host[q] = f;
            if (!e->pairwise) {
              // relaxed supernodes: the multifrontal plan factors the pattern with the chains'
              // explicit zeros (the reference-order mode keeps the bound pattern itself)
              const std::vector<int32_t> pad = cells::relaxed_padding(s, plan.relax_rows, plan.relax_zeros);
              if (!pad.empty()) {
                for (size_t x = 0; x < pad.size(); x += 2) {
                  rows.push_back(pad[x]);
                  cols.push_back(pad[x + 1]);
                }
                cells::Symbolic r = cells::bind(n, (int64_t)rows.size(), rows.data(), cols.data(), s.order.data());
                if (r.parent != s.parent) throw std::runtime_error("relaxed supernodes changed the elimination tree");
                s = std::move(r);
              }
            }
            e->shapes_host[q] = std::move(s);
            e->sn_host[q] = cells::supernodes(e->shapes_host[q], front_class, widen_class);
            cells::share_update_storage(e->sn_host[q]);
          } catch (const std::exception& ex) {
            bad_msg[q] = ex.what();
            if (bad_msg[q].empty()) bad_msg[q] = "symbolic analysis failed";
          }
        }
      });
    for (auto& t : pool_threads) t.join();
    for (const std::string& m : bad_msg)  // the first failing shape's message, whatever the timing
      if (!m.empty()) return fail(CELLS_EINVAL, m);
    // device plans, concatenated
    std::vector<ShapeDev> sd(rep.size());
    std::vector<int> pool;
    std::vector<unsigned long long> mask_pool;
    std::vector<int> slot_of;
    for (size_t q = 0; q < rep.size(); ++q) {
      const cells::Symbolic& s = e->shapes_host[q];
      ShapeDev d{};
      d.n = s.n;
      d.nnz = (int)s.nnz();
      d.levels = s.levels();
      int k = rep[q];
      // slot masks: for each row-structure pair (column kk, source column m), the slots of
      // column kk that column m's entries at rows >= kk reach
      d.o_mask = (long)mask_pool.size();
      {
        // the trailing columns that fit the factor's LDS budget
        int64_t budget = kFactorLdsDoubles, pt = s.nnz();
        for (int kk = s.n - 1; kk >= 0 && s.nnz() - s.column_start[kk] <= budget; --kk) pt = s.column_start[kk];
        d.lds_pt = (int)pt;
      }
      slot_of.assign(s.n, -1);
      for (int kk = 0; kk < s.n && e->pairwise; ++kk) {
        int64_t c0 = s.column_start[kk], c1 = s.column_start[kk + 1];
        if (c1 - c0 > 128) return fail(CELLS_EINVAL, "a factor column holds more than 128 entries");
        for (int64_t p2 = c0; p2 < c1; ++p2) slot_of[s.row_rank[p2]] = (int)(p2 - c0);
        for (int64_t x = s.row_ptr[kk]; x < s.row_ptr[kk + 1]; ++x) {
          int m = s.row_col[x];
          unsigned long long w2[2] = {0ull, 0ull};
          for (int64_t p2 = s.row_pos[x]; p2 < s.column_start[m + 1]; ++p2) {
            int slot = slot_of[s.row_rank[p2]];
            if (slot < 0) return fail(CELLS_EINVAL, "symbolic structure is not closed under the elimination tree");
            w2[slot >> 6] |= 1ull << (slot & 63);
          }
          mask_pool.push_back(w2[0]);
          mask_pool.push_back(w2[1]);
        }
        for (int64_t p2 = c0; p2 < c1; ++p2) slot_of[s.row_rank[p2]] = -1;
      }
      d.E = (int)(w->edge_off[k + 1] - w->edge_off[k]);
      d.o_cs = (long)pool.size();
      for (int64_t v : s.column_start) pool.push_back((int)v);
      d.o_entry = (long)pool.size();
      pool.insert(pool.end(), s.row_rank.begin(), s.row_rank.end());
      pool.insert(pool.end(), s.col_rank.begin(), s.col_rank.end());
      d.o_rowptr = (long)pool.size();
      for (int64_t v : s.row_ptr) pool.push_back((int)v);
      d.o_rowent = (long)pool.size();
      pool.insert(pool.end(), s.row_col.begin(), s.row_col.end());
      pool.insert(pool.end(), s.row_pos.begin(), s.row_pos.end());
      d.o_lev = (long)pool.size();
      pool.insert(pool.end(), s.level_ptr.begin(), s.level_ptr.end());
      // entries grouped by level: each level's columns in order, each column's entries in order
      std::vector<int> lev_ent_ptr(d.levels + 1, 0), lev_ent;
      lev_ent.reserve(d.nnz);
      for (int v = 0; v < d.levels; ++v) {
        for (int x = s.level_ptr[v]; x < s.level_ptr[v + 1]; ++x) {
          int col = s.level_col[x];
          for (int64_t p = s.column_start[col]; p < s.column_start[col + 1]; ++p) lev_ent.push_back((int)p);
        }
        lev_ent_ptr[v + 1] = (int)lev_ent.size();
      }
      pool.insert(pool.end(), lev_ent_ptr.begin(), lev_ent_ptr.end());
      d.o_levcol = (long)pool.size();
      pool.insert(pool.end(), s.level_col.begin(), s.level_col.end());
      d.o_levent = (long)pool.size();
      pool.insert(pool.end(), lev_ent.begin(), lev_ent.end());
      d.o_node = (long)pool.size();
      pool.insert(pool.end(), s.order.begin(), s.order.end());
      for (int i = 0; i < s.n; ++i) pool.push_back((int)s.column_start[s.rank[i]]);
      std::vector<int32_t> rows, cols;
      world_pattern(w, k, rows, cols);
      for (size_t q2 = 0; q2 < rows.size(); ++q2) pool.push_back((int)cells::position(s, rows[q2], cols[q2]));
      // the multifrontal schedule (indices into the pool are absolute, so a front needs only its
      // record and its world's buffers)
      const cells::Supernodes& sn = e->sn_host[q];
      const int ns = (int)sn.first.size();
      if (sn.update_total > 0x7fffffff) return fail(CELLS_EINVAL, "update matrices exceed 2^31 doubles");
      std::vector<int64_t> voff(ns + 1, 0);  // the solve's update vectors: |S| doubles per supernode
      for (int t = 0; t < ns; ++t) voff[t + 1] = voff[t] + (sn.rows[t] - sn.width[t]);
      d.vsz = voff[ns];
      while (pool.size() % kRecInts) pool.push_back(0);  // records on 64-byte boundaries
      d.o_sn = (long)pool.size();
      const long o_lev = d.o_sn + (long)kRecInts * ns;
      long o_child = o_lev + (long)kLatBounds * sn.levels;
      o_child += (4 - o_child % 4) % 4;
      const long o_rel = o_child + 4L * (long)sn.child.size();
      // each front's junction entries: (front-local packed position, world-local junction)
      std::vector<int> front_of(s.n);
      for (int t = 0; t < ns; ++t)
        for (int c = sn.first[t]; c < sn.first[t] + sn.width[t]; ++c) front_of[c] = t;
      std::vector<std::vector<int>> fedges(ns);
      for (size_t q2 = 0; q2 < rows.size(); ++q2) {
        const int64_t pe = cells::position(s, rows[q2], cols[q2]);
        const int t = front_of[s.col_rank[pe]];
        fedges[t].push_back((int)(pe - s.column_start[sn.first[t]]));
        fedges[t].push_back((int)q2);
      }
      long o_fedge = o_rel + (long)sn.rel.size();
      o_fedge += o_fedge % 2;  // int2 aligned
      std::vector<long> fstart(ns + 1, 0);
      for (int t = 0; t < ns; ++t) fstart[t + 1] = fstart[t] + (long)fedges[t].size();
      if (o_fedge + fstart[ns] > 0x7fffffffL) return fail(CELLS_EINVAL, "plans exceed 2^31 entries");
      for (int t = 0; t < ns; ++t) {
        if (sn.cls[t] > 7) return fail(CELLS_EINVAL, "a front of order above 128");
        if (sn.cls[t] == 7)
          d.big_tri = std::max(d.big_tri, sn.rows[t] * (sn.rows[t] + 1) / 2);
        const int64_t last = sn.first[t] + sn.width[t] - 1;
        int32_t r16[kRecInts] = {(int32_t)s.column_start[sn.first[t]], sn.width[t], sn.rows[t],
                                 (int32_t)sn.update_offset[t], (int32_t)(o_child / 4 + sn.child_ptr[t]),
                                 (int32_t)(o_child / 4 + sn.child_ptr[t + 1]), (int32_t)voff[t], sn.first[t],
                                 (int32_t)(d.o_entry + s.column_start[last] + 1), 0,
                                 (int32_t)(o_fedge + fstart[t]), (int32_t)(fedges[t].size() / 2)};
        pool.insert(pool.end(), r16, r16 + kRecInts);
      }
      d.sn_levels = sn.levels;
      d.o_snlev = (long)pool.size();
      {
        std::vector<int> bounds(kLatBounds * (size_t)sn.levels, 0);
        int t = 0;
        for (int L = 0; L < sn.levels; ++L) {
          for (int c = 0; c < kLatBounds - 1; ++c) {
            bounds[kLatBounds * L + c] = t;
            while (t < ns && sn.level[t] == L && sn.cls[t] == c) ++t;
          }
          bounds[kLatBounds * L + kLatBounds - 1] = t;
          if (t < ns && sn.level[t] == L) return fail(CELLS_EINVAL, "front schedule out of order");
        }
        pool.insert(pool.end(), bounds.begin(), bounds.end());
      }
      while ((long)pool.size() < o_child) pool.push_back(0);
      d.o_child = (long)pool.size();
      for (int32_t c : sn.child) {
        int32_t c4[4] = {sn.rows[c] - sn.width[c], (int32_t)(o_rel + sn.rel_ptr[c]), (int32_t)sn.update_offset[c],
                         (int32_t)voff[c]};
        pool.insert(pool.end(), c4, c4 + 4);
      }
      d.o_rel = (long)pool.size();
      if (d.o_rel != o_rel) return fail(CELLS_EINVAL, "plan layout mismatch");
      pool.insert(pool.end(), sn.rel.begin(), sn.rel.end());
      while ((long)pool.size() < o_fedge) pool.push_back(0);
      for (int t = 0; t < ns; ++t) pool.insert(pool.end(), fedges[t].begin(), fedges[t].end());
      d.usz = sn.update_total;
      sd[q] = d;
    }
    if (pool.size() > (size_t)0x7fffffff) return fail(CELLS_EINVAL, "plans exceed 2^31 entries");
    e->sd_host = sd;
    e->pool_host = pool;
    {
      int64_t most = 0;
      for (const ShapeDev& d : sd) most = std::max<int64_t>(most, d.nnz - d.lds_pt);
      e->factor_lds = sizeof(double) * (size_t)std::max<int64_t>(most, 1);
      // (N2's stage solves take their right-hand sides in HBM, at each cell's rank)
      e->y_in_lds = e->pairwise && e->method != CELLS_METHOD_N2 &&
                    (size_t)e->max_n * sizeof(double) <= kSolveLdsBytes ? 1 : 0;
      e->solve_lds = e->y_in_lds ? sizeof(double) * (size_t)std::max(e->max_n, 1) : 0;
    }
    {
      // each cell's place in its world's solve vector: the world's first cell plus the cell's rank
      std::vector<int> cell_rank(e->ncell);
      for (int k = 0; k < e->W; ++k) {
        const cells::Symbolic& sy = e->shapes_host[e->shape_of_host[k]];
        const int c0 = (int)w->cell_off[k];
        for (int i = 0; i < sy.n; ++i) cell_rank[c0 + i] = c0 + sy.rank[i];
      }
      int* d_rank = nullptr;
      if ((rc = upload(e.get(), &d_rank, cell_rank.data(), cell_rank.size()))) return rc;
      P.cell_rank = d_rank;
    }
    e->foff_host.resize(e->W + 1);
    e->foff_host[0] = 0;
    for (int k = 0; k < e->W; ++k) e->foff_host[k + 1] = e->foff_host[k] + sd[e->shape_of_host[k]].nnz;
    e->factor_total = e->foff_host[e->W];
    std::vector<long> uoffw(e->W + 1, 0), voffw(e->W + 1, 0), boffw(e->W + 1, 0);
    for (int k = 0; k < e->W; ++k) {
      uoffw[k + 1] = uoffw[k] + sd[e->shape_of_host[k]].usz;
      voffw[k + 1] = voffw[k] + sd[e->shape_of_host[k]].vsz;
      boffw[k + 1] = boffw[k] + sd[e->shape_of_host[k]].big_tri;
    }
    if ((rc = upload(e.get(), &e->shapes, sd.data(), sd.size())) ||
        (rc = upload(e.get(), &e->shape_of, e->shape_of_host.data(), e->shape_of_host.size())) ||
        (rc = upload(e.get(), &e->pool, pool.data(), pool.size())) ||
        (rc = upload(e.get(), &e->masks, mask_pool.data(), mask_pool.size())) ||
        (rc = upload(e.get(), &e->foff, e->foff_host.data(), e->foff_host.size())) ||
        (rc = alloc(e.get(), &e->fvals, std::max<long>(e->factor_total, 1))) ||
        (rc = upload(e.get(), &e->uoffw, uoffw.data(), uoffw.size())) ||
        (rc = alloc(e.get(), &e->ubuf, std::max<long>(uoffw[e->W], 1))) ||
        (rc = upload(e.get(), &e->voffw, voffw.data(), voffw.size())) ||
        (rc = alloc(e.get(), &e->vbuf, std::max<long>(voffw[e->W], 1))) ||
        (rc = upload(e.get(), &e->boffw, boffw.data(), boffw.size())) ||
        (rc = alloc(e.get(), &e->bigbuf, std::max<long>(boffw[e->W], 1))))
      return rc;
  }
  HIP_TRY(hipStreamCreateWithFlags(&e->capture_stream, hipStreamNonBlocking));
  *out = e.release();
  return CELLS_OK;
}

void cells_ensemble_destroy(cells_ensemble* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  (void)hipDeviceSynchronize();
  stationary_release(e);
  delete e;
}

int64_t cells_state_size(const cells_ensemble* e) { return e ? e->P.total : -1; }
int64_t cells_factor_size(const cells_ensemble* e) { return e ? e->factor_total : -1; }

int cells_world_positions(const cells_ensemble* e, int32_t w, int64_t* diag_pos, int64_t* edge_pos,
                          int64_t* factor_offset) {
  if (!e || w < 0 || w >= e->W) return fail(CELLS_EINVAL, "world out of range");
  const ShapeDev& d = e->sd_host[e->shape_of_host[w]];
  const int* node = e->pool_host.data() + d.o_node;
  if (diag_pos)
    for (int i = 0; i < d.n; ++i) diag_pos[i] = node[d.n + i];
  if (edge_pos)
    for (int q = 0; q < d.E; ++q) edge_pos[q] = node[2 * d.n + q];
  if (factor_offset) *factor_offset = e->foff_host[w];
  return CELLS_OK;
}

int cells_world_factor_positions(const cells_ensemble* e, int32_t w, const int32_t* rows, const int32_t* cols,
                                 int64_t count, int64_t* out) {
  if (!e || w < 0 || w >= e->W) return fail(CELLS_EINVAL, "world out of range");
  if (count < 0 || (count > 0 && (!rows || !cols || !out))) return fail(CELLS_EINVAL, "bad argument");
  const cells::Symbolic& s = e->shapes_host[e->shape_of_host[w]];
  for (int64_t x = 0; x < count; ++x) {
    if (rows[x] < 0 || rows[x] >= s.n || cols[x] < 0 || cols[x] >= s.n) return fail(CELLS_EINVAL, "node out of range");
    try {
      out[x] = cells::position(s, rows[x], cols[x]);
    } catch (const std::out_of_range&) {
      out[x] = -1;
    }
  }
  return CELLS_OK;
}

}  // extern "C"

namespace {

// One step's launches. world = -1: every world in each launch (batched); otherwise world k only.
// Records an event after each launch when timing (never under graph capture).
struct KernelTimer {
  std::vector<const char*> names;
  std::vector<hipEvent_t> events;
  hipStream_t stream = nullptr;
  void mark(const char* name) {
    hipEvent_t ev;
    (void)hipEventCreate(&ev);
    (void)hipEventRecord(ev, stream);
    names.push_back(name);
    events.push_back(ev);
  }
  ~KernelTimer() {
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  }
};

// The solve kernels' workgroup, LDS and mode.
dim3 solve_block(const cells_ensemble* e) { return dim3(e->pairwise ? kFactorBlock : kMfThreads); }
size_t solve_lds_bytes(const cells_ensemble* e) {
  return e->pairwise ? e->solve_lds : sizeof(double) * (size_t)kLatSvLds;
}
SolveArgs solve_args(const cells_ensemble* e) { return SolveArgs{e->pairwise ? 0 : 1, e->voffw, e->vbuf}; }

// A solve stage over worlds: launch(part) runs the stage kernel (part 0: right-hand side, solve
// and result in one workgroup per world).
template <typename Launch>
void solve_stage(cells_ensemble*, hipStream_t, int, const double*, Launch launch) {
  launch(0);
}

// Assemble and factor the stage matrix C + coef G of worlds [w0, w0 + nw).
void launch_factor(cells_ensemble* e, hipStream_t st, int w0, int nw, double coef) {
  if (e->pairwise)
    hipLaunchKernelGGL(k_factor, dim3(nw), dim3(kFactorBlock), e->factor_lds, st, e->P, e->shapes, e->shape_of,
                       e->pool, e->masks, e->foff, w0, coef, e->mdiag, e->msc, e->Gj, e->fvals, e->valid);
  else
    hipLaunchKernelGGL(k_factor_mf, dim3(nw), dim3(kMfThreads), sizeof(double) * kMfBlockLds, st, e->P, e->shapes,
                       e->shape_of, e->pool, e->foff, e->uoffw, w0, coef, e->mdiag, e->msc, e->Gj, e->fvals,
                       e->ubuf, e->bigbuf, e->boffw, e->valid);
}

// N2's step of worlds [w0, w0 + nw) between the water half-steps (models/surface_pump_v3.py:
// 3547-3955 imex_step under cells.solver.imex_ars222_implicit_conductances@1): the start's rates,
// the operator-time prediction and its slopes, the operator's factor, then the tableau's two
// stages (each: its known part and right-hand side, the operator solve, the linear parts and the
// stage state with its gates; the first's rates at the stage potentials less the linear parts)
// and the ledgers' tail.
template <typename Mark>
void launch_n2(cells_ensemble* e, hipStream_t st, double dt, const Ctl& c, int w0, int nw, double* A, double* B,
               Mark& mark) {
  const Params& P = e->P;
  const N2Buffers& n = e->n2;
  const int c0 = (int)e->cell_off[w0], c1 = (int)e->cell_off[w0 + nw];
  const int ed0 = (int)e->edge_off[w0], ed1 = (int)e->edge_off[w0 + nw];
  const int nc = c1 - c0, ne = ed1 - ed0;
  const int vol_next = P.water ? 1 : 0;
  const double gamma = kArsGamma, delta = kArsDelta;
  const double coef = gamma * dt, h = gamma * dt, theta = 0.5 * dt;
  // each stage's explicit fraction of the step, its tableau row summed in order
  const double span1 = ((gamma + 0.0) + 0.0) * dt, span2 = ((delta + (1.0 - delta)) + 0.0) * dt;
  // the start's rates
  if (ne) {
    hipLaunchKernelGGL(k_edges<false>, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, nullptr, vol_next,
                       ed0, ed1, c, n.E0, nullptr, nullptr, e->valid);
    mark("k_edges");
  }
  hipLaunchKernelGGL(k_cells<false>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, nullptr, vol_next,
                     c0, c1, c, n.E0, e->contrib1, nullptr, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr,
                     e->valid);
  mark("k_cells");
  // the operator-time prediction (its state in n.stage until the first stage) and its slopes
  hipLaunchKernelGGL(k_n2_predict, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, A, B, e->sel, vol_next, c0, nc,
                     ed0, ne, theta, n.E0, n.stage, n.phi0, n.target0, n.relax0, e->valid);
  if (ne)
    hipLaunchKernelGGL(k_n2_slopes_edges, dim3(blocks(ne)), dim3(kBlock), 0, st, P, n.stage, c, ed0, ne, e->gj,
                       e->Gj, e->valid);
  hipLaunchKernelGGL(k_n2_slopes_cells, dim3(blocks(nc)), dim3(kBlock), 0, st, P, n.stage, c, coef, c0, nc, e->Gj,
                     e->gm, e->mdiag, e->msc, n.diag, e->valid);
  mark("k_n2_predict");
  launch_factor(e, st, w0, nw, coef);
  mark("k_factor");
  // stage 0's explicit rates: the start's less the linear parts at the start's potentials
  const N2Src start{nullptr, vol_next}, at_stage{n.stage, 0};
  if (ne)
    hipLaunchKernelGGL(k_n2_remainder_edges, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, start, n.phi0,
                       e->gj, ed0, ne, n.E0);
  hipLaunchKernelGGL(k_n2_remainder_cells, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, start, n.phi0,
                     e->gm, e->gj, c0, nc, n.E0, e->contrib1);
  hipLaunchKernelGGL(k_n2_world_rates, dim3(nw), dim3(kBlock), 0, st, P, w0, e->contrib1, n.E0);
  mark("k_n2_remainder");
  auto solve = [&](double* phi) {
    hipLaunchKernelGGL((e->pairwise ? k_solve_stage<false> : k_solve_stage<true>), dim3(nw), solve_block(e),
                       solve_lds_bytes(e), st, P, e->shapes, e->shape_of, e->pool, e->foff, w0, e->msc, e->fvals,
                       n.E0, nullptr, phi, e->ysc, e->y_in_lds, solve_args(e), 0, 1);
    mark("k_solve_stage");
  };
  // stage 1: known = y + gamma dt E0; the operator's potentials; D1 and the stage state
  hipLaunchKernelGGL(k_n2_known, dim3(blocks(nc + ne + nw)), dim3(kBlock), 0, st, P, A, B, e->sel, vol_next, c0, nc,
                     ed0, ne, w0, nw, gamma * dt, n.E0, 0.0, nullptr, 0.0, nullptr, coef, e->gm, e->msc, n.known,
                     n.q, e->ysc);
  mark("k_n2_known");
  solve(e->dphi);
  hipLaunchKernelGGL(k_n2_stage_cells, dim3(blocks(nc)), dim3(kBlock), 0, st, P, n.known, e->dphi, e->gm, e->gj,
                     e->Gj, n.diag, n.q, coef, h, c0, nc, n.D1, n.cD1, e->sel, A, B, n.stage, e->res_atol,
                     e->res_rtol, e->valid);
  if (ne)
    hipLaunchKernelGGL(k_n2_stage_edges, dim3(blocks(ne)), dim3(kBlock), 0, st, P, n.known, e->dphi, e->gj,
                       n.target0, n.relax0, c, h, span1, ed0, ne, n.D1, e->sel, A, B, n.stage, e->valid);
  hipLaunchKernelGGL(k_n2_world_rates, dim3(nw), dim3(kBlock), 0, st, P, w0, n.cD1, n.D1);
  mark("k_n2_stage");
  // stage 1's explicit rates at the stage state and the operator's potentials, less the linear parts
  Params P1 = P;
  P1.phi = e->dphi;
  if (ne) {
    hipLaunchKernelGGL(k_edges<false>, dim3(blocks(ne)), dim3(kBlock), 0, st, P1, A, B, e->sel, n.stage, 0, ed0,
                       ed1, c, n.E1, nullptr, nullptr, e->valid);
    mark("k_edges");
  }
  hipLaunchKernelGGL(k_cells<false>, dim3(blocks(nc)), dim3(kBlock), 0, st, P1, A, B, e->sel, n.stage, 0, c0, c1, c,
                     n.E1, e->contrib2, nullptr, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, e->valid);
  mark("k_cells");
  if (ne)
    hipLaunchKernelGGL(k_n2_remainder_edges, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, at_stage,
                       e->dphi, e->gj, ed0, ne, n.E1);
  hipLaunchKernelGGL(k_n2_remainder_cells, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, at_stage,
                     e->dphi, e->gm, e->gj, c0, nc, n.E1, e->contrib2);
  hipLaunchKernelGGL(k_n2_world_rates, dim3(nw), dim3(kBlock), 0, st, P, w0, e->contrib2, n.E1);
  mark("k_n2_remainder");
  // stage 2: known = ((y + delta dt E0) + (1 - delta) dt E1) + (1 - gamma) dt D1; the new state
  hipLaunchKernelGGL(k_n2_known, dim3(blocks(nc + ne + nw)), dim3(kBlock), 0, st, P, A, B, e->sel, vol_next, c0, nc,
                     ed0, ne, w0, nw, delta * dt, n.E0, (1.0 - delta) * dt, n.E1, (1.0 - gamma) * dt, n.D1, coef,
                     e->gm, e->msc, n.known, n.q, e->ysc);
  mark("k_n2_known");
  solve(e->dphi2);
  hipLaunchKernelGGL(k_n2_stage_cells, dim3(blocks(nc)), dim3(kBlock), 0, st, P, n.known, e->dphi2, e->gm, e->gj,
                     e->Gj, n.diag, n.q, coef, h, c0, nc, n.D2, n.cD2, e->sel, A, B, nullptr, e->res_atol,
                     e->res_rtol, e->valid);
  if (ne)
    hipLaunchKernelGGL(k_n2_stage_edges, dim3(blocks(ne)), dim3(kBlock), 0, st, P, n.known, e->dphi2, e->gj,
                       n.target0, n.relax0, c, h, span2, ed0, ne, n.D2, e->sel, A, B, nullptr, e->valid);
  hipLaunchKernelGGL(k_n2_world_rates, dim3(nw), dim3(kBlock), 0, st, P, w0, n.cD2, n.D2);
  mark("k_n2_stage");
  hipLaunchKernelGGL(k_n2_tail, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, w0, dt, h, n.known, n.E0, n.E1,
                     n.D1, n.D2, e->valid);
  mark("k_n2_tail");
}

// The explicit tableaux' step of worlds [w0, w0 + nw) between the water half-steps: ARS(2,2,2)
// (its stages' rates with the membrane-rate refusal, each stage's capacitor potentials checked, the
// ledgers' tail) or RK4 (four slopes of every leaf, each trial's domain, the end's).
template <typename Mark>
void launch_explicit(cells_ensemble* e, hipStream_t st, double dt, const Ctl& c, int w0, int nw, double* A, double* B,
                     Mark& mark) {
  const Params& P = e->P;
  const int c0 = (int)e->cell_off[w0], c1 = (int)e->cell_off[w0 + nw];
  const int ed0 = (int)e->edge_off[w0], ed1 = (int)e->edge_off[w0 + nw];
  const int nc = c1 - c0, ne = ed1 - ed0;
  const int vol_next = P.water ? 1 : 0;
  const bool rk4 = e->method == CELLS_METHOD_RK4;
  const double rate_dt = rk4 ? 0.0 : dt;  // the explicit stages' membrane-rate refusal (IMEX only)
  const int all = nc + ne + nw;
  auto rates = [&](const double* src, int vn, double* R, double* contrib) {
    if (ne) {
      hipLaunchKernelGGL(k_edges<false>, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, src, vn, ed0, ed1, c,
                         R, nullptr, nullptr, e->valid);
      mark("k_edges");
    }
    hipLaunchKernelGGL(k_cells<false>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, src, vn, c0, c1, c, R,
                       contrib, nullptr, nullptr, 0.0, nullptr, nullptr, rate_dt, nullptr, nullptr, e->valid);
    mark("k_cells");
    if (rk4)
      hipLaunchKernelGGL(k_rk4_world_rates, dim3(nw), dim3(kBlock), 0, st, P, w0, contrib, R, e->valid);
    else
      hipLaunchKernelGGL(k_n2_world_rates, dim3(nw), dim3(kBlock), 0, st, P, w0, contrib, R);
  };
  if (!rk4) {
    const double gamma = kArsGamma, delta = kArsDelta;
    rates(nullptr, vol_next, e->K1, e->contrib1);
    hipLaunchKernelGGL(k_rk_combine, dim3(blocks(all)), dim3(kBlock), 0, st, P, A, B, e->sel, vol_next, c0, nc, ed0,
                       ne, w0, nw, gamma * dt, e->K1, 0.0, nullptr, e->Y1, 0, 1, 0, e->valid);
    mark("k_rk_combine");
    rates(e->Y1, 0, e->K2, e->contrib2);
    hipLaunchKernelGGL(k_rk_combine, dim3(blocks(all)), dim3(kBlock), 0, st, P, A, B, e->sel, vol_next, c0, nc, ed0,
                       ne, w0, nw, delta * dt, e->K1, (1.0 - delta) * dt, e->K2, nullptr, 0, 1, 1, e->valid);
    mark("k_rk_combine");
    hipLaunchKernelGGL(k_ars_tail, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, w0, dt, e->K1, e->K2, e->valid);
    mark("k_ars_tail");
    return;
  }
  double* k[4] = {e->K1, e->K2, e->K3, e->K4};
  double* contrib[2] = {e->contrib1, e->contrib2};
  const double factor[4] = {0.0, 0.5, 0.5, 1.0};
  for (int j = 0; j < 4; ++j) {
    if (j == 0) {
      rates(nullptr, 0, k[0], contrib[0]);
      continue;
    }
    hipLaunchKernelGGL(k_rk_combine, dim3(blocks(all)), dim3(kBlock), 0, st, P, A, B, e->sel, 0, c0, nc, ed0, ne,
                       w0, nw, factor[j] * dt, k[j - 1], 0.0, nullptr, e->Y1, 1, 0, 1, e->valid);
    mark("k_rk_combine");
    rates(e->Y1, 0, k[j], contrib[j % 2]);
  }
  hipLaunchKernelGGL(k_rk4_final, dim3(blocks(all)), dim3(kBlock), 0, st, P, A, B, e->sel, c0, nc, ed0, ne, w0, nw,
                     dt, k[0], k[1], k[2], k[3], e->valid);
  mark("k_rk4_final");
}

// A, B: the state buffer pair (a world's current state is A or B by sel[w]; the step writes every
// field of the other).
int launch_step(cells_ensemble* e, hipStream_t st, double dt, const Ctl& c, int world, double* A, double* B,
                KernelTimer* tm = nullptr) {
  auto mark = [&](const char* name) {
    if (tm) tm->mark(name);
  };
  const Params& P = e->P;
  int w0 = world < 0 ? 0 : world;
  int nw = world < 0 ? e->W : 1;
  if (e->checks_on && world <= 0) hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, st, e->step_counter);
  int c0 = (int)e->cell_off[w0], c1 = (int)e->cell_off[w0 + nw];
  int ed0 = (int)e->edge_off[w0], ed1 = (int)e->edge_off[w0 + nw];
  int nc = c1 - c0, ne = ed1 - ed0;
  if (P.water) {
    hipLaunchKernelGGL(k_water, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, w0, 0, 0.5 * dt, c, e->valid);
    mark("k_water");
  }
  int vol_next = P.water ? 1 : 0;
  if (e->method == CELLS_METHOD_EULER) {
    if (ne) {
      hipLaunchKernelGGL(k_edges<false>, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, nullptr, vol_next,
                         ed0, ed1, c, e->K1, nullptr, nullptr, e->valid);
      mark("k_edges");
    }
    hipLaunchKernelGGL(k_cells<false>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, nullptr, vol_next,
                       c0, c1, c, e->K1, e->contrib1, nullptr, nullptr, 0.0, nullptr, nullptr, dt, nullptr, nullptr,
                       e->valid);
    mark("k_cells");
    hipLaunchKernelGGL(k_euler_update, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, A, B, e->sel, c0, nc,
                       ed0, ne, dt, e->K1, e->valid);
    mark("k_euler_update");
    hipLaunchKernelGGL(k_world_ionic, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, w0, e->contrib1,
                       1.0 * dt, nullptr, 0.0, e->valid);
    mark("k_world_ionic");
  } else if (e->method == CELLS_METHOD_N2) {
    launch_n2(e, st, dt, c, w0, nw, A, B, mark);
  } else if (e->method == CELLS_METHOD_ARS222 || e->method == CELLS_METHOD_RK4) {
    launch_explicit(e, st, dt, c, w0, nw, A, B, mark);
  } else {
    const double gamma = 1.0 + 1.0 / std::sqrt(2.0);
    const double coef = gamma * dt;
    size_t lds = sizeof(double) * (size_t)std::max(e->max_n, 1);
    if (ne) {
      hipLaunchKernelGGL(k_edges<true>, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, nullptr, vol_next,
                         ed0, ed1, c, e->K1, e->gj, e->Gj, e->valid);
      mark("k_edges");
    }
    // the stage solves' right-hand sides written by the rate kernels into the solve vector (with
    // the vector in HBM; the reference-order mode's small worlds keep it in LDS and gather it)
    double* yrhs = e->y_in_lds ? nullptr : e->ysc;
    hipLaunchKernelGGL(k_cells<true>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, nullptr, vol_next,
                       c0, c1, c, e->K1, e->contrib1, e->gm, e->Gj, coef, e->mdiag, e->msc, 0.0, yrhs, nullptr,
                       e->valid);
    mark("k_cells");
    launch_factor(e, st, w0, nw, coef);
    mark("k_factor");
    solve_stage(e, st, nw, e->fvals, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_solve_stage<false> : k_solve_stage<true>), dim3(nw), solve_block(e), solve_lds_bytes(e), st, P, e->shapes, e->shape_of,
                       e->pool, e->foff, w0, e->msc, e->fvals, e->K1, nullptr, e->dphi, e->ysc, e->y_in_lds, solve_args(e), part,
                       yrhs ? 1 : 0);
  });
    mark("k_solve_stage");
    hipLaunchKernelGGL(k_ros2_stage1<false>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, c0, nc,
                       ed0, ne, dt, coef, e->gm, e->gj, e->Gj, e->dphi, e->res_atol, e->res_rtol, e->K1,
                       e->contrib1, e->Y1, e->valid);
    mark("k_ros2_stage1");
    if (ne) {
      hipLaunchKernelGGL(k_edges_ahead, dim3(blocks(ne)), dim3(kBlock), 0, st, P, A, B, e->sel, e->Y1, ed0, ed1,
                         c, e->K2, e->K1, dt, e->valid);
      mark("k_edges");
    }
    // stage 2 reads the stage-1 Jacobi scales (msc) for its right-hand side; it writes neither msc
    // nor mdiag (no slopes)
    hipLaunchKernelGGL(k_cells_ahead, dim3(blocks(nc)), dim3(kBlock), 0, st, P, A, B, e->sel, e->Y1, c0, c1, c,
                       e->K2, e->contrib2, e->msc, yrhs, e->K1, dt, e->valid);
    mark("k_cells");
    solve_stage(e, st, nw, e->fvals, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_solve_stage<false> : k_solve_stage<true>), dim3(nw), solve_block(e), solve_lds_bytes(e), st, P, e->shapes, e->shape_of,
                       e->pool, e->foff, w0, e->msc, e->fvals, e->K2, e->K1, e->dphi2, e->ysc, e->y_in_lds, solve_args(e), part,
                       yrhs ? 1 : 0);
  });
    mark("k_solve_stage");
    hipLaunchKernelGGL(k_ros2_final, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, A, B, e->sel, c0, nc,
                       ed0, ne, dt, coef, e->gm, e->gj, e->Gj, e->dphi2, e->res_atol, e->res_rtol, e->K1, e->K2,
                       e->contrib1, e->contrib2, e->dphi, e->valid);
    mark("k_ros2_final");
    hipLaunchKernelGGL(k_world_ionic, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, w0, e->contrib1,
                       1.5 * dt, e->contrib2, 0.5 * dt, e->valid);
    mark("k_world_ionic");
  }
  if (P.water) {
    hipLaunchKernelGGL(k_water, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, w0, 1, 0.5 * dt, c, e->valid);
    mark("k_water");
  }
  if (e->checks_on) {
    hipLaunchKernelGGL(k_step_checks, dim3(nw), dim3(kBlock), 0, st, P, A, B, e->sel, e->alive, e->valid,
                       e->step_counter, e->check_atol, w0, e->check_bits);
    mark("k_step_checks");
  }
  hipLaunchKernelGGL(k_commit, dim3(blocks(nw)), dim3(kBlock), 0, st, w0, nw, e->sel, e->alive, e->valid);
  mark("k_commit");
  return CELLS_OK;
}

}  // namespace

extern "C" int cells_step(cells_ensemble* e, void* stream, const double* y_in, double* y_out,
                          double dt, int32_t nsteps, const cells_control* control,
                          const int32_t* active, int32_t* ok_out, int32_t batched) {
  if (!e || !y_in || !y_out || !control) return fail(CELLS_EINVAL, "NULL argument");
  if (nsteps < 0) return fail(CELLS_EINVAL, "negative step count");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CELLS_EINVAL, "dt must be positive and finite");
  if (int rc = check_step_size(e, dt)) return rc;
  HIP_TRY(hipSetDevice(e->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Ctl c{control->membrane, control->junction, control->pump, control->calcium_pump, control->water};
  const Params& P = e->P;
  // The caller's output buffer is the pair's first state buffer (in place when y_out == y_in),
  // the ensemble's own the second: every step writes every field of a world's next state, so the
  // second needs no copy, and at the end only the worlds whose current state is the second one
  // are copied out.
  double* A = y_out;
  double* B = e->B;
  if (y_out != y_in) HIP_TRY(hipMemcpyAsync(A, y_in, sizeof(double) * P.total, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_init, dim3(blocks(e->W)), dim3(kBlock), 0, st, e->W, active, e->sel, e->alive, e->valid);
  HIP_TRY(hipMemsetAsync(e->check_bits, 0, sizeof(int) * e->W, st));
  if (e->checks_on) {
    HIP_TRY(hipMemcpyAsync(e->step_counter, &e->check_steps, sizeof(long long), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));  // the counter's host copy is read before it advances
    e->check_steps += nsteps;
  }
  HIP_TRY(hipGetLastError());
  if (batched) {
    cells_ensemble::GraphKey key;
    key.dt = dt;
    key.ctl = c;
    key.checks = e->checks_on;
    key.check_atol = e->checks_on ? e->check_atol : 0.0;
    key.res_atol = e->res_atol;
    key.res_rtol = e->res_rtol;
    key.A = A;
    auto& cache = e->graphs;
    size_t hit = cache.size();
    for (size_t q = 0; q < cache.size(); ++q)
      if (cache[q].first == key) hit = q;
    if (hit == cache.size() && nsteps > 0) {
      hipGraph_t g = nullptr;
      hipGraphExec_t exec = nullptr;
      HIP_TRY(hipStreamBeginCapture(e->capture_stream, hipStreamCaptureModeThreadLocal));
      launch_step(e, e->capture_stream, dt, c, -1, A, B);
      hipError_t cap = hipStreamEndCapture(e->capture_stream, &g);
      if (cap != hipSuccess) return fail(CELLS_EHIP, std::string("step capture: ") + hipGetErrorString(cap));
      hipError_t inst = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (inst != hipSuccess) return fail(CELLS_EHIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
      if (cache.size() == cells_ensemble::kGraphs) {  // the least recently used one goes
        HIP_TRY(hipDeviceSynchronize());  // it may still run on some stream
        HIP_TRY(hipGraphExecDestroy(cache.back().second));
        cache.pop_back();
      }
      cache.insert(cache.begin(), {key, exec});
      hit = 0;
    } else if (hit < cache.size() && hit > 0) {
      std::rotate(cache.begin(), cache.begin() + hit, cache.begin() + hit + 1);
      hit = 0;
    }
    if (nsteps > 0)
      for (int k = 0; k < nsteps; ++k) HIP_TRY(hipGraphLaunch(cache[hit].second, st));
  } else {
    for (int k = 0; k < nsteps; ++k)
      for (int w = 0; w < e->W; ++w) launch_step(e, st, dt, c, w, A, B);
  }
  HIP_TRY(hipGetLastError());
  long items = (long)P.ncell + P.nedge + P.W;
  hipLaunchKernelGGL(k_gather, dim3(blocks(items)), dim3(kBlock), 0, st, P, A, B, e->sel, y_out, e->alive,
                     ok_out);
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_kernel_times(cells_ensemble* e, void* stream, const double* y, double dt,
                                  const cells_control* control, int32_t reps, int32_t capacity,
                                  double* ms_out, char* names_out, int32_t* count_out) {
  if (!e || !y || !control || !ms_out || !names_out || !count_out || reps < 1)
    return fail(CELLS_EINVAL, "bad argument");
  HIP_TRY(hipSetDevice(e->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Ctl c{control->membrane, control->junction, control->pump, control->calcium_pump, control->water};
  const Params& P = e->P;
  HIP_TRY(hipMemcpyAsync(e->A, y, sizeof(double) * P.total, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(e->B, y, sizeof(double) * P.total, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_init, dim3(blocks(e->W)), dim3(kBlock), 0, st, e->W, nullptr, e->sel, e->alive, e->valid);
  std::vector<double> total;
  std::vector<const char*> names;
  for (int r = 0; r < reps; ++r) {
    KernelTimer tm;
    tm.stream = st;
    tm.mark("start");
    launch_step(e, st, dt, c, -1, e->A, e->B, &tm);
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    if (r == 0) {
      names.assign(tm.names.begin() + 1, tm.names.end());
      total.assign(names.size(), 0.0);
    }
    for (size_t k = 1; k < tm.events.size(); ++k) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, tm.events[k - 1], tm.events[k]));
      total[k - 1] += ms;
    }
  }
  int n = (int)std::min<size_t>(names.size(), (size_t)capacity);
  for (int k = 0; k < n; ++k) {
    ms_out[k] = total[k] / reps;
    std::snprintf(names_out + 32 * k, 32, "%s", names[k]);
  }
  *count_out = n;
  return CELLS_OK;
}

extern "C" int cells_rates(cells_ensemble* e, void* stream, const double* y,
                           const cells_control* control, double* rates_out, int32_t* valid_out) {
  if (!e || !y || !rates_out || !control) return fail(CELLS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(e->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Ctl c{control->membrane, control->junction, control->pump, control->calcium_pump, control->water};
  const Params& P = e->P;
  HIP_TRY(hipMemcpyAsync(e->A, y, sizeof(double) * P.total, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_init, dim3(blocks(e->W)), dim3(kBlock), 0, st, e->W, nullptr, e->sel, e->alive, e->valid);
  if (e->nedge)
    hipLaunchKernelGGL(k_edges<false>, dim3(blocks(e->nedge)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, nullptr, 0, 0,
                       e->nedge, c, rates_out, nullptr, nullptr, e->valid);
  hipLaunchKernelGGL(k_cells<false>, dim3(blocks(e->ncell)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, nullptr, 0, 0,
                     e->ncell, c, rates_out, e->contrib1, nullptr, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr,
                     nullptr, e->valid);
  hipLaunchKernelGGL(k_world_rates, dim3(e->W), dim3(kBlock), 0, st, P, e->contrib1, rates_out);
  if (valid_out) HIP_TRY(hipMemcpyAsync(valid_out, e->valid, sizeof(int) * e->W, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_energy_rates(cells_ensemble* e, void* stream, const double* y, const cells_control* control,
                                  double* out) {
  if (!e || !y || !control || !out) return fail(CELLS_EINVAL, "NULL argument");
  if (int rc = cells_rates(e, stream, y, control, e->K1, nullptr)) return rc;
  hipLaunchKernelGGL(k_energy_rates, dim3(e->W), dim3(kBlock), 0, static_cast<hipStream_t>(stream), e->P, e->A,
                     e->K1, out);
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_factor(cells_ensemble* e, void* stream, double* values) {
  if (!e || !values) return fail(CELLS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(e->device));
  if (e->pairwise)
    hipLaunchKernelGGL(k_factor_only, dim3(e->W), dim3(kFactorBlock), e->factor_lds, static_cast<hipStream_t>(stream),
                       e->shapes, e->shape_of, e->pool, e->masks, e->foff, values);
  else
    hipLaunchKernelGGL(k_factor_only_mf, dim3(e->W), dim3(kMfThreads), sizeof(double) * kMfBlockLds,
                       static_cast<hipStream_t>(stream), e->shapes, e->shape_of, e->pool, e->foff, e->uoffw,
                       values, e->ubuf, e->bigbuf, e->boffw);
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_solve(cells_ensemble* e, void* stream, const double* values, const double* rhs, double* x) {
  if (!e || !values || !rhs || !x) return fail(CELLS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(e->device));
  size_t lds = sizeof(double) * (size_t)std::max(e->max_n, 1);
  solve_stage(e, static_cast<hipStream_t>(stream), e->W, values, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_solve_only<false> : k_solve_only<true>), dim3(e->W), solve_block(e), solve_lds_bytes(e), static_cast<hipStream_t>(stream), e->shapes,
                     e->shape_of, e->pool, e->foff, e->P.cell_off, values, rhs, x, e->ysc, e->y_in_lds, solve_args(e), part);
  });
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_device_checks(cells_ensemble* e, void* stream, const double* previous,
                                   const double* y, double steps, double material_atol,
                                   int32_t* failed_out, double* ratios_out) {
  if (!e || !previous || !y || !failed_out || !ratios_out) return fail(CELLS_EINVAL, "NULL argument");
  if (!(steps >= 0) || !(material_atol >= 0)) return fail(CELLS_EINVAL, "steps and tolerance must be >= 0");
  HIP_TRY(hipSetDevice(e->device));
  hipLaunchKernelGGL(k_device_checks, dim3(e->W), dim3(kBlock), 0, static_cast<hipStream_t>(stream), e->P,
                     previous, y, steps, material_atol, failed_out, ratios_out);
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_step_checks(cells_ensemble* e, int32_t enabled, double material_atol, int64_t steps_before) {
  if (!e) return fail(CELLS_EINVAL, "NULL argument");
  if (!(material_atol >= 0) || steps_before < 0) return fail(CELLS_EINVAL, "tolerance and step count must be >= 0");
  HIP_TRY(hipSetDevice(e->device));
  e->checks_on = enabled != 0;
  e->check_atol = material_atol;
  e->check_steps = steps_before;
  return CELLS_OK;
}

extern "C" int cells_step_bounds(const cells_ensemble* e, double* rates_out) {
  if (!e || !rates_out) return fail(CELLS_EINVAL, "NULL argument");
  rates_out[0] = e->bound_explicit;
  rates_out[1] = e->bound_gate;
  rates_out[2] = e->bound_calcium;
  return CELLS_OK;
}

extern "C" int cells_set_stage_tolerance(cells_ensemble* e, double atol, double rtol) {
  if (!e) return fail(CELLS_EINVAL, "NULL argument");
  if (!(atol >= 0) || !(rtol >= 0) || !std::isfinite(atol) || !std::isfinite(rtol))
    return fail(CELLS_EINVAL, "tolerances must be finite and >= 0");
  e->res_atol = atol;
  e->res_rtol = rtol;
  return CELLS_OK;
}

extern "C" int cells_step_check_bits(cells_ensemble* e, void* stream, int32_t* bits_out) {
  if (!e || !bits_out) return fail(CELLS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipMemcpyAsync(bits_out, e->check_bits, sizeof(int) * e->W, hipMemcpyDeviceToDevice,
                         static_cast<hipStream_t>(stream)));
  return CELLS_OK;
}

namespace {

// The adjoint's workspace: the recomputed stage increments and the reverse sweep's cotangents.
// Every call writes each of these doubles before it reads it, except gmbar and gjbar, which the
// reverse sweep clears itself (k_fill), so no call depends on what an earlier one left.
struct AdjointWork {
  double *d1, *d2, *k1bar, *k2bar, *k1gbar, *g2bar, *gmbar, *gjbar, *dbar, *wv, *ca_bar, *cb_bar, *ev_bar;
};

// Lays the workspace of an ensemble out from `base` (each buffer on a 256-byte boundary); returns
// the doubles it spans. base may be NULL to ask for the size alone.
size_t adjoint_layout(const cells_ensemble* e, double* base, AdjointWork* w) {
  const size_t nc = e->ncell, ne = std::max(e->nedge, 1), S = e->S;
  size_t off = 0;
  auto take = [&](size_t count) {
    double* q = base ? base + off : nullptr;
    off += (count + 31) & ~(size_t)31;
    return q;
  };
  AdjointWork a;
  a.d1 = take(nc);
  a.d2 = take(nc);
  a.k1bar = take(S * nc);
  a.k2bar = take(S * nc);
  a.k1gbar = take(ne);
  a.g2bar = take(ne);
  a.gmbar = take(S * nc);
  a.gjbar = take(S * ne);
  a.dbar = take(nc);
  a.wv = take(nc);
  a.ca_bar = take(S * ne);
  a.cb_bar = take(S * ne);
  a.ev_bar = take(ne);
  if (w) *w = a;
  return off;
}

// The ensemble's own workspace, allocated at its first adjoint call.
int adjoint_workspace(cells_ensemble* e, AdjointWork* w) {
  if (!e->adj_base) {
    int rc = alloc(e, &e->adj_base, adjoint_layout(e, nullptr, nullptr));
    if (rc) return rc;
  }
  adjoint_layout(e, e->adj_base, w);
  return CELLS_OK;
}

// What the reverse sweep of one step reads: the step's input state, its stage-1 state, the factor
// of its stage matrix with the Jacobi scales, the route slopes at the input and both stage
// increments. Either the ensemble's own buffers after a recomputation or a tape slot.
struct StepRecord {
  const double *y, *y1, *fvals, *msc, *gm, *gj, *d1, *d2;
};

int check_adjoint(const cells_ensemble* e, double dt) {
  if (e->method != CELLS_METHOD_ROS2 || e->P.water || e->P.calcium || e->S != 3)
    return fail(CELLS_EINVAL, "the adjoint covers ROS2 on 3-ion worlds without water");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CELLS_EINVAL, "dt must be positive and finite");
  return CELLS_OK;
}

// ybar_in (+)= the pullback of ybar_out through the step recorded in r; parameter cotangents
// accumulate into pam_bar, cap_bar, paj_bar.
int reverse_sweep(cells_ensemble* e, hipStream_t st, const StepRecord& r, const AdjointWork& a,
                  const double* ybar_out, double dt, const Ctl& c, double* ybar_in, double* pam_bar,
                  double* cap_bar, double* paj_bar) {
  const Params& P = e->P;
  const double coef = (1.0 + 1.0 / std::sqrt(2.0)) * dt;
  const int nc = e->ncell, ne = e->nedge, W = e->W;
  long S = e->S;
  double* d2 = const_cast<double*>(r.d2);
  double* d1 = const_cast<double*>(r.d1);
  hipLaunchKernelGGL(k_adj_seed, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, ybar_out, dt, ybar_in, a.k1bar,
                     a.k2bar, a.k1gbar, a.g2bar);
  hipLaunchKernelGGL(k_fill, dim3(blocks(S * nc)), dim3(kBlock), 0, st, S * nc, a.gmbar, 0.0);
  if (ne) hipLaunchKernelGGL(k_fill, dim3(blocks(S * ne)), dim3(kBlock), 0, st, S * ne, a.gjbar, 0.0);
  // k2 = r2 + coef D delta2;  delta2 = M^-1 q(r2)
  hipLaunchKernelGGL(k_adj_dtrans, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, coef, r.gm, r.gj, a.k2bar,
                     r.d2, a.dbar, a.gmbar, a.gjbar);
  solve_stage(e, st, W, r.fvals, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_adj_solve<false> : k_adj_solve<true>), dim3(W), solve_block(e), solve_lds_bytes(e), st, P, e->shapes, e->shape_of, e->pool,
                     e->foff, coef, r.msc, r.fvals, a.dbar, d2, a.wv, a.k2bar, a.gmbar, a.gjbar, e->ysc,
                     e->y_in_lds, solve_args(e), part);
  });
  // r2 = f(Y1) - 2 k1
  hipLaunchKernelGGL(k_adj_axpy, dim3(blocks(S * nc + ne)), dim3(kBlock), 0, st, S * nc, a.k1bar, a.k2bar,
                     (long)ne, a.k1gbar, a.g2bar);
  double* gate_in = ybar_in + P.edge_base + (long)P.f_gate * P.nedge;
  if (ne)
    hipLaunchKernelGGL(k_edges_vjp, dim3(blocks(ne)), dim3(kBlock), 0, st, P, r.y1, c, a.k2bar, a.g2bar,
                       nullptr, a.ca_bar, a.cb_bar, a.ev_bar, gate_in, a.k1gbar, dt, paj_bar);
  hipLaunchKernelGGL(k_cells_vjp, dim3(blocks(nc)), dim3(kBlock), 0, st, P, r.y1, c, a.k2bar, nullptr,
                     a.ca_bar, a.cb_bar, a.ev_bar, ybar_in, a.k1bar, dt, pam_bar, cap_bar);
  // k1 = f(y) + coef D delta1;  delta1 = M^-1 q(f(y))
  hipLaunchKernelGGL(k_adj_dtrans, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, coef, r.gm, r.gj, a.k1bar,
                     r.d1, a.dbar, a.gmbar, a.gjbar);
  solve_stage(e, st, W, r.fvals, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_adj_solve<false> : k_adj_solve<true>), dim3(W), solve_block(e), solve_lds_bytes(e), st, P, e->shapes, e->shape_of, e->pool,
                     e->foff, coef, r.msc, r.fvals, a.dbar, d1, a.wv, a.k1bar, a.gmbar, a.gjbar, e->ysc,
                     e->y_in_lds, solve_args(e), part);
  });
  // f(y) and the slopes at y
  if (ne)
    hipLaunchKernelGGL(k_edges_vjp, dim3(blocks(ne)), dim3(kBlock), 0, st, P, r.y, c, a.k1bar, a.k1gbar,
                       a.gjbar, a.ca_bar, a.cb_bar, a.ev_bar, gate_in, nullptr, 0.0, paj_bar);
  hipLaunchKernelGGL(k_cells_vjp, dim3(blocks(nc)), dim3(kBlock), 0, st, P, r.y, c, a.k1bar, a.gmbar,
                     a.ca_bar, a.cb_bar, a.ev_bar, ybar_in, nullptr, 0.0, pam_bar, cap_bar);
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

void free_tape(cells_ensemble* e) {
  for (void* p : e->tape_allocations) (void)hipFree(p);
  e->tape_allocations.clear();
  e->tape_slots = 0;
  e->t_y = e->t_y1 = e->t_fv = e->t_msc = e->t_gm = e->t_gj = e->t_d1 = e->t_d2 = nullptr;
  e->t_valid = nullptr;
}

template <typename T>
int tape_alloc(cells_ensemble* e, T** dst, size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(CELLS_ENOMEM, "hipMalloc of the step tape failed");
  }
  e->tape_allocations.push_back(p);
  *dst = static_cast<T*>(p);
  return CELLS_OK;
}

// The reverse of one ROS2 step at y with its forward internals recomputed, on workspace a.
int step_vjp(cells_ensemble* e, void* stream, const double* y, const double* ybar_out, double dt,
             const cells_control* control, double* ybar_in, double* pam_bar, double* cap_bar,
             double* paj_bar, int32_t* valid_out, const AdjointWork& a) {
  int rc = CELLS_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Ctl c{control->membrane, control->junction, control->pump, control->calcium_pump, control->water};
  const Params& P = e->P;
  const double coef = (1.0 + 1.0 / std::sqrt(2.0)) * dt;
  const int nc = e->ncell, ne = e->nedge, W = e->W;
  HIP_TRY(hipMemcpyAsync(e->A, y, sizeof(double) * P.total, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_init, dim3(blocks(W)), dim3(kBlock), 0, st, W, nullptr, e->sel, e->alive, e->valid);
  // forward recomputation of the step's internals (the forward kernels, in the same order)
  if (ne)
    hipLaunchKernelGGL(k_edges<true>, dim3(blocks(ne)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, nullptr, 0, 0, ne,
                       c, e->K1, e->gj, e->Gj, e->valid);
  hipLaunchKernelGGL(k_cells<true>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, nullptr, 0, 0, nc, c,
                     e->K1, e->contrib1, e->gm, e->Gj, coef, e->mdiag, e->msc, 0.0, nullptr, nullptr, e->valid);
  launch_factor(e, st, 0, W, coef);
  solve_stage(e, st, W, e->fvals, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_solve_stage<false> : k_solve_stage<true>), dim3(W), solve_block(e), solve_lds_bytes(e), st, P, e->shapes, e->shape_of, e->pool,
                     e->foff, 0, e->msc, e->fvals, e->K1, nullptr, a.d1, e->ysc, e->y_in_lds, solve_args(e), part, 0);
  });
  hipLaunchKernelGGL(k_ros2_stage1<true>, dim3(blocks(nc + ne)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, 0, nc, 0,
                     ne, dt, coef, e->gm, e->gj, e->Gj, a.d1, e->res_atol, e->res_rtol, e->K1, e->contrib1, e->Y1,
                     e->valid);
  if (ne)
    hipLaunchKernelGGL(k_edges<false>, dim3(blocks(ne)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, e->Y1, 0, 0, ne, c,
                       e->K2, nullptr, nullptr, e->valid);
  hipLaunchKernelGGL(k_cells<false>, dim3(blocks(nc)), dim3(kBlock), 0, st, P, e->A, e->B, e->sel, e->Y1, 0, 0, nc, c,
                     e->K2, e->contrib2, nullptr, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, e->valid);
  solve_stage(e, st, W, e->fvals, [&](int part) {
    hipLaunchKernelGGL((e->pairwise ? k_solve_stage<false> : k_solve_stage<true>), dim3(W), solve_block(e), solve_lds_bytes(e), st, P, e->shapes, e->shape_of, e->pool,
                     e->foff, 0, e->msc, e->fvals, e->K2, e->K1, a.d2, e->ysc, e->y_in_lds, solve_args(e), part, 0);
  });
  StepRecord r{e->A, e->Y1, e->fvals, e->msc, e->gm, e->gj, a.d1, a.d2};
  if ((rc = reverse_sweep(e, st, r, a, ybar_out, dt, c, ybar_in, pam_bar, cap_bar, paj_bar))) return rc;
  if (valid_out) HIP_TRY(hipMemcpyAsync(valid_out, e->valid, sizeof(int) * W, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

}  // namespace

extern "C" int cells_step_vjp(cells_ensemble* e, void* stream, const double* y, const double* ybar_out,
                              double dt, const cells_control* control, double* ybar_in,
                              double* pam_bar, double* cap_bar, double* paj_bar, int32_t* valid_out) {
  if (!e || !y || !ybar_out || !control || !ybar_in || !pam_bar || !cap_bar || !paj_bar)
    return fail(CELLS_EINVAL, "NULL argument");
  int rc = check_adjoint(e, dt);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(e->device));
  AdjointWork a;
  if ((rc = adjoint_workspace(e, &a))) return rc;
  return step_vjp(e, stream, y, ybar_out, dt, control, ybar_in, pam_bar, cap_bar, paj_bar, valid_out, a);
}

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kScratchDevices] = {};  // per device: the scratch array's address on it

// The scratch array's device address on `device` (current), found at the first call there.
int scratch_address(int device, double** out) {
  if (device < 0 || device >= kScratchDevices) return fail(CELLS_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr)
      return fail(CELLS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return CELLS_OK;
}

}  // namespace

extern "C" int cells_step_vjp_noalloc(cells_ensemble* e, void* stream, const double* y, const double* ybar_out,
                                      double dt, const cells_control* control, double* ybar_in,
                                      double* pam_bar, double* cap_bar, double* paj_bar, int32_t* valid_out) {
  if (!e || !y || !ybar_out || !control || !ybar_in || !pam_bar || !cap_bar || !paj_bar)
    return fail(CELLS_EINVAL, "NULL argument");
  int rc = check_adjoint(e, dt);
  if (rc) return rc;
  const size_t need = adjoint_layout(e, nullptr, nullptr);
  if (need > kScratchDoubles)
    return fail(CELLS_EINVAL, "cells_step_vjp_noalloc: the ensemble's adjoint workspace needs " +
                                  std::to_string(need) + " doubles, above the " + std::to_string(kScratchDoubles) +
                                  "-double (64 MiB) scratch; use cells_step_vjp");
  HIP_TRY(hipSetDevice(e->device));
  double* base = nullptr;
  if ((rc = scratch_address(e->device, &base))) return rc;
  AdjointWork a;
  adjoint_layout(e, base, &a);
  return step_vjp(e, stream, y, ybar_out, dt, control, ybar_in, pam_bar, cap_bar, paj_bar, valid_out, a);
}

extern "C" int cells_tape_reserve(cells_ensemble* e, int32_t slots, int64_t* bytes_out) {
  if (!e) return fail(CELLS_EINVAL, "NULL argument");
  if (slots < 0) return fail(CELLS_EINVAL, "negative slot count");
  if (slots > 0 && (e->method != CELLS_METHOD_ROS2 || e->P.water || e->P.calcium || e->S != 3))
    return fail(CELLS_EINVAL, "the adjoint covers ROS2 on 3-ion worlds without water");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipDeviceSynchronize());
  free_tape(e);
  const Params& P = e->P;
  size_t n = slots, nc = e->ncell, ne = std::max(e->nedge, 1), S = e->S;
  size_t bytes = n * sizeof(double) * (2 * (size_t)P.total + (size_t)e->factor_total + nc * (3 + S) + S * ne) +
                 n * sizeof(int) * e->W;
  if (bytes_out) *bytes_out = (int64_t)bytes;
  if (slots == 0) return CELLS_OK;
  int rc;
  if ((rc = tape_alloc(e, &e->t_y, n * P.total)) || (rc = tape_alloc(e, &e->t_y1, n * P.total)) ||
      (rc = tape_alloc(e, &e->t_fv, n * e->factor_total)) || (rc = tape_alloc(e, &e->t_msc, n * nc)) ||
      (rc = tape_alloc(e, &e->t_gm, n * S * nc)) || (rc = tape_alloc(e, &e->t_gj, n * S * ne)) ||
      (rc = tape_alloc(e, &e->t_d1, n * nc)) || (rc = tape_alloc(e, &e->t_d2, n * nc)) ||
      (rc = tape_alloc(e, &e->t_valid, n * e->W))) {
    free_tape(e);
    return rc;
  }
  e->tape_slots = slots;
  return CELLS_OK;
}

extern "C" int cells_step_record(cells_ensemble* e, void* stream, const double* y_in, double* y_out,
                                 double dt, const cells_control* control, int32_t slot, int32_t* ok_out) {
  if (!e || !y_in || !y_out || !control) return fail(CELLS_EINVAL, "NULL argument");
  int rc = check_adjoint(e, dt);
  if (rc) return rc;
  if (slot < 0 || slot >= e->tape_slots) return fail(CELLS_EINVAL, "tape slot out of range");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CELLS_EINVAL, "dt must be positive and finite");
  if ((rc = check_step_size(e, dt))) return rc;
  HIP_TRY(hipSetDevice(e->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Params& P = e->P;
  size_t q = slot, nc = e->ncell, ne = std::max(e->nedge, 1), S = e->S, tot = P.total;
  auto copy = [&](double* dst, const double* src, size_t count) {
    return hipMemcpyAsync(dst + q * count, src, sizeof(double) * count, hipMemcpyDeviceToDevice, st);
  };
  HIP_TRY(copy(e->t_y, y_in, tot));  // before the step, which may run in place (y_out == y_in)
  rc = cells_step(e, stream, y_in, y_out, dt, 1, control, nullptr, ok_out, 1);
  if (rc) return rc;
  HIP_TRY(copy(e->t_y1, e->Y1, tot));
  HIP_TRY(copy(e->t_fv, e->fvals, e->factor_total));
  HIP_TRY(copy(e->t_msc, e->msc, nc));
  HIP_TRY(copy(e->t_gm, e->gm, S * nc));
  HIP_TRY(copy(e->t_gj, e->gj, S * ne));
  HIP_TRY(copy(e->t_d1, e->dphi, nc));
  HIP_TRY(copy(e->t_d2, e->dphi2, nc));
  HIP_TRY(hipMemcpyAsync(e->t_valid + q * e->W, e->valid, sizeof(int) * e->W, hipMemcpyDeviceToDevice, st));
  return CELLS_OK;
}

extern "C" int cells_step_vjp_recorded(cells_ensemble* e, void* stream, int32_t slot, const double* ybar_out,
                                       double dt, const cells_control* control, double* ybar_in,
                                       double* pam_bar, double* cap_bar, double* paj_bar, int32_t* valid_out) {
  if (!e || !ybar_out || !control || !ybar_in || !pam_bar || !cap_bar || !paj_bar)
    return fail(CELLS_EINVAL, "NULL argument");
  int rc = check_adjoint(e, dt);
  if (rc) return rc;
  if (slot < 0 || slot >= e->tape_slots) return fail(CELLS_EINVAL, "tape slot out of range");
  HIP_TRY(hipSetDevice(e->device));
  AdjointWork a;
  if ((rc = adjoint_workspace(e, &a))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Ctl c{control->membrane, control->junction, control->pump, control->calcium_pump, control->water};
  const Params& P = e->P;
  size_t q = slot, nc = e->ncell, ne = std::max(e->nedge, 1), S = e->S, tot = P.total;
  StepRecord r{e->t_y + q * tot, e->t_y1 + q * tot, e->t_fv + q * e->factor_total, e->t_msc + q * nc,
               e->t_gm + q * S * nc, e->t_gj + q * S * ne, e->t_d1 + q * nc, e->t_d2 + q * nc};
  if ((rc = reverse_sweep(e, st, r, a, ybar_out, dt, c, ybar_in, pam_bar, cap_bar, paj_bar))) return rc;
  if (valid_out)
    HIP_TRY(hipMemcpyAsync(valid_out, e->t_valid + q * e->W, sizeof(int) * e->W, hipMemcpyDeviceToDevice, st));
  return CELLS_OK;
}

extern "C" int cells_vmem_loss(cells_ensemble* e, void* stream, const double* y, const double* target,
                               double weight, double* loss_out, double* ybar) {
  if (!e || !y || !target || !loss_out) return fail(CELLS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(e->device));
  hipLaunchKernelGGL(k_vmem_loss, dim3(e->W), dim3(kBlock), 0, static_cast<hipStream_t>(stream), e->P, y, target,
                     weight, loss_out, ybar);
  HIP_TRY(hipGetLastError());
  return CELLS_OK;
}

extern "C" int cells_set_parameters(cells_ensemble* e, void* stream, const double* pam, const double* nak_cap,
                                    const double* paj) {
  if (!e) return fail(CELLS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(e->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  size_t nc = e->ncell, ne = e->nedge, S = e->S;
  if (pam) HIP_TRY(hipMemcpyAsync(const_cast<double*>(e->P.pam), pam, sizeof(double) * S * nc, hipMemcpyDefault, st));
  if (nak_cap) HIP_TRY(hipMemcpyAsync(const_cast<double*>(e->P.nak_cap), nak_cap, sizeof(double) * nc, hipMemcpyDefault, st));
  if (paj && ne) HIP_TRY(hipMemcpyAsync(const_cast<double*>(e->P.paj), paj, sizeof(double) * S * ne, hipMemcpyDefault, st));
  return CELLS_OK;
}

extern "C" int cells_plan_stats(const cells_ensemble* e, int64_t* nnz_total, double* flops_total,
                                int32_t* max_levels, int32_t* shapes) {
  if (!e) return fail(CELLS_EINVAL, "NULL argument");
  int64_t nnz = 0;
  double flops = 0.0;
  int32_t lev = 0;
  for (size_t q = 0; q < e->shapes_host.size(); ++q) lev = std::max(lev, e->shapes_host[q].levels());
  for (int w = 0; w < e->W; ++w) {
    int q = e->shape_of_host[w];
    nnz += e->exact_nnz_host[q];
    flops += e->exact_flops_host[q];
  }
  if (nnz_total) *nnz_total = nnz;
  if (flops_total) *flops_total = flops;
  if (max_levels) *max_levels = lev;
  if (shapes) *shapes = (int32_t)e->shapes_host.size();
  return CELLS_OK;
}

#include "stationary.inc"
#include "field.inc"

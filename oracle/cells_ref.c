/* Oracle (test infrastructure only): FP64 CPU restatement of the reference's tissue step for
 * lattices of cells on one held bath. Only tests, the smoke check and the benchmark's checker legs
 * (its CPU baseline and its parity checks, outside every timed region) load it; the product never
 * does.
 *
 * Reference files followed (paths under src/cells/ of the reference):
 *   laws/electrochemistry.py:11-48            F, R, bernoulli, electrochemical potential
 *   laws/junctions/selective.py:28-50         selective_flux (GHK route flux)
 *   laws/routes/patch_electrodiffusion.py:43-54  bernoulli_slope
 *   models/surface_pump_v3.py:2747-2755       route_slope (dJ/d dv)
 *   models/surface_pump_v3.py:2816-2837       conductance_laplacian (G = F' sum_s z_s slope_s)
 *   laws/pumps/nak_atp_mass_action_positive.py:55-149  Na/K-ATPase kinetics and domain
 *   laws/pumps/ca_atp_mass_action_positive.py:70-150, laws/reactions/er_calcium.py:30-50
 *                                             Ca-ATPase kinetics and domain
 *   laws/junctions/junction_voltage_relaxation.py:41-97  junction gate kinetics, rhs, domain
 *   numerics/surface_transport.py:298-386     transport: route fluxes, mu, dissipation, owners
 *   models/surface_pump_v2.py:370-450         explicit_rates with both pumps
 *   models/surface_pump.py:259-362            _combine, node_charge, electrode_increment, imex_step
 *   numerics/surface_imex.py:47-60,207-278    Euler tableau; Jacobi-scaled factor; solve_operator
 *   numerics/sparse_cholesky.py:474-514       factor / solve (pairwise folds, reciprocal pivots)
 *   operators/osmotic_water.py:127-221, numerics/osmotic_water.py:193-215  water half-steps
 *   models/pump_composite.py:1769-1815        Strang split: water half, ionic step, water half
 *   models/surface_tissue.py:269-420          charge_increment, capacitor_change, domain_valid
 * The ROS2 stage driver (Verwer, Spee, Blom, Hundsdorfer 1999, SIAM J. Sci. Comput. 20:1456,
 * gamma = 1 + 1/sqrt 2) is not the reference's: the reference has no Rosenbrock method. Its stage matrix is
 * C + gamma dt G with G the routes' slope-conductance Laplacian, as N2's operator
 * (surface_pump_v3.py:2758-2837).
 */
#include "cells_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define AVOGADRO 6.02214076e23
#define F_CONST (AVOGADRO * 1.602176634e-19)
#define R_CONST (AVOGADRO * 1.380649e-23)
#define MOL_PER_FMOL 1e-15
#define CHARGE_PER_FMOL (F_CONST * MOL_PER_FMOL)
#define MAX_ABS_X 50.0
#define NAK_MAX_AFFINITY 80.0
#define NAK_MAX_MU 1600.0
#define CA_MAX_AFFINITY 80.0
#define CA_MAX_VOLTAGE 0.15
#define GATE_MAX_VOLTAGE 0.120
#define STANDARD_STATE_mM 1000.0
#define WATER_MAX_RELAXATION 0.05
#define ROS2_GAMMA (1.0 + 1.0 / 1.4142135623730951)
#define ARS_GAMMA (1.0 - 1.0 / 1.4142135623730951)
#define ARS_DELTA (1.0 - 1.0 / (2.0 * ARS_GAMMA))

/* ---------------------------------------------------------------- point laws */

double ref_bernoulli(double x) {
    /* laws/electrochemistry.py:25-32 */
    if (fabs(x) < 1e-4) {
        double x2 = x * x;
        return 1.0 - x / 2.0 + x2 / 12.0 - (x2 * x2) / 720.0;
    }
    return x / expm1(x);
}

double ref_bernoulli_slope(double x) {
    /* laws/routes/patch_electrodiffusion.py:43-54 */
    if (fabs(x) < 1e-3) return -0.5 + x / 6.0 - (x * x * x) / 180.0;
    double em = exp(-fabs(x));
    if (x > 0) {
        double d = 1.0 - em;
        return (em * (1.0 - x) - em * em) / (d * d);
    }
    double d = em - 1.0;
    return (em - 1.0 - x * em) / (d * d);
}

double ref_bernoulli_derivative(double x) {
    /* numerics/tissue_field.py:186-190 (the field Newton's junction slope) */
    if (fabs(x) < 1e-4) return -0.5 + x / 6 - x * x * x / 180;
    double em1 = expm1(x);
    return (em1 - x * exp(x)) / (em1 * em1);
}

double ref_selective_flux(double c_from, double c_to, double dv, double z, double T, double pa,
                          double avail) {
    /* laws/junctions/selective.py:28-50 with residual 0, gate 1, multiplier avail. */
    double x = z * F_CONST * dv / (R_CONST * T);
    return pa * avail / MOL_PER_FMOL * (c_from * ref_bernoulli(-x) - c_to * ref_bernoulli(x));
}

double ref_route_slope(double c_from, double c_to, double dv, double z, double T, double pa,
                       double avail) {
    /* models/surface_pump_v3.py:2747-2755 */
    double k = z * F_CONST / (R_CONST * T);
    double x = k * dv;
    double pre = pa * avail / MOL_PER_FMOL;
    return pre * k * (-c_from * ref_bernoulli_slope(-x) - c_to * ref_bernoulli_slope(x));
}

double ref_logaddexp(double a, double b) {
    double m = a > b ? a : b;
    return m + log1p(exp(-fabs(a - b)));
}

static double nak_drive(const cells_world* w) {
    /* nak_atp_mass_action_positive.py:55-60 */
    return w->dg0 + R_CONST * w->temperature *
                        (log(w->adp) + log(w->pi) - log(w->atp) - log(STANDARD_STATE_mM));
}

void ref_nak_kinetics(const cells_world* w, double na_in, double k_out, double mu_na, double mu_k,
                      double cap, double avail, double* cycles, double* energy, double* drive) {
    /* nak_atp_mass_action_positive.py:63-98 */
    double T = w->temperature;
    double d = nak_drive(w);
    double e = d + (-3.0) * mu_na + 2.0 * mu_k;
    double a = e / (R_CONST * T);
    double ls = -ref_logaddexp(0.0, 3.0 * (log(w->km_na) - log(na_in))) -
                ref_logaddexp(0.0, 2.0 * (log(w->km_k) - log(k_out))) -
                ref_logaddexp(0.0, log(w->km_atp) - log(w->atp));
    int enabled = cap > 0 && avail > 0;
    double sc = enabled ? cap : 1.0, sa = enabled ? avail : 1.0;
    double mag = exp(log(sc) + log(sa) + ls + (a > 0.0 ? a : 0.0));
    double dep = a >= 0 ? expm1(-a) : -expm1(a);
    *cycles = enabled ? mag * dep : 0.0;
    *energy = e;
    *drive = d;
}

int ref_nak_domain(const cells_world* w, double na_in, double k_out, double mu_na, double mu_k,
                   double cap, double avail) {
    /* nak_atp_mass_action_positive.py:101-149 */
    double T = w->temperature;
    int ok = isfinite(na_in) && na_in > 0 && isfinite(k_out) && k_out > 0;
    ok = ok && T >= 200.0 && T <= 400.0;
    ok = ok && isfinite(cap) && cap >= 0 && isfinite(avail) && avail >= 0 && avail <= 1;
    double rt = R_CONST * T;
    double sn = mu_na / rt, sk = mu_k / rt;
    ok = ok && isfinite(sn) && fabs(sn) <= NAK_MAX_MU && isfinite(sk) && fabs(sk) <= NAK_MAX_MU;
    double aff = w->dg0 / rt + log(w->adp) + log(w->pi) - log(w->atp) - log(STANDARD_STATE_mM) +
                 (-3.0) * sn + 2.0 * sk;
    return ok && isfinite(aff) && fabs(aff) <= NAK_MAX_AFFINITY;
}

void ref_ca_kinetics(const cells_world* w, double ca_in, double ca_out, double v, double cap,
                     double avail, double* cycles, double* energy, double* drive) {
    /* ca_atp_mass_action_positive.py:70-99 -> plasma_calcium.pump_rates(v) ->
     * er_calcium.pump_rates(ci, co, -v) with delta_g0 - RT ln 1000 (er_calcium.py:30-50). */
    double T = w->temperature, rt = R_CONST * T;
    double dg0k = w->ca_dg0 - R_CONST * T * log(STANDARD_STATE_mM);
    double dk = dg0k + R_CONST * T * (log(w->adp) + log(w->pi) - log(w->atp));
    double e = dk + R_CONST * T * (log(ca_out) - log(ca_in)) + 2.0 * F_CONST * (-v);
    double aff = e / rt;
    double ls = -ref_logaddexp(0.0, log(w->ca_km_ca) - log(ca_in));
    ls -= ref_logaddexp(0.0, log(w->ca_km_atp) - log(w->atp));
    int enabled = cap > 0 && avail > 0;
    double mag = exp(log(enabled ? cap : 1.0) + log(enabled ? avail : 1.0) + ls +
                     (aff > 0.0 ? aff : 0.0));
    double dep = aff >= 0 ? expm1(-aff) : -expm1(aff);
    *cycles = enabled ? mag * dep : 0.0;
    *energy = e;
    *drive = w->ca_dg0 + R_CONST * T *
                             (log(w->adp) + log(w->pi) - log(w->atp) - log(STANDARD_STATE_mM));
}

static int ca_domain(const cells_world* w, double ca_in, double ca_out, double v, double cap,
                     double avail) {
    /* ca_atp_mass_action_positive.py:102-140 */
    double T = w->temperature, rt = R_CONST * T;
    int ok = isfinite(ca_in) && ca_in > 0 && isfinite(ca_out) && ca_out > 0;
    ok = ok && isfinite(cap) && cap >= 0 && isfinite(avail) && avail >= 0 && avail <= 1;
    ok = ok && isfinite(v) && fabs(v) <= CA_MAX_VOLTAGE;
    double aff = w->ca_dg0 / rt + log(w->adp) + log(w->pi) - log(w->atp) -
                 log(STANDARD_STATE_mM) + log(ca_out) - log(ca_in) - 2.0 * F_CONST * v / rt;
    return ok && isfinite(aff) && fabs(aff) <= CA_MAX_AFFINITY;
}

double ref_gate_rhs(double gate, double v, double threshold, double minimum) {
    /* junction_voltage_relaxation.py:41-59 */
    double shift = fabs(v) - threshold;
    double alpha = 1.3 * exp(-77.0 * shift);
    double log_beta = log(0.065) + 140.0 * shift;
    double beta = 20.0 * exp(-ref_logaddexp(0.0, -log_beta));
    return alpha * (1.0 - gate) - beta * (gate - minimum);
}

static int gate_domain(double gate, double v, double thr, double mn, double avail) {
    /* junction_voltage_relaxation.py:78-97 */
    return isfinite(thr) && thr >= 0 && thr <= GATE_MAX_VOLTAGE && isfinite(mn) && mn >= 0 &&
           mn <= 1 && isfinite(v) && fabs(v) <= GATE_MAX_VOLTAGE && isfinite(gate) &&
           gate >= mn && gate <= 1 && isfinite(avail) && avail >= 0 && avail <= 1;
}

/* ---------------------------------------------------------------- layout */

void ref_state_layout(const cells_world* w, cells_state_layout* L) {
    int32_t S = w->S, f = 0;
    memset(L, 0, sizeof(*L));
    L->dn = f; f += S;
    L->mtr = f; f += S;
    L->mdiss = f++;
    L->pcyc = f++;
    L->pdiss = f++;
    L->pwork = f++;
    L->ccyc = L->cdiss = L->cwork = -1;
    if (w->calcium) { L->ccyc = f++; L->cdiss = f++; L->cwork = f++; }
    L->dvol = L->wtr = L->wdiss = -1;
    if (w->water) { L->dvol = f++; L->wtr = f++; L->wdiss = f++; }
    L->cell_fields = f;
    f = 0;
    L->jtr = f; f += S;
    L->jdiss = f++;
    L->gate = f++;
    L->edge_fields = f;
    f = 0;
    L->sup = f; f += S;
    L->echg = f++;
    L->rwork = f++;
    L->ework = f++;
    L->wsup = L->wwork = -1;
    if (w->water) { L->wsup = f++; L->wwork = f++; }
    L->world_fields = f;
    L->cell_base = 0;
    L->edge_base = (int64_t)L->cell_fields * w->ncell;
    L->world_base = L->edge_base + (int64_t)L->edge_fields * w->nedge;
    L->total = L->world_base + (int64_t)L->world_fields * w->W;
}

#define CF(L, w, y, f, i) (y)[(L)->cell_base + (int64_t)(f) * (w)->ncell + (i)]
#define EF(L, w, y, f, e) (y)[(L)->edge_base + (int64_t)(f) * (w)->nedge + (e)]
#define WF(L, w, y, f, k) (y)[(L)->world_base + (int64_t)(f) * (w)->W + (k)]

static double cell_volume(const cells_world* w, const cells_state_layout* L, const double* y,
                          int64_t i) {
    return w->water ? w->vol0[i] + CF(L, w, y, L->dvol, i) : w->vol0[i];
}

static double cell_charge_increment(const cells_world* w, const cells_state_layout* L,
                                    const double* y, int64_t i) {
    /* surface_tissue._species_sum order: z0 n0 + z1 n1 + ... */
    double t = w->valence[0] * CF(L, w, y, L->dn, i);
    for (int s = 1; s < w->S; ++s) t = t + w->valence[s] * CF(L, w, y, L->dn + s, i);
    return CHARGE_PER_FMOL * t;
}

void ref_potentials(const cells_world* w, int32_t k, const double* y, double* phi) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double q = w->q0[i] + cell_charge_increment(w, &L, y, i);
        phi[i] = w->bath_V + q / w->cap[i];
    }
}

/* ---------------------------------------------------------------- explicit rates */

int ref_rates(const cells_world* w, int32_t k, const double* y, const double* phi,
              const cells_control* c, double* rt_out, double* gm, double* gj) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const int S = w->S;
    const double T = w->temperature, RT = R_CONST * T;
    int valid = 1;
    for (int s = 0; s < S; ++s)
        valid = valid && isfinite(w->bath_c[s]) && w->bath_c[s] > 0;
    const double ctl[5] = {c->membrane, c->junction, c->pump, c->calcium_pump, c->water};
    for (int q = 0; q < 5; ++q) valid = valid && isfinite(ctl[q]) && ctl[q] >= 0 && ctl[q] <= 1;

    /* junction routes (surface_transport.py:310-367, junction branch) */
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        double va = cell_volume(w, &L, y, a), vb = cell_volume(w, &L, y, b);
        double dv = phi[a] - phi[b];
        double g = EF(&L, w, y, L.gate, e);
        double avail = c->junction * g;
        double diss = 0.0;
        for (int s = 0; s < S; ++s) {
            double ca = (w->n0[(int64_t)s * w->ncell + a] + CF(&L, w, y, L.dn + s, a)) / va;
            double cb = (w->n0[(int64_t)s * w->ncell + b] + CF(&L, w, y, L.dn + s, b)) / vb;
            double z = w->valence[s];
            double pa = w->paj[(int64_t)s * w->nedge + e];
            valid = valid && isfinite(ca) && ca > 0 && isfinite(cb) && cb > 0;
            double J = ref_selective_flux(ca, cb, dv, z, T, pa, avail);
            double mu = RT * (log(ca) - log(cb)) + z * F_CONST * dv;
            diss = s == 0 ? J * mu : diss + J * mu;
            valid = valid && fabs(z * F_CONST * dv / RT) <= MAX_ABS_X && isfinite(J);
            EF(&L, w, rt_out, L.jtr + s, e) = J;
            if (gj) gj[(int64_t)s * w->nedge + e] = ref_route_slope(ca, cb, dv, z, T, pa, avail);
        }
        EF(&L, w, rt_out, L.jdiss, e) = 1e-3 * diss;
        valid = valid && isfinite(1e-3 * diss);
        EF(&L, w, rt_out, L.gate, e) = ref_gate_rhs(g, dv, w->gate_threshold, w->gate_minimum);
        valid = valid && gate_domain(g, dv, w->gate_threshold, w->gate_minimum, c->junction);
    }

    /* membrane routes, pumps, owners (surface_transport.py:310-375; surface_pump_v2.py:370-450) */
    double sup[4] = {0, 0, 0, 0};
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double vol = cell_volume(w, &L, y, i);
        double dv = phi[i] - w->bath_V;
        double ci[4], Jm[4], mu[4];
        double diss = 0.0;
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s], co = w->bath_c[s];
            ci[s] = (w->n0[(int64_t)s * w->ncell + i] + CF(&L, w, y, L.dn + s, i)) / vol;
            valid = valid && isfinite(ci[s]) && ci[s] > 0;
            double pa = w->pam[(int64_t)s * w->ncell + i];
            Jm[s] = ref_selective_flux(ci[s], co, dv, z, T, pa, c->membrane);
            mu[s] = RT * (log(ci[s]) - log(co)) + z * F_CONST * dv;
            diss = s == 0 ? Jm[s] * mu[s] : diss + Jm[s] * mu[s];
            valid = valid && fabs(z * F_CONST * dv / RT) <= MAX_ABS_X && isfinite(Jm[s]);
            if (gm) gm[(int64_t)s * w->ncell + i] = ref_route_slope(ci[s], co, dv, z, T, pa, c->membrane);
        }
        double cyc, energy, drive;
        ref_nak_kinetics(w, ci[0], w->bath_c[1], mu[0], mu[1], w->nak_cap[i], c->pump, &cyc,
                         &energy, &drive);
        valid = valid && ref_nak_domain(w, ci[0], w->bath_c[1], mu[0], mu[1], w->nak_cap[i], c->pump);
        double extra[4] = {3.0 * cyc, -2.0 * cyc, 0.0, 0.0}; /* -stoichiometry * cycles */
        if (S > 2) extra[2] = -0.0 * cyc;
        if (S > 3) extra[3] = -0.0 * cyc;
        double ccyc = 0, cenergy = 0, cdrive = 0;
        if (w->calcium) {
            ref_ca_kinetics(w, ci[3], w->bath_c[3], dv, w->ca_cap[i], c->calcium_pump, &ccyc,
                            &cenergy, &cdrive);
            valid = valid && ca_domain(w, ci[3], w->bath_c[3], dv, w->ca_cap[i], c->calcium_pump);
            /* extra = -stoichiometry*cycles - calcium_stoichiometry*ca_cycles, calcium -1 */
            for (int s = 0; s < S; ++s) extra[s] = extra[s] - (s == 3 ? -1.0 : 0.0) * ccyc;
        }
        /* owners: cell = (-membrane + sum(-junction at sender)) + sum(junction at receiver) */
        int64_t p0 = w->inc_ptr[i], p1 = w->inc_ptr[i + 1], pm = p0 + w->inc_out[i];
        for (int s = 0; s < S; ++s) {
            double out = 0.0, in = 0.0;
            for (int64_t p = p0; p < pm; ++p) out += -EF(&L, w, rt_out, L.jtr + s, w->inc_edge[p]);
            for (int64_t p = pm; p < p1; ++p) in += EF(&L, w, rt_out, L.jtr + s, w->inc_edge[p]);
            double cell = -Jm[s];
            cell += out;
            cell += in;
            CF(&L, w, rt_out, L.dn + s, i) = cell - extra[s];
            CF(&L, w, rt_out, L.mtr + s, i) = Jm[s];
            sup[s] -= Jm[s] + extra[s]; /* supply = -inflow at the held bath */
        }
        CF(&L, w, rt_out, L.mdiss, i) = 1e-3 * diss;
        CF(&L, w, rt_out, L.pcyc, i) = cyc;
        CF(&L, w, rt_out, L.pdiss, i) = -1e-3 * cyc * energy;
        CF(&L, w, rt_out, L.pwork, i) = -1e-3 * cyc * drive;
        if (w->calcium) {
            CF(&L, w, rt_out, L.ccyc, i) = ccyc;
            CF(&L, w, rt_out, L.cdiss, i) = -1e-3 * ccyc * cenergy;
            CF(&L, w, rt_out, L.cwork, i) = -1e-3 * ccyc * cdrive;
        }
        if (w->water) {
            CF(&L, w, rt_out, L.dvol, i) = 0.0;
            CF(&L, w, rt_out, L.wtr, i) = 0.0;
            CF(&L, w, rt_out, L.wdiss, i) = 0.0;
        }
        for (int f = 0; f < L.cell_fields; ++f) valid = valid && isfinite(CF(&L, w, rt_out, f, i));
    }
    for (int s = 0; s < S; ++s) WF(&L, w, rt_out, L.sup + s, k) = sup[s];
    WF(&L, w, rt_out, L.echg, k) = 0.0;
    WF(&L, w, rt_out, L.rwork, k) = 0.0;
    WF(&L, w, rt_out, L.ework, k) = 0.0;
    if (w->water) { WF(&L, w, rt_out, L.wsup, k) = 0.0; WF(&L, w, rt_out, L.wwork, k) = 0.0; }
    return valid;
}

/* ---------------------------------------------------------------- domain */

static int domain_valid(const cells_world* w, const cells_state_layout* L, int32_t k,
                        const double* y) {
    /* surface_tissue.py:391-420, surface_kinetic.py:422-451: finite fields, positive amounts,
     * gates in [minimum, 1]; water volumes strictly inside their bounds. */
    int ok = 1;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        for (int f = 0; f < L->cell_fields; ++f) ok = ok && isfinite(CF(L, w, y, f, i));
        for (int s = 0; s < w->S; ++s) {
            double n = w->n0[(int64_t)s * w->ncell + i] + CF(L, w, y, L->dn + s, i);
            ok = ok && isfinite(n) && n > 0;
        }
        if (w->water) {
            double v = cell_volume(w, L, y, i);
            ok = ok && v > w->water_vmin * w->vol0[i] && v < w->water_vmax * w->vol0[i];
        }
    }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        for (int f = 0; f < L->edge_fields; ++f) ok = ok && isfinite(EF(L, w, y, f, e));
        double g = EF(L, w, y, L->gate, e);
        ok = ok && g >= w->gate_minimum && g <= 1.0;
    }
    for (int f = 0; f < L->world_fields; ++f) ok = ok && isfinite(WF(L, w, y, f, k));
    return ok;
}

static void copy_world(const cells_world* w, const cells_state_layout* L, int32_t k,
                       const double* src, double* dst) {
    for (int f = 0; f < L->cell_fields; ++f)
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
            CF(L, w, dst, f, i) = CF(L, w, src, f, i);
    for (int f = 0; f < L->edge_fields; ++f)
        for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e)
            EF(L, w, dst, f, e) = EF(L, w, src, f, e);
    for (int f = 0; f < L->world_fields; ++f) WF(L, w, dst, f, k) = WF(L, w, src, f, k);
}

/* y_out = y + (wt*dt) * r on every cell and edge field of world k, and on the world fields
 * listed by with_world (the stage-integrated ones: bath supplies). */
static void combine(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y,
                    double h, const double* r, double* out) {
    for (int f = 0; f < L->cell_fields; ++f)
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
            CF(L, w, out, f, i) = CF(L, w, y, f, i) + h * CF(L, w, r, f, i);
    for (int f = 0; f < L->edge_fields; ++f)
        for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e)
            EF(L, w, out, f, e) = EF(L, w, y, f, e) + h * EF(L, w, r, f, e);
    for (int s = 0; s < w->S; ++s)
        WF(L, w, out, L->sup + s, k) = WF(L, w, y, L->sup + s, k) + h * WF(L, w, r, L->sup + s, k);
}

/* Electrode charge delivered over a step: the held bath's plate-charge change,
 * -(sum_i q_i(end) - sum_i q_i(start)) (surface_pump.py:273-289 with one held bath). */
static double electrode_increment(const cells_world* w, const cells_state_layout* L, int32_t k,
                                  const double* start, const double* end) {
    double qs = 0.0, qe = 0.0;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        qs += cell_charge_increment(w, L, start, i);
        qe += cell_charge_increment(w, L, end, i);
    }
    return -qe - -qs;
}

static void ionic_tail(const cells_world* w, const cells_state_layout* L, int32_t k,
                       const double* start, double* stage, const double* applied_supply) {
    /* surface_pump.py:346-359: electrode charge and work, reservoir work of the applied supply. */
    double electrode = electrode_increment(w, L, k, start, stage);
    double RT = R_CONST * w->temperature;
    double work = 0.0;
    for (int s = 0; s < w->S; ++s) {
        double mu = RT * log(w->bath_c[s]) + w->valence[s] * F_CONST * w->bath_V;
        work = s == 0 ? mu * applied_supply[s] : work + mu * applied_supply[s];
    }
    WF(L, w, stage, L->echg, k) = WF(L, w, start, L->echg, k) + electrode;
    WF(L, w, stage, L->rwork, k) = WF(L, w, start, L->rwork, k) + 1e-3 * work;
    WF(L, w, stage, L->ework, k) = WF(L, w, start, L->ework, k) + 1e12 * w->bath_V * electrode;
}

/* ---------------------------------------------------------------- water (Heun, h) */

static double water_psi_bath(const cells_world* w) {
    double total = 0.0;
    for (int s = 0; s < w->S; ++s) total += w->bath_c[s];
    return w->bath_pressure - R_CONST * w->temperature * total;
}

static double water_psi_cell(const cells_world* w, const cells_state_layout* L, const double* y,
                             int64_t i, double* osm_total) {
    /* numerics/osmotic_water.py:193-215: psi = K (V - Vref)/Vref + P0 - RT (sum N + X)/V */
    double v = cell_volume(w, L, y, i);
    double total = 0.0;
    for (int s = 0; s < w->S; ++s)
        total += w->n0[(int64_t)s * w->ncell + i] + CF(L, w, y, L->dn + s, i);
    total = total + w->imperm[i];
    if (osm_total) *osm_total = total;
    double rt = R_CONST * w->temperature;
    double osmotic = rt * total / v;
    double hydro = w->water_K * (v - w->vol0[i]) / w->vol0[i] + w->water_P0;
    return hydro - osmotic;
}

static int water_half(const cells_world* w, const cells_state_layout* L, int32_t k, double h,
                      const cells_control* c, double* y /* in/out */) {
    /* operators/osmotic_water.py:127-221: Heun over h at frozen amounts, cell -> bath flow. */
    double psib = water_psi_bath(w), rt = R_CONST * w->temperature;
    int valid = domain_valid(w, L, k, y);
    valid = valid && isfinite(c->water) && c->water >= 0 && c->water <= 1;
    /* Gershgorin rate: each cell row G (slope_cell + 0), the bath row the sum of them. */
    double maxrow = 0.0, bath_row = 0.0;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double total;
        (void)water_psi_cell(w, L, y, i, &total);
        double v = cell_volume(w, L, y, i);
        double slope = w->water_K / w->vol0[i] + rt * total / (v * v);
        double edge = w->lpa[i] * c->water * (slope + 0.0);
        if (edge > maxrow) maxrow = edge;
        bath_row += edge;
    }
    if (bath_row > maxrow) maxrow = bath_row;
    valid = valid && 2 * h * maxrow <= WATER_MAX_RELAXATION;
    int64_t c0 = w->cell_off[k], c1 = w->cell_off[k + 1];
    int64_t n = c1 - c0;
    double* r0 = (double*)malloc(sizeof(double) * 3 * n);
    double sup0 = 0.0, sup1 = 0.0;
    for (int64_t i = c0; i < c1; ++i) {
        double drive = water_psi_cell(w, L, y, i, NULL) - psib;
        double flow = w->lpa[i] * c->water * drive;
        r0[3 * (i - c0)] = -flow;
        r0[3 * (i - c0) + 1] = flow;
        r0[3 * (i - c0) + 2] = 1e-3 * flow * drive;
        sup0 += flow; /* bath supply = -(inflow into the bath) */
    }
    sup0 = -sup0;
    /* predicted state: only the water fields move */
    double* saved = (double*)malloc(sizeof(double) * 3 * n);
    for (int64_t i = c0; i < c1; ++i) {
        saved[3 * (i - c0)] = CF(L, w, y, L->dvol, i);
        saved[3 * (i - c0) + 1] = CF(L, w, y, L->wtr, i);
        saved[3 * (i - c0) + 2] = CF(L, w, y, L->wdiss, i);
        CF(L, w, y, L->dvol, i) = saved[3 * (i - c0)] + h * r0[3 * (i - c0)];
        CF(L, w, y, L->wtr, i) = saved[3 * (i - c0) + 1] + h * r0[3 * (i - c0) + 1];
        CF(L, w, y, L->wdiss, i) = saved[3 * (i - c0) + 2] + h * r0[3 * (i - c0) + 2];
    }
    double ws = WF(L, w, y, L->wsup, k), ww = WF(L, w, y, L->wwork, k);
    WF(L, w, y, L->wsup, k) = ws + h * sup0;
    WF(L, w, y, L->wwork, k) = ww + 1e-3 * psib * (h * sup0);
    valid = valid && domain_valid(w, L, k, y);
    for (int64_t i = c0; i < c1; ++i) {
        double drive = water_psi_cell(w, L, y, i, NULL) - psib;
        double flow = w->lpa[i] * c->water * drive;
        double heat = 1e-3 * flow * drive;
        sup1 += flow;
        double* a = &r0[3 * (i - c0)];
        a[0] = 0.5 * (a[0] + -flow);
        a[1] = 0.5 * (a[1] + flow);
        a[2] = 0.5 * (a[2] + heat);
    }
    sup1 = -sup1;
    double supm = 0.5 * (sup0 + sup1);
    for (int64_t i = c0; i < c1; ++i) {
        CF(L, w, y, L->dvol, i) = saved[3 * (i - c0)] + h * r0[3 * (i - c0)];
        CF(L, w, y, L->wtr, i) = saved[3 * (i - c0) + 1] + h * r0[3 * (i - c0) + 1];
        CF(L, w, y, L->wdiss, i) = saved[3 * (i - c0) + 2] + h * r0[3 * (i - c0) + 2];
    }
    double supplied = h * supm;
    WF(L, w, y, L->wsup, k) = ws + supplied;
    WF(L, w, y, L->wwork, k) = ww + 1e-3 * psib * supplied;
    valid = valid && domain_valid(w, L, k, y);
    free(r0);
    free(saved);
    return valid;
}

/* ---------------------------------------------------------------- IMEX Euler */

/* The explicit stage's membrane charge relaxation rate of one cell (1/s): models/surface_pump_v3.py
 * :4588-4638 membrane_rate with one patch per cell and no kinds (its patch slopes :4534-4585):
 * F^2/RT sum_s z_s^2 (A P)_s max(c_in, c_out) times the membrane availability, over the cell's
 * capacitance. The step refuses where dt times it exceeds EXPLICIT_STEP_BOUND = 1
 * (surface_pump_v3.py:3640-3642). */
/* ---------------------------------------------------------------- ledgers */

#define LEDGER_STEP_ROUNDING 2.5        /* surface_pump.py:113 */
#define COMPONENT_STEP_ROUNDING 2.5     /* surface_field.py:70 */
#define GLOBAL_CHARGE_ATOL_C 3e-25      /* surface_pump.py:96 */
#define GLOBAL_CHARGE_ROUNDING (256.0 * 2.220446049250313e-16) /* surface_pump.py:97 */
#define CHARGE_ATOL_C 3e-25             /* surface_field.py:47 */
#define CHARGE_RTOL 2e-12               /* surface_field.py:48 */
#define ENERGY_ATOL_PJ 2e-8             /* surface_tissue.py:29 */
#define WORK_FLOOR_PJ 3e-12             /* surface_pump.py:789-791 */
#define VOLUME_ATOL_PL 2e-12            /* osmotic_water.py:70 */
#define WATER_WORK_ATOL_PJ 3e-12        /* osmotic_water.py:73 */
#define DISSIPATION_ATOL 1e-12          /* surface_pump.py:865-866 */

static double ledger_ratio(double residual, double bound) {
    double m = fabs(residual);
    if (m == 0.0) return 0.0;
    double r = m / bound;
    return isnan(r) ? INFINITY : r;
}

int ref_ledgers(const cells_world* w, int32_t k, const double* prev, const double* y, double steps,
                double material_atol, double energy_rtol, double* ratios, double* terms) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const int S = w->S;
    const double RT = R_CONST * w->temperature, eps = 2.220446049250313e-16;
    const double rounding = LEDGER_STEP_ROUNDING * eps * steps;
    const double drive = nak_drive(w);
    const double ca_drive = w->ca_dg0 + RT * (log(w->adp) + log(w->pi) - log(w->atp) -
                                              log(STANDARD_STATE_mM));
    double r_local = 0, r_water = 0, r_cap = 0, r_met = 0;
    double tot[4] = {0}, totm[4] = {0}, hm[4] = {0}, hmm[4] = {0};
    double qsum = 0, qmag = 0, plate_sum = 0;
    double chem = 0, elec = 0, heat = 0, work = 0, water = 0, flows = 0;
    double dvs = 0, dvm = 0, trs = 0, trm = 0;
    int dissipation_ok = 1;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        const double cyc = CF(&L, w, y, L.pcyc, i);
        const double ccyc = w->calcium ? CF(&L, w, y, L.ccyc, i) : 0.0;
        double zq = 0, zm = 0;
        for (int s = 0; s < S; ++s) {
            /* balances' local ledger (surface_pump_checks.py:42-47): dn + membrane total + the
             * junction transfers out - in, the membrane total with the pumps' stoichiometry */
            double net = 0, netm = 0;
            for (int64_t q = w->inc_ptr[i]; q < w->inc_ptr[i + 1]; ++q) {
                int64_t e = w->inc_edge[q];
                double j = EF(&L, w, y, L.jtr + s, e);
                if (q < w->inc_ptr[i] + w->inc_out[i]) net -= j;
                else net += j;
                netm += fabs(j);
            }
            const double dn = CF(&L, w, y, L.dn + s, i), mtr = CF(&L, w, y, L.mtr + s, i);
            double pump = s == 0 ? 3.0 * cyc : s == 1 ? -2.0 * cyc : 0.0;
            if (s == 3 && w->calcium) pump += ccyc;
            const double h = mtr + pump, hmag = fabs(mtr) + fabs(pump);
            double r = ledger_ratio(dn + h - net, material_atol + rounding * (fabs(dn) + hmag + netm));
            if (!(r <= r_local)) r_local = r;
            tot[s] += dn;
            totm[s] += fabs(dn);
            hm[s] += h;
            hmm[s] += hmag;
            zq += w->valence[s] * dn;
            zm += fabs(w->valence[s] * dn);
            /* the mixing free energy at the reference volume (surface_pump_checks.py:68-78) */
            const double a = w->n0[(int64_t)s * w->ncell + i], v = w->vol0[i];
            chem += 1e-3 * RT * (dn * (log(a / v) - 1.0) + (a + dn) * log1p(dn / a));
        }
        const double q = CHARGE_PER_FMOL * zq, C = w->cap[i], q0 = w->q0[i];
        qsum += q;
        qmag += CHARGE_PER_FMOL * zm;
        /* the diagonal capacitor on the held bath (surface_tissue.py:358-393, balances' charge
         * ledger, surface_pump_checks.py:56-63): dp = q / C, plate charge C dp */
        const double dp = q / C, plate = C * dp;
        const double image = C * ((w->bath_V + q0 / C) - w->bath_V);
        double r = ledger_ratio(image - q0 + plate - q, CHARGE_ATOL_C + CHARGE_RTOL * fabs(q0 + q));
        if (!(r <= r_cap)) r_cap = r;
        plate_sum += plate;
        /* the capacitor energy 0.5 C dv (2 V0 + dv) (surface_pump_checks.py:79-81) */
        elec += 0.5e12 * (C * dp * (2.0 * (q0 / C) + dp));
        const double md = CF(&L, w, y, L.mdiss, i), pd = CF(&L, w, y, L.pdiss, i);
        const double pw = CF(&L, w, y, L.pwork, i);
        heat += md + pd;
        work += pw;
        flows += fabs(md) + fabs(pd) + fabs(pw);
        /* metabolic work against cycles x drive (surface_pump_v2.py:700-707) */
        r = ledger_ratio(pw + 1e-3 * drive * cyc,
                         WORK_FLOOR_PJ + rounding * (fabs(pw) + fabs(1e-3 * drive * cyc)));
        if (!(r <= r_met)) r_met = r;
        int heat_fields[4] = {L.mdiss, L.pdiss, L.cdiss, L.wdiss};
        if (w->calcium) {
            const double cd = CF(&L, w, y, L.cdiss, i), cw = CF(&L, w, y, L.cwork, i);
            heat += cd;
            work += cw;
            flows += fabs(cd) + fabs(cw);
            r = ledger_ratio(cw + 1e-3 * ca_drive * ccyc,
                             WORK_FLOOR_PJ + rounding * (fabs(cw) + fabs(1e-3 * ca_drive * ccyc)));
            if (!(r <= r_met)) r_met = r;
        }
        for (int f = 0; f < 4; ++f) {
            const int fi = heat_fields[f];
            if (fi < 0) continue;
            const double v = CF(&L, w, y, fi, i), u = CF(&L, w, prev, fi, i);
            if (v < -DISSIPATION_ATOL || v - u < -DISSIPATION_ATOL) dissipation_ok = 0;
        }
        if (w->water) {
            /* the water ledgers and water's energy terms (osmotic_water.py:254-295) */
            const double dv = CF(&L, w, y, L.dvol, i), tr = CF(&L, w, y, L.wtr, i);
            const double wd = CF(&L, w, y, L.wdiss, i);
            r = ledger_ratio(dv + tr, VOLUME_ATOL_PL + rounding * (fabs(dv) + fabs(tr)));
            if (!(r <= r_water)) r_water = r;
            dvs += dv;
            dvm += fabs(dv);
            trs += tr;
            trm += fabs(tr);
            double total = 0;
            for (int s = 0; s < S; ++s)
                total += w->n0[(int64_t)s * w->ncell + i] + CF(&L, w, y, L.dn + s, i);
            total += w->imperm[i];
            const double vr = w->vol0[i], stiff = w->water_K / vr;
            const double pressure = stiff * (vr - vr) + w->water_P0;
            water += -1e-3 * RT * (total * log1p(dv / vr)) + 1e-3 * (pressure * dv + 0.5 * stiff * dv * dv);
            heat += wd;
            flows += fabs(wd);
        }
    }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        const double v = EF(&L, w, y, L.jdiss, e), u = EF(&L, w, prev, L.jdiss, e);
        if (v < -DISSIPATION_ATOL || v - u < -DISSIPATION_ATOL) dissipation_ok = 0;
        heat += v;
        flows += fabs(v);
    }
    double r_tot = 0, r_out = 0, supq = 0, supm = 0, rw_expect = 0, rw_mag = 0;
    for (int s = 0; s < S; ++s) {
        const double sup = WF(&L, w, y, L.sup + s, k);
        double r = ledger_ratio(tot[s] - sup, material_atol + rounding * (totm[s] + fabs(sup)));
        if (!(r <= r_tot)) r_tot = r;
        /* the held bath: its increment 0, less the membranes' totals, less its supply */
        r = ledger_ratio(0.0 - hm[s] - sup, material_atol + rounding * (hmm[s] + fabs(sup)));
        if (!(r <= r_out)) r_out = r;
        supq += w->valence[s] * sup;
        supm += fabs(w->valence[s] * sup);
        const double mu = RT * log(w->bath_c[s]) + w->valence[s] * F_CONST * w->bath_V;
        rw_expect += mu * sup;
        rw_mag += fabs(mu) * fabs(sup);
    }
    const double echg = WF(&L, w, y, L.echg, k), rwork = WF(&L, w, y, L.rwork, k);
    const double ework = WF(&L, w, y, L.ework, k);
    const double moved = CHARGE_PER_FMOL * supm + fabs(echg);
    const double grounded = COMPONENT_STEP_ROUNDING * eps * steps * qmag;
    const double base = GLOBAL_CHARGE_ATOL_C + GLOBAL_CHARGE_ROUNDING * moved;
    const double gbound = fmax(base + rounding * moved, base + grounded);
    const double r_gq = ledger_ratio(echg + CHARGE_PER_FMOL * supq, gbound);
    /* the electrode: its charge less the bath's plate charge, -sum C dp */
    const double r_el = ledger_ratio(echg - -plate_sum, CHARGE_ATOL_C + CHARGE_RTOL * fabs(echg));
    work += rwork + ework;
    flows += fabs(rwork) + fabs(ework);
    const double r_rw = ledger_ratio(rwork - 1e-3 * rw_expect,
                                     WORK_FLOOR_PJ + rounding * (fabs(rwork) + 1e-3 * rw_mag));
    const double ew_expect = 1e12 * w->bath_V * echg;
    const double r_ew = ledger_ratio(ework - ew_expect, WORK_FLOOR_PJ + rounding * (fabs(ework) + fabs(ew_expect)));
    double r_bw = 0;
    if (w->water) {
        const double ws = WF(&L, w, y, L.wsup, k), ww = WF(&L, w, y, L.wwork, k);
        double r = ledger_ratio(dvs - ws, VOLUME_ATOL_PL + rounding * (dvm + fabs(ws)));
        if (!(r <= r_water)) r_water = r;
        r = ledger_ratio(ws + trs, VOLUME_ATOL_PL + rounding * (trm + fabs(ws)));
        if (!(r <= r_water)) r_water = r;
        const double bw = 1e-3 * water_psi_bath(w) * ws;
        r_bw = ledger_ratio(ww - bw, WATER_WORK_ATOL_PJ + rounding * (fabs(ww) + fabs(bw)));
        work += ww;
        flows += fabs(ww);
    }
    const double energy = chem + elec + heat - work + water;
    const double r_en = ledger_ratio(energy, ENERGY_ATOL_PJ + energy_rtol * flows);
    const double rr[12] = {r_local, r_tot, r_gq, r_water, r_out, r_cap, r_el, r_en, r_rw, r_ew, r_met, r_bw};
    const int bit[12] = {2, 4, 8, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192};
    int bits = domain_valid(w, &L, k, y) ? 0 : 1;
    if (!dissipation_ok) bits |= 16;
    for (int j = 0; j < 12; ++j) {
        ratios[j] = rr[j];
        if (!(rr[j] <= 1.0)) bits |= bit[j];
    }
    if (terms) {
        terms[0] = chem;
        terms[1] = elec;
        terms[2] = heat;
        terms[3] = work;
        terms[4] = water;
        terms[5] = energy;
        terms[6] = flows;
    }
    return bits;
}

double ref_membrane_rate(const cells_world* w, const double* y, int64_t i, const cells_control* c) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const double vol = cell_volume(w, &L, y, i);
    const double RT = R_CONST * w->temperature;
    double t = 0.0;
    for (int s = 0; s < w->S; ++s) {
        const double ci = (w->n0[(int64_t)s * w->ncell + i] + CF(&L, w, y, L.dn + s, i)) / vol;
        const double term = w->valence[s] * w->valence[s] * w->pam[(int64_t)s * w->ncell + i] *
                            (ci > w->bath_c[s] ? ci : w->bath_c[s]);
        t = s == 0 ? term : t + term;
    }
    return F_CONST * F_CONST / RT * (t * c->membrane) / w->cap[i];
}

int ref_step_euler(const cells_world* w, const double* y, double* out, double dt,
                   const cells_control* c, int32_t* valid_out) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    double* r = (double*)calloc((size_t)L.total, sizeof(double));
    double* phi = (double*)malloc(sizeof(double) * (size_t)(w->ncell ? w->ncell : 1));
    double* start = (double*)malloc(sizeof(double) * (size_t)L.total);
    int all = 1;
    for (int32_t k = 0; k < w->W; ++k) {
        int valid = 1;
        copy_world(w, &L, k, y, start);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, start) && valid;
        /* surface_pump.py:292-362 with no faces: the stage is y + dt r(start potentials) */
        ref_potentials(w, k, start, phi);
        valid = ref_rates(w, k, start, phi, c, r, NULL, NULL) && valid;
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
            valid = valid && dt * ref_membrane_rate(w, start, i, c) <= 1.0;
        copy_world(w, &L, k, start, out);
        combine(w, &L, k, start, 1.0 * dt, r, out);
        double applied[4];
        for (int s = 0; s < w->S; ++s) applied[s] = dt * WF(&L, w, r, L.sup + s, k);
        ionic_tail(w, &L, k, start, out, applied);
        /* stage potentials (the diagonal operator solve) and their residual */
        ref_potentials(w, k, out, phi);
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
            double q = w->q0[i] + cell_charge_increment(w, &L, out, i);
            double cphi = w->cap[i] * (phi[i] - w->bath_V);
            double scale = fabs(q) > fabs(cphi) ? fabs(q) : fabs(cphi);
            valid = valid && isfinite(phi[i]) && fabs(cphi - q) <= 3e-25 + 2e-12 * scale;
        }
        valid = valid && domain_valid(w, &L, k, out);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, out) && valid;
        if (!valid) copy_world(w, &L, k, y, out);
        valid_out[k] = valid;
        all = all && valid;
    }
    free(r);
    free(phi);
    free(start);
    return all;
}

/* The stage potentials' residual of world k's state (the diagonal operator solve of a lattice on
 * its held bath, surface_imex.py:269-278 with no faces): each cell's C (phi - V_bath) against its
 * node charge, 3e-25 + 2e-12 of the larger. */
static int stage_potentials_ok(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y,
                               double* phi) {
    int ok = 1;
    ref_potentials(w, k, y, phi);
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double q = w->q0[i] + cell_charge_increment(w, L, y, i);
        double cphi = w->cap[i] * (phi[i] - w->bath_V);
        double scale = fabs(q) > fabs(cphi) ? fabs(q) : fabs(cphi);
        ok = ok && isfinite(phi[i]) && fabs(cphi - q) <= 3e-25 + 2e-12 * scale;
    }
    return ok;
}

/* ARS(2,2,2) (numerics/surface_imex.py:53-57, cells.solver.imex_ars222_face_electromigration@1) as
 * models/surface_pump_v3.py:3547-3955 imex_step runs it for cells on one held bath: with no faces
 * the implicit part is empty and the operator is the capacitors alone, so each stage is its known
 * part (stage 1: y + gamma dt E0; stage 2: (y + delta dt E0) + (1 - delta) dt E1), the stage
 * potentials the capacitor solve's, every explicit stage checked against the membrane rate
 * (:3640-3642, :3717-3737), the gates moved by the tableau. */
int ref_step_ars222(const cells_world* w, const double* y, double* out, double dt, const cells_control* c,
                    int32_t* valid_out) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const double gamma = ARS_GAMMA, delta = ARS_DELTA;
    const size_t T = (size_t)L.total;
    double* E0 = (double*)calloc(T, sizeof(double));
    double* E1 = (double*)calloc(T, sizeof(double));
    double* stage = (double*)malloc(sizeof(double) * T);
    double* start = (double*)malloc(sizeof(double) * T);
    double* phi = (double*)malloc(sizeof(double) * (size_t)(w->ncell ? w->ncell : 1));
    int all = 1;
    for (int32_t k = 0; k < w->W; ++k) {
        int valid = 1;
        copy_world(w, &L, k, y, start);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, start) && valid;
        ref_potentials(w, k, start, phi);
        valid = ref_rates(w, k, start, phi, c, E0, NULL, NULL) && valid;
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
            valid = valid && dt * ref_membrane_rate(w, start, i, c) <= 1.0;
        /* stage 1 */
        copy_world(w, &L, k, start, stage);
        combine(w, &L, k, start, gamma * dt, E0, stage);
        valid = stage_potentials_ok(w, &L, k, stage, phi) && valid;
        valid = ref_rates(w, k, stage, phi, c, E1, NULL, NULL) && valid;
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
            valid = valid && dt * ref_membrane_rate(w, stage, i, c) <= 1.0;
        /* stage 2: the new state */
        copy_world(w, &L, k, start, out);
        combine(w, &L, k, start, delta * dt, E0, out);
        combine(w, &L, k, out, (1.0 - delta) * dt, E1, out);
        valid = stage_potentials_ok(w, &L, k, out, phi) && valid;
        double applied[4];
        for (int s2 = 0; s2 < w->S; ++s2) {
            double ex = (0.0 + delta * WF(&L, w, E0, L.sup + s2, k)) + (1.0 - delta) * WF(&L, w, E1, L.sup + s2, k);
            applied[s2] = dt * (ex + 0.0);
        }
        for (int f = 0; f < L.world_fields; ++f)
            if (f < L.sup || f >= L.sup + w->S) WF(&L, w, out, f, k) = WF(&L, w, start, f, k);
        ionic_tail(w, &L, k, start, out, applied);
        valid = valid && domain_valid(w, &L, k, out);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, out) && valid;
        if (!valid) copy_world(w, &L, k, y, out);
        valid_out[k] = valid;
        all = all && valid;
    }
    free(E0); free(E1); free(stage); free(start); free(phi);
    return all;
}

/* The reference's right-hand side of every leaf (models/surface_tissue.py:328-347
 * derivative_from_field, with the pumps): the explicit rates, the electrode current (the held
 * bath's plate-charge rate, minus the cells' charge rate), the reservoir work rate
 * 1e-3 sum_s mu_s supply_s and the electrode work rate 1e12 V_bath current. */
static int full_rhs(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y,
                    const cells_control* c, double* phi, double* r) {
    ref_potentials(w, k, y, phi);
    ref_rates(w, k, y, phi, c, r, NULL, NULL);
    double current = 0.0;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) current += cell_charge_increment(w, L, r, i);
    current = -current;
    const double RT = R_CONST * w->temperature;
    double work = 0.0;
    for (int s = 0; s < w->S; ++s) {
        double mu = RT * log(w->bath_c[s]) + w->valence[s] * F_CONST * w->bath_V;
        work += mu * WF(L, w, r, L->sup + s, k);
    }
    WF(L, w, r, L->echg, k) = current;
    WF(L, w, r, L->rwork, k) = 1e-3 * work;
    WF(L, w, r, L->ework, k) = 1e12 * w->bath_V * current;
    int ok = 1;
    for (int f = 0; f < L->cell_fields; ++f)
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) ok = ok && isfinite(CF(L, w, r, f, i));
    for (int f = 0; f < L->edge_fields; ++f)
        for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) ok = ok && isfinite(EF(L, w, r, f, e));
    for (int f = 0; f < L->world_fields; ++f) ok = ok && isfinite(WF(L, w, r, f, k));
    return ok;
}

/* trial = y + (f dt) r on every field of world k (numerics/fixed_step.py's increment) */
static void rk4_trial(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y, double h,
                      const double* r, double* trial) {
    for (int f = 0; f < L->cell_fields; ++f)
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
            CF(L, w, trial, f, i) = CF(L, w, y, f, i) + h * CF(L, w, r, f, i);
    for (int f = 0; f < L->edge_fields; ++f)
        for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e)
            EF(L, w, trial, f, e) = EF(L, w, y, f, e) + h * EF(L, w, r, f, e);
    for (int f = 0; f < L->world_fields; ++f) WF(L, w, trial, f, k) = WF(L, w, y, f, k) + h * WF(L, w, r, f, k);
}

/* Classical RK4 with refusal at every trial and at the endpoint (numerics/fixed_step.py:25-71
 * rk4_checked_step, the step of an empty tableau, models/surface_pump_v3.py:3989-3998): each
 * trial's domain, each slope finite, the end's domain. Worlds without water (the reference's
 * water runs RK4 on a coupled right-hand side). */
int ref_step_rk4(const cells_world* w, const double* y, double* out, double dt, const cells_control* c,
                 int32_t* valid_out) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const size_t T = (size_t)L.total;
    double* kk[4];
    for (int j = 0; j < 4; ++j) kk[j] = (double*)calloc(T, sizeof(double));
    double* trial = (double*)malloc(sizeof(double) * T);
    double* phi = (double*)malloc(sizeof(double) * (size_t)(w->ncell ? w->ncell : 1));
    const double factor[4] = {0.0, 0.5, 0.5, 1.0};
    int all = 1;
    for (int32_t k = 0; k < w->W; ++k) {
        int valid = !w->water;
        copy_world(w, &L, k, y, trial);
        for (int j = 0; j < 4; ++j) {
            if (j > 0) rk4_trial(w, &L, k, y, factor[j] * dt, kk[j - 1], trial);
            int ok = domain_valid(w, &L, k, trial);
            if (ok) ok = full_rhs(w, &L, k, trial, c, phi, kk[j]);
            valid = valid && ok;
        }
        for (int f = 0; f < L.cell_fields; ++f)
            for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
                CF(&L, w, out, f, i) = CF(&L, w, y, f, i) + (dt / 6.0) * (((CF(&L, w, kk[0], f, i) +
                    2.0 * CF(&L, w, kk[1], f, i)) + 2.0 * CF(&L, w, kk[2], f, i)) + CF(&L, w, kk[3], f, i));
        for (int f = 0; f < L.edge_fields; ++f)
            for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e)
                EF(&L, w, out, f, e) = EF(&L, w, y, f, e) + (dt / 6.0) * (((EF(&L, w, kk[0], f, e) +
                    2.0 * EF(&L, w, kk[1], f, e)) + 2.0 * EF(&L, w, kk[2], f, e)) + EF(&L, w, kk[3], f, e));
        for (int f = 0; f < L.world_fields; ++f)
            WF(&L, w, out, f, k) = WF(&L, w, y, f, k) + (dt / 6.0) * (((WF(&L, w, kk[0], f, k) +
                2.0 * WF(&L, w, kk[1], f, k)) + 2.0 * WF(&L, w, kk[2], f, k)) + WF(&L, w, kk[3], f, k));
        valid = valid && domain_valid(w, &L, k, out);
        if (!valid) copy_world(w, &L, k, y, out);
        valid_out[k] = valid;
        all = all && valid;
    }
    for (int j = 0; j < 4; ++j) free(kk[j]);
    free(trial); free(phi);
    return all;
}

/* The readouts' energy rates of world k at y (models/surface_pump_v3.py:2221-2262 readouts, cells
 * on one held bath without water): out[8] = chemical, capacitor, dissipation, reservoir power,
 * electrode power, metabolic work rate, energy balance, electrode current. */
int ref_energy_rates(const cells_world* w, int32_t k, const double* y, const cells_control* c, double* out) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const int S = w->S;
    const double RT = R_CONST * w->temperature;
    double* r = (double*)calloc((size_t)L.total, sizeof(double));
    double* phi = (double*)malloc(sizeof(double) * (size_t)(w->ncell ? w->ncell : 1));
    ref_potentials(w, k, y, phi);
    int valid = ref_rates(w, k, y, phi, c, r, NULL, NULL);
    double chem = 0, cap = 0, heat = 0, work = 0, plate = 0;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double vol = cell_volume(w, &L, y, i), ch = 0.0;
        for (int s = 0; s < S; ++s) {
            double ci = (w->n0[(int64_t)s * w->ncell + i] + CF(&L, w, y, L.dn + s, i)) / vol;
            ch = s == 0 ? log(ci) * CF(&L, w, r, L.dn + s, i) : ch + log(ci) * CF(&L, w, r, L.dn + s, i);
        }
        chem += ch;
        double qdot = cell_charge_increment(w, &L, r, i), C = w->cap[i];
        double dvm = qdot / C;
        cap += C * (phi[i] - w->bath_V) * dvm;
        plate += C * dvm;
        heat += CF(&L, w, r, L.mdiss, i) + CF(&L, w, r, L.pdiss, i);
        work += CF(&L, w, r, L.pwork, i);
        if (w->calcium) {
            heat += CF(&L, w, r, L.cdiss, i);
            work += CF(&L, w, r, L.cwork, i);
        }
    }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) heat += EF(&L, w, r, L.jdiss, e);
    double reservoir = 0.0;
    for (int s = 0; s < S; ++s) {
        double mu = RT * log(w->bath_c[s]) + w->valence[s] * F_CONST * w->bath_V;
        reservoir += mu * WF(&L, w, r, L.sup + s, k);
    }
    reservoir *= 1e-3;
    double current = -plate;
    out[0] = 1e-3 * RT * chem;
    out[1] = 1e12 * cap;
    out[2] = heat;
    out[3] = reservoir;
    out[4] = 1e12 * w->bath_V * current;
    out[5] = work;
    out[6] = out[0] + out[1] + out[2] - (out[3] + out[4]) - out[5];
    out[7] = current;
    free(r);
    free(phi);
    return valid;
}

/* ---------------------------------------------------------------- sparse factor and solve */

typedef struct {
    int64_t* ptr; /* [n+1] row structures: columns m < r holding row r, ascending */
    int64_t* col;
    int64_t* pos;
} rowstruct;

static void build_rows(const cells_symbolic* s, rowstruct* rs) {
    int64_t n = s->n;
    rs->ptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < n; ++k)
        for (int64_t p = s->column_start[k] + 1; p < s->column_start[k + 1]; ++p)
            rs->ptr[s->row_rank[p] + 1]++;
    for (int64_t r = 0; r < n; ++r) rs->ptr[r + 1] += rs->ptr[r];
    int64_t m = rs->ptr[n] ? rs->ptr[n] : 1;
    rs->col = (int64_t*)malloc(sizeof(int64_t) * (size_t)m);
    rs->pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)m);
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t r = 0; r < n; ++r) fill[r] = rs->ptr[r];
    for (int64_t k = 0; k < n; ++k) /* ascending k, so each row's columns come out ascending */
        for (int64_t p = s->column_start[k] + 1; p < s->column_start[k + 1]; ++p) {
            int64_t r = s->row_rank[p];
            rs->col[fill[r]] = k;
            rs->pos[fill[r]] = p;
            fill[r]++;
        }
    free(fill);
}

static void free_rows(rowstruct* rs) {
    free(rs->ptr);
    free(rs->col);
    free(rs->pos);
}

static double pairwise(double* t, int64_t m) {
    /* sparse_cholesky._pairwise_steps / fold: adjacent pairs step by step, +0.0 when empty */
    if (m == 0) return 0.0;
    while (m > 1) {
        int64_t o = 0;
        for (int64_t a = 0; a < m; a += 2) t[o++] = a + 1 < m ? t[a] + t[a + 1] : t[a];
        m = o;
    }
    return t[0];
}

void ref_factor(const cells_symbolic* s, double* v) {
    /* sparse_cholesky.py:474-490 in column order (each level's columns depend only on earlier
     * ones, so column order gives the level order's bits). */
    rowstruct rs;
    build_rows(s, &rs);
    int64_t n = s->n;
    int64_t maxc = 1;
    for (int64_t r = 0; r < n; ++r)
        if (rs.ptr[r + 1] - rs.ptr[r] > maxc) maxc = rs.ptr[r + 1] - rs.ptr[r];
    double* terms = (double*)malloc(sizeof(double) * (size_t)maxc);
    double* acc = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    for (int64_t k = 0; k < n; ++k) {
        int64_t c0 = s->column_start[k], c1 = s->column_start[k + 1];
        for (int64_t p = c0; p < c1; ++p) {
            int64_t r = s->row_rank[p];
            int64_t m = 0;
            int64_t a = rs.ptr[k], ae = rs.ptr[k + 1], b = rs.ptr[r], be = rs.ptr[r + 1];
            while (a < ae && b < be) {
                if (rs.col[a] == rs.col[b]) {
                    terms[m++] = v[rs.pos[b]] * v[rs.pos[a]];
                    ++a;
                    ++b;
                } else if (rs.col[a] < rs.col[b]) {
                    ++a;
                } else {
                    ++b;
                }
            }
            acc[p - c0] = v[p] - pairwise(terms, m);
        }
        double root = sqrt(acc[0]);
        v[c0] = root;
        for (int64_t p = c0 + 1; p < c1; ++p) v[p] = acc[p - c0] * (1.0 / root);
    }
    free(terms);
    free(acc);
    free_rows(&rs);
}

void ref_solve(const cells_symbolic* s, const double* f, const double* rhs, double* x) {
    /* sparse_cholesky.py:493-514, the device form (times each pivot's reciprocal). */
    rowstruct rs;
    build_rows(s, &rs);
    int64_t n = s->n;
    double* y = (double*)malloc(sizeof(double) * (size_t)n); /* by rank */
    int64_t maxc = 1;
    for (int64_t r = 0; r < n; ++r) {
        if (rs.ptr[r + 1] - rs.ptr[r] > maxc) maxc = rs.ptr[r + 1] - rs.ptr[r];
        if (s->column_start[r + 1] - s->column_start[r] > maxc)
            maxc = s->column_start[r + 1] - s->column_start[r];
    }
    double* t = (double*)malloc(sizeof(double) * (size_t)maxc);
    for (int64_t k = 0; k < n; ++k) y[k] = rhs[s->order[k]];
    for (int64_t k = 0; k < n; ++k) {
        int64_t m = 0;
        for (int64_t q = rs.ptr[k]; q < rs.ptr[k + 1]; ++q) t[m++] = f[rs.pos[q]] * y[rs.col[q]];
        double known = y[k] - pairwise(t, m);
        y[k] = known * (1.0 / f[s->column_start[k]]);
    }
    for (int64_t k = n - 1; k >= 0; --k) {
        int64_t m = 0;
        for (int64_t p = s->column_start[k] + 1; p < s->column_start[k + 1]; ++p)
            t[m++] = f[p] * y[s->row_rank[p]];
        double known = y[k] - pairwise(t, m);
        y[k] = known * (1.0 / f[s->column_start[k]]);
    }
    for (int64_t k = 0; k < n; ++k) x[s->order[k]] = y[k];
    free(y);
    free(t);
    free_rows(&rs);
}

/* ---------------------------------------------------------------- ROS2 */

/* The W-method's correction (d f / d phi) dphi on every field the route slopes reach: membrane
 * and junction transfers, cell amounts and the bath supplies. */
static void correct(const cells_world* w, const cells_state_layout* L, int32_t k, double coef,
                    const double* gm, const double* gj, const double* dphi, double* kv) {
    const int S = w->S;
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        double d = dphi[w->snd[e]] - dphi[w->rcv[e]];
        for (int s = 0; s < S; ++s)
            EF(L, w, kv, L->jtr + s, e) += coef * (gj[(int64_t)s * w->nedge + e] * d);
    }
    double sup[4] = {0, 0, 0, 0};
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        int64_t p0 = w->inc_ptr[i], p1 = w->inc_ptr[i + 1], pm = p0 + w->inc_out[i];
        for (int s = 0; s < S; ++s) {
            double dm = coef * (gm[(int64_t)s * w->ncell + i] * dphi[i]);
            double out = 0.0, in = 0.0;
            for (int64_t p = p0; p < pm; ++p) {
                int64_t e = w->inc_edge[p];
                out += -(coef * (gj[(int64_t)s * w->nedge + e] * (dphi[w->snd[e]] - dphi[w->rcv[e]])));
            }
            for (int64_t p = pm; p < p1; ++p) {
                int64_t e = w->inc_edge[p];
                in += coef * (gj[(int64_t)s * w->nedge + e] * (dphi[w->snd[e]] - dphi[w->rcv[e]]));
            }
            double cell = -dm;
            cell += out;
            cell += in;
            CF(L, w, kv, L->dn + s, i) += cell;
            CF(L, w, kv, L->mtr + s, i) += dm;
            sup[s] -= dm;
        }
    }
    for (int s = 0; s < S; ++s) WF(L, w, kv, L->sup + s, k) += sup[s];
}

/* The stage solve's charge residual (numerics/surface_imex.py:269-278 solve_operator's
 * (C + a L) phi - q, checked as models/surface_pump.py:325-330 checks it, CHARGE_ATOL_C 3e-25 and
 * CHARGE_RTOL 2e-12 of numerics/surface_field.py:47-48) of world k's stage x = dphi with
 * right-hand side q (world-local), the matrix's unscaled diagonal diag and junction slopes Gj.
 * The stage equation is in rate form, where q can vanish at a cell whose neighbours' terms do
 * not, so each row's scale is the largest of |q|, |M x| and the sum of the product's absolute
 * terms. */
static int stage_residual(const cells_world* w, int32_t k, double coef, const double* diag,
                          const double* Gj, const double* dphi, const double* q) {
    const int64_t c0 = w->cell_off[k], e0 = w->edge_off[k];
    int ok = 1;
    for (int64_t i = c0; i < w->cell_off[k + 1]; ++i) {
        double off = 0.0, offmag = 0.0;
        for (int64_t p = w->inc_ptr[i]; p < w->inc_ptr[i + 1]; ++p) {
            const int64_t e = w->inc_edge[p];
            const int64_t other = w->snd[e] == i ? w->rcv[e] : w->snd[e];
            const double t = Gj[e - e0] * dphi[other];
            off += t;
            offmag += fabs(t);
        }
        const double mx = diag[i - c0] * dphi[i] - coef * off;
        const double mag = fabs(diag[i - c0] * dphi[i]) + fabs(coef) * offmag;
        double scale = fabs(q[i - c0]) > fabs(mx) ? fabs(q[i - c0]) : fabs(mx);
        scale = scale > mag ? scale : mag;
        ok = ok && isfinite(mx) && fabs(mx - q[i - c0]) <= 3e-25 + 2e-12 * scale;
    }
    return ok;
}

static void charge_rate(const cells_world* w, const cells_state_layout* L, int32_t k,
                        const double* kv, double* q) {
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
        q[i - w->cell_off[k]] = cell_charge_increment(w, L, kv, i);
}

int ref_step_ros2(const cells_world* w, const cells_symbolic* sym, const double* y, double* out,
                  double dt, const cells_control* c, int32_t* valid_out) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const int S = w->S;
    const double coef = ROS2_GAMMA * dt;
    double* start = (double*)malloc(sizeof(double) * (size_t)L.total);
    double* k1 = (double*)calloc((size_t)L.total, sizeof(double));
    double* k2 = (double*)calloc((size_t)L.total, sizeof(double));
    double* y1 = (double*)malloc(sizeof(double) * (size_t)L.total);
    double* phi = (double*)malloc(sizeof(double) * (size_t)(w->ncell ? w->ncell : 1));
    double* gm = (double*)malloc(sizeof(double) * (size_t)(S * (w->ncell ? w->ncell : 1)));
    double* gj = (double*)malloc(sizeof(double) * (size_t)(S * (w->nedge ? w->nedge : 1)));
    int all = 1;
    for (int32_t k = 0; k < w->W; ++k) {
        const cells_symbolic* s = &sym[k];
        int64_t c0 = w->cell_off[k], n = w->cell_off[k + 1] - c0;
        int64_t e0 = w->edge_off[k];
        int valid = 1;
        copy_world(w, &L, k, y, start);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, start) && valid;
        ref_potentials(w, k, start, phi);
        valid = ref_rates(w, k, start, phi, c, k1, gm, gj) && valid;
        /* stage matrix M = C + gamma dt G on the lattice pattern, Jacobi-scaled and symmetrised
         * as surface_imex._factor_sparse (:228-259) */
        double* Gm = (double*)malloc(sizeof(double) * (size_t)n);
        double* Gj = (double*)malloc(sizeof(double) * (size_t)(w->edge_off[k + 1] - e0 + 1));
        for (int64_t i = 0; i < n; ++i) {
            double t = w->valence[0] * gm[c0 + i];
            for (int q = 1; q < S; ++q) t = t + w->valence[q] * gm[(int64_t)q * w->ncell + c0 + i];
            Gm[i] = CHARGE_PER_FMOL * t;
        }
        for (int64_t e = e0; e < w->edge_off[k + 1]; ++e) {
            double t = w->valence[0] * gj[e];
            for (int q = 1; q < S; ++q) t = t + w->valence[q] * gj[(int64_t)q * w->nedge + e];
            Gj[e - e0] = CHARGE_PER_FMOL * t;
        }
        double* diag = (double*)malloc(sizeof(double) * (size_t)n);
        double* sc = (double*)malloc(sizeof(double) * (size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            int64_t gi = c0 + i;
            int64_t p0 = w->inc_ptr[gi], p1 = w->inc_ptr[gi + 1];
            double g = Gm[i];
            for (int64_t p = p0; p < p1; ++p) g += Gj[w->inc_edge[p] - e0];
            diag[i] = w->cap[gi] + coef * g;
            valid = valid && diag[i] > 0;
            sc[i] = 1.0 / sqrt(diag[i] > 0 ? diag[i] : 1.0);
        }
        double* fv = (double*)calloc((size_t)s->nnz, sizeof(double));
        for (int64_t i = 0; i < n; ++i) {
            double v = diag[i];
            fv[s->diag_pos[i]] = (v * sc[i] * sc[i] + v * sc[i] * sc[i]) / 2;
        }
        /* each off-diagonal entry is the sum of the conductances of every junction between its
         * pair of cells, in junction order (numerics/surface_field.py:413-418), before the
         * scaling (surface_imex.py:248-250) */
        double* off = (double*)calloc((size_t)s->nnz, sizeof(double));
        for (int64_t e = e0; e < w->edge_off[k + 1]; ++e)
            off[s->edge_pos[e - e0]] += -(coef * Gj[e - e0]);
        for (int64_t e = e0; e < w->edge_off[k + 1]; ++e) {
            double v = off[s->edge_pos[e - e0]];
            int64_t a = w->snd[e] - c0, b = w->rcv[e] - c0;
            fv[s->edge_pos[e - e0]] = (v * sc[a] * sc[b] + v * sc[b] * sc[a]) / 2;
        }
        free(off);
        ref_factor(s, fv);
        for (int64_t p = 0; p < s->nnz; ++p) valid = valid && isfinite(fv[p]);
        double* q = (double*)malloc(sizeof(double) * (size_t)n);
        double* q0 = (double*)malloc(sizeof(double) * (size_t)n);
        double* x = (double*)malloc(sizeof(double) * (size_t)n);
        double* dphi = (double*)calloc((size_t)w->ncell, sizeof(double));
        /* stage 1 */
        charge_rate(w, &L, k, k1, q);
        memcpy(q0, q, sizeof(double) * (size_t)n);
        for (int64_t i = 0; i < n; ++i) q[i] = q[i] * sc[i];
        ref_solve(s, fv, q, x);
        for (int64_t i = 0; i < n; ++i) dphi[c0 + i] = x[i] * sc[i];
        valid = valid && stage_residual(w, k, coef, diag, Gj, dphi, q0);
        correct(w, &L, k, coef, gm, gj, dphi, k1);
        combine(w, &L, k, start, dt, k1, y1);
        /* stage 2: f(y + dt k1) - 2 k1 */
        ref_potentials(w, k, y1, phi);
        valid = ref_rates(w, k, y1, phi, c, k2, NULL, NULL) && valid;
        combine(w, &L, k, k2, -2.0, k1, k2);
        charge_rate(w, &L, k, k2, q);
        memcpy(q0, q, sizeof(double) * (size_t)n);
        for (int64_t i = 0; i < n; ++i) q[i] = q[i] * sc[i];
        ref_solve(s, fv, q, x);
        for (int64_t i = 0; i < n; ++i) dphi[c0 + i] = x[i] * sc[i];
        valid = valid && stage_residual(w, k, coef, diag, Gj, dphi, q0);
        correct(w, &L, k, coef, gm, gj, dphi, k2);
        /* y + (3/2) dt k1 + (1/2) dt k2 */
        copy_world(w, &L, k, start, out);
        combine(w, &L, k, start, 1.5 * dt, k1, out);
        combine(w, &L, k, out, 0.5 * dt, k2, out);
        for (int f = 0; f < L.world_fields; ++f)
            if (f < L.sup || f >= L.sup + S) WF(&L, w, out, f, k) = WF(&L, w, start, f, k);
        double applied[4];
        for (int q2 = 0; q2 < S; ++q2)
            applied[q2] = (1.5 * dt) * WF(&L, w, k1, L.sup + q2, k) + (0.5 * dt) * WF(&L, w, k2, L.sup + q2, k);
        ionic_tail(w, &L, k, start, out, applied);
        valid = valid && domain_valid(w, &L, k, out);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, out) && valid;
        if (!valid) copy_world(w, &L, k, y, out);
        valid_out[k] = valid;
        all = all && valid;
        free(Gm); free(Gj); free(diag); free(sc); free(fv); free(q); free(q0); free(x); free(dphi);
    }
    free(start); free(k1); free(k2); free(y1); free(phi); free(gm); free(gj);
    return all;
}

/* ================================================================ N2
 *
 * cells.solver.imex_ars222_implicit_conductances@1 (numerics/surface_imex.py:38-79: ARS(2,2,2)'s
 * arrays with the routes' operator time 0.5), as models/surface_pump_v3.py:3547-3955 imex_step runs
 * it for cells on one held bath (no faces, channels or kinds): the routes' slope conductances W,
 * frozen at the operator time (operator_time_conductances :2961-3028), make the implicit operator
 * C + gamma dt G (conductance_laplacian :2816-2837); the explicit stages carry the remainder
 * J - W dv (explicit_remainder :2905-2927), the implicit ones the linear parts W dv
 * (implicit_rates :2659-2700, linear_route_fluxes :2843-2857, route_amount_rates :2860-2872) with
 * their heat at the stage (implicit_route_heat :2930-2950); the junction gates take their
 * exponential update at each stage (stage_junction_gates :3127-3156, exponential_gates.advance
 * numerics/exponential_gates.py:44-61). */

#define GATE_SERIES_LIMIT 1e-3

/* junction_voltage_relaxation.py:35-48 kinetics: the target and the relaxation rate */
static void junction_kinetics(double v, double thr, double mn, double* target, double* relax) {
    double shift = fabs(v) - thr;
    double alpha = 1.3 * exp(-77.0 * shift);
    double log_beta = log(0.065) + 140.0 * shift;
    double beta = 20.0 * exp(-ref_logaddexp(0.0, -log_beta));
    double r = alpha + beta;
    *relax = r;
    *target = (alpha + beta * mn) / r;
}

/* exponential_gates.py:37-42 lag_weight: 1 - (1 - e^-z) / z, its series below 1e-3 */
double ref_lag_weight(double z) {
    if (z < GATE_SERIES_LIMIT) return z * (0.5 - z * (1.0 / 6 - z * (1.0 / 24 - z / 120)));
    return (expm1(-z) + z) / z;
}

/* exponential_gates.py:44-51 advance */
double ref_gate_advance(double gate, double target_start, double tau_start, double target_end,
                        double tau_end, double dt) {
    double z = 0.5 * dt * (1 / tau_start + 1 / tau_end);
    return gate + (target_start - gate) * -expm1(-z) + (target_end - target_start) * ref_lag_weight(z);
}

/* The routes' slopes W (conductance_weights :2758-2813) of world k at state y and potentials phi:
 * gm [S][ncell] (membranes, at the control's membrane availability) and gj [S][nedge] (junctions,
 * at the control's junction availability times the state's gate). Returns their finiteness. */
static int route_slopes(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y,
                        const double* phi, const cells_control* c, double* gm, double* gj) {
    const int S = w->S;
    const double T = w->temperature;
    int ok = 1;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double vol = cell_volume(w, L, y, i);
        for (int s = 0; s < S; ++s) {
            double ci = (w->n0[(int64_t)s * w->ncell + i] + CF(L, w, y, L->dn + s, i)) / vol;
            double g = ref_route_slope(ci, w->bath_c[s], phi[i] - w->bath_V, w->valence[s], T,
                                       w->pam[(int64_t)s * w->ncell + i], c->membrane);
            gm[(int64_t)s * w->ncell + i] = g;
            ok = ok && isfinite(g);
        }
    }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        double va = cell_volume(w, L, y, a), vb = cell_volume(w, L, y, b);
        double avail = c->junction * EF(L, w, y, L->gate, e);
        for (int s = 0; s < S; ++s) {
            double ca = (w->n0[(int64_t)s * w->ncell + a] + CF(L, w, y, L->dn + s, a)) / va;
            double cb = (w->n0[(int64_t)s * w->ncell + b] + CF(L, w, y, L->dn + s, b)) / vb;
            double g = ref_route_slope(ca, cb, phi[a] - phi[b], w->valence[s], T,
                                       w->paj[(int64_t)s * w->nedge + e], avail);
            gj[(int64_t)s * w->nedge + e] = g;
            ok = ok && isfinite(g);
        }
    }
    return ok;
}

/* The linear parts W dv at potentials phi of world k, as rates r (every other field zero): amounts
 * (route_amount_rates: the junctions into a cell, less those out of it, less its membrane's),
 * membrane and junction transfers, the bath supply (minus the membranes' inflow) and, from the
 * state `at` (its concentrations) with phi, their heat (route_heat_rates :2890-2902). */
static void linear_parts(const cells_world* w, const cells_state_layout* L, int32_t k, const double* gm,
                         const double* gj, const double* phi, const double* at, double* r) {
    const int S = w->S;
    const double RT = R_CONST * w->temperature;
    for (int f = 0; f < L->cell_fields; ++f)
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) CF(L, w, r, f, i) = 0.0;
    for (int f = 0; f < L->edge_fields; ++f)
        for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) EF(L, w, r, f, e) = 0.0;
    for (int f = 0; f < L->world_fields; ++f) WF(L, w, r, f, k) = 0.0;
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        double dj = phi[a] - phi[b];
        double va = cell_volume(w, L, at, a), vb = cell_volume(w, L, at, b);
        double heat = 0.0;
        for (int s = 0; s < S; ++s) {
            double j = gj[(int64_t)s * w->nedge + e] * dj;
            EF(L, w, r, L->jtr + s, e) = j;
            double ca = (w->n0[(int64_t)s * w->ncell + a] + CF(L, w, at, L->dn + s, a)) / va;
            double cb = (w->n0[(int64_t)s * w->ncell + b] + CF(L, w, at, L->dn + s, b)) / vb;
            double mu = RT * (log(ca) - log(cb)) + w->valence[s] * F_CONST * dj;
            heat = s == 0 ? j * mu : heat + j * mu;
        }
        EF(L, w, r, L->jdiss, e) = 1e-3 * heat;
    }
    double inflow[4] = {0, 0, 0, 0};
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        int64_t p0 = w->inc_ptr[i], p1 = w->inc_ptr[i + 1], pm = p0 + w->inc_out[i];
        double dv = phi[i] - w->bath_V;
        double vol = cell_volume(w, L, at, i);
        double heat = 0.0;
        for (int s = 0; s < S; ++s) {
            double m = gm[(int64_t)s * w->ncell + i] * dv;
            double in = 0.0, out = 0.0;
            for (int64_t p = pm; p < p1; ++p) in += EF(L, w, r, L->jtr + s, w->inc_edge[p]);
            for (int64_t p = p0; p < pm; ++p) out += EF(L, w, r, L->jtr + s, w->inc_edge[p]);
            CF(L, w, r, L->dn + s, i) = (in - out) - m;
            CF(L, w, r, L->mtr + s, i) = m;
            inflow[s] += m;
            double ci = (w->n0[(int64_t)s * w->ncell + i] + CF(L, w, at, L->dn + s, i)) / vol;
            double mu = RT * (log(ci) - log(w->bath_c[s])) + w->valence[s] * F_CONST * dv;
            heat = s == 0 ? m * mu : heat + m * mu;
        }
        CF(L, w, r, L->mdiss, i) = 1e-3 * heat;
    }
    for (int s = 0; s < S; ++s) WF(L, w, r, L->sup + s, k) = -inflow[s];
}

/* rates - lin on the fields the linear parts reach (explicit_remainder) */
static void explicit_remainder(const cells_world* w, const cells_state_layout* L, int32_t k, const double* lin,
                      double* rates) {
    const int S = w->S;
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        for (int s = 0; s < S; ++s) {
            CF(L, w, rates, L->dn + s, i) = CF(L, w, rates, L->dn + s, i) - CF(L, w, lin, L->dn + s, i);
            CF(L, w, rates, L->mtr + s, i) = CF(L, w, rates, L->mtr + s, i) - CF(L, w, lin, L->mtr + s, i);
        }
        CF(L, w, rates, L->mdiss, i) = CF(L, w, rates, L->mdiss, i) - CF(L, w, lin, L->mdiss, i);
    }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        for (int s = 0; s < S; ++s)
            EF(L, w, rates, L->jtr + s, e) = EF(L, w, rates, L->jtr + s, e) - EF(L, w, lin, L->jtr + s, e);
        EF(L, w, rates, L->jdiss, e) = EF(L, w, rates, L->jdiss, e) - EF(L, w, lin, L->jdiss, e);
    }
    for (int s = 0; s < S; ++s) WF(L, w, rates, L->sup + s, k) = WF(L, w, rates, L->sup + s, k) - WF(L, w, lin, L->sup + s, k);
}

/* The operator's stage: potentials phi (cells of world k) of the node charges of `known`, solving
 * (C + coef G) phi = q on the cells with the bath held at V_bath (solve_operator, surface_imex.py:
 * 269-278), and the residual check |(C + coef G) phi - q| <= 3e-25 + 2e-12 max(|q|, |M phi|). */
static int operator_stage(const cells_world* w, const cells_state_layout* L, int32_t k, const cells_symbolic* s,
                          double coef, const double* Gm, const double* Gj, const double* diag, const double* sc,
                          const double* fv, const double* known, double* phi) {
    const int64_t c0 = w->cell_off[k], n = w->cell_off[k + 1] - c0, e0 = w->edge_off[k];
    double* q = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double* rhs = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double* x = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        q[i] = w->q0[c0 + i] + cell_charge_increment(w, L, known, c0 + i);
        double held = -(w->cap[c0 + i] + coef * Gm[i]) * w->bath_V; /* M's bath column at V_bath */
        rhs[i] = (q[i] - held) * sc[i];
    }
    ref_solve(s, fv, rhs, x);
    for (int64_t i = 0; i < n; ++i) phi[c0 + i] = x[i] * sc[i];
    int ok = 1;
    for (int64_t i = 0; i < n; ++i) {
        int64_t gi = c0 + i;
        double off = 0.0;
        for (int64_t p = w->inc_ptr[gi]; p < w->inc_ptr[gi + 1]; ++p) {
            int64_t e = w->inc_edge[p];
            int64_t other = w->snd[e] == gi ? w->rcv[e] : w->snd[e];
            off += Gj[e - e0] * phi[other];
        }
        double held = -(w->cap[gi] + coef * Gm[i]) * w->bath_V;
        double mx = (diag[i] * phi[gi] - coef * off) + held;
        double scale = fabs(q[i]) > fabs(mx) ? fabs(q[i]) : fabs(mx);
        ok = ok && isfinite(phi[gi]) && isfinite(mx) && fabs(mx - q[i]) <= 3e-25 + 2e-12 * scale;
    }
    free(q); free(rhs); free(x);
    return ok;
}

/* y_out = y + (h r) on every cell, edge and supply field (the reference's _combine, one term) */
static void combine_into(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y,
                         double h, const double* r, double* out) {
    combine(w, L, k, y, h, r, out);
}

int ref_step_n2(const cells_world* w, const cells_symbolic* sym, const double* y, double* out,
                double dt, const cells_control* c, int32_t* valid_out) {
    cells_state_layout L;
    ref_state_layout(w, &L);
    const int S = w->S;
    const double gamma = ARS_GAMMA, delta = ARS_DELTA, coef = gamma * dt, theta = 0.5 * dt;
    /* the tableau's rows (surface_imex.py:47-57) and each row's explicit fraction, summed in order */
    const double ae[3][3] = {{0, 0, 0}, {gamma, 0, 0}, {delta, 1.0 - delta, 0}};
    const double ai[3][3] = {{0, 0, 0}, {0, gamma, 0}, {0, 1.0 - gamma, gamma}};
    const size_t T = (size_t)L.total;
    double* start = (double*)malloc(sizeof(double) * T);
    double* ahead = (double*)malloc(sizeof(double) * T);
    double* E0 = (double*)calloc(T, sizeof(double));
    double* E1 = (double*)calloc(T, sizeof(double));
    double* D1 = (double*)calloc(T, sizeof(double));
    double* D2 = (double*)calloc(T, sizeof(double));
    double* lin = (double*)calloc(T, sizeof(double));
    double* known = (double*)malloc(sizeof(double) * T);
    double* stage = (double*)malloc(sizeof(double) * T);
    const size_t nc = (size_t)(w->ncell ? w->ncell : 1), ne = (size_t)(w->nedge ? w->nedge : 1);
    double* phi0 = (double*)malloc(sizeof(double) * nc);
    double* phi = (double*)malloc(sizeof(double) * nc);
    double* gm = (double*)malloc(sizeof(double) * (size_t)S * nc);
    double* gj = (double*)malloc(sizeof(double) * (size_t)S * ne);
    double* target0 = (double*)malloc(sizeof(double) * ne);
    double* relax0 = (double*)malloc(sizeof(double) * ne);
    int all = 1;
    for (int32_t k = 0; k < w->W; ++k) {
        const cells_symbolic* s = &sym[k];
        const int64_t c0 = w->cell_off[k], n = w->cell_off[k + 1] - c0, e0 = w->edge_off[k];
        const int64_t e1 = w->edge_off[k + 1];
        int valid = 1;
        copy_world(w, &L, k, y, start);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, start) && valid;
        /* the step's start: its potentials and rates */
        ref_potentials(w, k, start, phi0);
        valid = ref_rates(w, k, start, phi0, c, E0, NULL, NULL) && valid;
        /* the operator-time prediction (operator_time_conductances): amounts ahead by theta times
         * the start's rates, each junction gate by its exponential relaxation at the start's
         * junction voltage, the potentials of the predicted charges */
        copy_world(w, &L, k, start, ahead);
        for (int s2 = 0; s2 < S; ++s2)
            for (int64_t i = c0; i < c0 + n; ++i)
                CF(&L, w, ahead, L.dn + s2, i) = CF(&L, w, start, L.dn + s2, i) + theta * CF(&L, w, E0, L.dn + s2, i);
        for (int64_t e = e0; e < e1; ++e) {
            junction_kinetics(phi0[w->snd[e]] - phi0[w->rcv[e]], w->gate_threshold, w->gate_minimum,
                              &target0[e], &relax0[e]);
            double g = EF(&L, w, start, L.gate, e);
            EF(&L, w, ahead, L.gate, e) = g + (target0[e] - g) * -expm1(-relax0[e] * theta);
        }
        for (int64_t i = c0; i < c0 + n; ++i)
            for (int s2 = 0; s2 < S; ++s2) {
                double ci = w->n0[(int64_t)s2 * w->ncell + i] + CF(&L, w, ahead, L.dn + s2, i);
                valid = valid && isfinite(ci) && ci > 0;
            }
        ref_potentials(w, k, ahead, phi);
        for (int64_t i = c0; i < c0 + n; ++i) valid = valid && isfinite(phi[i]);
        valid = route_slopes(w, &L, k, ahead, phi, c, gm, gj) && valid;
        /* the operator C + gamma dt G, Jacobi-scaled and symmetrised (surface_imex.py:228-259) */
        double* Gm = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
        double* Gj = (double*)malloc(sizeof(double) * (size_t)(e1 - e0 + 1));
        for (int64_t i = 0; i < n; ++i) {
            double t = w->valence[0] * gm[c0 + i];
            for (int q = 1; q < S; ++q) t = t + w->valence[q] * gm[(int64_t)q * w->ncell + c0 + i];
            Gm[i] = CHARGE_PER_FMOL * t;
        }
        for (int64_t e = e0; e < e1; ++e) {
            double t = w->valence[0] * gj[e];
            for (int q = 1; q < S; ++q) t = t + w->valence[q] * gj[(int64_t)q * w->nedge + e];
            Gj[e - e0] = CHARGE_PER_FMOL * t;
        }
        double* diag = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
        double* sc = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
        for (int64_t i = 0; i < n; ++i) {
            int64_t gi = c0 + i;
            double g = Gm[i];
            for (int64_t p = w->inc_ptr[gi]; p < w->inc_ptr[gi + 1]; ++p) g += Gj[w->inc_edge[p] - e0];
            diag[i] = w->cap[gi] + coef * g;
            valid = valid && diag[i] > 0;
            sc[i] = 1.0 / sqrt(diag[i] > 0 ? diag[i] : 1.0);
        }
        double* fv = (double*)calloc((size_t)s->nnz, sizeof(double));
        for (int64_t i = 0; i < n; ++i) {
            double v = diag[i];
            fv[s->diag_pos[i]] = (v * sc[i] * sc[i] + v * sc[i] * sc[i]) / 2;
        }
        for (int64_t e = e0; e < e1; ++e) {
            double v = -(coef * Gj[e - e0]);
            int64_t a = w->snd[e] - c0, b = w->rcv[e] - c0;
            fv[s->edge_pos[e - e0]] = (v * sc[a] * sc[b] + v * sc[b] * sc[a]) / 2;
        }
        ref_factor(s, fv);
        for (int64_t p = 0; p < s->nnz; ++p) valid = valid && isfinite(fv[p]);
        /* stage 0's explicit rates: the start's less the linear parts at the start */
        linear_parts(w, &L, k, gm, gj, phi0, start, lin);
        explicit_remainder(w, &L, k, lin, E0);
        double* Es[2] = {E0, E1};
        double* Ds[3] = {NULL, D1, D2};
        for (int i = 1; i < 3; ++i) {
            /* known = state + dt (explicit terms, then implicit terms), one term at a time */
            copy_world(w, &L, k, start, known);
            for (int j = 0; j < i; ++j) combine_into(w, &L, k, known, ae[i][j] * dt, Es[j], known);
            for (int j = 1; j < i; ++j) combine_into(w, &L, k, known, ai[i][j] * dt, Ds[j], known);
            valid = operator_stage(w, &L, k, s, coef, Gm, Gj, diag, sc, fv, known, phi) && valid;
            /* the linear parts at the stage potentials, their heat at the stage they reach */
            linear_parts(w, &L, k, gm, gj, phi, known, Ds[i]);
            copy_world(w, &L, k, known, stage);
            combine_into(w, &L, k, known, ai[i][i] * dt, Ds[i], stage);
            linear_parts(w, &L, k, gm, gj, phi, stage, lin);
            for (int64_t i2 = c0; i2 < c0 + n; ++i2) CF(&L, w, Ds[i], L.mdiss, i2) = CF(&L, w, lin, L.mdiss, i2);
            for (int64_t e = e0; e < e1; ++e) EF(&L, w, Ds[i], L.jdiss, e) = EF(&L, w, lin, L.jdiss, e);
            copy_world(w, &L, k, known, stage);
            combine_into(w, &L, k, known, ai[i][i] * dt, Ds[i], stage);
            /* the junction gates' exponential update from the start to the stage */
            const double fraction = (ae[i][0] + ae[i][1]) + ae[i][2];
            for (int64_t e = e0; e < e1; ++e) {
                double v = phi[w->snd[e]] - phi[w->rcv[e]], t1, r1;
                junction_kinetics(v, w->gate_threshold, w->gate_minimum, &t1, &r1);
                double g = ref_gate_advance(EF(&L, w, start, L.gate, e), target0[e], 1.0 / relax0[e], t1,
                                            1.0 / r1, fraction * dt);
                EF(&L, w, stage, L.gate, e) = g;
                valid = valid && gate_domain(g, v, w->gate_threshold, w->gate_minimum, c->junction);
            }
            if (i < 2) {
                valid = ref_rates(w, k, stage, phi, c, E1, NULL, NULL) && valid;
                linear_parts(w, &L, k, gm, gj, phi, stage, lin);
                explicit_remainder(w, &L, k, lin, E1);
            }
        }
        /* the tail: the stage is the new state; electrode and reservoir ledgers from the applied
         * supply (the last row's weights) */
        copy_world(w, &L, k, stage, out);
        for (int f = 0; f < L.world_fields; ++f)
            if (f < L.sup || f >= L.sup + S) WF(&L, w, out, f, k) = WF(&L, w, start, f, k);
        double applied[4];
        for (int q2 = 0; q2 < S; ++q2) {
            double ex = (0.0 + ae[2][0] * WF(&L, w, E0, L.sup + q2, k)) + ae[2][1] * WF(&L, w, E1, L.sup + q2, k);
            double im = (0.0 + ai[2][1] * WF(&L, w, D1, L.sup + q2, k)) + ai[2][2] * WF(&L, w, D2, L.sup + q2, k);
            applied[q2] = dt * (ex + im);
        }
        ionic_tail(w, &L, k, start, out, applied);
        valid = valid && domain_valid(w, &L, k, out);
        if (w->water) valid = water_half(w, &L, k, 0.5 * dt, c, out) && valid;
        if (!valid) copy_world(w, &L, k, y, out);
        valid_out[k] = valid;
        all = all && valid;
        free(Gm); free(Gj); free(diag); free(sc); free(fv);
    }
    free(start); free(ahead); free(E0); free(E1); free(D1); free(D2); free(lin); free(known); free(stage);
    free(phi0); free(phi); free(gm); free(gj); free(target0); free(relax0);
    return all;
}

/* ================================================================ adjoint of the ROS2 step
 *
 * The reference has no adjoint (it differentiates only forward, analysis/stationary.py:505-517);
 * this is the exact reverse of ref_step_ros2's arithmetic restricted to the fields the dynamics
 * read: the amounts dN [S][n] and the junction gates [E] (3-ion worlds without water; the
 * ledgers never feed back). Parameter cotangents are taken with respect to the membrane
 * area x permeability pam [S][n], the Na/K capacity [n] and the junction area x permeability
 * paj [S][E]. It is pinned by central finite differences of ref_step_ros2 (tests).
 */

static double bernoulli_second(double x) {
    /* B''(x) of B(x) = x / (e^x - 1): the Bernoulli-number series below |x| = 0.1, else
     * e^x (x (e^x + 1) - 2 (e^x - 1)) / (e^x - 1)^3 in the form that stays finite for large |x|. */
    if (fabs(x) < 0.1) {
        double x2 = x * x;
        return 1.0 / 6.0 - x2 / 60.0 + x2 * x2 / 1008.0 - x2 * x2 * x2 / 21600.0 +
               x2 * x2 * x2 * x2 / 532224.0;
    }
    if (x > 0) {
        double em = exp(-x), d = 1.0 - em;
        return em * (x * (1.0 + em) - 2.0 * d) / (d * d * d);
    }
    double e = exp(x), d = e - 1.0;
    return e * (x * (e + 1.0) - 2.0 * d) / (d * d * d);
}

double ref_bernoulli_second(double x) { return bernoulli_second(x); }

typedef struct {
    double *cbar, *vbar; /* [S][n], [n] cotangents of concentrations and potentials */
} local_bars;

/* Cotangents of f_N (amount rates) and f_g (gate rates) at y pulled back to concentrations,
 * potentials, gates and parameters; adds. */
static void rates_vjp(const cells_world* w, const cells_state_layout* L, int32_t k, const double* y,
                      const double* phi, const cells_control* c, const double* fN, const double* fg,
                      const double* gmbar, const double* gjbar, double* cbar, double* vbar,
                      double* gatebar, double* pam_bar, double* cap_bar, double* paj_bar) {
    const int S = w->S;
    const double T = w->temperature, RT = R_CONST * T;
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        double va = cell_volume(w, L, y, a), vb = cell_volume(w, L, y, b);
        double dv = phi[a] - phi[b];
        double g = EF(L, w, y, L->gate, e);
        double vsum = 0.0;
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s], kk = z * F_CONST / RT, x = kk * dv;
            double ca = (w->n0[(int64_t)s * w->ncell + a] + CF(L, w, y, L->dn + s, a)) / va;
            double cb = (w->n0[(int64_t)s * w->ncell + b] + CF(L, w, y, L->dn + s, b)) / vb;
            double pa = w->paj[(int64_t)s * w->nedge + e];
            double unit = c->junction / MOL_PER_FMOL; /* J = pa g unit (...) */
            double Bm = ref_bernoulli(-x), Bp = ref_bernoulli(x);
            double Sm = ref_bernoulli_slope(-x), Sp = ref_bernoulli_slope(x);
            double Qm = bernoulli_second(-x), Qp = bernoulli_second(x);
            double A = pa * g * unit;
            double core = ca * Bm - cb * Bp;
            double gj = A * kk * (-ca * Sm - cb * Sp);
            /* J_e leaves a (sender) and enters b */
            double Jbar = (fN ? fN[(int64_t)s * w->ncell + b] - fN[(int64_t)s * w->ncell + a] : 0.0);
            double Gbar = gjbar ? gjbar[(int64_t)s * w->nedge + e] : 0.0;
            cbar[(int64_t)s * w->ncell + a] += Jbar * A * Bm + Gbar * (-A * kk * Sm);
            cbar[(int64_t)s * w->ncell + b] += Jbar * (-A * Bp) + Gbar * (-A * kk * Sp);
            double dv_bar = Jbar * gj + Gbar * (A * kk * kk * (ca * Qm - cb * Qp));
            vsum += dv_bar;
            gatebar[e] += Jbar * (pa * unit * core) + Gbar * (pa * unit * kk * (-ca * Sm - cb * Sp));
            if (paj_bar)
                paj_bar[(int64_t)s * w->nedge + e] += Jbar * (g * unit * core) +
                                                      Gbar * (g * unit * kk * (-ca * Sm - cb * Sp));
        }
        if (fg) {
            double shift = fabs(dv) - w->gate_threshold;
            double alpha = 1.3 * exp(-77.0 * shift);
            double log_beta = log(0.065) + 140.0 * shift;
            double beta = 20.0 * exp(-ref_logaddexp(0.0, -log_beta));
            double gb = fg[e];
            gatebar[e] += gb * -(alpha + beta);
            double sgn = dv > 0 ? 1.0 : (dv < 0 ? -1.0 : 0.0);
            double dal = -77.0 * alpha, dbe = 140.0 * beta * (1.0 - beta / 20.0);
            vsum += gb * (dal * (1.0 - g) - dbe * (g - w->gate_minimum)) * sgn;
        }
        vbar[a] += vsum;
        vbar[b] -= vsum;
    }
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double vol = cell_volume(w, L, y, i);
        double dv = phi[i] - w->bath_V;
        double ci[4];
        for (int s = 0; s < S; ++s)
            ci[s] = (w->n0[(int64_t)s * w->ncell + i] + CF(L, w, y, L->dn + s, i)) / vol;
        double unit = c->membrane / MOL_PER_FMOL;
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s], kk = z * F_CONST / RT, x = kk * dv, co = w->bath_c[s];
            double pa = w->pam[(int64_t)s * w->ncell + i];
            double A = pa * unit;
            double Bm = ref_bernoulli(-x), Bp = ref_bernoulli(x);
            double Sm = ref_bernoulli_slope(-x), Sp = ref_bernoulli_slope(x);
            double Qm = bernoulli_second(-x), Qp = bernoulli_second(x);
            double gm = A * kk * (-ci[s] * Sm - co * Sp);
            double Jbar = fN ? -fN[(int64_t)s * w->ncell + i] : 0.0; /* f_N -= J_m */
            double Gbar = gmbar ? gmbar[(int64_t)s * w->ncell + i] : 0.0;
            cbar[(int64_t)s * w->ncell + i] += Jbar * A * Bm + Gbar * (-A * kk * Sm);
            vbar[i] += Jbar * gm + Gbar * (A * kk * kk * (ci[s] * Qm - co * Qp));
            if (pam_bar)
                pam_bar[(int64_t)s * w->ncell + i] += Jbar * (unit * (ci[s] * Bm - co * Bp)) +
                                                      Gbar * (unit * kk * (-ci[s] * Sm - co * Sp));
        }
        if (fN) {
            /* Na/K-ATPase: f_N,Na -= 3 cyc, f_N,K += 2 cyc */
            double cyc_bar = -3.0 * fN[i] + 2.0 * fN[w->ncell + i];
            double cap = w->nak_cap[i], av = c->pump;
            if (cap > 0 && av > 0 && cyc_bar != 0.0) {
                double mu_na = RT * (log(ci[0]) - log(w->bath_c[0])) + w->valence[0] * F_CONST * dv;
                double mu_k = RT * (log(ci[1]) - log(w->bath_c[1])) + w->valence[1] * F_CONST * dv;
                double e = nak_drive(w) + (-3.0) * mu_na + 2.0 * mu_k, a = e / RT;
                double r = pow(w->km_na / ci[0], 3.0);
                double s1 = 1.0 / (1.0 + r);
                double s2 = 1.0 / (1.0 + pow(w->km_k / w->bath_c[1], 2.0));
                double s3 = 1.0 / (1.0 + w->km_atp / w->atp);
                double sat = s1 * s2 * s3;
                double ea = exp(a);
                double abar = cyc_bar * (-cap * av * sat * ea);
                double mu_na_bar = abar * (-3.0 / RT), mu_k_bar = abar * (2.0 / RT);
                double ds1 = (3.0 * r / ci[0]) * s1 * s1;
                cbar[i] += cyc_bar * cap * av * ds1 * s2 * s3 * (1.0 - ea) + mu_na_bar * RT / ci[0];
                cbar[w->ncell + i] += mu_k_bar * RT / ci[1];
                vbar[i] += (mu_na_bar * w->valence[0] + mu_k_bar * w->valence[1]) * F_CONST;
                if (cap_bar) cap_bar[i] += cyc_bar * av * sat * (1.0 - ea);
            }
        }
    }
}

/* dN cotangents from concentration and potential cotangents (amounts over volume; potentials
 * q / C with q = q0 + F' sum_s z_s dN). */
static void pull_to_amounts(const cells_world* w, const cells_state_layout* L, int32_t k,
                            const double* y, const double* cbar, const double* vbar, double* Nbar) {
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double vol = cell_volume(w, L, y, i);
        for (int s = 0; s < w->S; ++s)
            Nbar[(int64_t)s * w->ncell + i] +=
                cbar[(int64_t)s * w->ncell + i] / vol + vbar[i] * CHARGE_PER_FMOL * w->valence[s] / w->cap[i];
    }
}

/* u -> D^T u: cotangent on the potential correction from a cotangent u [S][n] on the amounts. */
static void dtrans(const cells_world* w, int32_t k, const double* gm, const double* gj,
                   const double* u, double* out /* [n] global index */) {
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) {
        double t = 0.0;
        for (int s = 0; s < w->S; ++s) t += -gm[(int64_t)s * w->ncell + i] * u[(int64_t)s * w->ncell + i];
        out[i] = t;
    }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        double t = 0.0;
        for (int s = 0; s < w->S; ++s)
            t += gj[(int64_t)s * w->nedge + e] * (u[(int64_t)s * w->ncell + b] - u[(int64_t)s * w->ncell + a]);
        out[a] += t;
        out[b] -= t;
    }
}

/* slope cotangents of d(D(g) delta)/dg applied to u. */
static void dslopes(const cells_world* w, int32_t k, const double* delta, const double* u,
                    double coef, double* gmbar, double* gjbar) {
    for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i)
        for (int s = 0; s < w->S; ++s)
            gmbar[(int64_t)s * w->ncell + i] += coef * (u[(int64_t)s * w->ncell + i] * -delta[i]);
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        for (int s = 0; s < w->S; ++s)
            gjbar[(int64_t)s * w->nedge + e] +=
                coef * ((delta[a] - delta[b]) * (u[(int64_t)s * w->ncell + b] - u[(int64_t)s * w->ncell + a]));
    }
}

/* The solve's cotangent: w = M^-1 dbar (M symmetric), then rhs cotangent q^T w onto the amounts'
 * rate cotangent, and M's cotangent -w delta^T onto the slopes (through G = F' sum z slope). */
static void solve_vjp(const cells_world* w, const cells_symbolic* s, int32_t k, const double* fv,
                      const double* sc, double coef, const double* dbar, const double* delta,
                      double* rate_bar, double* gmbar, double* gjbar) {
    int64_t c0 = w->cell_off[k], n = w->cell_off[k + 1] - c0;
    double* rhs = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
    double* x = (double*)malloc(sizeof(double) * (size_t)n);
    double* wv = (double*)calloc((size_t)w->ncell, sizeof(double));
    for (int64_t i = 0; i < n; ++i) rhs[i] = dbar[c0 + i] * sc[i];
    ref_solve(s, fv, rhs, x);
    for (int64_t i = 0; i < n; ++i) wv[c0 + i] = x[i] * sc[i];
    for (int64_t i = c0; i < c0 + n; ++i)
        for (int q = 0; q < w->S; ++q) {
            rate_bar[(int64_t)q * w->ncell + i] += CHARGE_PER_FMOL * w->valence[q] * wv[i];
            gmbar[(int64_t)q * w->ncell + i] += CHARGE_PER_FMOL * w->valence[q] * (-coef * wv[i] * delta[i]);
        }
    for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) {
        int64_t a = w->snd[e], b = w->rcv[e];
        double t = -coef * (wv[a] - wv[b]) * (delta[a] - delta[b]);
        for (int q = 0; q < w->S; ++q) gjbar[(int64_t)q * w->nedge + e] += CHARGE_PER_FMOL * w->valence[q] * t;
    }
    free(rhs);
    free(x);
    free(wv);
}

int ref_ros2_vjp(const cells_world* w, const cells_symbolic* sym, const double* y, double dt,
                 const cells_control* c, const double* ybar_out, double* ybar_in, double* pam_bar,
                 double* cap_bar, double* paj_bar) {
    if (w->water || w->calcium) return 0;
    cells_state_layout L;
    ref_state_layout(w, &L);
    const int S = w->S;
    const double coef = ROS2_GAMMA * dt;
    const int64_t nc = w->ncell, ne = w->nedge;
    double* k1 = (double*)calloc((size_t)L.total, sizeof(double));
    double* f1 = (double*)calloc((size_t)L.total, sizeof(double));
    double* y1 = (double*)malloc(sizeof(double) * (size_t)L.total);
    double* phi = (double*)malloc(sizeof(double) * (size_t)nc);
    double* phi1 = (double*)malloc(sizeof(double) * (size_t)nc);
    double* gm = (double*)malloc(sizeof(double) * (size_t)(S * nc));
    double* gj = (double*)malloc(sizeof(double) * (size_t)(S * (ne ? ne : 1)));
    double* d1 = (double*)calloc((size_t)nc, sizeof(double));
    double* d2 = (double*)calloc((size_t)nc, sizeof(double));
    double* gmbar = (double*)calloc((size_t)(S * nc), sizeof(double));
    double* gjbar = (double*)calloc((size_t)(S * (ne ? ne : 1)), sizeof(double));
    double* k1bar = (double*)calloc((size_t)(S * nc), sizeof(double));
    double* k1gbar = (double*)calloc((size_t)(ne ? ne : 1), sizeof(double));
    double* r2bar = (double*)calloc((size_t)(S * nc), sizeof(double));
    double* cbar = (double*)calloc((size_t)(S * nc), sizeof(double));
    double* vbar = (double*)calloc((size_t)nc, sizeof(double));
    double* dbar = (double*)calloc((size_t)nc, sizeof(double));
    double* x1bar = (double*)calloc((size_t)(S * nc), sizeof(double));
    double* g1bar = (double*)calloc((size_t)(ne ? ne : 1), sizeof(double));
    int ok = 1;
    for (int32_t k = 0; k < w->W; ++k) {
        const cells_symbolic* s = &sym[k];
        int64_t c0 = w->cell_off[k], n = w->cell_off[k + 1] - c0, e0 = w->edge_off[k];
        int64_t e1 = w->edge_off[k + 1];
        /* ---- forward recomputation (ref_step_ros2's arithmetic) */
        ref_potentials(w, k, y, phi);
        ok = ref_rates(w, k, y, phi, c, k1, gm, gj) && ok;
        double* Gj = (double*)malloc(sizeof(double) * (size_t)(e1 - e0 + 1));
        double* sc = (double*)malloc(sizeof(double) * (size_t)n);
        double* fv = (double*)calloc((size_t)s->nnz, sizeof(double));
        for (int64_t e = e0; e < e1; ++e) {
            double t = w->valence[0] * gj[e];
            for (int q = 1; q < S; ++q) t = t + w->valence[q] * gj[(int64_t)q * ne + e];
            Gj[e - e0] = CHARGE_PER_FMOL * t;
        }
        for (int64_t i = 0; i < n; ++i) {
            int64_t gi = c0 + i;
            double t = w->valence[0] * gm[gi];
            for (int q = 1; q < S; ++q) t = t + w->valence[q] * gm[(int64_t)q * nc + gi];
            double g = CHARGE_PER_FMOL * t;
            for (int64_t p = w->inc_ptr[gi]; p < w->inc_ptr[gi + 1]; ++p) g += Gj[w->inc_edge[p] - e0];
            double d = w->cap[gi] + coef * g;
            sc[i] = 1.0 / sqrt(d > 0 ? d : 1.0);
            fv[s->diag_pos[i]] = (d * sc[i] * sc[i] + d * sc[i] * sc[i]) / 2;
        }
        for (int64_t e = e0; e < e1; ++e) {
            double v = -(coef * Gj[e - e0]);
            int64_t a = w->snd[e] - c0, b = w->rcv[e] - c0;
            fv[s->edge_pos[e - e0]] = (v * sc[a] * sc[b] + v * sc[b] * sc[a]) / 2;
        }
        ref_factor(s, fv);
        double* q = (double*)malloc(sizeof(double) * (size_t)n);
        double* x = (double*)malloc(sizeof(double) * (size_t)n);
        charge_rate(w, &L, k, k1, q);
        for (int64_t i = 0; i < n; ++i) q[i] = q[i] * sc[i];
        ref_solve(s, fv, q, x);
        for (int64_t i = 0; i < n; ++i) d1[c0 + i] = x[i] * sc[i];
        correct(w, &L, k, coef, gm, gj, d1, k1);
        combine(w, &L, k, y, dt, k1, y1);
        ref_potentials(w, k, y1, phi1);
        ok = ref_rates(w, k, y1, phi1, c, f1, NULL, NULL) && ok;
        /* r2 = f1 - 2 k1 (amounts) and its solve */
        for (int64_t i = 0; i < n; ++i) {
            int64_t gi = c0 + i;
            double t = w->valence[0] * (CF(&L, w, f1, L.dn, gi) + -2.0 * CF(&L, w, k1, L.dn, gi));
            for (int qq = 1; qq < S; ++qq)
                t = t + w->valence[qq] * (CF(&L, w, f1, L.dn + qq, gi) + -2.0 * CF(&L, w, k1, L.dn + qq, gi));
            q[i] = CHARGE_PER_FMOL * t * sc[i];
        }
        ref_solve(s, fv, q, x);
        for (int64_t i = 0; i < n; ++i) d2[c0 + i] = x[i] * sc[i];
        /* ---- reverse */
        for (int64_t i = c0; i < c0 + n; ++i)
            for (int qq = 0; qq < S; ++qq) {
                double xb = CF(&L, w, ybar_out, L.dn + qq, i);
                CF(&L, w, ybar_in, L.dn + qq, i) = xb;
                k1bar[(int64_t)qq * nc + i] = 1.5 * dt * xb;
                r2bar[(int64_t)qq * nc + i] = 0.5 * dt * xb; /* k2bar, then r2bar */
            }
        for (int64_t e = e0; e < e1; ++e) {
            double gb = EF(&L, w, ybar_out, L.gate, e);
            EF(&L, w, ybar_in, L.gate, e) = gb;
            k1gbar[e] = 1.5 * dt * gb;
            g1bar[e] = 0.5 * dt * gb; /* k2's gate cotangent = r2's = f1's */
        }
        /* k2 = r2 + coef D delta2: delta2bar = coef D^T k2bar; slopes */
        dtrans(w, k, gm, gj, r2bar, dbar);
        for (int64_t i = c0; i < c0 + n; ++i) dbar[i] *= coef;
        dslopes(w, k, d2, r2bar, coef, gmbar, gjbar);
        /* delta2 = M^-1 q(r2): r2bar += q^T w2; slopes from M */
        solve_vjp(w, s, k, fv, sc, coef, dbar, d2, r2bar, gmbar, gjbar);
        /* r2 = f1 - 2 k1 */
        for (int64_t i = c0; i < c0 + n; ++i)
            for (int qq = 0; qq < S; ++qq) k1bar[(int64_t)qq * nc + i] += -2.0 * r2bar[(int64_t)qq * nc + i];
        for (int64_t e = e0; e < e1; ++e) k1gbar[e] += -2.0 * g1bar[e];
        /* f1 = f(y1): pull r2bar (amount rates) and g1bar (gate rates) back to y1 */
        for (int64_t i = c0; i < c0 + n; ++i) {
            vbar[i] = 0.0;
            for (int qq = 0; qq < S; ++qq) { cbar[(int64_t)qq * nc + i] = 0.0; x1bar[(int64_t)qq * nc + i] = 0.0; }
        }
        double* gate1bar = (double*)calloc((size_t)(ne ? ne : 1), sizeof(double));
        rates_vjp(w, &L, k, y1, phi1, c, r2bar, g1bar, NULL, NULL, cbar, vbar, gate1bar, pam_bar,
                  cap_bar, paj_bar);
        pull_to_amounts(w, &L, k, y1, cbar, vbar, x1bar);
        /* y1 = y + dt k1 */
        for (int64_t i = c0; i < c0 + n; ++i)
            for (int qq = 0; qq < S; ++qq) {
                CF(&L, w, ybar_in, L.dn + qq, i) += x1bar[(int64_t)qq * nc + i];
                k1bar[(int64_t)qq * nc + i] += dt * x1bar[(int64_t)qq * nc + i];
            }
        for (int64_t e = e0; e < e1; ++e) {
            EF(&L, w, ybar_in, L.gate, e) += gate1bar[e];
            k1gbar[e] += dt * gate1bar[e];
        }
        free(gate1bar);
        /* k1 = f(y) + coef D delta1 */
        dtrans(w, k, gm, gj, k1bar, dbar);
        for (int64_t i = c0; i < c0 + n; ++i) dbar[i] *= coef;
        dslopes(w, k, d1, k1bar, coef, gmbar, gjbar);
        solve_vjp(w, s, k, fv, sc, coef, dbar, d1, k1bar, gmbar, gjbar);
        /* f(y) and the slopes at y */
        for (int64_t i = c0; i < c0 + n; ++i) {
            vbar[i] = 0.0;
            for (int qq = 0; qq < S; ++qq) { cbar[(int64_t)qq * nc + i] = 0.0; x1bar[(int64_t)qq * nc + i] = 0.0; }
        }
        double* gate0bar = (double*)calloc((size_t)(ne ? ne : 1), sizeof(double));
        rates_vjp(w, &L, k, y, phi, c, k1bar, k1gbar, gmbar, gjbar, cbar, vbar, gate0bar, pam_bar,
                  cap_bar, paj_bar);
        pull_to_amounts(w, &L, k, y, cbar, vbar, x1bar);
        for (int64_t i = c0; i < c0 + n; ++i)
            for (int qq = 0; qq < S; ++qq) CF(&L, w, ybar_in, L.dn + qq, i) += x1bar[(int64_t)qq * nc + i];
        for (int64_t e = e0; e < e1; ++e) EF(&L, w, ybar_in, L.gate, e) += gate0bar[e];
        free(gate0bar);
        for (int64_t i = c0; i < c0 + n; ++i)
            for (int qq = 0; qq < S; ++qq) gmbar[(int64_t)qq * nc + i] = 0.0;
        for (int64_t e = e0; e < e1; ++e)
            for (int qq = 0; qq < S; ++qq) gjbar[(int64_t)qq * ne + e] = 0.0;
        free(Gj); free(sc); free(fv); free(q); free(x);
    }
    free(k1); free(f1); free(y1); free(phi); free(phi1); free(gm); free(gj); free(d1); free(d2);
    free(gmbar); free(gjbar); free(k1bar); free(k1gbar); free(r2bar); free(cbar); free(vbar);
    free(dbar); free(x1bar); free(g1bar);
    return ok;
}

/* ---------------------------------------------------------------- spatial worlds
 *
 * A lattice of cells in an extracellular finite-volume grid, stepped by the reference's IMEX
 * Euler with implicit face electromigration (models/surface_pump.py:292-362, tableau
 * numerics/surface_imex.py:59-61): explicit membrane, junction, pump, gate and face-diffusion
 * rates at the step-start potentials (models/surface_pump.py:158-240, face_weights
 * surface_imex.py:123-132 with logmean numerics/extracellular_field.py:48-62); the stage charge;
 * the operator (C + dt L) phi = q on every node but the held ring, Jacobi-scaled and factored
 * (surface_imex.py:228-259, the face Laplacian :142-162), its solve and charge residual
 * (:269-278, models/surface_pump.py:325-330); the face drift with the operator's weights
 * (surface_imex.py:135-139, models/surface_pump.py:243-256); the applied face flux's heat and the
 * reservoir work (models/surface_pump.py:333-359); domain_valid and where(valid, new, old).
 * The step-start potentials are the field solve at the state (ref_field_start below, the
 * reference's readouts), or, in the carried form, the previous accepted step's operator
 * potentials (the state's node potentials; a cell's is its node's plus its charge over its
 * capacitance). */

void ref_logmean(double left, double right, double* mean, double* log_ratio) {
    /* numerics/extracellular_field.py:48-62 */
    double middle = 0.5 * left + 0.5 * right;
    double difference = left - right;
    double ratio = (0.5 * difference) / middle;
    int near = fabs(ratio) < 0.5;
    double lg = near ? 2 * atanh(ratio) : log(left) - log(right);
    int small = fabs(ratio) < 1e-4;
    double square = ratio * ratio;
    *mean = small ? middle * (1 - square / 3 - 4 * square * square / 45) : difference / lg;
    *log_ratio = lg;
}

void ref_field_layout(const cells_world* w, const cells_field_grid* g, cells_field_layout* L) {
    int32_t S = w->S, f = 0;
    memset(L, 0, sizeof(*L));
    L->dn = f; f += S;
    L->mtr = f; f += S;
    L->mdiss = f++;
    L->pcyc = f++;
    L->pdiss = f++;
    L->pwork = f++;
    L->cell_fields = f;
    f = 0;
    L->jtr = f; f += S;
    L->jdiss = f++;
    L->gate = f++;
    L->edge_fields = f;
    f = 0;
    L->dno = f; f += S;
    L->phi = f++;
    L->node_fields = f;
    f = 0;
    L->ftr = f; f += S;
    L->fheat = f++;
    L->face_fields = f;
    f = 0;
    L->sup = f; f += S;
    L->rwork = f++;
    L->world_fields = f;
    L->nodes = (int64_t)g->side * g->side;
    L->faces = 2 * (int64_t)g->side * (g->side - 1);
    L->edge_base = (int64_t)L->cell_fields * w->ncell;
    L->node_base = L->edge_base + (int64_t)L->edge_fields * w->nedge;
    L->face_base = L->node_base + (int64_t)L->node_fields * L->nodes * w->W;
    L->world_base = L->face_base + (int64_t)L->face_fields * L->faces * w->W;
    L->total = L->world_base + (int64_t)L->world_fields * w->W;
}

#define FCF(L, w, y, f, i) (y)[(int64_t)(f) * (w)->ncell + (i)]
#define FEF(L, w, y, f, e) (y)[(L)->edge_base + (int64_t)(f) * (w)->nedge + (e)]
#define FNF(L, w, y, f, k, g) (y)[(L)->node_base + ((int64_t)(f) * (w)->W + (k)) * (L)->nodes + (g)]
#define FFF(L, w, y, f, k, q) (y)[(L)->face_base + ((int64_t)(f) * (w)->W + (k)) * (L)->faces + (q)]
#define FWF(L, w, y, f, k) (y)[(L)->world_base + (int64_t)(f) * (w)->W + (k)]

static void face_nodes(int32_t side, int64_t q, int64_t* a, int64_t* b) {
    int64_t nh = (int64_t)side * (side - 1);
    if (q < nh) {
        int64_t r = q / (side - 1), c = q % (side - 1);
        *a = r * side + c;
        *b = *a + 1;
    } else {
        int64_t v = q - nh;
        *a = v;
        *b = v + side;
    }
}

static int on_ring(int32_t side, int64_t node) {
    int64_t r = node / side, c = node % side;
    return r == 0 || c == 0 || r == side - 1 || c == side - 1;
}

/* Interior nodes in row-major order: the operator system's nodes after the cells. */
static int64_t interior_index(int32_t side, int64_t node) {
    int64_t r = node / side, c = node % side;
    return (r - 1) * (side - 2) + (c - 1);
}

/* World k's slices of every field, src -> dst. */
static void copy_field_world(const cells_world* w, const cells_field_layout* L, int32_t k,
                             const double* src, double* dst) {
    for (int f = 0; f < L->cell_fields; ++f)
        for (int64_t i = w->cell_off[k]; i < w->cell_off[k + 1]; ++i) FCF(L, w, dst, f, i) = FCF(L, w, src, f, i);
    for (int f = 0; f < L->edge_fields; ++f)
        for (int64_t e = w->edge_off[k]; e < w->edge_off[k + 1]; ++e) FEF(L, w, dst, f, e) = FEF(L, w, src, f, e);
    for (int f = 0; f < L->node_fields; ++f)
        for (int64_t n = 0; n < L->nodes; ++n) FNF(L, w, dst, f, k, n) = FNF(L, w, src, f, k, n);
    for (int f = 0; f < L->face_fields; ++f)
        for (int64_t q = 0; q < L->faces; ++q) FFF(L, w, dst, f, k, q) = FFF(L, w, src, f, k, q);
    for (int f = 0; f < L->world_fields; ++f) FWF(L, w, dst, f, k) = FWF(L, w, src, f, k);
}

static double field_charge(const cells_world* w, const double* y_dn, int64_t stride, int64_t i) {
    double t = w->valence[0] * y_dn[i];
    for (int s = 1; s < w->S; ++s) t = t + w->valence[s] * y_dn[(int64_t)s * stride + i];
    return CHARGE_PER_FMOL * t;
}

/* ---- the field solve at a state (numerics/surface_field.py:614-855)
 *
 * The capacitor components of a spatial world are each cell with its grid node (joined by the
 * membrane capacitor) and every other interior node alone; the ring nodes are held at potential 0
 * (reservoir clamps), so every component floats. The capacitor solve fixes each component's
 * representative (its node) at 0: base = q_i / C_i at a cell, 0 at the nodes (_capacitor_solve,
 * :598-602). The offsets, one per interior node, solve Kirchhoff current continuity: each
 * component's net outflow through its faces (log-mean electrodiffusion, numerics/
 * surface_transport.py:340-353) and its cell's junctions (GHK) is zero. Start: the faces' linear
 * balance at the base (:668-672); then the damped Newton with the junctions' slopes (:737-775):
 * at most 16 iterations, 12 halvings, a trial accepted when its largest residual falls or its norm
 * max |r| / (1e-23 + 2e-12 max current) reaches 0.01. The offset systems (the face conductance
 * Laplacian plus the junction slopes on the interior nodes) are scaled by their largest entry as
 * _offset_solve (:604-612) scales them, and solved with the sparse Cholesky on ks, world k's
 * pattern of interior nodes with edge_pos = each face's entry (-1 where it touches the ring), then
 * each junction's (world-local order). */

#define NEWTON_TARGET 0.01
#define CURRENT_ATOL_A 1e-23
#define CURRENT_RTOL 2e-12
#define NEWTON_ITERATIONS 16
#define NEWTON_BACKTRACKS 12

typedef struct {
    const cells_world* w;
    const cells_field_grid* g;
    const cells_symbolic* ks;
    const cells_control* c;
    int32_t k, side;
    int64_t c0, nc, e0, ne, NN, NF, ni;
    double T, RT;
    double *ci, *co, *base, *open, *phi, *fcur, *jcur, *res, *slope, *cond, *mean, *lg;
} field_newton;

static double newton_evaluate(field_newton* K, const double* offset, int* valid_out, double* resmax_out) {
    const cells_world* w = K->w;
    const int S = w->S;
    const int32_t side = K->side;
    int valid = 1;
    for (int64_t n = 0; n < K->NN; ++n)
        K->phi[K->nc + n] = K->base[K->nc + n] + (on_ring(side, n) ? 0.0 : offset[interior_index(side, n)]);
    for (int64_t i = 0; i < K->nc; ++i)
        K->phi[i] = K->base[i] + offset[interior_index(side, K->g->cell_node[K->c0 + i])];
    for (int64_t x = 0; x < K->nc + K->NN; ++x) valid = valid && isfinite(K->phi[x]);
    /* membrane routes: inside each component, checked for the transport's domain only */
    for (int64_t i = 0; i < K->nc; ++i) {
        int64_t node = K->g->cell_node[K->c0 + i];
        double dv = K->phi[i] - K->phi[K->nc + node];
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s];
            double J = ref_selective_flux(K->ci[(int64_t)s * K->nc + i], K->co[(int64_t)s * K->NN + node], dv, z,
                                          K->T, w->pam[(int64_t)s * w->ncell + K->c0 + i], K->c->membrane);
            valid = valid && fabs(z * F_CONST * dv / K->RT) <= MAX_ABS_X && isfinite(J);
        }
    }
    double jmax = 0.0, fmax = 0.0;
    for (int64_t j = 0; j < K->ne; ++j) {
        int64_t e = K->e0 + j, a = w->snd[e] - K->c0, b = w->rcv[e] - K->c0;
        double dv = K->phi[a] - K->phi[b];
        double cur = 0.0, sl = 0.0;
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s], pa = w->paj[(int64_t)s * w->nedge + e];
            double ca = K->ci[(int64_t)s * K->nc + a], cb = K->ci[(int64_t)s * K->nc + b];
            double J = ref_selective_flux(ca, cb, dv, z, K->T, pa, K->open[j]);
            cur = s == 0 ? z * J : cur + z * J;
            double x = z * F_CONST * dv / K->RT;
            double d = z * z * pa * (-ca * ref_bernoulli_derivative(-x) - cb * ref_bernoulli_derivative(x));
            sl = s == 0 ? d : sl + d;
            valid = valid && fabs(x) <= MAX_ABS_X && isfinite(J);
        }
        K->jcur[j] = CHARGE_PER_FMOL * cur;
        K->slope[j] = F_CONST * F_CONST / K->RT * K->open[j] * sl;
        valid = valid && isfinite(K->slope[j]) && K->slope[j] >= 0;
        if (fabs(K->jcur[j]) > jmax) jmax = fabs(K->jcur[j]); /* a NaN current fails isfinite(J) */
    }
    for (int64_t q = 0; q < K->NF; ++q) {
        int64_t a, b;
        face_nodes(side, q, &a, &b);
        double dv = K->phi[K->nc + a] - K->phi[K->nc + b];
        double cur = 0.0;
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s];
            double ca = K->co[(int64_t)s * K->NN + a], cb = K->co[(int64_t)s * K->NN + b];
            double mu = K->RT * (log(ca) - log(cb)) + z * F_CONST * dv;
            double J = K->g->face_factor[s] * K->mean[(int64_t)s * K->NF + q] * mu / (K->RT * MOL_PER_FMOL);
            cur = s == 0 ? z * J : cur + z * J;
            valid = valid && fabs(z * F_CONST * dv / K->RT) <= MAX_ABS_X && isfinite(J);
        }
        K->fcur[q] = CHARGE_PER_FMOL * cur;
        if (fabs(K->fcur[q]) > fmax) fmax = fabs(K->fcur[q]);
    }
    /* each component's net outflow: face_incidence @ face currents + junction_incidence @ junction currents */
    for (int64_t x = 0; x < K->ni; ++x) K->res[x] = 0.0;
    for (int64_t q = 0; q < K->NF; ++q) {
        int64_t a, b;
        face_nodes(side, q, &a, &b);
        if (!on_ring(side, a)) K->res[interior_index(side, a)] += K->fcur[q];
        if (!on_ring(side, b)) K->res[interior_index(side, b)] -= K->fcur[q];
    }
    for (int64_t j = 0; j < K->ne; ++j) {
        int64_t e = K->e0 + j;
        K->res[interior_index(side, K->g->cell_node[w->snd[e]])] += K->jcur[j];
        K->res[interior_index(side, K->g->cell_node[w->rcv[e]])] -= K->jcur[j];
    }
    double rmax = 0.0;
    for (int64_t x = 0; x < K->ni; ++x)
        if (fabs(K->res[x]) > rmax) rmax = fabs(K->res[x]);
    double scale = fmax > jmax ? fmax : jmax;
    double norm = rmax / (CURRENT_ATOL_A + CURRENT_RTOL * scale);
    valid = valid && isfinite(norm);
    *valid_out = valid;
    *resmax_out = rmax;
    return norm;
}

/* _offset_solve: the face conductance Laplacian on the interior nodes (+ the junction slopes when
 * slope is given) scaled by its largest entry; x = matrix^-1 rhs. */
static void newton_offset_solve(field_newton* K, const double* slope, const double* rhs, double* x) {
    const cells_symbolic* ks = K->ks;
    const int32_t side = K->side;
    double* v = (double*)calloc((size_t)ks->nnz, sizeof(double));
    double* diag = (double*)calloc((size_t)K->ni, sizeof(double));
    double* b = (double*)malloc(sizeof(double) * (size_t)K->ni);
    for (int64_t q = 0; q < K->NF; ++q) {
        int64_t a, bb;
        face_nodes(side, q, &a, &bb);
        if (!on_ring(side, a)) diag[interior_index(side, a)] += K->cond[q];
        if (!on_ring(side, bb)) diag[interior_index(side, bb)] += K->cond[q];
        if (ks->edge_pos[q] >= 0) v[ks->edge_pos[q]] -= K->cond[q];
    }
    if (slope)
        for (int64_t j = 0; j < K->ne; ++j) {
            int64_t e = K->e0 + j;
            diag[interior_index(side, K->g->cell_node[K->w->snd[e]])] += slope[j];
            diag[interior_index(side, K->g->cell_node[K->w->rcv[e]])] += slope[j];
            v[ks->edge_pos[K->NF + j]] -= slope[j];
        }
    for (int64_t xx = 0; xx < K->ni; ++xx) v[ks->diag_pos[xx]] = diag[xx];
    double scale = 0.0;
    for (int64_t p = 0; p < ks->nnz; ++p)
        if (fabs(v[p]) > scale) scale = fabs(v[p]);
    if (!(scale > 0)) scale = 1.0;
    for (int64_t p = 0; p < ks->nnz; ++p) v[p] = v[p] / scale;
    for (int64_t xx = 0; xx < K->ni; ++xx) b[xx] = rhs[xx] / scale;
    ref_factor(ks, v);
    ref_solve(ks, v, b, x);
    free(v);
    free(diag);
    free(b);
}

int ref_field_start(const cells_world* w, const cells_field_grid* g, const cells_symbolic* ks, int32_t k,
                    const double* y, const cells_control* c, double* phi, int32_t* iterations) {
    cells_field_layout L;
    ref_field_layout(w, g, &L);
    const int S = w->S;
    field_newton K;
    K.w = w; K.g = g; K.ks = ks; K.c = c; K.k = k; K.side = g->side;
    K.c0 = w->cell_off[k]; K.nc = w->cell_off[k + 1] - K.c0;
    K.e0 = w->edge_off[k]; K.ne = w->edge_off[k + 1] - K.e0;
    K.NN = L.nodes; K.NF = L.faces; K.ni = (int64_t)(g->side - 2) * (g->side - 2);
    K.T = w->temperature; K.RT = R_CONST * K.T;
    K.ci = (double*)malloc(sizeof(double) * (size_t)(S * K.nc));
    K.co = (double*)malloc(sizeof(double) * (size_t)(S * K.NN));
    K.base = (double*)malloc(sizeof(double) * (size_t)(K.nc + K.NN));
    K.open = (double*)malloc(sizeof(double) * (size_t)(K.ne > 0 ? K.ne : 1));
    K.phi = phi;
    K.fcur = (double*)malloc(sizeof(double) * (size_t)K.NF);
    K.jcur = (double*)malloc(sizeof(double) * (size_t)(K.ne > 0 ? K.ne : 1));
    K.res = (double*)malloc(sizeof(double) * (size_t)K.ni);
    K.slope = (double*)malloc(sizeof(double) * (size_t)(K.ne > 0 ? K.ne : 1));
    K.cond = (double*)malloc(sizeof(double) * (size_t)K.NF);
    K.mean = (double*)malloc(sizeof(double) * (size_t)(S * K.NF));
    K.lg = (double*)malloc(sizeof(double) * (size_t)(S * K.NF));
    double* offset = (double*)malloc(sizeof(double) * (size_t)K.ni);
    double* proposal = (double*)malloc(sizeof(double) * (size_t)K.ni);
    double* delta = (double*)malloc(sizeof(double) * (size_t)K.ni);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)K.ni);
    int valid = 1;
    for (int s = 0; s < S; ++s) {
        for (int64_t i = 0; i < K.nc; ++i) {
            double v = (w->n0[(int64_t)s * w->ncell + K.c0 + i] + FCF(&L, w, y, L.dn + s, K.c0 + i)) / w->vol0[K.c0 + i];
            K.ci[(int64_t)s * K.nc + i] = v;
            valid = valid && isfinite(v) && v > 0;
        }
        for (int64_t n = 0; n < K.NN; ++n) {
            double v = (g->node_n0[(int64_t)s * K.NN + n] + FNF(&L, w, y, L.dno + s, k, n)) / g->node_volume;
            K.co[(int64_t)s * K.NN + n] = v;
            valid = valid && isfinite(v) && v > 0;
        }
    }
    /* base: the capacitor solve with each component's node fixed at 0 */
    for (int64_t n = 0; n < K.NN; ++n) K.base[K.nc + n] = 0.0;
    for (int64_t i = 0; i < K.nc; ++i) {
        double q = w->q0[K.c0 + i] + field_charge(w, &FCF(&L, w, y, L.dn, 0), w->ncell, K.c0 + i);
        K.base[i] = q / w->cap[K.c0 + i];
    }
    for (int64_t j = 0; j < K.ne; ++j) K.open[j] = c->junction * FEF(&L, w, y, L.gate, K.e0 + j);
    /* the faces' conductances and their currents at the base (dv = 0 between nodes) */
    for (int64_t q = 0; q < K.NF; ++q) {
        int64_t a, b;
        face_nodes(g->side, q, &a, &b);
        double G = 0.0, f0 = 0.0;
        for (int s = 0; s < S; ++s) {
            double z = w->valence[s], m, lg;
            ref_logmean(K.co[(int64_t)s * K.NN + a], K.co[(int64_t)s * K.NN + b], &m, &lg);
            K.mean[(int64_t)s * K.NF + q] = m;
            K.lg[(int64_t)s * K.NF + q] = lg;
            double weight = g->face_factor[s] * m;
            G = s == 0 ? z * z * weight : G + z * z * weight;
            f0 = s == 0 ? z * weight * lg : f0 + z * weight * lg;
        }
        K.cond[q] = F_CONST * F_CONST / K.RT * G;
        K.fcur[q] = F_CONST * f0;
    }
    for (int64_t x = 0; x < K.ni; ++x) rhs[x] = 0.0;
    for (int64_t q = 0; q < K.NF; ++q) {
        int64_t a, b;
        face_nodes(g->side, q, &a, &b);
        if (!on_ring(g->side, a)) rhs[interior_index(g->side, a)] -= K.fcur[q];
        if (!on_ring(g->side, b)) rhs[interior_index(g->side, b)] += K.fcur[q];
    }
    newton_offset_solve(&K, NULL, rhs, offset);
    int ok;
    double rmax;
    double norm = newton_evaluate(&K, offset, &ok, &rmax);
    int accepted = ok && valid;
    int32_t count = 0;
    while (count < NEWTON_ITERATIONS && norm > NEWTON_TARGET && accepted) {
        norm = newton_evaluate(&K, offset, &ok, &rmax);
        for (int64_t x = 0; x < K.ni; ++x) rhs[x] = -K.res[x];
        newton_offset_solve(&K, K.slope, rhs, delta);
        double before = rmax;
        int step = 0;
        accepted = 0;
        while (step < NEWTON_BACKTRACKS && !accepted) {
            double f = pow(0.5, step);
            for (int64_t x = 0; x < K.ni; ++x) proposal[x] = offset[x] + delta[x] * f;
            norm = newton_evaluate(&K, proposal, &ok, &rmax);
            accepted = ok && (rmax < before || norm <= NEWTON_TARGET);
            ++step;
        }
        ++count;
        if (accepted) memcpy(offset, proposal, sizeof(double) * (size_t)K.ni);
    }
    norm = newton_evaluate(&K, offset, &ok, &rmax);
    int converged = valid && ok && accepted && norm <= NEWTON_TARGET;
    /* the cells' capacitor rows (solved rows; each node's row carries its component's charge) */
    for (int64_t i = 0; i < K.nc; ++i) {
        double q = w->q0[K.c0 + i] + field_charge(w, &FCF(&L, w, y, L.dn, 0), w->ncell, K.c0 + i);
        double cq = w->cap[K.c0 + i] * (phi[i] - phi[K.nc + g->cell_node[K.c0 + i]]);
        double sc = fabs(q) > fabs(cq) ? fabs(q) : fabs(cq);
        converged = converged && fabs(cq - q) <= 3e-25 + 2e-12 * sc;
    }
    if (iterations) *iterations = count;
    free(K.ci); free(K.co); free(K.base); free(K.open); free(K.fcur); free(K.jcur); free(K.res);
    free(K.slope); free(K.cond); free(K.mean); free(K.lg);
    free(offset); free(proposal); free(delta); free(rhs);
    return converged;
}

int ref_field_step(const cells_world* w, const cells_field_grid* g, const cells_symbolic* sym,
                   const cells_symbolic* ksym, const double* y, double* out, double dt, const cells_control* c,
                   int32_t refresh, double* op_weight, double* op_scale, double* op_factor, int32_t* valid_out) {
    cells_field_layout L;
    ref_field_layout(w, g, &L);
    const int S = w->S;
    const int32_t side = g->side;
    const int64_t NN = L.nodes, NF = L.faces, ni = (int64_t)(side - 2) * (side - 2);
    const double T = w->temperature, RT = R_CONST * T;
    double* r = (double*)calloc((size_t)L.total, sizeof(double));
    double* known = (double*)malloc(sizeof(double) * (size_t)L.total);
    double* inflow = (double*)calloc((size_t)S * NN, sizeof(double));
    double* drift = (double*)calloc((size_t)S * NF, sizeof(double));
    memcpy(known, y, sizeof(double) * (size_t)L.total);
    memcpy(out, y, sizeof(double) * (size_t)L.total);
    int all = 1;
    int64_t fpos = 0;
    for (int32_t k = 0; k < w->W; ++k) {
        const int64_t c0 = w->cell_off[k], c1 = w->cell_off[k + 1], nc = c1 - c0;
        const int64_t e0 = w->edge_off[k], e1 = w->edge_off[k + 1];
        const cells_symbolic* sk = &sym[k];
        const int64_t nsys = nc + ni;
        double* fk = op_factor + fpos;
        fpos += sk->nnz;
        double* wk = op_weight + (int64_t)k * S * NF;
        double* sck = op_scale + (int64_t)k * nsys;
        int valid = 1;
        const double ctl[5] = {c->membrane, c->junction, c->pump, c->calcium_pump, c->water};
        for (int q = 0; q < 5; ++q) valid = valid && isfinite(ctl[q]) && ctl[q] >= 0 && ctl[q] <= 1;
        /* concentrations */
        double* co = (double*)malloc(sizeof(double) * (size_t)S * NN);
        for (int s = 0; s < S; ++s)
            for (int64_t n = 0; n < NN; ++n) {
                double amount = g->node_n0[(int64_t)s * NN + n] + FNF(&L, w, y, L.dno + s, k, n);
                co[(int64_t)s * NN + n] = amount / g->node_volume;
                valid = valid && isfinite(co[(int64_t)s * NN + n]) && co[(int64_t)s * NN + n] > 0;
            }
        /* step-start potentials: the field solve at y (ksym given), else the state's node
         * potentials, cells their node's plus q / C */
        double* phi = (double*)malloc(sizeof(double) * (size_t)(nc + NN));
        if (ksym) {
            valid = ref_field_start(w, g, &ksym[k], k, y, c, phi, NULL) && valid;
        } else {
            for (int64_t n = 0; n < NN; ++n) phi[nc + n] = FNF(&L, w, y, L.phi, k, n);
            for (int64_t i = c0; i < c1; ++i) {
                double q = w->q0[i] + field_charge(w, &FCF(&L, w, y, L.dn, 0), w->ncell, i);
                phi[i - c0] = phi[nc + g->cell_node[i]] + q / w->cap[i];
            }
        }
        /* junction routes (surface_transport.py:310-367) */
        for (int64_t e = e0; e < e1; ++e) {
            int64_t a = w->snd[e], b = w->rcv[e];
            double dv = phi[a - c0] - phi[b - c0];
            double gt = FEF(&L, w, y, L.gate, e);
            double avail = c->junction * gt;
            double diss = 0.0;
            for (int s = 0; s < S; ++s) {
                double ca = (w->n0[(int64_t)s * w->ncell + a] + FCF(&L, w, y, L.dn + s, a)) / w->vol0[a];
                double cb = (w->n0[(int64_t)s * w->ncell + b] + FCF(&L, w, y, L.dn + s, b)) / w->vol0[b];
                double z = w->valence[s];
                double pa = w->paj[(int64_t)s * w->nedge + e];
                valid = valid && isfinite(ca) && ca > 0 && isfinite(cb) && cb > 0;
                double J = ref_selective_flux(ca, cb, dv, z, T, pa, avail);
                double mu = RT * (log(ca) - log(cb)) + z * F_CONST * dv;
                diss = s == 0 ? J * mu : diss + J * mu;
                valid = valid && fabs(z * F_CONST * dv / RT) <= MAX_ABS_X && isfinite(J);
                FEF(&L, w, r, L.jtr + s, e) = J;
            }
            FEF(&L, w, r, L.jdiss, e) = 1e-3 * diss;
            FEF(&L, w, r, L.gate, e) = ref_gate_rhs(gt, dv, w->gate_threshold, w->gate_minimum);
            valid = valid && gate_domain(gt, dv, w->gate_threshold, w->gate_minimum, c->junction);
        }
        /* membrane routes into the cells' nodes, Na/K-ATPase, owner sums */
        for (int s = 0; s < S; ++s)
            for (int64_t n = 0; n < NN; ++n) inflow[(int64_t)s * NN + n] = 0.0;
        for (int64_t i = c0; i < c1; ++i) {
            int64_t node = g->cell_node[i];
            double dv = phi[i - c0] - phi[nc + node];
            double ci[4], Jm[4], mu[4];
            double diss = 0.0;
            for (int s = 0; s < S; ++s) {
                double z = w->valence[s], cout = co[(int64_t)s * NN + node];
                ci[s] = (w->n0[(int64_t)s * w->ncell + i] + FCF(&L, w, y, L.dn + s, i)) / w->vol0[i];
                valid = valid && isfinite(ci[s]) && ci[s] > 0;
                double pa = w->pam[(int64_t)s * w->ncell + i];
                Jm[s] = ref_selective_flux(ci[s], cout, dv, z, T, pa, c->membrane);
                mu[s] = RT * (log(ci[s]) - log(cout)) + z * F_CONST * dv;
                diss = s == 0 ? Jm[s] * mu[s] : diss + Jm[s] * mu[s];
                valid = valid && fabs(z * F_CONST * dv / RT) <= MAX_ABS_X && isfinite(Jm[s]);
            }
            double cyc, energy, drive;
            double kout = co[1 * NN + node];
            ref_nak_kinetics(w, ci[0], kout, mu[0], mu[1], w->nak_cap[i], c->pump, &cyc, &energy, &drive);
            valid = valid && ref_nak_domain(w, ci[0], kout, mu[0], mu[1], w->nak_cap[i], c->pump);
            double extra[4] = {3.0 * cyc, -2.0 * cyc, -0.0 * cyc, -0.0 * cyc};
            int64_t p0 = w->inc_ptr[i], p1 = w->inc_ptr[i + 1], pm = p0 + w->inc_out[i];
            for (int s = 0; s < S; ++s) {
                double outs = 0.0, ins = 0.0;
                for (int64_t p = p0; p < pm; ++p) outs += -FEF(&L, w, r, L.jtr + s, w->inc_edge[p]);
                for (int64_t p = pm; p < p1; ++p) ins += FEF(&L, w, r, L.jtr + s, w->inc_edge[p]);
                double cell = -Jm[s];
                cell += outs;
                cell += ins;
                FCF(&L, w, r, L.dn + s, i) = cell - extra[s];
                FCF(&L, w, r, L.mtr + s, i) = Jm[s];
                inflow[(int64_t)s * NN + node] += Jm[s] + extra[s];
            }
            FCF(&L, w, r, L.mdiss, i) = 1e-3 * diss;
            FCF(&L, w, r, L.pcyc, i) = cyc;
            FCF(&L, w, r, L.pdiss, i) = -1e-3 * cyc * energy;
            FCF(&L, w, r, L.pwork, i) = -1e-3 * cyc * drive;
        }
        /* faces: weights and explicit diffusion (surface_imex.py:123-132) */
        for (int64_t q = 0; q < NF; ++q) {
            int64_t a, b;
            face_nodes(side, q, &a, &b);
            for (int s = 0; s < S; ++s) {
                double mean, lg;
                ref_logmean(co[(int64_t)s * NN + a], co[(int64_t)s * NN + b], &mean, &lg);
                double weight = g->face_factor[s] * mean / MOL_PER_FMOL;
                double lr = log(co[(int64_t)s * NN + a]) - log(co[(int64_t)s * NN + b]);
                double d = weight * lr;
                valid = valid && isfinite(d);
                FFF(&L, w, r, L.ftr + s, k, q) = d;
                inflow[(int64_t)s * NN + a] -= d;
                inflow[(int64_t)s * NN + b] += d;
                if (refresh) wk[(int64_t)s * NF + q] = weight;
            }
        }
        /* node rates: the ring's inflow is its supply (surface_transport.py reservoirs) */
        double supply[4] = {0, 0, 0, 0};
        for (int64_t n = 0; n < NN; ++n) {
            int ring = on_ring(side, n);
            for (int s = 0; s < S; ++s) {
                FNF(&L, w, r, L.dno + s, k, n) = ring ? 0.0 : inflow[(int64_t)s * NN + n];
                if (ring) supply[s] -= inflow[(int64_t)s * NN + n];
            }
        }
        /* the operator at the step start (Euler's operator time 0), when refreshed */
        double* cond = (double*)malloc(sizeof(double) * (size_t)NF);
        for (int64_t q = 0; q < NF; ++q) {
            double t = 0.0;
            for (int s = 0; s < S; ++s)
                t = s == 0 ? w->valence[s] * w->valence[s] * wk[(int64_t)s * NF + q]
                           : t + w->valence[s] * w->valence[s] * wk[(int64_t)s * NF + q];
            cond[q] = CHARGE_PER_FMOL * F_CONST / RT * t;
        }
        if (refresh) {
            /* M = C + dt L on the system nodes (cells, then interior nodes); scale 1/sqrt(diag) */
            double* diag = (double*)calloc((size_t)nsys, sizeof(double));
            for (int64_t i = c0; i < c1; ++i) {
                diag[i - c0] += w->cap[i];
                diag[nc + interior_index(side, g->cell_node[i])] += w->cap[i];
            }
            for (int64_t q = 0; q < NF; ++q) {
                int64_t a, b;
                face_nodes(side, q, &a, &b);
                if (!on_ring(side, a)) diag[nc + interior_index(side, a)] += dt * cond[q];
                if (!on_ring(side, b)) diag[nc + interior_index(side, b)] += dt * cond[q];
            }
            for (int64_t x = 0; x < nsys; ++x) {
                valid = valid && diag[x] > 0;
                sck[x] = 1.0 / sqrt(diag[x] > 0 ? diag[x] : 1.0);
            }
            for (int64_t p = 0; p < sk->nnz; ++p) fk[p] = 0.0;
            for (int64_t x = 0; x < nsys; ++x) fk[sk->diag_pos[x]] = (diag[x] * sck[x] * sck[x] + diag[x] * sck[x] * sck[x]) / 2;
            for (int64_t i = c0; i < c1; ++i) {
                int64_t xn = nc + interior_index(side, g->cell_node[i]);
                double v = -w->cap[i], si = sck[i - c0], sn = sck[xn];
                fk[sk->edge_pos[i - c0]] = (v * si * sn + v * sn * si) / 2;
            }
            for (int64_t q = 0; q < NF; ++q) {
                int64_t p = sk->edge_pos[nc + q];
                if (p < 0) continue;
                int64_t a, b;
                face_nodes(side, q, &a, &b);
                double v = -(dt * cond[q]);
                double sa = sck[nc + interior_index(side, a)], sb = sck[nc + interior_index(side, b)];
                fk[p] = (v * sa * sb + v * sb * sa) / 2;
            }
            free(diag);
            ref_factor(sk, fk);
            for (int64_t p = 0; p < sk->nnz; ++p) valid = valid && isfinite(fk[p]);
        }
        /* known = y + dt r (every field but the potentials) */
        for (int f = 0; f < L.cell_fields; ++f)
            for (int64_t i = c0; i < c1; ++i) FCF(&L, w, known, f, i) += dt * FCF(&L, w, r, f, i);
        for (int f = 0; f < L.edge_fields; ++f)
            for (int64_t e = e0; e < e1; ++e) FEF(&L, w, known, f, e) += dt * FEF(&L, w, r, f, e);
        for (int s = 0; s < S; ++s)
            for (int64_t n = 0; n < NN; ++n) FNF(&L, w, known, L.dno + s, k, n) += dt * FNF(&L, w, r, L.dno + s, k, n);
        for (int s = 0; s < S; ++s)
            for (int64_t q = 0; q < NF; ++q) FFF(&L, w, known, L.ftr + s, k, q) += dt * FFF(&L, w, r, L.ftr + s, k, q);
        /* stage charges and the operator solve */
        double* q = (double*)malloc(sizeof(double) * (size_t)nsys);
        double* x = (double*)malloc(sizeof(double) * (size_t)nsys);
        for (int64_t i = c0; i < c1; ++i)
            q[i - c0] = w->q0[i] + field_charge(w, &FCF(&L, w, known, L.dn, 0), w->ncell, i);
        for (int64_t n = 0; n < NN; ++n) {
            if (on_ring(side, n)) continue;
            double t = w->valence[0] * FNF(&L, w, known, L.dno, k, n);
            for (int s = 1; s < S; ++s) t = t + w->valence[s] * FNF(&L, w, known, L.dno + s, k, n);
            q[nc + interior_index(side, n)] = g->node_q0[(int64_t)k * NN + n] + CHARGE_PER_FMOL * t;
        }
        double* rhs = (double*)malloc(sizeof(double) * (size_t)nsys);
        for (int64_t xx = 0; xx < nsys; ++xx) rhs[xx] = q[xx] * sck[xx];
        ref_solve(sk, fk, rhs, x);
        double* pot = (double*)malloc(sizeof(double) * (size_t)(nc + NN));
        for (int64_t xx = 0; xx < nc; ++xx) pot[xx] = x[xx] * sck[xx];
        for (int64_t n = 0; n < NN; ++n)
            pot[nc + n] = on_ring(side, n) ? 0.0 : x[nc + interior_index(side, n)] * sck[nc + interior_index(side, n)];
        /* charge residual (C + dt L) phi - q on the system nodes */
        double* mphi = (double*)calloc((size_t)nsys, sizeof(double));
        for (int64_t i = c0; i < c1; ++i) {
            int64_t n = g->cell_node[i], xn = nc + interior_index(side, n);
            double cv = w->cap[i] * (pot[i - c0] - pot[nc + n]);
            mphi[i - c0] += cv;
            mphi[xn] -= cv;
        }
        for (int64_t qq = 0; qq < NF; ++qq) {
            int64_t a, b;
            face_nodes(side, qq, &a, &b);
            double cur = dt * cond[qq] * (pot[nc + a] - pot[nc + b]);
            if (!on_ring(side, a)) mphi[nc + interior_index(side, a)] += cur;
            if (!on_ring(side, b)) mphi[nc + interior_index(side, b)] -= cur;
        }
        for (int64_t xx = 0; xx < nsys; ++xx) {
            double scale = fabs(q[xx]) > fabs(mphi[xx]) ? fabs(q[xx]) : fabs(mphi[xx]);
            valid = valid && isfinite(x[xx]) && fabs(mphi[xx] - q[xx]) <= 3e-25 + 2e-12 * scale;
        }
        /* drift with the operator's weights; stage = known + dt drift */
        double dsup[4] = {0, 0, 0, 0};
        for (int s = 0; s < S; ++s)
            for (int64_t n = 0; n < NN; ++n) inflow[(int64_t)s * NN + n] = 0.0;
        for (int64_t qq = 0; qq < NF; ++qq) {
            int64_t a, b;
            face_nodes(side, qq, &a, &b);
            double dvf = pot[nc + a] - pot[nc + b];
            for (int s = 0; s < S; ++s) {
                double fl = wk[(int64_t)s * NF + qq] * (w->valence[s] * F_CONST / RT) * dvf;
                drift[(int64_t)s * NF + qq] = fl;
                inflow[(int64_t)s * NN + a] -= fl;
                inflow[(int64_t)s * NN + b] += fl;
            }
        }
        copy_field_world(w, &L, k, known, out);
        for (int64_t n = 0; n < NN; ++n) {
            int ring = on_ring(side, n);
            for (int s = 0; s < S; ++s) {
                if (ring) dsup[s] -= inflow[(int64_t)s * NN + n];
                else FNF(&L, w, out, L.dno + s, k, n) += dt * inflow[(int64_t)s * NN + n];
            }
            FNF(&L, w, out, L.phi, k, n) = pot[nc + n];
        }
        for (int64_t qq = 0; qq < NF; ++qq) {
            double heat = 0.0;
            for (int s = 0; s < S; ++s) {
                double wt = wk[(int64_t)s * NF + qq];
                double fl = FFF(&L, w, r, L.ftr + s, k, qq) + drift[(int64_t)s * NF + qq];
                FFF(&L, w, out, L.ftr + s, k, qq) += dt * drift[(int64_t)s * NF + qq];
                double term = wt > 0 ? fl * fl / wt : 0.0;
                heat = s == 0 ? term : heat + term;
            }
            FFF(&L, w, out, L.fheat, k, qq) += dt * (1e-3 * RT * heat);
        }
        double rw = 0.0;
        for (int s = 0; s < S; ++s) {
            double applied = dt * supply[s] + dt * dsup[s];
            FWF(&L, w, out, L.sup + s, k) += applied;
            double mu = RT * log(w->bath_c[s]);
            rw = s == 0 ? mu * applied : rw + mu * applied;
        }
        FWF(&L, w, out, L.rwork, k) += 1e-3 * rw;
        /* domain_valid: finite fields, positive amounts, gates in [minimum, 1] */
        for (int64_t i = c0; i < c1; ++i) {
            for (int f = 0; f < L.cell_fields; ++f) valid = valid && isfinite(FCF(&L, w, out, f, i));
            for (int s = 0; s < S; ++s) {
                double nn = w->n0[(int64_t)s * w->ncell + i] + FCF(&L, w, out, L.dn + s, i);
                valid = valid && isfinite(nn) && nn > 0;
            }
        }
        for (int64_t e = e0; e < e1; ++e) {
            for (int f = 0; f < L.edge_fields; ++f) valid = valid && isfinite(FEF(&L, w, out, f, e));
            double gt = FEF(&L, w, out, L.gate, e);
            valid = valid && gt >= w->gate_minimum && gt <= 1.0;
        }
        for (int64_t n = 0; n < NN; ++n)
            for (int s = 0; s < S; ++s) {
                double nn = g->node_n0[(int64_t)s * NN + n] + FNF(&L, w, out, L.dno + s, k, n);
                valid = valid && isfinite(nn) && nn > 0 && isfinite(FNF(&L, w, out, L.phi, k, n));
            }
        for (int64_t qq = 0; qq < NF; ++qq)
            for (int f = 0; f < L.face_fields; ++f) valid = valid && isfinite(FFF(&L, w, out, f, k, qq));
        for (int f = 0; f < L.world_fields; ++f) valid = valid && isfinite(FWF(&L, w, out, f, k));
        if (!valid) copy_field_world(w, &L, k, y, out); /* the world keeps its old state */
        valid_out[k] = valid;
        all = all && valid;
        free(co); free(phi); free(cond); free(q); free(x); free(rhs); free(pot); free(mphi);
    }
    free(r);
    free(known);
    free(inflow);
    free(drift);
    return all;
}

/* ---- the field solve's energy rates (numerics/surface_field.py:792-835)
 *
 * At spatial world k's state y with the field's potentials phi (its cells', then its nodes'):
 * transport alone at phi (numerics/surface_transport.py:298-386: each cell's membrane into its
 * node, the junctions, the faces by log-mean electrodiffusion), then the field's energy terms:
 * chemical 1e-3 RT (sum ln ci cell_rate + sum ln co outside_rate), the outside rate being zero on
 * the held ring (its supply the ring's net outflow); the patch charge C V_m, the capacitor energy
 * 0.5e12 sum C V_m^2 and rate 1e12 sum C V_m dV_m/dt, dV_m/dt from the capacitor solve of the ionic
 * charge rate (each cell with its node one component whose node is the representative, so
 * dV_m/dt = the cell's charge rate over C); the reservoir power 1e-3 sum (RT ln co + z F phi)
 * supply over the ring; the electrode power 1e12 sum phi I over the ring with I the ring's
 * capacitor current (the ring holds no membrane: zero); the dissipation, every route's 1e-3 J mu;
 * and the balance chemical + capacitor + dissipation - reservoir - electrode.
 * out: chemical, capacitor, dissipation, reservoir power, electrode power, energy balance,
 * electrode current, capacitor energy, membrane heat, junction heat, face heat. */
int ref_field_energy_rates(const cells_world* w, const cells_field_grid* g, int32_t k, const double* y,
                           const cells_control* c, const double* phi, double* out) {
    cells_field_layout L;
    ref_field_layout(w, g, &L);
    const int S = w->S;
    const double T = w->temperature, RT = R_CONST * T;
    const int64_t c0 = w->cell_off[k], nc = w->cell_off[k + 1] - c0;
    const int64_t e0 = w->edge_off[k], ne = w->edge_off[k + 1] - e0;
    const int64_t NN = L.nodes, NF = L.faces;
    double* ci = (double*)malloc(sizeof(double) * (size_t)(S * nc));
    double* co = (double*)malloc(sizeof(double) * (size_t)(S * NN));
    double* cr = (double*)calloc((size_t)(S * nc), sizeof(double));  /* cell rates */
    double* orr = (double*)calloc((size_t)(S * NN), sizeof(double)); /* outside rates (all nodes) */
    for (int s = 0; s < S; ++s) {
        for (int64_t i = 0; i < nc; ++i)
            ci[s * nc + i] = (w->n0[(int64_t)s * w->ncell + c0 + i] + FCF(&L, w, y, L.dn + s, c0 + i)) / w->vol0[c0 + i];
        for (int64_t t = 0; t < NN; ++t)
            co[s * NN + t] = (g->node_n0[(int64_t)s * NN + t] + FNF(&L, w, y, L.dno + s, k, t)) / g->node_volume;
    }
    const double* pc = phi;
    const double* pn = phi + nc;
    double mheat = 0.0, jheat = 0.0, fheat = 0.0;
    int valid = 1;
    /* membranes: cell -> its node */
    double* vm = (double*)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
    for (int64_t i = 0; i < nc; ++i) {
        const int64_t t = g->cell_node[c0 + i];
        const double dv = pc[i] - pn[t];
        vm[i] = dv;
        double h = 0.0;
        for (int s = 0; s < S; ++s) {
            const double z = w->valence[s];
            const double J = ref_selective_flux(ci[s * nc + i], co[s * NN + t], dv, z, T,
                                                w->pam[(int64_t)s * w->ncell + c0 + i], c->membrane);
            const double mu = RT * (log(ci[s * nc + i]) - log(co[s * NN + t])) + z * F_CONST * dv;
            h += J * mu;
            cr[s * nc + i] -= J;
            orr[s * NN + t] += J;
        }
        mheat += 1e-3 * h;
    }
    /* junctions: sender -> receiver */
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t a = w->snd[e0 + e] - c0, b = w->rcv[e0 + e] - c0;
        const double dv = pc[a] - pc[b];
        const double avail = c->junction * FEF(&L, w, y, L.gate, e0 + e);
        double h = 0.0;
        for (int s = 0; s < S; ++s) {
            const double z = w->valence[s];
            const double J = ref_selective_flux(ci[s * nc + a], ci[s * nc + b], dv, z, T,
                                                w->paj[(int64_t)s * w->nedge + e0 + e], avail);
            const double mu = RT * (log(ci[s * nc + a]) - log(ci[s * nc + b])) + z * F_CONST * dv;
            h += J * mu;
            cr[s * nc + a] -= J;
            cr[s * nc + b] += J;
        }
        jheat += 1e-3 * h;
    }
    /* faces: sender -> receiver, log-mean electrodiffusion */
    for (int64_t q = 0; q < NF; ++q) {
        int64_t a, b;
        face_nodes(g->side, q, &a, &b);
        const double dv = pn[a] - pn[b];
        double h = 0.0;
        for (int s = 0; s < S; ++s) {
            const double z = w->valence[s];
            const double l = co[s * NN + a], r = co[s * NN + b];
            const double mu = RT * (log(l) - log(r)) + z * F_CONST * dv;
            double mean, lg;
            ref_logmean(l, r, &mean, &lg);
            const double J = g->face_factor[s] * mean * mu / (RT * MOL_PER_FMOL);
            h += J * mu;
            orr[s * NN + a] -= J;
            orr[s * NN + b] += J;
        }
        fheat += 1e-3 * h;
    }
    double chem_c = 0.0, chem_o = 0.0, reservoir = 0.0, cap_e = 0.0, cap_r = 0.0;
    for (int64_t i = 0; i < nc; ++i) {
        double zq = 0.0;
        for (int s = 0; s < S; ++s) {
            chem_c += log(ci[s * nc + i]) * cr[s * nc + i];
            zq += w->valence[s] * cr[s * nc + i];
        }
        const double C = w->cap[c0 + i];
        const double patch = C * vm[i];
        const double rate = CHARGE_PER_FMOL * zq / C; /* the capacitor solve: the node is fixed */
        cap_e += patch * vm[i];
        cap_r += patch * rate;
    }
    for (int64_t t = 0; t < NN; ++t) {
        const int ring = on_ring(g->side, t);
        for (int s = 0; s < S; ++s) {
            if (ring) {
                const double supply = -orr[s * NN + t];
                reservoir += (RT * log(co[s * NN + t]) + w->valence[s] * F_CONST * pn[t]) * supply;
            } else {
                chem_o += log(co[s * NN + t]) * orr[s * NN + t];
            }
        }
    }
    const double chemical = 1e-3 * RT * (chem_c + chem_o);
    const double capacitor = 1e12 * cap_r, capacitor_energy = 0.5e12 * cap_e;
    const double reservoir_power = 1e-3 * reservoir, electrode_power = 0.0, electrode_current = 0.0;
    const double dissipation = (mheat + jheat) + fheat;
    out[0] = chemical;
    out[1] = capacitor;
    out[2] = dissipation;
    out[3] = reservoir_power;
    out[4] = electrode_power;
    out[5] = chemical + capacitor + dissipation - reservoir_power - electrode_power;
    out[6] = electrode_current;
    out[7] = capacitor_energy;
    out[8] = mheat;
    out[9] = jheat;
    out[10] = fheat;
    for (int j = 0; j < 11; ++j) valid = valid && isfinite(out[j]);
    free(ci); free(co); free(cr); free(orr); free(vm);
    return valid;
}

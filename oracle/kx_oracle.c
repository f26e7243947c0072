/*
 * TEST INFRASTRUCTURE ONLY: the CPU oracle. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only to check GPU results against (or as
 * the timed CPU baseline); the product path never calls it.
 *
 * Plain-C, data-driven restatement of the reference's FP64 hot path for ANY mechanism,
 * reading the dense tables the generator writes (kinetix_amd/gen/tables.py):
 *
 *   kxo_species_rates      kinetix_species_rates, rolled form
 *                          (kinetix/core/reaction_rates.py:616-1333; arithmetic of
 *                          :1282-1288 Arrhenius, :770-822 special cases, :854-865
 *                          efficiencies, :887-1009 three-body/Lindemann/Troe,
 *                          :1011-1062 SRI, :1064-1175 PLOG, :1186-1231 Gibbs and reverse
 *                          rates, :1234-1261 net rates and wdot update)
 *   kxo_enthalpy_RT        kinetix/core/thermodynamics.py:17-38, utils/write_utils.py:145-207
 *   kxo_heat_capacity_R    kinetix/core/thermodynamics.py:41-62
 *   kxo_conductivity       kinetix/core/mix_transport.py:261-294
 *   kxo_viscosity          kinetix/core/mix_transport.py:297-357
 *   kxo_diffusivity        kinetix/core/mix_transport.py:414-471 (symmetric pair loop)
 *   kxo_production_rates   benchmark/okl/productionRates.okl:10-62
 *   kxo_transport          benchmark/okl/transportProps.okl:11-47
 *   kxo_thermo             benchmark/okl/thermoCoeffs.okl:10-39
 *
 * Pinning (what checks this restatement): golden vectors written by the reference's own
 * compiled GRI-3.0 routines (oracle/_ref, tests/golden/), the reference's Cantera
 * known-answer states (kinetix/ci_data/gri30.*.cantera, tolerances of
 * benchmark/src/bk.cpp:148,198-199,258), and invariants (mass and element conservation,
 * inert species, detailed balance at equilibrium, single-reaction closed forms, pure-species
 * transport limits). See tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KXO_CFLOAT_MIN 1e-300   /* benchmark/src/kinetix.cpp:245 */
#define KXO_MAXSP 1024

enum { KXO_ELEMENTARY = 0, KXO_THREE_BODY, KXO_LINDEMANN, KXO_TROE, KXO_SRI, KXO_PLOG };

typedef struct {
  int n_species, n_active, n_reactions;
  double R;                 /* gas constant [J/mol/K] */
  double one_atm_R;         /* P_atm / R  (C0 = one_atm_R * rcpT) */
  double R_one_atm;         /* R / P_atm  (1/C0 = R_one_atm * T) */
  const double *W, *rcpW;   /* [N] */
  const double *g_lo, *g_hi;   /* [N*7] Gibbs coefficients (a0, a1', a2', a3', a4', a5, a6) */
  const double *h_lo, *h_hi;   /* [N*6] enthalpy coefficients (a0, a1/2, a2/3, a3/4, a4/5, a5) */
  const double *cp_lo, *cp_hi; /* [N*5] heat capacity coefficients */
  const double *T_mid;         /* [N] */
  /* reactions */
  const int *kind, *reversible, *collider, *has_eff;   /* [R] */
  const double *lnA, *A, *beta, *Ta;                  /* [R]  k (or k_inf)  */
  const double *lnA0, *A0, *beta0, *Ta0;              /* [R]  k_0 (falloff)  */
  const int *nu_f, *nu_r;                             /* [R*N] */
  const double *eff;                                  /* [R*N] */
  const double *troe;   /* [R*4]: A, 1/(T3+1e-300), 1/(T1+1e-300), T2 (negative: absent) */
  const double *sri;    /* [R*5]: a, b, 1/(c+1e-300), d, e */
  const int *plog_off;  /* [R+1] offsets into the plog tables */
  const double *plog_P, *plog_lnP;            /* [n_plog_points] */
  const int *plog_nk;                         /* [n_plog_points] rates per pressure */
  const int *plog_koff;                       /* [n_plog_points] offset into plog_k */
  const double *plog_k;                       /* [n_plog_rates*3]: lnA, beta, Ta (A > 0) */
  const double *plog_A;                       /* [n_plog_rates] A */
  /* transport */
  const double *cond, *visc;   /* [N*5] */
  const double *diff;          /* [N*N*5] symmetric */
  const double *C1, *C2;       /* [N*N] Wilke coefficients */
  int rcp_diff;
} kxo_mech;

static double arrhenius(double A, double lnA, double beta, double Ta, double lnT, double T,
                        double T2, double rcpT)
{
  /* reaction_rates.py:648-663 grouping and the special cases of :770-822 */
  if (Ta == 0.0) {
    if (beta == 0.0) return A;
    if (beta == -2.0) return A * (rcpT * rcpT);
    if (beta == -1.0) return A * rcpT;
    if (beta == 1.0) return A * T;
    if (beta == 2.0) return A * T2;
  }
  const double blogT = beta * lnT;
  const double E_RT = Ta * rcpT;
  const double diff = blogT - E_RT;
  return exp(lnA + diff);
}

static double plog_rate(const kxo_mech *m, int pt, double lnT, double T, double T2, double rcpT)
{
  double k = 0.0;
  for (int j = 0; j < m->plog_nk[pt]; ++j) {
    const int o = m->plog_koff[pt] + j;
    k += arrhenius(m->plog_A[o], m->plog_k[3 * o], m->plog_k[3 * o + 1], m->plog_k[3 * o + 2],
                   lnT, T, T2, rcpT);
  }
  return k;
}

static double third_body(const kxo_mech *m, int r, const double *C, double Cm)
{
  const int N = m->n_species;
  if (m->has_eff[r]) {
    /* eff = Cm + sum_k (eps_k - 1) C_k  (reaction_rates.py:854-865) */
    double e = Cm;
    const double *eps = m->eff + (size_t)r * N;
    for (int k = 0; k < N; ++k) {
      if (eps[k] == 1.0) continue;
      if (eps[k] == 2.0) e += C[k];
      else if (eps[k] == 0.0) e -= C[k];
      else e += (eps[k] - 1.0) * C[k];
    }
    return e;
  }
  if (m->collider[r] >= 0) return C[m->collider[r]];
  return Cm;
}

/* Forward rate coefficient of reaction r, including third-body and falloff corrections. */
static double forward_rate(const kxo_mech *m, int r, double lnT, double T, double T2,
                           double rcpT, double P, double lnP, const double *C, double Cm)
{
  const int kind = m->kind[r];
  if (kind == KXO_PLOG) {
    /* reaction_rates.py:1064-1175: log-linear interpolation between bracketing pressures,
     * clamped to the end pressures outside the table. */
    const int o0 = m->plog_off[r], o1 = m->plog_off[r + 1];
    const int np = o1 - o0;
    if (P <= m->plog_P[o0]) return plog_rate(m, o0, lnT, T, T2, rcpT);
    if (P > m->plog_P[o1 - 1]) return plog_rate(m, o1 - 1, lnT, T, T2, rcpT);
    for (int i = 0; i + 1 < np; ++i) {
      const double p1 = m->plog_P[o0 + i], p2 = m->plog_P[o0 + i + 1];
      if (P > p1 && P <= p2) {
        const double k1 = plog_rate(m, o0 + i, lnT, T, T2, rcpT);
        const double k2 = plog_rate(m, o0 + i + 1, lnT, T, T2, rcpT);
        const double rcp_dlnp = 1.0 / (m->plog_lnP[o0 + i + 1] - m->plog_lnP[o0 + i]);
        return exp(log(k1) + (log(k2) - log(k1)) * (lnP - m->plog_lnP[o0 + i]) * rcp_dlnp);
      }
    }
    return plog_rate(m, o1 - 1, lnT, T, T2, rcpT);
  }
  const double kinf = arrhenius(m->A[r], m->lnA[r], m->beta[r], m->Ta[r], lnT, T, T2, rcpT);
  if (kind == KXO_ELEMENTARY) return kinf;
  if (kind == KXO_THREE_BODY) return kinf * third_body(m, r, C, Cm);
  double k0 = arrhenius(m->A0[r], m->lnA0[r], m->beta0[r], m->Ta0[r], lnT, T, T2, rcpT);
  k0 *= third_body(m, r, C, Cm);
  if (kind == KXO_LINDEMANN) {
    /* reaction_rates.py:904-922 */
    k0 /= (1 + k0 / (kinf + KXO_CFLOAT_MIN));
    return k0;
  }
  const double Pr = k0 / (kinf + KXO_CFLOAT_MIN);
  const double logPr = log10(Pr + KXO_CFLOAT_MIN);
  double F;
  if (kind == KXO_TROE) {
    /* reaction_rates.py:924-1009 */
    const double *t = m->troe + 4 * (size_t)r;
    const double term3 = (t[3] >= 0.0) ? exp(-t[3] * rcpT) : 0.0;
    const double logFcent = log10((1. - t[0]) * exp(-T * t[1]) + t[0] * exp(-T * t[2]) + term3);
    const double troe_c = -0.4 - 0.67 * logFcent;
    const double troe_n = 0.75 - 1.27 * logFcent;
    const double troe = (troe_c + logPr) / (troe_n - 0.14 * (troe_c + logPr) + KXO_CFLOAT_MIN);
    F = pow(10, logFcent / (1. + troe * troe));
  } else {
    /* SRI, reaction_rates.py:1011-1062 */
    const double *s = m->sri + 5 * (size_t)r;
    F = s[3] * pow(T, s[4]) * pow(s[0] * exp(-s[1] * rcpT) + exp(-s[2] * T), 1. / (1. + logPr * logPr));
  }
  double k = Pr / (1. + Pr);
  k *= F;
  return k * kinf;
}

static void gibbs_exp(const kxo_mech *m, double lnT, double T, double T2, double T3, double T4,
                      double rcpT, double *eg)
{
  for (int k = 0; k < m->n_species; ++k) {
    const double *a = (T > m->T_mid[k] ? m->g_hi : m->g_lo) + 7 * (size_t)k;
    const double g = a[0] * (1 - lnT) + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4 +
                     a[5] * rcpT - a[6];
    eg[k] = exp(g);
  }
}

/* Net rate of progress of every reaction (q_r); wdot optional (accumulated, +=). */
static void net_rates(const kxo_mech *m, double lnT, double T, double T2, double T3, double T4,
                      double rcpT, double P, double lnP, const double *C, double *q, double *qf)
{
  const int N = m->n_species;
  double eg[KXO_MAXSP], rcp_eg[KXO_MAXSP];
  double Cm = 0;
  for (int k = 0; k < N; ++k) Cm += C[k];
  gibbs_exp(m, lnT, T, T2, T3, T4, rcpT, eg);
  for (int k = 0; k < N; ++k) rcp_eg[k] = 1. / eg[k];
  const double C0 = m->one_atm_R * rcpT;
  const double rcpC0 = m->R_one_atm * T;
  for (int r = 0; r < m->n_reactions; ++r) {
    const int *nf = m->nu_f + (size_t)r * N, *nr = m->nu_r + (size_t)r * N;
    const double kf = forward_rate(m, r, lnT, T, T2, rcpT, P, lnP, C, Cm);
    double fwd = 1.0, rev = 1.0;
    int first_f = 1, first_r = 1;
    for (int k = 0; k < N; ++k) {
      for (int c = 0; c < nf[k]; ++c) { fwd = first_f ? C[k] : fwd * C[k]; first_f = 0; }
      for (int c = 0; c < nr[k]; ++c) { rev = first_r ? C[k] : rev * C[k]; first_r = 0; }
    }
    if (qf) qf[r] = kf * fwd;
    if (!m->reversible[r]) { q[r] = kf * fwd; continue; }
    /* reverse coefficient = 1/Kc: product of exp(g) over products, 1/exp(g) over
     * reactants (species-index order), times C0^-dnu (reaction_rates.py:1212-1231) */
    double kr = 1.0;
    int first = 1, dnu = 0;
    for (int k = 0; k < N; ++k) {
      const int net = nr[k] - nf[k];
      dnu += net;
      for (int c = 0; c < -net; ++c) { kr = first ? rcp_eg[k] : kr * rcp_eg[k]; first = 0; }
      for (int c = 0; c < net; ++c) { kr = first ? eg[k] : kr * eg[k]; first = 0; }
    }
    for (int c = 0; c < abs(dnu); ++c) kr *= (dnu < 0 ? C0 : rcpC0);
    q[r] = kf * (fwd - kr * rev);
  }
}

void kxo_species_rates(const kxo_mech *m, double lnT, double T, double T2, double T3, double T4,
                       double rcpT, double P, double lnP, const double *C, double *wdot)
{
  const int N = m->n_species;
  double *q = (double *)malloc(sizeof(double) * (size_t)m->n_reactions);
  net_rates(m, lnT, T, T2, T3, T4, rcpT, P, lnP, C, q, NULL);
  for (int r = 0; r < m->n_reactions; ++r) {
    const int *nf = m->nu_f + (size_t)r * N, *nr = m->nu_r + (size_t)r * N;
    for (int k = 0; k < N; ++k) {
      const int net = nr[k] - nf[k];
      if (net == 1) wdot[k] += q[r];
      else if (net == -1) wdot[k] += -q[r];
      else if (net != 0) wdot[k] += net * q[r];
    }
  }
  free(q);
}

void kxo_enthalpy_RT(const kxo_mech *m, double T, double T2, double T3, double T4, double rcpT,
                     double *h_RT)
{
  for (int k = 0; k < m->n_species; ++k) {
    const double *a = (T > m->T_mid[k] ? m->h_hi : m->h_lo) + 6 * (size_t)k;
    h_RT[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4 + a[5] * rcpT;
  }
}

void kxo_heat_capacity_R(const kxo_mech *m, double T, double T2, double T3, double T4, double *cp_R)
{
  for (int k = 0; k < m->n_species; ++k) {
    const double *a = (T > m->T_mid[k] ? m->cp_hi : m->cp_lo) + 5 * (size_t)k;
    cp_R[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
  }
}

double kxo_conductivity(const kxo_mech *m, double rcpMbar, double lnT, double lnT2, double lnT3,
                        double lnT4, const double *Xi)
{
  (void)rcpMbar;
  double sum1 = 0., sum2 = 0.;
  for (int k = 0; k < m->n_species; ++k) {
    const double *b = m->cond + 5 * (size_t)k;
    const double lambda_k = b[0] + b[1] * lnT + b[2] * lnT2 + b[3] * lnT3 + b[4] * lnT4;
    sum1 += Xi[k] * lambda_k;
    sum2 += Xi[k] / lambda_k;
  }
  return 0.5 * (sum1 + 1. / sum2);
}

double kxo_viscosity(const kxo_mech *m, double lnT, double lnT2, double lnT3, double lnT4,
                     const double *Xi)
{
  const int N = m->n_species;
  double eta[KXO_MAXSP], rcp_eta[KXO_MAXSP];
  for (int k = 0; k < N; ++k) {
    const double *a = m->visc + 5 * (size_t)k;
    eta[k] = a[0] + a[1] * lnT + a[2] * lnT2 + a[3] * lnT3 + a[4] * lnT4;
    rcp_eta[k] = 1. / eta[k];
  }
  double vis = 0.;
  for (int k = 0; k < N; ++k) {
    double s = 0.;
    for (int j = 0; j < N; ++j) {
      const size_t idx = (size_t)N * k + j;
      const double sqrt_Phi_kj = m->C1[idx] + m->C2[idx] * eta[k] * rcp_eta[j];
      s += Xi[j] * sqrt_Phi_kj * sqrt_Phi_kj;
    }
    vis += Xi[k] * eta[k] * eta[k] / s;
  }
  return vis;
}

void kxo_diffusivity(const kxo_mech *m, double Mbar, double p, double TsqrT, double lnT,
                     double lnT2, double lnT3, double lnT4, const double *Xi, double *Dkm)
{
  const int N = m->n_species;
  double sums1[KXO_MAXSP], sums2[KXO_MAXSP];
  for (int k = 0; k < N; ++k) sums1[k] = sums2[k] = 0.;
  for (int k = 1; k < N; ++k) {
    for (int j = 0; j < k; ++j) {
      const double *d = m->diff + ((size_t)k * N + j) * 5;
      const double poly = d[0] + d[1] * lnT + d[2] * lnT2 + d[3] * lnT3 + d[4] * lnT4;
      const double rcp_Dkj = m->rcp_diff ? poly : 1 / poly;
      sums1[k] += Xi[j] * rcp_Dkj;
      sums2[j] += Xi[k] * rcp_Dkj;
    }
  }
  for (int k = 0; k < N; ++k) {
    const double s = sums1[k] + sums2[k];
    Dkm[k] = TsqrT * (Mbar - m->W[k] * Xi[k]) / (p * Mbar * s);
  }
}

/* ---- kernel-level SoA entry points ---------------------------------------------------- */

static void state_prologue(const kxo_mech *m, int id, int offsetT, int offset,
                           const double *state, double *wrk1, double *rcpMbar_out)
{
  double rcpMbar = 0;
  for (int k = 0; k < m->n_species; ++k) {
    const double Yi = fmax(0.0, state[id + offsetT + (size_t)k * offset]);
    wrk1[k] = Yi * m->rcpW[k];
    rcpMbar += wrk1[k];
  }
  *rcpMbar_out = rcpMbar;
}

/* productionRates.okl:10-62. p_state (per-state pressure [Pa]) may be NULL, then pressure_
 * [Pa] is used for all states. scale (optional, same layout as rates) receives
 * W_k * sum_r |nu_kr q_r|, the condition scale of each species' rate. */
void kxo_production_rates(const kxo_mech *m, int n, int offsetT, int offset, double pressure_,
                          const double *p_state, const double *state, double *rates,
                          double *scale, double Tref)
{
  const int N = m->n_species;
  double wrk1[KXO_MAXSP], wrk2[KXO_MAXSP], sc[KXO_MAXSP];
  double *q = (double *)malloc(sizeof(double) * (size_t)(m->n_reactions + 1));
  for (int id = 0; id < n; ++id) {
    const double P = p_state ? p_state[id] : pressure_;
    const double pressure_R = P / m->R;
    const double T = Tref * state[id];
    const double rcpT = 1 / T;
    const double logT = log(T);
    const double T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;
    const double logP = log(P);
    double rcpMbar;
    state_prologue(m, id, offsetT, offset, state, wrk1, &rcpMbar);
    const double Mbar = 1 / rcpMbar;
    const double rho = pressure_R * rcpT * Mbar;
    for (int k = 0; k < N; ++k) { wrk1[k] = wrk1[k] * rho; wrk2[k] = 0; sc[k] = 0; }
    net_rates(m, logT, T, T2, T3, T4, rcpT, P, logP, wrk1, q, NULL);
    for (int r = 0; r < m->n_reactions; ++r) {
      const int *nf = m->nu_f + (size_t)r * N, *nr = m->nu_r + (size_t)r * N;
      for (int k = 0; k < N; ++k) {
        const int net = nr[k] - nf[k];
        if (net == 1) wrk2[k] += q[r];
        else if (net == -1) wrk2[k] += -q[r];
        else if (net != 0) wrk2[k] += net * q[r];
        sc[k] += fabs((double)net * q[r]);
      }
    }
    for (int k = 0; k < N; ++k) {
      rates[id + offsetT + (size_t)k * offset] = m->W[k] * wrk2[k];
      if (scale) scale[id + offsetT + (size_t)k * offset] = m->W[k] * sc[k];
    }
    kxo_enthalpy_RT(m, T, T2, T3, T4, rcpT, wrk1);
    double sum_h_RT = 0, sum_abs = 0;
    for (int k = 0; k < N; ++k) {
      sum_h_RT += wrk2[k] * wrk1[k];
      sum_abs += sc[k] * fabs(wrk1[k]);
    }
    rates[id] = -m->R * T * sum_h_RT;
    if (scale) scale[id] = m->R * T * sum_abs;
  }
  free(q);
}

/* Net rates of progress q_r [mol/m^3/s] of every reaction at each state (R x n, row-major by
 * reaction), for diagnostics and the equilibrium tests. */
void kxo_reaction_rates(const kxo_mech *m, int n, int offsetT, int offset, double pressure_,
                        const double *p_state, const double *state, double *q_out, double *qf_out,
                        double Tref)
{
  /* net rates q_r (and, if qf_out, the forward rates k_f prod C^nu') per reaction */
  const int N = m->n_species;
  double wrk1[KXO_MAXSP];
  double *q = (double *)malloc(sizeof(double) * (size_t)(m->n_reactions + 1));
  double *qf = qf_out ? (double *)malloc(sizeof(double) * (size_t)(m->n_reactions + 1)) : NULL;
  for (int id = 0; id < n; ++id) {
    const double P = p_state ? p_state[id] : pressure_;
    const double T = Tref * state[id];
    const double rcpT = 1 / T;
    double rcpMbar;
    state_prologue(m, id, offsetT, offset, state, wrk1, &rcpMbar);
    const double rho = P / m->R * rcpT / rcpMbar;
    for (int k = 0; k < N; ++k) wrk1[k] *= rho;
    net_rates(m, log(T), T, T * T, T * T * T, T * T * T * T, rcpT, P, log(P), wrk1, q, qf);
    for (int r = 0; r < m->n_reactions; ++r) {
      q_out[(size_t)r * n + id] = q[r];
      if (qf_out) qf_out[(size_t)r * n + id] = qf[r];
    }
  }
  free(q);
  free(qf);
}

/* transportProps.okl:11-47; pressure is the value the reference passes (normalised p, it
 * cancels in rho*D). If p_state is given, Dkm_out (optional) receives D_km [m^2/s] at the
 * per-state pressure [Pa]. */
void kxo_transport(const kxo_mech *m, int n, int offsetT, int offset, double pressure,
                   const double *p_state, const double *state, double *conductivity,
                   double *viscosity, double *rhoD, double *Dkm_out, double Tref)
{
  const int N = m->n_species;
  double wrk1[KXO_MAXSP], wrk2[KXO_MAXSP];
  for (int id = 0; id < n; ++id) {
    const double p = p_state ? p_state[id] : pressure;
    const double T = Tref * state[id];
    const double lnT = log(T);
    const double rcpT = 1 / T;
    const double sqrT = sqrt(T);
    const double lnT2 = lnT * lnT, lnT3 = lnT * lnT * lnT, lnT4 = lnT * lnT * lnT * lnT;
    double rcpMbar;
    state_prologue(m, id, offsetT, offset, state, wrk1, &rcpMbar);
    const double Mbar = 1 / rcpMbar;
    for (int k = 0; k < N; ++k) wrk1[k] *= Mbar;
    conductivity[id] = sqrT * kxo_conductivity(m, rcpMbar, lnT, lnT2, lnT3, lnT4, wrk1);
    viscosity[id] = sqrT * kxo_viscosity(m, lnT, lnT2, lnT3, lnT4, wrk1);
    kxo_diffusivity(m, Mbar, p, T * sqrT, lnT, lnT2, lnT3, lnT4, wrk1, wrk2);
    const double rho = p / m->R * rcpT * Mbar;
    for (int k = 0; k < N; ++k) {
      rhoD[(size_t)k * offset + id] = rho * wrk2[k];
      if (Dkm_out) Dkm_out[(size_t)k * offset + id] = wrk2[k];
    }
  }
}

/* thermoCoeffs.okl:10-39 */
void kxo_thermo(const kxo_mech *m, int n, int offsetT, int offset, double pressure_R,
                const double *state, double *rho, double *cp, double *rhoCp, double Tref)
{
  const int N = m->n_species;
  double wrk1[KXO_MAXSP], cp_R[KXO_MAXSP];
  for (int id = 0; id < n; ++id) {
    const double T = Tref * state[id];
    const double rcpT = 1 / T;
    const double T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;
    double rcpMbar;
    state_prologue(m, id, offsetT, offset, state, wrk1, &rcpMbar);
    const double Mbar = 1 / rcpMbar;
    const double rho_ = pressure_R * rcpT * Mbar;
    rho[id] = rho_;
    kxo_heat_capacity_R(m, T, T2, T3, T4, cp_R);
    double mean_cp_R = 0;
    for (int k = 0; k < N; ++k) {
      cp[(size_t)k * offset + id] = cp_R[k] * m->R * m->rcpW[k];
      mean_cp_R += cp_R[k] * wrk1[k] * Mbar;
    }
    rhoCp[id] = rho_ * (mean_cp_R * m->R * rcpMbar);
  }
}

// TEST INFRASTRUCTURE ONLY (oracle): never part of the product path.
//
// Driver around the reference's own pre-generated GRI-3.0 FP64 routines
// (examples/KinetiX/mechanisms/gri/{mech.h,rates.cpp,enthalpy_RT.cpp,heat_capacity_R.cpp,
// conductivity.cpp,viscosity.cpp,diffusivity.cpp}). Those files are pulled in with
// `-include` by oracle/Makefile straight from the reference tree; nothing of them is
// copied here. This file only restates the per-state kernel wrappers around them:
//   productionRates  : benchmark/okl/productionRates.okl:10-62
//   transport        : benchmark/okl/transportProps.okl:11-47
//   thermoCoeffs     : benchmark/okl/thermoCoeffs.okl:10-39
// and adds a std::thread fan-out over contiguous state chunks (the reference runs one
// SERIAL rank per core on n/C states, benchmark/src/bk.cpp:443).
//
// The example rates.cpp predates the P/lnP arguments (rates.cpp:2); GRI-3.0 has no PLOG
// reaction, so the 8-argument form computes the same thing.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace {

const double kR = 1.380649e-23 * 6.02214076e23;   // benchmark/src/kinetix.cpp:37

void rates_range(int lo, int hi, int offsetT, int offset, double pressure_R,
                 const double* state, double* rates, double Tref)
{
  for (int id = lo; id < hi; ++id) {
    const double T = Tref * state[id];
    const double rcpT = 1 / T;
    const double logT = log(T);
    const double T2 = T * T;
    const double T3 = T * T * T;
    const double T4 = T * T * T * T;
    double wrk1[n_species];
    double wrk2[n_species];
    double rcpMbar = 0;
    for (int k = 0; k < n_species; k++) {
      const double Yi = fmax(0.0, state[id + offsetT + k * offset]);
      wrk1[k] = Yi * kinetix_rcp_molar_mass[k];
      rcpMbar += wrk1[k];
    }
    const double Mbar = 1 / rcpMbar;
    const double rho = pressure_R * rcpT * Mbar;
    for (int k = 0; k < n_species; k++) {
      wrk1[k] = wrk1[k] * rho;
      wrk2[k] = 0;
    }
    kinetix_species_rates(logT, T, T2, T3, T4, rcpT, wrk1, wrk2);
    for (int k = 0; k < n_species; k++)
      rates[id + offsetT + k * offset] = kinetix_molar_mass[k] * wrk2[k];
    kinetix_enthalpy_RT(T, T2, T3, T4, rcpT, wrk1);
    double sum_h_RT = 0;
    for (int k = 0; k < n_species; k++) sum_h_RT += wrk2[k] * wrk1[k];
    rates[id] = -kR * T * sum_h_RT;
  }
}

void transport_range(int lo, int hi, int offsetT, int offset, double pressure,
                     const double* state, double* conductivity, double* viscosity,
                     double* rhoD, double Tref)
{
  for (int id = lo; id < hi; ++id) {
    const double T = Tref * state[id];
    const double lnT = log(T);
    const double rcpT = 1 / T;
    const double sqrT = sqrt(T);
    const double lnT2 = lnT * lnT;
    const double lnT3 = lnT * lnT * lnT;
    const double lnT4 = lnT * lnT * lnT * lnT;
    double wrk1[n_species];
    double wrk2[n_species];
    double rcpMbar = 0;
    for (int k = 0; k < n_species; k++) {
      const double Yi = fmax(0.0, state[id + offsetT + k * offset]);
      wrk1[k] = Yi * kinetix_rcp_molar_mass[k];
      rcpMbar += wrk1[k];
    }
    const double Mbar = 1 / rcpMbar;
    for (int k = 0; k < n_species; k++) wrk1[k] *= Mbar;
    conductivity[id] = sqrT * kinetix_conductivity(rcpMbar, lnT, lnT2, lnT3, lnT4, wrk1);
    viscosity[id] = sqrT * kinetix_viscosity(lnT, lnT2, lnT3, lnT4, wrk1);
    kinetix_diffusivity(Mbar, pressure, T * sqrT, lnT, lnT2, lnT3, lnT4, wrk1, wrk2);
    const double rho = pressure / kR * rcpT * Mbar;
    for (int k = 0; k < n_species; k++) rhoD[k * offset + id] = rho * wrk2[k];
  }
}

void thermo_range(int lo, int hi, int offsetT, int offset, double pressure_R,
                  const double* state, double* rho, double* cp, double* rhoCp, double Tref)
{
  for (int id = lo; id < hi; ++id) {
    const double T = Tref * state[id];
    const double rcpT = 1 / T;
    const double T2 = T * T;
    const double T3 = T * T * T;
    const double T4 = T * T * T * T;
    double wrk1[n_species];
    double rcpMbar = 0;
    for (int k = 0; k < n_species; k++) {
      const double Yi = fmax(0.0, state[id + offsetT + k * offset]);
      wrk1[k] = Yi * kinetix_rcp_molar_mass[k];
      rcpMbar += wrk1[k];
    }
    const double Mbar = 1 / rcpMbar;
    const double rho_ = pressure_R * rcpT * Mbar;
    rho[id] = rho_;
    double cp_R[n_species];
    kinetix_molar_heat_capacity_R(T, T2, T3, T4, cp_R);
    double mean_cp_R = 0;
    for (int k = 0; k < n_species; k++) {
      cp[k * offset + id] = cp_R[k] * kR * kinetix_rcp_molar_mass[k];
      mean_cp_R += cp_R[k] * wrk1[k] * Mbar;
    }
    rhoCp[id] = rho_ * (mean_cp_R * kR * rcpMbar);
  }
}

template <class F>
void fan_out(int n, int nthreads, F f)
{
  if (nthreads <= 1 || n < 2 * nthreads) { f(0, n); return; }
  std::vector<std::thread> pool;
  const int chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    const int lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo < hi) pool.emplace_back(f, lo, hi);
  }
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

int kxref_n_species() { return n_species; }
int kxref_n_active_species() { return n_active_species; }
int kxref_n_reactions() { return n_reactions; }
const char* kxref_species_names() { return species_names; }

// pressure_R = p [Pa] / R as in benchmark/src/kinetix.cpp:802.
void kxref_production_rates(int n, int offsetT, int offset, double pressure_R,
                            const double* state, double* rates, double Tref, int nthreads)
{
  fan_out(n, nthreads, [=](int lo, int hi) {
    rates_range(lo, hi, offsetT, offset, pressure_R, state, rates, Tref);
  });
}

void kxref_transport(int n, int offsetT, int offset, double pressure, const double* state,
                     double* conductivity, double* viscosity, double* rhoD, double Tref,
                     int nthreads)
{
  fan_out(n, nthreads, [=](int lo, int hi) {
    transport_range(lo, hi, offsetT, offset, pressure, state, conductivity, viscosity, rhoD, Tref);
  });
}

void kxref_thermo(int n, int offsetT, int offset, double pressure_R, const double* state,
                  double* rho, double* cp, double* rhoCp, double Tref, int nthreads)
{
  fan_out(n, nthreads, [=](int lo, int hi) {
    thermo_range(lo, hi, offsetT, offset, pressure_R, state, rho, cp, rhoCp, Tref);
  });
}

}  // extern "C"

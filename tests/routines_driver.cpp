// Test driver for the generated routine files (kinetix_amd/mechs/<m>/routines), compiled for
// the CPU with the reference example's macros (examples/KinetiX/Makefile:4-26) and -I to the
// routines directory; structured as the reference's productionRates and transport kernels
// (benchmark/okl/productionRates.okl:10-62, transportProps.okl:11-47), one state at a time.
#include <cmath>

#include "mech.h"
#include "rates.cpp"
#include "enthalpy_RT.cpp"
#include "heat_capacity_R.cpp"
#include "conductivity.cpp"
#include "viscosity.cpp"
#include "diffusivity.cpp"

extern "C" int rt_n_species() { return n_species; }

// SoA state (T/Tref in row 0, Y_k in rows 1..N, n states per row), rates as the kernels write
extern "C" void rt_production_rates(int n, double pressure_R, double pressure, const double* state,
                                    double* rates, double R)
{
  for (int id = 0; id < n; ++id) {
    const cfloat T = state[id], rcpT = 1 / T, lnT = log(T);
    const cfloat T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;
    cfloat c[n_species], wdot[n_species];
    cfloat rcpMbar = 0;
    for (int k = 0; k < n_species; ++k) {
      c[k] = fmax(0.0, state[id + n + k * n]) * kinetix_rcp_molar_mass[k];
      rcpMbar += c[k];
    }
    const cfloat rho = pressure_R * rcpT / rcpMbar;
    for (int k = 0; k < n_species; ++k) { c[k] *= rho; wdot[k] = 0; }
    kinetix_species_rates(lnT, T, T2, T3, T4, rcpT, pressure, log(pressure), c, wdot);
    kinetix_enthalpy_RT(T, T2, T3, T4, rcpT, c);
    cfloat sum_h = 0;
    for (int k = 0; k < n_species; ++k) {
      rates[id + n + k * n] = kinetix_molar_mass[k] * wdot[k];
      sum_h += wdot[k] * c[k];
    }
    rates[id] = -R * T * sum_h;
  }
}

extern "C" void rt_heat_capacity_R(int n, const double* state, double* cp_R)
{
  for (int id = 0; id < n; ++id) {
    const cfloat T = state[id];
    cfloat cp[n_species];
    kinetix_molar_heat_capacity_R(T, T * T, T * T * T, T * T * T * T, cp);
    for (int k = 0; k < n_species; ++k) cp_R[k * n + id] = cp[k];
  }
}

extern "C" void rt_transport(int n, double pressure, const double* state, double* cond,
                             double* visc, double* rhoD, double R)
{
  for (int id = 0; id < n; ++id) {
    const cfloat T = state[id], lnT = log(T), rcpT = 1 / T, sqrT = sqrt(T);
    const cfloat lnT2 = lnT * lnT, lnT3 = lnT * lnT * lnT, lnT4 = lnT * lnT * lnT * lnT;
    cfloat X[n_species], D[n_species];
    cfloat rcpMbar = 0;
    for (int k = 0; k < n_species; ++k) {
      X[k] = fmax(0.0, state[id + n + k * n]) * kinetix_rcp_molar_mass[k];
      rcpMbar += X[k];
    }
    const cfloat Mbar = 1 / rcpMbar;
    for (int k = 0; k < n_species; ++k) X[k] *= Mbar;
    cond[id] = sqrT * kinetix_conductivity(rcpMbar, lnT, lnT2, lnT3, lnT4, X);
    visc[id] = sqrT * kinetix_viscosity(lnT, lnT2, lnT3, lnT4, X);
    kinetix_diffusivity(Mbar, pressure, T * sqrT, lnT, lnT2, lnT3, lnT4, X, D);
    const cfloat rho = pressure / R * rcpT * Mbar;
    for (int k = 0; k < n_species; ++k) rhoD[k * n + id] = rho * D[k];
  }
}

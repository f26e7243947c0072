/*
 * kinetix_mi.h -- C ABI of the MI355X-native (gfx950) KinetiX hot path.
 *
 * Drop-in for the reference's host API benchmark/src/kinetix.hpp:17-107 (namespace
 * kinetix, C++ with occa::device/occa::memory arguments). Here every argument is a plain
 * pointer or size: device buffers are HIP device pointers, the stream is a hipStream_t
 * passed as void* (NULL = the default stream). Every entry point returns an int status
 * (KX_OK = 0, negative on error) and never throws across the ABI; launches are
 * asynchronous on the given stream, as the reference's OCCA launches are
 * (benchmark/src/bk.cpp:697-716 synchronises with device.finish()).
 *
 * Device-side layout (benchmark/src/bk.cpp:606-648, benchmark/okl/productionRates.okl:1-8):
 * structure of arrays, FP64,
 *     state[id]                       = T / T_ref
 *     state[id + offsetT + k*offset]  = Y_k               (k = 0 .. n_species-1)
 * with offsetT >= n_states and offset >= n_states (offsetT = offset = n_states for a
 * dense array; larger values address a sub-view of a bigger SoA buffer, the state-sharding
 * idiom of the reference's offset arguments).
 */
#ifndef KINETIX_MI_H
#define KINETIX_MI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  KX_OK = 0,
  KX_ERR_INVALID_ARG = -1,
  KX_ERR_NOT_INITIALIZED = -2,   /* handle unknown or finalised */
  KX_ERR_NOT_BUILT = -3,         /* kx_build not called (reference: assert(initialized)) */
  KX_ERR_HIP = -4,               /* a HIP runtime call failed; kx_last_error() has the text */
  KX_ERR_IO = -5,                /* mechanism artefacts missing or unreadable */
  KX_ERR_NO_KERNEL = -6,         /* the code object lacks the requested kernel variant */
  KX_ERR_TOO_LARGE = -7          /* request exceeds a documented size limit */
};

typedef int kx_mech_t;           /* handle of one loaded mechanism on one device */

/* Human-readable text of the last error on this thread. */
const char *kx_last_error(void);

/*
 * kinetix::init (kinetix.cpp:481-599): load one mechanism's generated artefacts
 * (mech_dir/kernels.hsaco + mech_dir/mech.txt, written by `python -m kinetix_amd.gen`)
 * onto HIP device `device_id`. The reference runs its generator and JIT-compiles OKL here;
 * this build generates and compiles ahead of time, so init only loads the code object.
 */
int kx_init(const char *mech_dir, int device_id, kx_mech_t *out);

/*
 * The host-application hook of kinetix::init (kinetixBuildKernel_t, kinetix.hpp:11-13;
 * kinetix.cpp:499-502, 542): the reference lets a host code supply the kernel build; here
 * the application supplies the built artefacts from memory instead of a directory -- the
 * text of mech.txt (mech_txt_bytes bytes, not NUL-terminated necessarily) and the gfx950
 * code object (as written to kernels.hsaco; it must stay valid for the call only). For
 * codes that embed the mechanism in their binary or run the generator in their own build.
 */
int kx_init_image(const char *mech_txt, size_t mech_txt_bytes, const void *code_object,
                  size_t code_object_bytes, int device_id, kx_mech_t *out);

/* kinetix::build (kinetix.cpp:609-784): reference pressure [Pa], temperature [K] and mass
 * fractions (n_species values); the state arrays hold T/T_ref and p arguments are p/p_ref. */
int kx_build(kx_mech_t m, double ref_pressure, double ref_temperature,
             const double *ref_mass_fractions);

int kx_finalize(kx_mech_t m);

/*
 * The generated routine files (mechs/<m>/routines: mech.h, rates.cpp, enthalpy_RT.cpp,
 * heat_capacity_R.cpp, conductivity.cpp, viscosity.cpp, diffusivity.cpp -- the reference's
 * device-routine ABI, reaction_rates.py:1271-1273, examples/KinetiX/kinetix_test.hpp) called
 * from kernels structured as the reference's productionRates and transport kernels
 * (benchmark/okl/productionRates.okl:10-62, transportProps.okl:11-47): the form a host code
 * that includes those files runs. Same arguments and outputs as kx_production_rates and
 * kx_mixture_avg_transport, FP64; slower than those fused kernels.
 */
int kx_production_rates_routines(kx_mech_t m, int n_states, int offsetT, int offset,
                                 double pressure, const double *d_state, double *d_rates,
                                 void *stream);
int kx_mixture_avg_transport_routines(kx_mech_t m, int n_states, int offsetT, int offset,
                                      double pressure, const double *d_state, double *d_viscosity,
                                      double *d_conductivity, double *d_rhoD, void *stream);

/* kinetix::productionRates (kinetix.cpp:786-813) -> benchmark/okl/productionRates.okl.
 * pressure = p/p_ref (scalar for all states). rates: (n_species+1) rows in the state
 * layout; rates[id] = heat release rate [W/m^3], rates[id+offsetT+k*offset] = W_k wdot_k
 * [kg/m^3/s]. FP64 throughout. */
int kx_production_rates(kx_mech_t m, int n_states, int offsetT, int offset, double pressure,
                        const double *d_state, double *d_rates, void *stream);

/* Mixed-precision variant of kx_production_rates (no reference counterpart; the nearest is
 * the reference's "fpmix" kernels, kinetix.cpp:254-281, which run ALL kinetics in FP32):
 * Arrhenius/falloff/third-body in FP32, equilibrium constants, net forward-minus-reverse
 * rates and the per-species sums in FP64. Same layout and units as kx_production_rates.
 * Runs the fastest correct mixed kernel, kx_bk1_mixed_valu: the stoichiometric
 * accumulation q.nu as FP64 FMAs on the vector ALUs (2.1-2.2x the FP64 kernel on MI355X at
 * GRI-3.0, relative L2 error per species ~1.3e-6 against FP64). */
int kx_production_rates_mixed(kx_mech_t m, int n_states, int offsetT, int offset,
                              double pressure, const double *d_state, double *d_rates,
                              void *stream);

/* The same kernel as kx_production_rates_mixed, under its explicit name. */
int kx_production_rates_mixed_valu(kx_mech_t m, int n_states, int offsetT, int offset,
                                   double pressure, const double *d_state, double *d_rates,
                                   void *stream);

/* Same precision split, with q.nu accumulated on FP32 MFMA (v_mfma_f32_16x16x4_f32, FP32
 * accumulation per reaction block, blocks folded in FP64) over the wave's 64 states. Slower
 * than kx_production_rates_mixed on MI355X (the q values make an LDS round trip); kept for
 * the MFMA-accumulation form of the mixed kinetics. */
int kx_production_rates_mixed_mfma(kx_mech_t m, int n_states, int offsetT, int offset,
                                   double pressure, const double *d_state, double *d_rates,
                                   void *stream);

/* kinetix::mixtureAvgTransportProps (kinetix.cpp:815-841) -> transportProps.okl.
 * Argument order follows the reference API: viscosity [Pa s], conductivity [W/m/K],
 * rhoD (n_species rows, rho*D_km [kg/m/s], at [k*offset + id]). pressure is passed
 * through as the reference does (it cancels in rho*D). */
int kx_mixture_avg_transport(kx_mech_t m, int n_states, int offsetT, int offset,
                             double pressure, const double *d_state, double *d_viscosity,
                             double *d_conductivity, double *d_rhoD, void *stream);

/* Same arguments, outputs and FP64 arithmetic as kx_mixture_avg_transport, with the O(N^2)
 * Wilke viscosity sum (kinetix_viscosity, mix_transport.py:337-345) computed as three GEMMs
 * (C1^2, 2 C1 C2, C2^2 against x_j, x_j/eta_j, x_j/eta_j^2) on FP64 MFMA
 * (v_mfma_f64_16x16x4_f64) over the states of each wavefront. Straight-line mechanisms only
 * (KX_ERR_NO_KERNEL for the loop-form ones). */
int kx_mixture_avg_transport_mfma(kx_mech_t m, int n_states, int offsetT, int offset,
                                  double pressure, const double *d_state, double *d_viscosity,
                                  double *d_conductivity, double *d_rhoD, void *stream);

/* Per-state-pressure variant (no reference counterpart: the reference takes one scalar
 * p): d_p holds p [Pa] per state; if d_Dkm is non-NULL it also receives the mixture-averaged
 * diffusion coefficients D_km [m^2/s] at that pressure. */
int kx_mixture_avg_transport_pstate(kx_mech_t m, int n_states, int offsetT, int offset,
                                    const double *d_p, const double *d_state,
                                    double *d_viscosity, double *d_conductivity,
                                    double *d_rhoD, double *d_Dkm, void *stream);

/* kinetix::thermodynamicProps (kinetix.cpp:843-871) -> thermoCoeffs.okl: rho [kg/m^3],
 * cp_k [J/kg/K] (n_species rows), rho*cp [J/m^3/K]. */
int kx_thermodynamic_props(kx_mech_t m, int n_states, int offsetT, int offset,
                           double pressure, const double *d_state, double *d_rho,
                           double *d_cp, double *d_rhoCp, void *stream);

/* Queries (kinetix.hpp:91-103; metadata kernels of benchmark/okl/mech.okl). */
int kx_n_species(kx_mech_t m);
int kx_n_active_species(kx_mech_t m);
int kx_n_reactions(kx_mech_t m);
/* Copies the k-th species name (NUL-terminated) into buf; returns its length or < 0. */
int kx_species_name(kx_mech_t m, int k, char *buf, int buflen);
int kx_species_index(kx_mech_t m, const char *name);       /* -1 if absent */
/* molecularWeights(): W_k / Mbar_ref (kinetix.cpp:885-892); needs kx_build. */
int kx_molecular_weights(kx_mech_t m, double *out);
/* Molar masses W_k [kg/mol] (mech.h kinetix_molar_mass). */
int kx_molar_masses(kx_mech_t m, double *out);
double kx_ref_pressure(kx_mech_t m);
double kx_ref_temperature(kx_mech_t m);
double kx_ref_mean_molecular_weight(kx_mech_t m);
/* refMassFractions() (kinetix.hpp:100, kinetix.cpp:605): copies the n_species reference mass
 * fractions given to kx_build into out; needs kx_build. */
int kx_ref_mass_fractions(kx_mech_t m, double *out);

/* Diagnostics (no reference counterpart): runs one wave of v_mfma_f32_16x16x4_f32 on
 * closed-form 16x4 and 4x16 operands laid out as the generated mixed-precision kernels
 * assume, and writes the 16x16 product row-major to d_out (256 floats). */
int kx_selftest_mfma_layout(float *d_out, void *stream);

/* Diagnostics (no reference counterpart): one wave of v_mfma_f64_16x16x4_f64 on closed-form
 * operands (mode 0: every operand entry used; mode 1: D[i][j] = 16 i + j), each lane's four
 * raw accumulators written to d_out[4 * lane + r] (256 doubles), so that the host can check
 * the lane layout the FP64-MFMA transport kernel assumes. */
int kx_selftest_mfma_f64_layout(int mode, double *d_out, void *stream);

/* Diagnostics (no reference counterpart): d_out[i] = the transport kernels' FP64 reciprocal
 * (v_rcp_f64 + two Newton steps, kinetix_amd/csrc/kx_rcp.h) of d_in[i], i < n. */
int kx_selftest_rcp(int n, const double *d_in, double *d_out, void *stream);

/* Learned transport surrogate (no reference counterpart): see kinetix_mi_surrogate.h. */

#ifdef __cplusplus
}
#endif
#endif /* KINETIX_MI_H */

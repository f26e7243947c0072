// Host side of the C ABI declared in include/kinetix_mi.h.
//
// The reference's host API (benchmark/src/kinetix.cpp) JIT-builds OCCA kernels per
// mechanism and keeps process-global state. Here each mechanism is an ahead-of-time
// compiled gfx950 code object loaded with hipModuleLoad, held behind an integer handle;
// launches go straight to hipModuleLaunchKernel on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/kinetix_mi.h"
#include "../../include/kinetix_mi_surrogate.h"

namespace {

constexpr double kR = 1.380649e-23 * 6.02214076e23;   // kinetix.cpp:37

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
  g_err = msg;
  return code;
}

#define KX_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(KX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

struct Mech {
  int device = 0;
  hipModule_t module = nullptr;
  std::map<std::string, hipFunction_t> fn;
  std::map<std::string, unsigned> block;     // each kernel's launch bound (threads per block)
  std::map<std::string, int> lanes;          // lanes per state of kernels that share a state
  int queue_lanes = 1;                        // lanes per queued state of kx_bk2_fp64_queue
  long long queue_grid_cap = 0;               // its workgroups resident at once (strided form)
  int n_species = 0, n_active = 0, n_reactions = 0;
  std::vector<std::string> names;
  std::vector<double> W;
  std::string dir;
  bool built = false;
  double ref_p = 0, ref_T = 0, ref_Mbar = 0;
  std::vector<double> ref_Y;
};

std::mutex g_mu;
std::map<int, Mech*> g_mechs;
int g_next = 1;

// ---- device scratch for kx_sur_transport_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends. Host code must not pass `scratch` itself to a kernel (in
// host code the name is a stand-in, not a device address): kx_init resolves each device's
// address once with hipGetSymbolAddress, with that device current, and the entry point uses
// only that address.
constexpr size_t kScratchBytes = KX_SUR_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// with `device` current
int scratch_setup(int device)
{
  if (device < 0 || device >= kMaxDevices) return fail(KX_ERR_INVALID_ARG, "device id out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device]) return KX_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(KX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return KX_OK;
}

Mech* get(kx_mech_t h)
{
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_mechs.find(h);
  return it == g_mechs.end() ? nullptr : it->second;
}

// mech.txt: one "key values..." line each for n_species, n_active_species, n_reactions,
// species (names) and molar_mass (kg/mol), in any order
int parse_meta(Mech* m, std::istream& f, const std::string& what)
{
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string key;
    is >> key;
    if (key == "n_species") is >> m->n_species;
    else if (key == "n_active_species") is >> m->n_active;
    else if (key == "n_reactions") is >> m->n_reactions;
    else if (key == "species") { std::string s; while (is >> s) m->names.push_back(s); }
    else if (key == "molar_mass") { double w; while (is >> w) m->W.push_back(w); }
  }
  if (m->n_species <= 0 || (int)m->names.size() != m->n_species || (int)m->W.size() != m->n_species)
    return fail(KX_ERR_IO, "malformed " + what);
  return KX_OK;
}

int read_meta(Mech* m, const std::string& path)
{
  std::ifstream f(path);
  if (!f) return fail(KX_ERR_IO, "cannot open " + path);
  return parse_meta(m, f, path);
}

// Launch helper: packs the kernel arguments; 1-D grid in workgroups of the kernel's own
// launch bound (the generator picks it per kernel). n_lanes counts states (or queue slots),
// one lane each; a kernel that exports its lanes per state gets that many per state.
int launch(Mech* m, const char* name, long long n_lanes, void* stream, void** args)
{
  auto it = m->fn.find(name);
  if (it == m->fn.end() || !it->second)
    return fail(KX_ERR_NO_KERNEL, std::string("kernel ") + name + " not in " + m->dir);
  if (n_lanes == 0) return KX_OK;
  KX_HIP(hipSetDevice(m->device));
  const unsigned block = m->block.at(name);
  auto l = m->lanes.find(name);
  if (l != m->lanes.end()) n_lanes *= l->second;
  const long long g = (n_lanes + (long long)block - 1) / block;
  if (g > 0x7fffffffLL) return fail(KX_ERR_INVALID_ARG, "grid too large");
  const unsigned grid = (unsigned)g;
  KX_HIP(hipModuleLaunchKernel(it->second, grid, 1, 1, block, 1, 1, 0, (hipStream_t)stream,
                               args, nullptr));
  return KX_OK;
}

// An explicit grid of workgroups of the kernel's launch bound.
int launch_grid(Mech* m, const char* name, long long grid, void* stream, void** args)
{
  auto it = m->fn.find(name);
  if (it == m->fn.end() || !it->second)
    return fail(KX_ERR_NO_KERNEL, std::string("kernel ") + name + " not in " + m->dir);
  if (grid == 0) return KX_OK;
  if (grid < 0 || grid > 0x7fffffffLL) return fail(KX_ERR_INVALID_ARG, "grid out of range");
  KX_HIP(hipSetDevice(m->device));
  KX_HIP(hipModuleLaunchKernel(it->second, (unsigned)grid, 1, 1, m->block.at(name), 1, 1, 0,
                               (hipStream_t)stream, args, nullptr));
  return KX_OK;
}

int check_layout(Mech* m, int n, int offsetT, int offset, const void* a, const void* b)
{
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (!m->built) return fail(KX_ERR_NOT_BUILT, "kx_build has not been called");
  if (n < 0) return fail(KX_ERR_INVALID_ARG, "n_states < 0");
  if (n > 0 && (!a || !b)) return fail(KX_ERR_INVALID_ARG, "null device pointer");
  if (offsetT < n || offset < n)
    return fail(KX_ERR_INVALID_ARG, "offsetT and offset must be >= n_states");
  return KX_OK;
}

}  // namespace

extern "C" {

const char* kx_last_error(void) { return g_err.c_str(); }

// Registers the module's kernels with their launch bounds and publishes the handle; takes
// ownership of m (deleted on failure).
static int finish_init(Mech* m, kx_mech_t* out)
{
  for (const char* k : {"kx_bk1_fp64", "kx_bk1_mixed", "kx_bk1_mixed_valu", "kx_bk2_fp64",
                        "kx_bk2_fp64_pstate", "kx_bk2_fp64_mfma", "kx_thermo_fp64", "kx_sur_features",
                        "kx_sur_targets", "kx_sur_outputs", "kx_sur_infer",
                        "kx_sur_queue_init", "kx_bk2_fp64_queue", "kx_sur_decide",
                        "kx_bk1_routines", "kx_bk2_routines"}) {
    hipFunction_t f = nullptr;
    if (hipModuleGetFunction(&f, m->module, k) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    int bound = 0;
    if (hipFuncGetAttribute(&bound, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, f) != hipSuccess ||
        bound <= 0 || bound > 1024 || bound % 64) {
      (void)hipGetLastError();
      (void)hipModuleUnload(m->module);
      delete m;
      return fail(KX_ERR_HIP, std::string("no usable launch bound for ") + k);
    }
    m->fn[k] = f;
    m->block[k] = (unsigned)bound;
    // a kernel that spreads each state over several lanes exports how many
    hipDeviceptr_t lp = nullptr;
    size_t lbytes = 0;
    if (hipModuleGetGlobal(&lp, &lbytes, m->module, (std::string(k) + "_lanes_per_state").c_str()) ==
            hipSuccess &&
        lbytes == sizeof(int)) {
      int lanes = 0;
      if (hipMemcpyDtoH(&lanes, lp, sizeof(int)) != hipSuccess || lanes < 1 || bound % lanes) {
        (void)hipGetLastError();
        (void)hipModuleUnload(m->module);
        delete m;
        return fail(KX_ERR_HIP, std::string("unusable lanes per state for ") + k);
      }
      m->lanes[k] = lanes;
    } else {
      (void)hipGetLastError();
    }
  }
  // the queue kernel may share each state among several lanes; it then exports how many
  {
    hipDeviceptr_t p = nullptr;
    size_t bytes = 0;
    if (hipModuleGetGlobal(&p, &bytes, m->module, "kx_bk2q_lanes") == hipSuccess &&
        bytes == sizeof(int)) {
      int lanes = 0;
      if (hipMemcpyDtoH(&lanes, p, sizeof(int)) != hipSuccess || lanes < 1 ||
          !m->block.count("kx_bk2_fp64_queue") || m->block["kx_bk2_fp64_queue"] % lanes) {
        (void)hipGetLastError();
        (void)hipModuleUnload(m->module);
        delete m;
        return fail(KX_ERR_HIP, "unusable kx_bk2q_lanes");
      }
      m->queue_lanes = lanes;
      // the strided queue kernel runs on a grid capped at the workgroups resident at once
      int ncu = 0, per_cu = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device) !=
              hipSuccess || ncu <= 0)
        ncu = 1;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
              &per_cu, m->fn["kx_bk2_fp64_queue"], (int)m->block["kx_bk2_fp64_queue"], 0) !=
              hipSuccess || per_cu <= 0)
        per_cu = 1;
      (void)hipGetLastError();
      m->queue_grid_cap = (long long)ncu * per_cu;
    } else {
      (void)hipGetLastError();
    }
  }
  std::lock_guard<std::mutex> lk(g_mu);
  *out = g_next++;
  g_mechs[*out] = m;
  return KX_OK;
}

int kx_init(const char* mech_dir, int device_id, kx_mech_t* out)
{
  if (!mech_dir || !out) return fail(KX_ERR_INVALID_ARG, "null argument");
  Mech* m = new Mech();
  m->dir = mech_dir;
  m->device = device_id;
  int rc = read_meta(m, m->dir + "/mech.txt");
  if (rc) { delete m; return rc; }
  hipError_t e = hipSetDevice(device_id);
  if (e == hipSuccess) e = hipModuleLoad(&m->module, (m->dir + "/kernels.hsaco").c_str());
  if (e != hipSuccess) {
    delete m;
    return fail(e == hipErrorFileNotFound ? KX_ERR_IO : KX_ERR_HIP,
                std::string("hipModuleLoad(") + mech_dir + "/kernels.hsaco): " + hipGetErrorString(e));
  }
  if ((rc = scratch_setup(device_id))) { (void)hipModuleUnload(m->module); delete m; return rc; }
  return finish_init(m, out);
}

int kx_init_image(const char* mech_txt, size_t mech_txt_bytes, const void* code_object,
                  size_t code_object_bytes, int device_id, kx_mech_t* out)
{
  if (!mech_txt || !code_object || !out || code_object_bytes == 0)
    return fail(KX_ERR_INVALID_ARG, "null argument");
  Mech* m = new Mech();
  m->dir = "<image>";
  m->device = device_id;
  std::istringstream meta(std::string(mech_txt, mech_txt_bytes));
  int rc = parse_meta(m, meta, "mechanism image text");
  if (rc) { delete m; return rc; }
  hipError_t e = hipSetDevice(device_id);
  if (e == hipSuccess) e = hipModuleLoadData(&m->module, code_object);
  if (e != hipSuccess) {
    delete m;
    return fail(KX_ERR_HIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
  }
  if ((rc = scratch_setup(device_id))) { (void)hipModuleUnload(m->module); delete m; return rc; }
  return finish_init(m, out);
}

int kx_build(kx_mech_t h, double ref_pressure, double ref_temperature, const double* refY)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (!refY || !(ref_pressure > 0) || !(ref_temperature > 0))
    return fail(KX_ERR_INVALID_ARG, "reference state must be positive with mass fractions");
  double sum = 0.;
  m->ref_Y.assign(refY, refY + m->n_species);
  for (int k = 0; k < m->n_species; ++k) sum += refY[k] / m->W[k];
  if (!(sum > 0)) return fail(KX_ERR_INVALID_ARG, "reference mass fractions sum to zero");
  m->ref_Mbar = 1. / sum;
  m->ref_p = ref_pressure;
  m->ref_T = ref_temperature;
  m->built = true;
  return KX_OK;
}

int kx_finalize(kx_mech_t h)
{
  Mech* m = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mechs.find(h);
    if (it == g_mechs.end()) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
    m = it->second;
    g_mechs.erase(it);
  }
  if (m->module) { (void)hipSetDevice(m->device); (void)hipModuleUnload(m->module); }
  delete m;
  return KX_OK;
}

static int production_rates_impl(const char* kernel, kx_mech_t h, int n, int offsetT,
                                 int offset, double pressure, const double* d_state,
                                 double* d_rates, void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_rates);
  if (rc) return rc;
  // kinetix.cpp:802-803
  double pressure_R = pressure * m->ref_p / kR;
  double pressure_ = pressure * m->ref_p;
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  const double* s = d_state;
  double* r = d_rates;
  void* args[] = {&n, &oT, &o, &pressure_R, &pressure_, &s, &r, &Tref};
  return launch(m, kernel, n, stream, args);
}

int kx_production_rates(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                        const double* d_state, double* d_rates, void* stream)
{
  return production_rates_impl("kx_bk1_fp64", h, n, offsetT, offset, pressure, d_state, d_rates,
                               stream);
}

int kx_production_rates_mixed(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                              const double* d_state, double* d_rates, void* stream)
{
  return production_rates_impl("kx_bk1_mixed_valu", h, n, offsetT, offset, pressure, d_state,
                               d_rates, stream);
}

int kx_production_rates_mixed_mfma(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                                   const double* d_state, double* d_rates, void* stream)
{
  return production_rates_impl("kx_bk1_mixed", h, n, offsetT, offset, pressure, d_state, d_rates,
                               stream);
}

int kx_production_rates_mixed_valu(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                                   const double* d_state, double* d_rates, void* stream)
{
  return production_rates_impl("kx_bk1_mixed_valu", h, n, offsetT, offset, pressure, d_state,
                               d_rates, stream);
}

int kx_production_rates_routines(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                                 const double* d_state, double* d_rates, void* stream)
{
  return production_rates_impl("kx_bk1_routines", h, n, offsetT, offset, pressure, d_state, d_rates,
                               stream);
}

int kx_mixture_avg_transport_routines(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                                      const double* d_state, double* d_visc, double* d_cond,
                                      double* d_rhoD, void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_rhoD);
  if (rc) return rc;
  if (n > 0 && (!d_visc || !d_cond)) return fail(KX_ERR_INVALID_ARG, "null device pointer");
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  const double* s = d_state;
  void* args[] = {&n, &oT, &o, &pressure, &s, &d_cond, &d_visc, &d_rhoD, &Tref};
  return launch(m, "kx_bk2_routines", n, stream, args);
}

static int transport_impl(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                          const double* d_p, const double* d_state, double* d_visc,
                          double* d_cond, double* d_rhoD, double* d_Dkm, void* stream,
                          const int* n_dev = nullptr)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_rhoD);
  if (rc) return rc;
  if (n > 0 && (!d_visc || !d_cond)) return fail(KX_ERR_INVALID_ARG, "null device pointer");
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  const double* s = d_state;
  if (d_Dkm) {
    // per-state pressure, D_km output (conductivity, viscosity and rho*D do not depend on p)
    const double* p = d_p;
    void* args[] = {&n, &oT, &o, &p, &s, &d_cond, &d_visc, &d_rhoD, &d_Dkm, &Tref, &n_dev};
    return launch(m, "kx_bk2_fp64_pstate", n, stream, args);
  }
  void* args[] = {&n, &oT, &o, &pressure, &s, &d_cond, &d_visc, &d_rhoD, &Tref, &n_dev};
  return launch(m, "kx_bk2_fp64", n, stream, args);
}

int kx_mixture_avg_transport(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                             const double* d_state, double* d_visc, double* d_cond,
                             double* d_rhoD, void* stream)
{
  return transport_impl(h, n, offsetT, offset, pressure, nullptr, d_state, d_visc, d_cond, d_rhoD,
                        nullptr, stream);
}

int kx_mixture_avg_transport_mfma(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                                  const double* d_state, double* d_visc, double* d_cond,
                                  double* d_rhoD, void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_rhoD);
  if (rc) return rc;
  if (n > 0 && (!d_visc || !d_cond)) return fail(KX_ERR_INVALID_ARG, "null device pointer");
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  const double* s = d_state;
  const int* n_dev = nullptr;
  void* args[] = {&n, &oT, &o, &pressure, &s, &d_cond, &d_visc, &d_rhoD, &Tref, &n_dev};
  return launch(m, "kx_bk2_fp64_mfma", n, stream, args);
}

int kx_mixture_avg_transport_pstate(kx_mech_t h, int n, int offsetT, int offset,
                                    const double* d_p, const double* d_state, double* d_visc,
                                    double* d_cond, double* d_rhoD, double* d_Dkm, void* stream)
{
  if (n > 0 && !d_p) return fail(KX_ERR_INVALID_ARG, "null pressure array");
  return transport_impl(h, n, offsetT, offset, 1.0, d_p, d_state, d_visc, d_cond, d_rhoD, d_Dkm,
                        stream);
}

int kx_thermodynamic_props(kx_mech_t h, int n, int offsetT, int offset, double pressure,
                           const double* d_state, double* d_rho, double* d_cp, double* d_rhoCp,
                           void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_cp);
  if (rc) return rc;
  if (n > 0 && (!d_rho || !d_rhoCp)) return fail(KX_ERR_INVALID_ARG, "null device pointer");
  double pressure_R = pressure * m->ref_p / kR;   // kinetix.cpp:858
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  const double* s = d_state;
  void* args[] = {&n, &oT, &o, &pressure_R, &s, &d_rho, &d_cp, &d_rhoCp, &Tref};
  return launch(m, "kx_thermo_fp64", n, stream, args);
}

// ---- learned transport surrogate: mechanism-specific kernels of the generated code object
int kx_sur_features(kx_mech_t h, int n, int offsetT, int offset, const double* d_p,
                    const double* d_state, const float* d_norm, float* d_xin, float* d_base,
                    int* d_flags, void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_p);
  if (rc) return rc;
  if (n > 0 && (!d_norm || !d_xin || !d_base || !d_flags))
    return fail(KX_ERR_INVALID_ARG, "null device pointer");
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  void* args[] = {&n, &oT, &o, &d_p, &d_state, &d_norm, &d_xin, &d_base, &d_flags, &Tref};
  return launch(m, "kx_sur_features", n, stream, args);
}

int kx_sur_targets(kx_mech_t h, int n, int offset, const double* d_visc, const double* d_cond,
                   const double* d_Dkm, const float* d_base, const float* d_ostat, float* d_tgt,
                   void* stream)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (n < 0 || offset < n) return fail(KX_ERR_INVALID_ARG, "bad size/offset");
  if (n > 0 && (!d_visc || !d_cond || !d_Dkm || !d_base || !d_ostat || !d_tgt))
    return fail(KX_ERR_INVALID_ARG, "null device pointer");
  long long o = offset;
  void* args[] = {&n, &o, &d_visc, &d_cond, &d_Dkm, &d_base, &d_ostat, &d_tgt};
  return launch(m, "kx_sur_targets", n, stream, args);
}

int kx_sur_outputs(kx_mech_t h, int n, int offsetT, int offset, const double* d_p,
                   const double* d_state, const float* d_base, const float* d_resid,
                   const float* d_ostat, const int* d_flags, double* d_visc, double* d_cond,
                   double* d_rhoD, double* d_Dkm, void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_p);
  if (rc) return rc;
  if (n > 0 && (!d_base || !d_resid || !d_ostat || !d_flags || !d_visc || !d_cond || !d_rhoD))
    return fail(KX_ERR_INVALID_ARG, "null device pointer");
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  void* args[] = {&n, &oT, &o, &d_p, &d_state, &d_base, &d_resid, &d_ostat, &d_flags,
                  &d_visc, &d_cond, &d_rhoD, &d_Dkm, &Tref};
  return launch(m, "kx_sur_outputs", n, stream, args);
}

int kx_sur_infer(kx_mech_t h, int n, int offsetT, int offset, const double* d_p,
                 const double* d_state, const float* d_norm, const float* d_ostat,
                 const uint16_t* d_wfrag, const float* d_params, int* d_flags, double* d_visc,
                 double* d_cond, double* d_rhoD, double* d_Dkm, int* d_fb_queue, int* d_fb_count,
                 void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_p);
  if (rc) return rc;
  if (n > 0 && (!d_norm || !d_ostat || !d_wfrag || !d_params || !d_flags || !d_visc || !d_cond ||
                !d_rhoD))
    return fail(KX_ERR_INVALID_ARG, "null device pointer");
  if ((reinterpret_cast<uintptr_t>(d_wfrag) | reinterpret_cast<uintptr_t>(d_params)) & 15)
    return fail(KX_ERR_INVALID_ARG, "weights and parameters must be 16-byte aligned");
  if ((d_fb_queue == nullptr) != (d_fb_count == nullptr))
    return fail(KX_ERR_INVALID_ARG, "fallback queue and its counter go together");
  long long oT = offsetT, o = offset;
  double Tref = m->ref_T;
  void* args[] = {&n, &oT, &o, &d_p, &d_state, &d_norm, &d_ostat, &d_wfrag, &d_params, &d_flags,
                  &d_visc, &d_cond, &d_rhoD, &d_Dkm, &Tref, &d_fb_queue, &d_fb_count};
  return launch(m, "kx_sur_infer", n, stream, args);
}

// ---- audited surrogate transport: one call, caller-provided workspace, no allocation
namespace {

struct SurPlan {
  long long n_audit, audit_max;
  size_t counters, flags, queue, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

SurPlan sur_plan(int n, int audit_stride, int audit_offset)
{
  SurPlan P{};
  const long long stride = audit_stride > 0 ? audit_stride : 0;
  const long long a0 = stride ? (long long)(audit_offset < 0 ? 0 : audit_offset) % stride : 0;
  P.audit_max = stride ? (n + stride - 1) / stride : 0;
  P.n_audit = stride && n > a0 ? (n - a0 + stride - 1) / stride : 0;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = align256(o + bytes); return at; };
  P.counters = take(4 * sizeof(int));
  P.flags = take((size_t)n * sizeof(int));
  P.queue = take((size_t)(P.audit_max + n) * sizeof(int));
  P.total = o;
  return P;
}

}  // namespace

size_t kx_sur_workspace_bytes(kx_mech_t h, int n, int audit_stride, int with_dkm)
{
  Mech* m = get(h);
  if (!m || n < 0) return 0;
  (void)with_dkm;   // the plain pass writes D_km in place: no staging either way
  return sur_plan(n, audit_stride, 0).total;
}

int kx_sur_transport(kx_mech_t h, int n, int offsetT, int offset, const double* d_p,
                     const double* d_state, const float* d_norm, const float* d_ostat,
                     const uint16_t* d_wfrag, const float* d_params, int audit_stride,
                     int audit_offset, double audit_tol, double audit_tol_rhoD,
                     double max_audit_fail, double* d_visc,
                     double* d_cond, double* d_rhoD, double* d_Dkm, void* d_work,
                     size_t work_bytes, kx_sur_stats* stats, void* stream)
{
  Mech* m = get(h);
  int rc = check_layout(m, n, offsetT, offset, d_state, d_p);
  if (rc) return rc;
  if (n > 0 && (!d_norm || !d_ostat || !d_wfrag || !d_params || !d_visc || !d_cond || !d_rhoD))
    return fail(KX_ERR_INVALID_ARG, "null device pointer");
  if (n > 0 && !d_work) return fail(KX_ERR_INVALID_ARG, "null workspace");
  const SurPlan P = sur_plan(n, audit_stride, audit_offset);
  if (P.audit_max + (long long)n > INT32_MAX)
    return fail(KX_ERR_TOO_LARGE, "too many states for one call");
  if (work_bytes < P.total)
    return fail(KX_ERR_TOO_LARGE, "workspace smaller than kx_sur_workspace_bytes()");
  kx_sur_stats st{};
  st.n_states = n;
  if (n == 0) { if (stats) *stats = st; return KX_OK; }
  char* w = static_cast<char*>(d_work);
  int* counters = reinterpret_cast<int*>(w + P.counters);
  int* flags = reinterpret_cast<int*>(w + P.flags);
  int* queue = reinterpret_cast<int*>(w + P.queue);
  const long long stride = audit_stride > 0 ? audit_stride : 0;
  const long long a0 = stride ? (long long)(audit_offset < 0 ? 0 : audit_offset) % stride : 0;
  int n_audit = (int)P.n_audit;

  // 1. audit ids and counters
  {
    void* args[] = {&n_audit, (void*)&a0, (void*)&stride, &queue, &counters};
    if ((rc = launch(m, "kx_sur_queue_init", std::max(n_audit, 4), stream, args))) return rc;
  }
  // 2. surrogate for every state; out-of-hull ids appended after the audit ids
  int* fbq = queue + n_audit;
  if ((rc = kx_sur_infer(h, n, offsetT, offset, d_p, d_state, d_norm, d_ostat, d_wfrag, d_params,
                         flags, d_visc, d_cond, d_rhoD, d_Dkm, fbq, counters, stream)))
    return rc;
  // 3. plain kernel over the audit samples + queued out-of-hull states, in place: it reads
  //    the queued states, compares mu, lambda and rho*D of the audit entries with the surrogate's,
  //    and overwrites the outputs with its own. The queue length is read on the device; the
  //    grid covers the longest possible queue (n_audit + n), and the workgroups past the
  //    actual end return at once, after the working ones (the first) were dispatched.
  {
    long long oT = offsetT, o = offset;
    double Tref = m->ref_T;
    void* args[] = {&n_audit, &queue, &counters, &flags, &audit_tol, &audit_tol_rhoD, &oT, &o,
                    (void*)&d_p,
                    (void*)&d_state, &d_cond, &d_visc, &d_rhoD, &d_Dkm, &Tref};
    const long long longest = (long long)n_audit + n;
    if (m->queue_lanes > 1) {
      // the cooperative form strides over the queue: a grid of the workgroups resident at
      // once, or fewer when even the longest possible queue needs fewer
      const long long per_wg = m->block.at("kx_bk2_fp64_queue") / m->queue_lanes;
      const long long grid = std::min((longest + per_wg - 1) / per_wg, m->queue_grid_cap);
      if ((rc = launch_grid(m, "kx_bk2_fp64_queue", grid, stream, args))) return rc;
    } else if ((rc = launch(m, "kx_bk2_fp64_queue", longest, stream, args))) {
      return rc;
    }
  }
  // 4. whole-call fallback, decided on the device: counters[3] = n when the failed audits
  //    exceed max_audit_fail * n_audit, else 0; the plain kernel for all states runs with
  //    that device-side count (every workgroup returns at once when it is 0)
  {
    void* args[] = {&n, &n_audit, &max_audit_fail, &counters};
    if ((rc = launch(m, "kx_sur_decide", 1, stream, args))) return rc;
  }
  if ((rc = transport_impl(h, n, offsetT, offset, 1.0, d_p, d_state, d_visc, d_cond, d_rhoD,
                           d_Dkm, stream, counters + 3)))
    return rc;
  // 5. statistics: the only host read, and the only wait, of the call
  if (stats) {
    int host_counters[4] = {0, 0, 0, 0};
    KX_HIP(hipMemcpyAsync(host_counters, counters, sizeof(host_counters), hipMemcpyDeviceToHost,
                          (hipStream_t)stream));
    KX_HIP(hipStreamSynchronize((hipStream_t)stream));
    st.out_of_hull = host_counters[0];
    st.audited = n_audit;
    st.audit_failed = host_counters[1];
    st.whole_call_fallback = host_counters[3] != 0;
    st.fallback = st.whole_call_fallback ? n : st.out_of_hull + st.audit_failed;
    *stats = st;
  }
  return KX_OK;
}

int kx_sur_transport_noalloc(kx_mech_t h, int n, int offsetT, int offset, const double* d_p,
                             const double* d_state, const float* d_norm, const float* d_ostat,
                             const uint16_t* d_wfrag, const float* d_params, int audit_stride,
                             int audit_offset, double audit_tol, double audit_tol_rhoD,
                             double max_audit_fail, double* d_visc, double* d_cond,
                             double* d_rhoD, double* d_Dkm, kx_sur_stats* stats, void* stream)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (n < 0) return fail(KX_ERR_INVALID_ARG, "n_states < 0");
  const size_t need = sur_plan(n, audit_stride, audit_offset).total;
  if (need > kScratchBytes)
    return fail(KX_ERR_TOO_LARGE, "kx_sur_transport_noalloc: " + std::to_string(n) +
                " states at audit stride " + std::to_string(audit_stride) + " need " +
                std::to_string(need) + " bytes of scratch, more than KX_SUR_SCRATCH_BYTES (" +
                std::to_string(kScratchBytes) + "); use kx_sur_transport with a workspace");
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ptr = (m->device >= 0 && m->device < kMaxDevices) ? g_scratch_ptr[m->device] : nullptr;
  }
  if (!ptr) return fail(KX_ERR_NOT_INITIALIZED, "no scratch address for this device");
  return kx_sur_transport(h, n, offsetT, offset, d_p, d_state, d_norm, d_ostat, d_wfrag, d_params,
                          audit_stride, audit_offset, audit_tol, audit_tol_rhoD, max_audit_fail,
                          d_visc, d_cond, d_rhoD, d_Dkm, ptr, kScratchBytes, stats, stream);
}

int kx_n_species(kx_mech_t h) { Mech* m = get(h); return m ? m->n_species : KX_ERR_NOT_INITIALIZED; }
int kx_n_active_species(kx_mech_t h) { Mech* m = get(h); return m ? m->n_active : KX_ERR_NOT_INITIALIZED; }
int kx_n_reactions(kx_mech_t h) { Mech* m = get(h); return m ? m->n_reactions : KX_ERR_NOT_INITIALIZED; }

int kx_species_name(kx_mech_t h, int k, char* buf, int buflen)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (k < 0 || k >= m->n_species || !buf || buflen <= 0) return fail(KX_ERR_INVALID_ARG, "bad index");
  const std::string& s = m->names[k];
  if ((int)s.size() >= buflen) return fail(KX_ERR_INVALID_ARG, "buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

int kx_species_index(kx_mech_t h, const char* name)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (!name) return -1;
  for (int k = 0; k < m->n_species; ++k)
    if (m->names[k] == name) return k;
  return -1;
}

int kx_molecular_weights(kx_mech_t h, double* out)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (!m->built) return fail(KX_ERR_NOT_BUILT, "kx_build has not been called");
  if (!out) return fail(KX_ERR_INVALID_ARG, "null output");
  for (int k = 0; k < m->n_species; ++k) out[k] = m->W[k] / m->ref_Mbar;
  return KX_OK;
}

int kx_molar_masses(kx_mech_t h, double* out)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (!out) return fail(KX_ERR_INVALID_ARG, "null output");
  for (int k = 0; k < m->n_species; ++k) out[k] = m->W[k];
  return KX_OK;
}

double kx_ref_pressure(kx_mech_t h) { Mech* m = get(h); return m ? m->ref_p : NAN; }
double kx_ref_temperature(kx_mech_t h) { Mech* m = get(h); return m ? m->ref_T : NAN; }
double kx_ref_mean_molecular_weight(kx_mech_t h) { Mech* m = get(h); return m ? m->ref_Mbar : NAN; }

int kx_ref_mass_fractions(kx_mech_t h, double* out)
{
  Mech* m = get(h);
  if (!m) return fail(KX_ERR_NOT_INITIALIZED, "unknown mechanism handle");
  if (!m->built) return fail(KX_ERR_NOT_BUILT, "kx_build has not been called");
  if (!out) return fail(KX_ERR_INVALID_ARG, "null output");
  std::copy(m->ref_Y.begin(), m->ref_Y.end(), out);
  return KX_OK;
}

}  // extern "C"

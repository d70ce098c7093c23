// REGIME model extension of the host API (REGIME.cc): the source and the
// per-cell fields come from the engine (mgpu_source_extension,
// mgpu_regime_fields) after uploading the given host state.
#include "REGIME.h"

#include <cstring>
#include <stdexcept>

#include "gpu_link.h"

using method_host::check;
using method_host::engine;

namespace {
double *zeros(size_t n) { return n ? new double[n]() : nullptr; }
}  // namespace

REGIME::REGIME()
    : dfxdw(nullptr), dfydw(nullptr), dfzdw(nullptr), dwdsb(nullptr), E(nullptr), q(nullptr), K(nullptr),
      Mx(nullptr), My(nullptr), Mz(nullptr), fbx(nullptr), fby(nullptr), fbz(nullptr), diffuX(nullptr),
      diffuY(nullptr), diffuZ(nullptr), alpha(nullptr) {
  sourceExists = true;
}

REGIME::REGIME(Data *data, FluxMethod *fluxMethod) : ModelExtension(data), fluxMethod(fluxMethod) {
  sourceExists = true;
  const size_t n = (size_t)data->Nx * data->Ny * data->Nz;
  dfxdw = zeros(n * 8 * 12);
  dfydw = zeros(n * 8 * 12);
  dfzdw = zeros(n * 8 * 12);
  dwdsb = zeros(n * 12 * 3);
  E = zeros(n * 3);
  q = zeros(n);
  K = zeros(n * 3);
  Mx = zeros(n * 8 * 3);
  My = zeros(n * 8 * 3);
  Mz = zeros(n * 8 * 3);
  fbx = zeros(n * 3);
  fby = zeros(n * 3);
  fbz = zeros(n * 3);
  diffuX = zeros(n * 8);
  diffuY = zeros(n * 8);
  diffuZ = zeros(n * 8);
  alpha = zeros(n);
  delete[] data->sourceExtension;
  data->sourceExtension = zeros(n * 9);  // REGIME.cc:59
}

REGIME::~REGIME() {
  double *all[] = {dfxdw, dfydw, dfzdw, dwdsb, E, q, K, Mx, My, Mz, fbx, fby, fbz, diffuX, diffuY, diffuZ, alpha};
  for (double *p : all) delete[] p;
  if (data) {
    delete[] data->sourceExtension;
    data->sourceExtension = nullptr;
  }
}

void REGIME::sourceExtension(double *cons, double *prims, double *aux, double *source) {
  if (data->gpu.model != MGPU_SRMHD) throw std::invalid_argument("REGIME extends the ideal SRMHD model only");
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, prims, aux), "REGIME::sourceExtension");
  check(mgpu_set_model_extension(e, MGPU_EXT_REGIME), "REGIME::sourceExtension");
  check(mgpu_source_extension(e, source), "REGIME::sourceExtension");
  data->gpu.extensionApplied = MGPU_EXT_REGIME;
  data->gpu.deviceCurrent = false;
}

void REGIME::fields(double *prims, double *Eo, double *qo, double *ao, double *dwo, double *Ko) {
  if (data->gpu.model != MGPU_SRMHD) throw std::invalid_argument("REGIME extends the ideal SRMHD model only");
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, nullptr, prims, nullptr), "REGIME fields");
  check(mgpu_set_model_extension(e, MGPU_EXT_REGIME), "REGIME fields");
  check(mgpu_regime_fields(e, Eo, qo, ao, dwo, Ko), "REGIME fields");
  data->gpu.extensionApplied = MGPU_EXT_REGIME;
  data->gpu.deviceCurrent = false;
}

// REGIME::set_vars (REGIME.cc:141-224): E = -v x B, q, alpha.
void REGIME::set_vars(double *, double *prims, double *) { fields(prims, E, q, alpha, nullptr, nullptr); }

// REGIME::set_dwdsb (REGIME.cc:226-290).
void REGIME::set_dwdsb(double *, double *prims, double *) { fields(prims, nullptr, nullptr, nullptr, dwdsb, nullptr); }

// REGIME::set_K (REGIME.cc:293-370).
void REGIME::set_K(double *, double *prims, double *) { fields(prims, nullptr, nullptr, nullptr, nullptr, K); }

namespace {
[[noreturn]] void in_kernels(const char *what) {
  throw std::logic_error(std::string(what) + ": the per-axis REGIME matrices are formed inside the engine's kernels");
}
}  // namespace

void REGIME::set_Dx(double *, double *, double *) { in_kernels("REGIME::set_Dx"); }
void REGIME::set_Dy(double *, double *, double *) { in_kernels("REGIME::set_Dy"); }
void REGIME::set_Dz(double *, double *, double *) { in_kernels("REGIME::set_Dz"); }
void REGIME::set_dfxdw(double *, double *, double *) { in_kernels("REGIME::set_dfxdw"); }
void REGIME::set_dfydw(double *, double *, double *) { in_kernels("REGIME::set_dfydw"); }
void REGIME::set_dfzdw(double *, double *, double *) { in_kernels("REGIME::set_dfzdw"); }

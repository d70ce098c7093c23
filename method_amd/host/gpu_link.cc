#include "gpu_link.h"

#include <cstring>
#include <stdexcept>
#include <string>

namespace method_host {

namespace {

mgpu_config config_of(const Data *d) {
  mgpu_config c;
  std::memset(&c, 0, sizeof c);
  c.nx = d->nx;
  c.ny = d->ny;
  c.nz = d->nz;
  c.Ng = d->Ng;
  c.xmin = d->xmin;
  c.xmax = d->xmax;
  c.ymin = d->ymin;
  c.ymax = d->ymax;
  c.zmin = d->nz > 0 ? d->zmin : 0.0;
  c.zmax = d->nz > 0 ? d->zmax : 1.0;
  c.gamma = d->gamma;
  c.sigma = d->sigma;
  c.cp = d->cp;
  c.cfl = d->cfl;
  c.model = d->gpu.model;
  c.weno_order = d->gpu.wenoOrder ? d->gpu.wenoOrder : 3;
  c.integrator = d->gpu.integrator;
  for (int a = 0; a < 3; ++a) c.bc[a] = d->gpu.bc[a];
  c.quirks = d->gpu.quirks;
  c.device = d->gpu.device;
  return c;
}

}  // namespace

// the model extension an integrator asked for, given to the engine once
mgpu_engine *with_extension(Data *d, mgpu_engine *e) {
  if (d->gpu.extension != d->gpu.extensionApplied) {
    check(mgpu_set_model_extension(e, d->gpu.extension), "mgpu_set_model_extension");
    d->gpu.extensionApplied = d->gpu.extension;
  }
  return e;
}

int check(int rc, const char *what) {
  if (rc < 0) throw std::runtime_error(std::string(what) + ": " + mgpu_last_error());
  return rc;
}

mgpu_engine *engine(Data *d) {
  if (d->gpu.model < 0) throw std::runtime_error("Must set model before using the GPU engine");
  mgpu_config c = config_of(d);
  if (d->gpu.engine) {
    // engine configuration is fixed at creation; rebuild if the objects around Data changed it
    if (std::memcmp(&c, &d->gpu.created, sizeof c) == 0) return with_extension(d, d->gpu.engine);
    // the state only the old engine holds goes back to the host arrays first
    if (d->gpu.hostStale) check(mgpu_download_state(d->gpu.engine, d->cons, d->prims, d->aux), "engine rebuild");
    release(d);
  }
  mgpu_engine *e = nullptr;
  check(mgpu_create(&c, &e), "mgpu_create");
  d->gpu.engine = e;
  d->gpu.created = c;
  d->gpu.deviceCurrent = false;
  d->gpu.hostStale = false;
  d->gpu.extensionApplied = MGPU_EXT_NONE;  // a new engine has none
  return with_extension(d, e);
}

void release(Data *d) {
  if (d->gpu.engine) mgpu_destroy(d->gpu.engine);
  d->gpu.engine = nullptr;
  d->gpu.deviceCurrent = false;
  d->gpu.hostStale = false;
}

}  // namespace method_host

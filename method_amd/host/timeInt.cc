// Time integrators of the host API (timeInt.h:117-132, RK2.cc:31-100).
#include <cstdio>
#include <stdexcept>

#include "REGIME.h"
#include "RK2.h"
#include "RKPlus.h"
#include "rkSplit.h"
#include "SSP2.h"
#include "fluxVectorSplitting.h"
#include "gpu_link.h"

using method_host::check;
using method_host::engine;

void TimeIntegrator::finalise(double *cons, double *prims, double *aux) {
  try {
    model->getPrimitiveVars(cons, prims, aux);
  } catch (const std::exception &e) {
    printf("Time integration raises exception with following message:\n%s\n", e.what());
    throw;
  }
  model->finalise(cons, prims, aux);
  bcs->apply(cons, prims, aux);
}

RK2::RK2(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : TimeIntegrator(data, model, bcs, fluxMethod, ext) {
  engineIntegrator = data->gpu.integrator = MGPU_RK2;
}

void TimeIntegratorBase::bind() {
  if (engineIntegrator >= 0) data->gpu.integrator = engineIntegrator;
  data->gpu.extension = engineExtension;
  if (bcs) bcs->bind();
  if (FVS *f = dynamic_cast<FVS *>(fluxMethod))
    if (f->weno) data->gpu.wenoOrder = f->weno->order;
}

namespace {

// One engine step of whichever integrator the engine was created with, on the
// device-resident state. The host arrays are uploaded first unless the device
// already holds them, and copied back afterwards unless the caller
// (Simulation::evolve) defers that.
void engine_step(TimeIntegrator *ti, double *cons, double *prims, double *aux, double dt, const char *what) {
  Data *d = ti->data;
  if (dt <= 0) dt = d->dt;
  ti->bind();
  mgpu_engine *e = engine(d);
  const bool own = cons == d->cons && prims == d->prims && aux == d->aux;
  if (!(own && d->gpu.deviceCurrent)) check(mgpu_upload_state(e, cons, prims, aux), what);
  mgpu_stats st;
  const int rc = check(mgpu_step(e, dt, &st), what);
  if (rc > 0) printf("Smart guessing did not work for %d cells\n", rc);
  if (own && !d->gpu.mirrorEachStep) {
    d->gpu.deviceCurrent = true;
    d->gpu.hostStale = true;
    return;
  }
  check(mgpu_download_state(e, cons, prims, aux), what);
  d->gpu.deviceCurrent = own;
  d->gpu.hostStale = false;
}

}  // namespace

void RK2::step(double *cons, double *prims, double *aux, double dt) { engine_step(this, cons, prims, aux, dt, "RK2::step"); }

SSP2::SSP2(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod) : TimeIntegrator(data, model, bcs, fluxMethod) {
  engineIntegrator = data->gpu.integrator = MGPU_SSP2;
}

void SSP2::step(double *cons, double *prims, double *aux, double dt) {
  engine_step(this, cons, prims, aux, dt, "SSP2::step");
}

namespace {
// The engine's extension for a model extension handed to RKSplit / RKPlus,
// whose steps add its source (rkSplit.cc:15-25, RKPlus.cc:26-38).
int extension_kind(ModelExtension *ext) {
  if (!ext || !ext->sourceExists) return MGPU_EXT_NONE;
  if (dynamic_cast<REGIME *>(ext)) return MGPU_EXT_REGIME;
  throw std::invalid_argument("model extension: the GPU engine provides REGIME only");
}
}  // namespace

RKSplit::RKSplit(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : RK2(data, model, bcs, fluxMethod, ext) {
  engineIntegrator = data->gpu.integrator = MGPU_RKSPLIT;
  engineExtension = data->gpu.extension = extension_kind(ext);
}

// RKSplit::setSource (rkSplit.cc:10-26): the model's source plus the
// extension's over the interior cells.
void RKSplit::setSource(double *cons, double *prims, double *aux) {
  Data *d = data;
  model->sourceTerm(cons, prims, aux, d->source);
  if (modelExtension && modelExtension->sourceExists) {
    modelExtension->sourceExtension(cons, prims, aux, d->sourceExtension);
    for (int var = 0; var < d->Ncons; ++var)
      for (int i = d->is; i < d->ie; ++i)
        for (int j = d->js; j < d->je; ++j)
          for (int k = d->ks; k < d->ke; ++k) d->source[ID(var, i, j, k)] += d->sourceExtension[ID(var, i, j, k)];
  }
}

RKPlus::RKPlus(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : TimeIntegrator(data, model, bcs, fluxMethod, ext) {
  engineExtension = data->gpu.extension = extension_kind(ext);
  fluxCont = new double[(size_t)data->Nx * data->Ny * data->Nz * data->Ncons]();
}

RKPlus::~RKPlus() { delete[] fluxCont; }

void RKPlus::rhs(double *cons, double *prims, double *aux, double *rhsVec) {
  Data *d = data;
  fluxMethod->F(cons, prims, aux, d->f, fluxCont);
  model->sourceTerm(cons, prims, aux, d->source);
  if (modelExtension && modelExtension->sourceExists) {
    modelExtension->sourceExtension(cons, prims, aux, d->sourceExtension);
    for (int var = 0; var < d->Ncons; ++var)
      for (int i = d->is; i < d->ie; ++i)
        for (int j = d->js; j < d->je; ++j)
          for (int k = d->ks; k < d->ke; ++k) d->source[ID(var, i, j, k)] += d->sourceExtension[ID(var, i, j, k)];
  }
  for (int var = 0; var < d->Ncons; ++var)
    for (int i = d->is; i < d->ie; ++i)
      for (int j = d->js; j < d->je; ++j)
        for (int k = d->ks; k < d->ke; ++k) rhsVec[ID(var, i, j, k)] = d->source[ID(var, i, j, k)] - fluxCont[ID(var, i, j, k)];
}

void RKPlus::step(double *cons, double *prims, double *aux, double dt) {
  engine_step(this, cons, prims, aux, dt, "RKPlus::step");
}

RK2B::RK2B(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : RKPlus(data, model, bcs, fluxMethod, ext) {
  engineIntegrator = data->gpu.integrator = MGPU_RK2B;
}

RK3::RK3(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : RKPlus(data, model, bcs, fluxMethod, ext) {
  engineIntegrator = data->gpu.integrator = MGPU_RK3;
}

RK4::RK4(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : RKPlus(data, model, bcs, fluxMethod, ext) {
  engineIntegrator = data->gpu.integrator = MGPU_RK4;
}

RK4_10::RK4_10(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext)
    : RKPlus(data, model, bcs, fluxMethod, ext) {
  engineIntegrator = data->gpu.integrator = MGPU_RK4_10;
}

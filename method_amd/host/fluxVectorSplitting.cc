// FVS and WENO objects of the host API (fluxVectorSplitting.cc:58-131, weno.cc:7-20).
#include "fluxVectorSplitting.h"

#include <cstdio>
#include <stdexcept>

#include "gpu_link.h"

WenoBase::WenoBase(Data *data, int order) : data(data), order(order) {
  shift = (order + 1) / 2;
  checkSufficientGhostZones();
  data->gpu.wenoOrder = order;
}

void WenoBase::checkSufficientGhostZones() {
  if (data->Ng < shift + 1) {
    printf("This order Weno reconstruction requires at least %d ghost zones, you have %d...", shift + 1, data->Ng);
    throw std::invalid_argument("You must increase number of boundary cells.");
  }
}

// The face fluxes of the given arrays (fluxVectorSplitting.cc:4-56), formed on
// the device from copies of cons and f; the resident state is not touched.
void FVS::fluxReconstruction(double *cons, double *prims, double *aux, double *f, double *frecon, int dir,
                             int vars) {
  (void)prims;
  (void)aux;
  if (weno) data->gpu.wenoOrder = weno->order;
  mgpu_engine *e = method_host::engine(data);
  method_host::check(mgpu_flux_reconstruction(e, cons, f, dir, vars, frecon), "FVS::fluxReconstruction");
}

void FVS::F(double *cons, double *prims, double *aux, double *f, double *fnet) {
  (void)aux;
  (void)f;
  if (weno) data->gpu.wenoOrder = weno->order;
  mgpu_engine *e = method_host::engine(data);
  method_host::check(mgpu_upload_state(e, cons, prims, nullptr), "FVS::F");
  method_host::check(mgpu_fnet(e, fnet), "FVS::F");
  data->gpu.deviceCurrent = false;
}

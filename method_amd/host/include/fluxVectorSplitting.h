// Flux vector splitting with WENO reconstruction (Project/CPU/Include/
// fluxVectorSplitting.h, fluxVectorSplitting.cc:4-131), run by the GPU engine.
#ifndef FLUXVECTORSPLITTING_H
#define FLUXVECTORSPLITTING_H

#include "flux.h"
#include "weno.h"

class FVS : public FluxMethod {
 public:
  WenoBase *weno;
  // The engine reconstructs with this object's order, whatever other WENO
  // objects the program holds.
  FVS(Data *data, WenoBase *weno, Model *model) : FluxMethod(data, model), weno(weno) {
    if (weno) data->gpu.wenoOrder = weno->order;
  }
  virtual ~FVS() {}
  // The face fluxes at i - 1/2 along dir of the given host arrays
  // (fluxVectorSplitting.cc:4-56), formed on the device by
  // mgpu_flux_reconstruction; the resident state is not touched.
  void fluxReconstruction(double *cons, double *prims, double *aux, double *f, double *frecon, int dir,
                          int vars = -1);
  void F(double *cons, double *prims, double *aux, double *f, double *fnet);
};

#endif

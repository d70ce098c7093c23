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
  // Reconstructed face fluxes are never formed as a separate array by the
  // engine; this throws std::logic_error.
  void fluxReconstruction(double *cons, double *prims, double *aux, double *f, double *frecon, int dir,
                          int vars = -1);
  void F(double *cons, double *prims, double *aux, double *f, double *fnet);
};

#endif

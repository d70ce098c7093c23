// Operator-split RK2 (Project/CPU/Include/rkSplit.h, Src/rkSplit.cc:31-53):
// RK2::step, then cons += dt source and finalise. The whole step runs in the
// engine (MGPU_RKSPLIT): both RK2 stages, the source half-step and the second
// cons2prim on the device-resident state.
#ifndef RKSPLIT_H
#define RKSPLIT_H

#include "RK2.h"

class RKSplit : public RK2 {
 public:
  RKSplit(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
  virtual ~RKSplit() {}
  // Model::sourceTerm of the given state into data->source (rkSplit.cc:6-28).
  void setSource(double *cons, double *prims, double *aux);
};

#endif

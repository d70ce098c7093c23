// Second-order IMEX SSP scheme (Project/CPU/Include/SSP2.h, SSP2.cc:53-173) as
// one engine step: both implicit stages, the flux sweeps between them and the
// final combination run on the GPU with the state resident in HBM. Used with
// the resistive model.
#ifndef SSP2_H
#define SSP2_H

#include "timeInt.h"

class SSP2 : public TimeIntegrator {
 public:
  SSP2(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod);
  virtual ~SSP2() {}
  void step(double *cons, double *prims, double *aux, double dt = 0);
};

#endif

// Shu-Osher RK2 (Project/CPU/Include/RK2.h, RK2.cc:31-100) as one engine step:
// both stages, their flux sweeps, updates and cons2prim run on the GPU with
// the state resident in HBM.
#ifndef RK2_H
#define RK2_H

#include "timeInt.h"

class RK2 : public TimeIntegrator {
 public:
  RK2(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
  virtual ~RK2() {}
  void step(double *cons, double *prims, double *aux, double dt = 0);
};

#endif

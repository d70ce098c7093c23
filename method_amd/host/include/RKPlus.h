// Explicit Runge-Kutta family with rhs = source - F (Project/CPU/Include/RKPlus.h,
// Src/RKPlus.cc): RK2B, RK3, RK4 (five-stage) and RK4_10 (ten-stage). A step
// runs in the engine (MGPU_RK2B .. MGPU_RK4_10): every stage's flux sweeps,
// source, combination and cons2prim on the device-resident state. rhs() is
// the reference's public RKPlus::rhs on host arrays.
#ifndef RKPLUS_H
#define RKPLUS_H

#include "timeInt.h"

class RKPlus : public TimeIntegrator {
 public:
  double *fluxCont;  // net flux of the last rhs() call (RKPlus.h:39)
  RKPlus(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
  virtual ~RKPlus();
  // rhsVec = source - F over the interior (RKPlus.cc:15-50)
  virtual void rhs(double *cons, double *prims, double *aux, double *rhsVec);
  void step(double *cons, double *prims, double *aux, double dt = 0);
};

class RK2B : public RKPlus {
 public:
  RK2B(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
};

class RK3 : public RKPlus {
 public:
  RK3(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
};

class RK4 : public RKPlus {
 public:
  RK4(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
};

class RK4_10 : public RKPlus {
 public:
  RK4_10(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr);
};

#endif

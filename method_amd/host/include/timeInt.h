// Time integrator interface of the reference (Project/CPU/Include/timeInt.h:18-134).
#ifndef TIMEINT_H
#define TIMEINT_H

#include "boundaryConds.h"
#include "flux.h"
#include "model.h"
#include "simData.h"

class ModelExtension;

class TimeIntegratorBase {
 public:
  Data *data;
  Model *model;
  Bcs *bcs;
  FluxMethod *fluxMethod;
  ModelExtension *modelExtension;
  TimeIntegratorBase(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr)
      : data(data), model(model), bcs(bcs), fluxMethod(fluxMethod), modelExtension(ext) {}
  virtual ~TimeIntegratorBase() {}
  virtual void step(double *cons, double *prims, double *aux, double dt = 0) = 0;
  virtual void finalise(double *cons, double *prims, double *aux) = 0;
  // Makes this integrator, its boundary conditions, its flux method's
  // reconstruction order and its model extension the engine's configuration
  // (each step and Simulation::set call it).
  void bind();

 protected:
  int engineIntegrator = -1;            // MGPU_RK2, MGPU_SSP2, ...
  int engineExtension = MGPU_EXT_NONE;  // MGPU_EXT_*
};

class TimeIntegrator : public TimeIntegratorBase {
 public:
  TimeIntegrator(Data *data, Model *model, Bcs *bcs, FluxMethod *fluxMethod, ModelExtension *ext = nullptr)
      : TimeIntegratorBase(data, model, bcs, fluxMethod, ext) {}
  virtual ~TimeIntegrator() {}
  // cons2prim, model finalise, boundaries (timeInt.h:117-132), on the GPU.
  void finalise(double *cons, double *prims, double *aux);
};

#endif

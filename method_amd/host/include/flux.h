// Numerical flux interface of the reference (Project/CPU/Include/flux.h:17-75).
#ifndef FLUX_H
#define FLUX_H

#include "model.h"
#include "simData.h"

class FluxMethod {
 public:
  Data *data;
  Model *model;
  FluxMethod(Data *data, Model *model) : data(data), model(model) {}
  virtual ~FluxMethod() {}
  virtual void F(double *cons, double *prims, double *aux, double *f, double *fnet) = 0;
  virtual void fluxReconstruction(double *cons, double *prims, double *aux, double *f, double *frecon, int dir,
                                  int vars) = 0;
};

#endif

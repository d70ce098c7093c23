// Model extensions (Project/CPU/Include/modelExtension.h:14-59): a source
// and/or flux added to the model's. The engine provides REGIME (REGIME.h).
#ifndef MODELEXTENSION_H
#define MODELEXTENSION_H

#include "simData.h"

class ModelExtension {
 public:
  Data *data = nullptr;
  bool sourceExists = false;
  bool fluxExists = false;
  ModelExtension() {}
  ModelExtension(Data *data) : data(data) {}
  virtual ~ModelExtension() {}
  virtual void sourceExtension(double *cons, double *prims, double *aux, double *source) {}
  virtual void fluxExtension(double *cons, double *prims, double *aux, double *flux) {}
};

#endif

// Physics model interface of the reference (Project/CPU/Include/model.h:15-135).
#ifndef MODEL_H
#define MODEL_H

#include "simData.h"

class Model {
 public:
  Data *data;
  int Ncons, Nprims, Naux;
  Model() : data(nullptr) {}
  Model(Data *data) : data(data) {}
  virtual ~Model() {}
  virtual void sourceTermSingleCell(double *cons, double *prims, double *aux, double *source, int i = -1, int j = -1,
                                    int k = -1) = 0;
  virtual void sourceTerm(double *cons, double *prims, double *aux, double *source) = 0;
  virtual void getPrimitiveVarsSingleCell(double *cons, double *prims, double *aux, int i = -1, int j = -1,
                                          int k = -1) = 0;
  virtual void getPrimitiveVars(double *cons, double *prims, double *aux) = 0;
  virtual void primsToAll(double *cons, double *prims, double *aux) = 0;
  virtual void fluxVector(double *cons, double *prims, double *aux, double *f, int dir) = 0;
  virtual void finalise(double *cons, double *prims, double *aux) {}
};

#endif

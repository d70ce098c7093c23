// Ideal SRMHD model (Project/CPU/Include/srmhd.h, srmhd.cc): 9 cons, 8 prims,
// 13 aux. Every full-grid method runs on the GPU engine.
#ifndef SRMHD_H
#define SRMHD_H

#include "model.h"

class SRMHD : public Model {
 public:
  int smartGuesses;
  SRMHD();
  SRMHD(Data *data);
  virtual ~SRMHD();
  void fluxVector(double *cons, double *prims, double *aux, double *f, const int dir);
  void sourceTermSingleCell(double *cons, double *prims, double *aux, double *source, int i = -1, int j = -1,
                            int k = -1);
  void sourceTerm(double *cons, double *prims, double *aux, double *source);
  void getPrimitiveVarsSingleCell(double *cons, double *prims, double *aux, int i = -1, int j = -1, int k = -1);
  void getPrimitiveVars(double *cons, double *prims, double *aux);
  void primsToAll(double *cons, double *prims, double *aux);
  void finalise(double *cons, double *prims, double *aux) {}
};

#endif

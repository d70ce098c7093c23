// Resistive SRMHD model (Project/CPU/Include/srrmhd.h, srrmhd.cc): 14 cons,
// 11 prims, 17 aux. Every full-grid method runs on the GPU engine.
#ifndef SRRMHD_H
#define SRRMHD_H

#include "model.h"

class SRRMHD : public Model {
 public:
  SRRMHD();
  SRRMHD(Data *data);
  virtual ~SRRMHD() {}
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

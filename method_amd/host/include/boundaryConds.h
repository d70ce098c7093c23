// Boundary conditions of the reference (Project/CPU/Include/boundaryConds.h):
// each class records its per-axis rule for the engine, which applies it by
// index remapping inside its kernels; apply() on host arrays fills their ghost
// cells through the engine.
#ifndef BOUNDARYCONDS_H
#define BOUNDARYCONDS_H

#include "simData.h"

class Bcs {
 public:
  Data *data;
  Bcs(Data *data) : data(data) { data->bcsSet = 1; }
  virtual ~Bcs() {}
  virtual void apply(double *cons, double *prims = nullptr, double *aux = nullptr);
  // Makes this object's rule the one the engine applies (the integrator and
  // Simulation::set call it, so the rule in force is the handed object's).
  void bind();

 protected:
  void setAxes(int px, int py, int pz);
  int axisRule[3] = {MGPU_BC_OUTFLOW, MGPU_BC_OUTFLOW, MGPU_BC_OUTFLOW};
};

// Outflow on every face (boundaryConds.cc:5-145).
class Outflow : public Bcs {
 public:
  Outflow(Data *data) : Bcs(data) { setAxes(0, 0, 0); }
};

// Periodic on every face (boundaryConds.cc:590-730).
class Periodic : public Bcs {
 public:
  Periodic(Data *data) : Bcs(data) { setAxes(1, 1, 1); }
};

// Periodic in x, outflow in y and z (boundaryConds.cc:449-588).
class Flow : public Bcs {
 public:
  Flow(Data *data) : Bcs(data) { setAxes(1, 0, 0); }
};

// Any per-axis combination, e.g. periodic x and z with outflow y, which no
// reference class provides.
class PerAxisBcs : public Bcs {
 public:
  PerAxisBcs(Data *data, int periodicX, int periodicY, int periodicZ) : Bcs(data) {
    setAxes(periodicX, periodicY, periodicZ);
  }
};

#endif

// WENO reconstruction objects (Project/CPU/Include/weno.h). The engine fuses
// reconstruction into its flux sweeps, so these objects carry the order and
// the ghost-zone check (weno.cc:7-20); the per-cell helpers are not provided.
#ifndef WENO_H
#define WENO_H

#include <stdexcept>

#include "simData.h"

class WenoBase {
 public:
  Data *data;
  int order, shift;
  WenoBase(Data *data, int order);
  virtual ~WenoBase() {}
  virtual void checkSufficientGhostZones();
};

class Weno3 : public WenoBase {
 public:
  Weno3(Data *data) : WenoBase(data, 3) {}
};

class Weno5 : public WenoBase {
 public:
  Weno5(Data *data) : WenoBase(data, 5) {}
};

#endif

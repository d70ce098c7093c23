// Simulation driver of the reference (Project/CPU/Include/simulation.h:23-118).
// evolve() keeps the state on the GPU between steps and refreshes the host
// arrays when it returns; updateTime() called on its own refreshes them after
// every step, as the reference leaves them current.
#ifndef SIMULATION_H
#define SIMULATION_H

#include "boundaryConds.h"
#include "flux.h"
#include "initFunc.h"
#include "model.h"
#include "platformEnv.h"
#include "saveData.h"
#include "simData.h"
#include "timeInt.h"

class Simulation {
 public:
  Data *data;
  InitialFunc *init;
  Model *model;
  TimeIntegrator *timeInt;
  Bcs *bcs;
  FluxMethod *fluxMethod;
  SaveData *save;
  PlatformEnv *env;
  Simulation(Data *data, PlatformEnv *env);
  virtual ~Simulation();
  void set(InitialFunc *init, Model *model, TimeIntegrator *timeInt, Bcs *bcs, FluxMethod *fluxMethod,
           SaveData *save = nullptr);
  void updateTime();
  void evolve(bool output = false, int safety = -1);
  // Copies the device state into d->cons/prims/aux if the device is ahead.
  void syncHost();
};

#endif

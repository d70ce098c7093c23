// Simulation driver of the host API (simulation.cc:8-171).
#include "simulation.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "gpu_link.h"

Simulation::Simulation(Data *data, PlatformEnv *env)
    : data(data), init(nullptr), model(nullptr), timeInt(nullptr), bcs(nullptr), fluxMethod(nullptr), save(nullptr),
      env(env) {
  Data *d = data;
  const int Ntot = d->Nx * d->Ny * d->Nz;
  if (env->rank == 0) {
    if (d->Ncons == 0) throw std::runtime_error("Must set model before constructing simulation");
    if (d->bcsSet != 1)
      throw std::runtime_error(
          "Must construct boundary condition class before implementing simulation. Need to set domain decomposition "
          "parameters including periodicity.");
  }
  d->cons = (double *)malloc(sizeof(double) * Ntot * d->Ncons);
  d->f = (double *)malloc(sizeof(double) * Ntot * d->Ncons);
  d->fnet = (double *)malloc(sizeof(double) * Ntot * d->Ncons);
  d->source = (double *)malloc(sizeof(double) * Ntot * d->Ncons);
  d->prims = (double *)malloc(sizeof(double) * Ntot * d->Nprims);
  d->aux = (double *)malloc(sizeof(double) * Ntot * d->Naux);
  d->x = (double *)malloc(sizeof(double) * d->Nx);
  d->y = (double *)malloc(sizeof(double) * d->Ny);
  d->z = (double *)malloc(sizeof(double) * d->Nz);
  d->dx = (d->xmax - d->xmin) / (d->nx);
  d->dy = (d->ymax - d->ymin) / (d->ny);
  d->dz = (d->zmax - d->zmin) / (d->nz);
  d->iters = 0;
  d->alphaX = d->alphaY = d->alphaZ = 1.0;
  const double dtX = d->cfl * d->dx / (d->alphaX * std::sqrt(d->dims));
  const double dtY = d->cfl * d->dy / (d->alphaY * std::sqrt(d->dims));
  const double dtZ = d->cfl * d->dz / (d->alphaZ * std::sqrt(d->dims));
  d->dt = (dtX < dtY && dtX < dtZ) ? dtX : ((dtY < dtZ) ? dtY : dtZ);
  d->memSet = 1;
  for (int i = 0; i < d->Nx; i++) d->x[i] = d->xmin + (i + 0.5 - d->Ng) * d->dx;
  for (int j = 0; j < d->Ny; j++) d->y[j] = d->ymin + (j + 0.5 - d->Ng) * d->dy;
  for (int k = 0; k < d->Nz; k++) d->z[k] = d->zmin + (k + 0.5 - d->Ng) * d->dz;
}

Simulation::~Simulation() {
  free(data->cons);
  free(data->f);
  free(data->fnet);
  free(data->source);
  free(data->prims);
  free(data->aux);
  free(data->x);
  free(data->y);
  free(data->z);
  method_host::release(data);
}

void Simulation::set(InitialFunc *init, Model *model, TimeIntegrator *timeInt, Bcs *bcs, FluxMethod *fluxMethod,
                     SaveData *save) {
  Data *d = data;
  this->init = init;
  this->model = model;
  this->timeInt = timeInt;
  this->bcs = bcs;
  this->fluxMethod = fluxMethod;
  this->save = save;
  // the engine runs with the handed objects' settings
  this->bcs->bind();
  this->timeInt->bind();
  this->model->primsToAll(d->cons, d->prims, d->aux);
  this->bcs->apply(d->cons, d->prims, d->aux);
}

void Simulation::updateTime() {
  Data *d = data;
  if (env->rank == 0 && (d->iters % d->reportItersPeriod == 0)) printf("t = %f\n", d->t);
  const double dtX = d->cfl * d->dx / (d->alphaX * std::sqrt(d->dims));
  const double dtY = d->cfl * d->dy / (d->alphaY * std::sqrt(d->dims));
  const double dtZ = d->cfl * d->dz / (d->alphaZ * std::sqrt(d->dims));
  d->dt = (dtX <= dtY && dtX <= dtZ) ? dtX : ((dtY < dtZ) ? dtY : dtZ);
  if (d->iters < 15) d->dt *= 0.1;
  if (d->iters >= 15 && d->iters < 30) d->dt *= 0.25;
  if (d->t + d->dt > d->endTime) d->dt = d->endTime - d->t;
  this->timeInt->step(d->cons, d->prims, d->aux);
  d->t += d->dt;
  d->iters++;
}

void Simulation::syncHost() {
  Data *d = data;
  if (d->gpu.engine && d->gpu.hostStale) {
    method_host::check(mgpu_download_state(d->gpu.engine, d->cons, d->prims, d->aux), "Simulation::syncHost");
    d->gpu.hostStale = false;
  }
}

void Simulation::evolve(bool output, int safety) {
  Data *d = data;
  const bool mirror = d->gpu.mirrorEachStep;
  d->gpu.mirrorEachStep = false;
  if (output && save) save->saveVar("rho", 1);
  try {
    while (d->t < d->endTime) {
      updateTime();
      if ((output && save && d->iters % d->frameSkip == 0) || (safety > 0 && d->iters % safety == 0)) {
        syncHost();
        if (output && save && d->iters % d->frameSkip == 0) save->saveVar("rho", 1);
        if (safety > 0 && d->iters % safety == 0) {
          save->saveAll(true);
          if (env->rank == 0) printf("Data saved...\n");
        }
      }
    }
  } catch (...) {
    d->gpu.mirrorEachStep = mirror;
    syncHost();
    throw;
  }
  d->gpu.mirrorEachStep = mirror;
  syncHost();
  if (output && save) save->saveVar("rho", 1);
  if (env->rank == 0) printf("\n");
}

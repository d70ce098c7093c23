// Simulation data of the reference API (Project/CPU/Include/simData.h:49-233):
// grid sizes with ghost cells, constants and the host state arrays, laid out
// [var][i][j][k] with k fastest. `gpu` carries what the GPU engine needs to
// know about the objects built around this Data (model, WENO order,
// boundaries, integrator) and the engine handle itself.
#ifndef SIMDATA_H
#define SIMDATA_H

#include <string>
#include <vector>

#include "method_gpu.h"
#include "platformEnv.h"

#define ID(variable, idx, jdx, kdx) \
  ((variable) * (d->Nx) * (d->Ny) * (d->Nz) + (idx) * (d->Ny) * (d->Nz) + (jdx) * (d->Nz) + (kdx))

struct GpuLink {
  int model = -1;       // MGPU_SRMHD / MGPU_SRRMHD, set by the model's constructor
  int wenoOrder = 0;    // set by WenoBase
  int bc[3] = {1, 1, 1};// set by the Bcs subclass
  int integrator = 0;   // set by the time integrator
  int quirks = 1;       // reproduce the reference's flux-index quirks
  int extension = 0;    // MGPU_EXT_*, set by an integrator handed a model extension
  int extensionApplied = 0;   // extension the current engine was given
  int device = 0;
  mgpu_engine *engine = nullptr;
  mgpu_config created;         // configuration the engine was created with
  bool deviceCurrent = false;  // device state equals d->cons/prims/aux
  bool hostStale = false;      // device ahead of the host arrays
  bool mirrorEachStep = true;  // copy the state back to the host after each step
};

class Data {
 public:
  int nx, ny, nz;
  double xmin, xmax, ymin, ymax, zmin, zmax, endTime, cfl;
  int Ng;
  double gamma, sigma;
  int memSet, bcsSet, Ncons, Nprims, Naux;
  double cp, mu1, mu2;
  int frameSkip, reportItersPeriod;
  bool functionalSigma;
  double gam;
  double *cons, *prims, *aux, *f, *fnet, *source, *sourceExtension, *x, *y, *z;
  double alphaX, alphaY, alphaZ, t = -1, dt, dx, dy, dz;
  int iters, Nx, Ny, Nz, Ntot;
  std::vector<std::string> consLabels, primsLabels, auxLabels;
  int dims, is, js, ks, ie, je, ke;
  std::vector<double> optionalSimArgs;
  std::vector<std::string> optionalSimArgNames;
  int nOptionalSimArgs = 0;
  GpuLink gpu;

  int id(int var, int i, int j, int k) { return var * Nx * Ny * Nz + i * Ny * Nz + j * Nz + k; }
  double sigmaFunc(double *cons, double *prims, double *aux, int i = -1, int j = -1, int k = -1);
  void initData(PlatformEnv *env, int nOptionalSimArgs = 0,
                std::vector<double> optionalSimArgs = std::vector<double>(),
                std::vector<std::string> optionalSimArgNames = std::vector<std::string>());

  Data(int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
       double endTime, PlatformEnv *env, double cfl = 0.5, int Ng = 4, double gamma = 5.0 / 3.0,
       double sigma = 1e3, double cp = 0.1, double mu1 = -1.0e4, double mu2 = 1.0e4, int frameskip = 10,
       int reportItersPeriod = 1, bool funtionalSigma = false, double gam = 12);
  ~Data();
};

#endif

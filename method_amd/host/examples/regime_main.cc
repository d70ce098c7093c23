// REGIME driver written against the host API as the reference's REGIME
// example mains are (Examples/REGIME/CurrentSheetSerial/main.cc,
// Examples/REGIME/BrioWuSerial/main.cc), for a fixed number of steps.
//
//   regime_main <currentsheet|briowu|kh2d> <out.bin> <steps> [prims_in.bin]
//
// currentsheet: 128 cells on [-3,3], gamma 2, sigma 1e4, cfl 0.2, WENO3,
//   Outflow, CurrentSheetSingleFluid, RKSplit + REGIME.
// briowu: 128 cells on [0,1], gamma 2, sigma 50, cfl 0.15, WENO3, Outflow,
//   BrioWuSingleFluid, RKSplit + REGIME.
// kh2d: 32x32 on [-0.5,0.5]x[-1,1], gamma 4/3, sigma 100, cfl 0.4, WENO5,
//   Flow, primitives from prims_in.bin, RK2B + REGIME.
// Writes cons, prims and aux (reference layout, float64) after the steps.
// With steps = 0 it writes, for the initial state, REGIME::sourceExtension
// (Ncons x cells) followed by the fields set_vars, set_K and set_dwdsb fill:
// E (3), q, alpha, K (3), dwdsb (36), each over the padded cells.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "REGIME.h"
#include "RKPlus.h"
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "rkSplit.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
#include "weno.h"

static void put(FILE *o, const double *p, size_t n) { fwrite(p, sizeof(double), n, o); }

int main(int argc, char *argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s currentsheet|briowu|kh2d out.bin steps [prims_in.bin]\n", argv[0]);
    return 2;
  }
  const std::string which = argv[1];
  const int steps = atoi(argv[3]);
  const bool kh = which == "kh2d";
  if (!kh && which != "currentsheet" && which != "briowu") {
    fprintf(stderr, "unknown case %s\n", argv[1]);
    return 2;
  }
  if (kh && argc < 5) {
    fprintf(stderr, "kh2d needs prims_in.bin\n");
    return 2;
  }
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data *dp = nullptr;
  if (which == "currentsheet")
    dp = new Data(128, 0, 0, -3.0, 3.0, -1.0, 1.0, -1.5, 1.5, 7.0, &env, 0.2, 4, 2.0, 10000);
  else if (which == "briowu")
    dp = new Data(128, 0, 0, 0.0, 1.0, -1.0, 1.0, -1.5, 1.5, 0.4, &env, 0.15, 4, 2.0, 50);
  else
    dp = new Data(32, 32, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0, 1e10, &env, 0.4, 4, 4.0 / 3.0, 100.0);
  std::unique_ptr<Data> owner(dp);  // outlives the objects below, which use it
  Data &data = *dp;
  SRMHD model(&data);
  Weno3 weno3(&data);
  Weno5 weno5(&data);
  FVS fluxMethod(&data, kh ? (WenoBase *)&weno5 : (WenoBase *)&weno3, &model);
  REGIME modelExtension(&data, &fluxMethod);
  Outflow outflow(&data);
  Flow flow(&data);
  Bcs *bcs = kh ? (Bcs *)&flow : (Bcs *)&outflow;
  Simulation sim(&data, &env);
  InitialFunc *init = nullptr;
  std::vector<double> prims;
  if (which == "currentsheet") {
    init = new CurrentSheetSingleFluid(&data);
  } else if (which == "briowu") {
    init = new BrioWuSingleFluid(&data);
  } else {
    prims.resize((size_t)data.Nprims * data.Nx * data.Ny * data.Nz);
    FILE *f = fopen(argv[4], "rb");
    if (!f || fread(prims.data(), sizeof(double), prims.size(), f) != prims.size()) {
      fprintf(stderr, "cannot read %s\n", argv[4]);
      return 1;
    }
    fclose(f);
    init = new FromArray(&data, prims.data());
  }
  TimeIntegrator *timeInt = kh ? (TimeIntegrator *)new RK2B(&data, &model, bcs, &fluxMethod, &modelExtension)
                               : (TimeIntegrator *)new RKSplit(&data, &model, bcs, &fluxMethod, &modelExtension);
  sim.set(init, &model, timeInt, bcs, &fluxMethod);
  const size_t n = (size_t)data.Nx * data.Ny * data.Nz;
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 1;
  if (steps == 0) {
    modelExtension.sourceExtension(data.cons, data.prims, data.aux, data.sourceExtension);
    modelExtension.set_vars(data.cons, data.prims, data.aux);
    modelExtension.set_K(data.cons, data.prims, data.aux);
    modelExtension.set_dwdsb(data.cons, data.prims, data.aux);
    put(o, data.sourceExtension, n * data.Ncons);
    put(o, modelExtension.E, 3 * n);
    put(o, modelExtension.q, n);
    put(o, modelExtension.alpha, n);
    put(o, modelExtension.K, 3 * n);
    put(o, modelExtension.dwdsb, 36 * n);
  } else {
    for (int s = 0; s < steps; ++s) sim.updateTime();
    put(o, data.cons, n * data.Ncons);
    put(o, data.prims, n * data.Nprims);
    put(o, data.aux, n * data.Naux);
  }
  fclose(o);
  printf("regime_main %s: %d steps, t = %.17g\n", argv[1], steps, data.t);
  delete timeInt;
  delete init;
  return 0;
}

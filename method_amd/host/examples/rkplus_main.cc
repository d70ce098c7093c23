// RKPlus driver written against the host API as the reference's example
// mains are (Project/CPU/Src/main.cc:18-81), with one of RK2B, RK3, RK4 or
// RK4_10 (RKPlus.h) as the integrator.
//
//   rkplus_main <rk2b|rk3|rk4|rk4_10> <prims_in.bin> <out.bin> <steps>
//
// Ideal SRMHD on 32x32 over [-0.5,0.5]x[-1,1] (gamma 4/3, WENO3, cfl 0.4,
// cp 0.1, periodic x / outflow y) from the given primitives; writes cons,
// prims and aux (reference layout, float64) after the steps. With steps = 0
// it writes RKPlus::rhs of the initial state instead.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "RKPlus.h"
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
#include "weno.h"

int main(int argc, char *argv[]) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s rk2b|rk3|rk4|rk4_10 prims_in.bin out.bin steps\n", argv[0]);
    return 2;
  }
  const int steps = atoi(argv[4]);
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data data(32, 32, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0, 1e10, &env, 0.4, 4, 4.0 / 3.0, 0.0, 0.1, -1.0e4, 1.0e4, 10,
            1000000);
  SRMHD model(&data);
  Weno3 weno(&data);
  FVS fluxMethod(&data, &weno, &model);
  Flow bcs(&data);
  Simulation sim(&data, &env);
  std::vector<double> prims((size_t)data.Nprims * data.Nx * data.Ny * data.Nz);
  FILE *f = fopen(argv[2], "rb");
  if (!f || fread(prims.data(), sizeof(double), prims.size(), f) != prims.size()) {
    fprintf(stderr, "cannot read %s\n", argv[2]);
    return 1;
  }
  fclose(f);
  FromArray init(&data, prims.data());
  RKPlus *timeInt = nullptr;
  if (!strcmp(argv[1], "rk2b"))
    timeInt = new RK2B(&data, &model, &bcs, &fluxMethod);
  else if (!strcmp(argv[1], "rk3"))
    timeInt = new RK3(&data, &model, &bcs, &fluxMethod);
  else if (!strcmp(argv[1], "rk4"))
    timeInt = new RK4(&data, &model, &bcs, &fluxMethod);
  else if (!strcmp(argv[1], "rk4_10"))
    timeInt = new RK4_10(&data, &model, &bcs, &fluxMethod);
  else {
    fprintf(stderr, "unknown integrator %s\n", argv[1]);
    return 2;
  }
  sim.set(&init, &model, timeInt, &bcs, &fluxMethod);
  const size_t n = (size_t)data.Nx * data.Ny * data.Nz;
  FILE *o = fopen(argv[3], "wb");
  if (!o) return 1;
  if (steps == 0) {
    std::vector<double> rhs(n * data.Ncons, 0.0);
    timeInt->rhs(data.cons, data.prims, data.aux, rhs.data());
    fwrite(rhs.data(), sizeof(double), rhs.size(), o);
  } else {
    for (int s = 0; s < steps; ++s) sim.updateTime();
    fwrite(data.cons, sizeof(double), n * data.Ncons, o);
    fwrite(data.prims, sizeof(double), n * data.Nprims, o);
    fwrite(data.aux, sizeof(double), n * data.Naux, o);
  }
  fclose(o);
  printf("rkplus_main %s: %d steps, t = %.17g\n", argv[1], steps, data.t);
  delete timeInt;
  return 0;
}

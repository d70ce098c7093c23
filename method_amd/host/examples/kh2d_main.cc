// Kelvin-Helmholtz driver written against the host API exactly as the
// reference's example mains are (Examples/KelvinHelmholtz/*/main.cc,
// Project/CPU/Src/main.cc:18-81): Data, model, WENO, FVS, boundaries,
// Simulation, initial data, integrator, set, then updateTime calls.
//
//   kh2d_main <out.bin> [steps] [prims_in.bin]
//
// Runs ideal SRMHD on 64x64 over [-0.5,0.5]x[-1,1] (gamma 4/3, WENO3, RK2,
// cfl 0.4, periodic x / outflow y) and writes cons, prims and aux (reference
// layout, float64) after the last step. Initial primitives come from
// prims_in.bin when given, else from KHShearLayers(seed 1234).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "RK2.h"
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
#include "weno.h"

int main(int argc, char *argv[]) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s out.bin [steps] [prims_in.bin]\n", argv[0]);
    return 2;
  }
  const int steps = argc > 2 ? atoi(argv[2]) : 10;
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data data(64, 64, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0, 1e10, &env, 0.4, 4, 4.0 / 3.0, 0.0, 0.1, -1.0e4, 1.0e4, 10,
            1000000);
  SRMHD model(&data);
  Weno3 weno(&data);
  FVS fluxMethod(&data, &weno, &model);
  Flow bcs(&data);
  Simulation sim(&data, &env);
  InitialFunc *init = nullptr;
  std::vector<double> prims;
  if (argc > 3) {
    prims.resize((size_t)data.Nprims * data.Nx * data.Ny * data.Nz);
    FILE *f = fopen(argv[3], "rb");
    if (!f || fread(prims.data(), sizeof(double), prims.size(), f) != prims.size()) {
      fprintf(stderr, "cannot read %s\n", argv[3]);
      return 1;
    }
    fclose(f);
    init = new FromArray(&data, prims.data());
  } else {
    init = new KHShearLayers(&data, 1234);
  }
  RK2 timeInt(&data, &model, &bcs, &fluxMethod);
  sim.set(init, &model, &timeInt, &bcs, &fluxMethod);
  for (int s = 0; s < steps; ++s) sim.updateTime();
  const size_t n = (size_t)data.Nx * data.Ny * data.Nz;
  FILE *o = fopen(argv[1], "wb");
  if (!o) return 1;
  fwrite(data.cons, sizeof(double), n * data.Ncons, o);
  fwrite(data.prims, sizeof(double), n * data.Nprims, o);
  fwrite(data.aux, sizeof(double), n * data.Naux, o);
  fclose(o);
  printf("kh2d_main: %d steps, t = %.17g\n", steps, data.t);
  delete init;
  return 0;
}

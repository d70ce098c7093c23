// The reference's serial Kelvin-Helmholtz example set-up
// (Examples/KelvinHelmholtz/SingleFluidIdealSerial/main.cc) written against
// the host API: KHInstabilitySingleFluid with a magnetic field, ideal SRMHD,
// WENO3, FVS, periodic boundaries, RKSplit, and SerialSaveData::saveAll at the
// end (text files under ./Data/Final).
//
//   kh_example_main [n] [steps]
//
// n x n cells (default 64); steps updateTime calls, or evolve() to the end
// time 3.0 when steps is omitted or negative.
#include <cstdlib>

#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "rkSplit.h"
#include "serialEnv.h"
#include "serialSaveData.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
#include "weno.h"

int main(int argc, char *argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 64;
  const int steps = argc > 2 ? std::atoi(argv[2]) : -1;

  const int Ng = 4, nz = 0;
  const double xmin = -0.5, xmax = 0.5, ymin = -1.0, ymax = 1.0, zmin = 0.0, zmax = 1.0;
  const double endTime = 3.0, cfl = 0.6, gamma = 4.0 / 3.0, sigma = 10, cp = 1.0, mu1 = -100, mu2 = 100;
  const int frameSkip = 10, reportItersPeriod = 50;

  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data data(n, n, nz, xmin, xmax, ymin, ymax, zmin, zmax, endTime, &env, cfl, Ng, gamma, sigma, cp, mu1, mu2,
            frameSkip, reportItersPeriod);
  SRMHD model(&data);
  Weno3 weno(&data);
  FVS fluxMethod(&data, &weno, &model);
  Periodic bcs(&data);
  Simulation sim(&data, &env);
  KHInstabilitySingleFluid init(&data, 1);
  RKSplit timeInt(&data, &model, &bcs, &fluxMethod);
  SerialSaveData save(&data, &env);

  sim.set(&init, &model, &timeInt, &bcs, &fluxMethod, &save);
  if (steps < 0) {
    sim.evolve(false);
  } else {
    for (int s = 0; s < steps; ++s) sim.updateTime();
  }
  save.saveAll();
  return 0;
}

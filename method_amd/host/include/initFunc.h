// Initial data (Project/CPU/Include/initFunc.h). The base class zeroes every
// state array (initFunc.cc:10-40); subclasses write d->prims.
#ifndef INITFUNC_H
#define INITFUNC_H

#include <cstdint>

#include "simData.h"

class InitialFunc {
 public:
  Data *data;
  InitialFunc(Data *data);
  virtual ~InitialFunc() {}
};

// Primitives copied from a caller's array over the whole padded grid.
class FromArray : public InitialFunc {
 public:
  FromArray(Data *data, const double *prims);
};

// Ideal SRMHD Kelvin-Helmholtz shear layers at |y| = 0.5 with seeded random
// transverse velocity (the workload variant of initFunc.cc:512-563):
// rho = rho_in inside, rho_out outside via tanh((|y|-0.5)/width),
// vx = shear tanh((|y|-0.5)/width), vy (and vz in 3-D) = amp (2u - 1) with
// u = splitmix64 counter stream of the seed, p = pressure, Bz = bz.
class KHShearLayers : public InitialFunc {
 public:
  KHShearLayers(Data *data, uint64_t seed = 1234, double shear = 0.5, double bz = 0.1, double width = 0.01,
                double rhoIn = 0.55, double rhoOut = 0.45, double amp = 0.01, double pressure = 1.0);
};

// The reference's single-fluid Kelvin-Helmholtz state (initFunc.cc:512-563):
// shear layers at y = +-0.5 of width a = 0.01 with rho = 0.55 +- 0.45 tanh,
// vx = +-0.5 tanh, a vy bump 0.1 * 0.5 sin(2 pi x) exp(-(y -+ 0.5)^2 / 0.1^2),
// p = 1, and Bz = 0.1 when mag is non-zero (ideal E = -v x B for resistive
// models). Requires gamma = 4/3 and the domain [-0.5, 0.5] x [-1, 1].
class KHInstabilitySingleFluid : public InitialFunc {
 public:
  KHInstabilitySingleFluid(Data *data, int mag = 0);
};

// Uniform draw in [0,1) of the counter stream used by KHShearLayers.
double initUniform01(uint64_t seed, uint64_t index, int stream);

#endif

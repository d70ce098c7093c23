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

// Orszag-Tang vortex (initFunc.cc:229-271) on [0,1]^2: rho = 25/(36 pi),
// p = 5/(12 pi), v = 0.5 (-sin 2 pi y, sin 2 pi x), B = (-sin 2 pi y, sin 4 pi x)
// / sqrt(4 pi); in 3-D the reference's form with a z modulation.
class OTVortexSingleFluid : public InitialFunc {
 public:
  OTVortexSingleFluid(Data *data);
};

// Brio-Wu shock tube along axis dir (initFunc.cc:373-510): left state
// rho = p = 1, right state rho = 0.125, p = 0.1, with the transverse field
// +-0.5 (By for dir 0, Bz for dir 1, Bx for dir 2).
class BrioWuSingleFluid : public InitialFunc {
 public:
  BrioWuSingleFluid(Data *data, int dir = 0);
};

// Current sheet (initFunc.cc:191-226): rho = 1, p = 50, at rest, with one
// field component erf(s sqrt(sigma) / 2) across the sheet at s = 0 (By along
// x for direction 0, Bz along y for 1, Bx along z for 2). Needs x in [-3, 3]
// and gamma = 2.
class CurrentSheetSingleFluid : public InitialFunc {
 public:
  CurrentSheetSingleFluid(Data *data, int direction = 0);
};

// Kelvin-Helmholtz with two interfaces at y = 0.25 and 0.75 displaced by ten
// random Fourier modes (initFunc.cc:612-680): rho = 1 / 0.1 and vx = +-0.5
// outside / between them, p = 1, Bz = 0.1 when mag is non-zero. The mode
// amplitudes and phases are drawn with std::mt19937(seed) and
// std::uniform_real_distribution<>(0, 1) in the reference's order.
class KHRandomInstabilitySingleFluid : public InitialFunc {
 public:
  KHRandomInstabilitySingleFluid(Data *data, int mag = 0, int seed = 10);
};

// Uniform draw in [0,1) of the counter stream used by KHShearLayers.
double initUniform01(uint64_t seed, uint64_t index, int stream);

#endif

// Initial data of the host API (initFunc.cc:10-40 and the workload variants).
#include "initFunc.h"

#include <cmath>
#include <cstdio>
#include <random>
#include <cstring>
#include <stdexcept>

InitialFunc::InitialFunc(Data *data) : data(data) {
  Data *d = data;
  if (d->memSet != 1)
    throw std::runtime_error("Must construct simulation class before implementing initial state. Need to allocate arrays.");
  const size_t n = (size_t)d->Nx * d->Ny * d->Nz;
  std::memset(d->cons, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->f, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->fnet, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->source, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->prims, 0, sizeof(double) * n * d->Nprims);
  std::memset(d->aux, 0, sizeof(double) * n * d->Naux);
  d->gpu.deviceCurrent = false;
}

FromArray::FromArray(Data *data, const double *prims) : InitialFunc(data) {
  std::memcpy(data->prims, prims, sizeof(double) * (size_t)data->Nx * data->Ny * data->Nz * data->Nprims);
}

// splitmix64 finaliser (Steele, Lea & Flood 2014).
static uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double initUniform01(uint64_t seed, uint64_t index, int stream) {
  const uint64_t key = seed + 0x9E3779B97F4A7C15ULL * (2 * index + (uint64_t)stream + 1);
  return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

KHShearLayers::KHShearLayers(Data *data, uint64_t seed, double shear, double bz, double width, double rhoIn,
                             double rhoOut, double amp, double pressure)
    : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims != 8) throw std::invalid_argument("KHShearLayers is an ideal SRMHD (8 primitive) initial state");
  const double hs = 0.5 * (rhoIn + rhoOut), hd = 0.5 * (rhoIn - rhoOut);
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        const double t = std::tanh((std::fabs(d->y[j]) - 0.5) / width);
        const uint64_t n = ((uint64_t)i * d->Ny + j) * d->Nz + k;
        d->prims[ID(0, i, j, k)] = hs - hd * t;
        d->prims[ID(1, i, j, k)] = shear * t;
        d->prims[ID(2, i, j, k)] = amp * (2.0 * initUniform01(seed, n, 0) - 1.0);
        d->prims[ID(3, i, j, k)] = d->dims == 3 ? amp * (2.0 * initUniform01(seed, n, 1) - 1.0) : 0.0;
        d->prims[ID(4, i, j, k)] = pressure;
        d->prims[ID(7, i, j, k)] = bz;
      }
}

KHInstabilitySingleFluid::KHInstabilitySingleFluid(Data *data, int mag) : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims > 15)
    throw std::invalid_argument(
        "Trying to implement a single fluid initial state on incorrect model.\n\tModel has wrong number of primitive "
        "variables to be single fluid model.");
  if (d->gamma != 4.0 / 3.0) throw std::invalid_argument("Expected the index gamma = 4/3\n");
  if (d->xmin != -0.5 || d->xmax != 0.5) throw std::invalid_argument("Domain has incorrect values. Expected x E [-0.5, 0.5]\n");
  if (d->ymin != -1.0 || d->ymax != 1.0) throw std::invalid_argument("Domain has incorrect values. Expected y E [-1.0, 1.0]\n");
  const double pi = 3.14159265358979323846;
  const double sig = 0.1, vShear = 0.5, A0 = 0.1, a = 0.01, rho0 = 0.55, rho1 = 0.45, B0 = 0.1;
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        double *P = d->prims;
        P[ID(4, i, j, k)] = 1.0;
        if (mag) P[ID(7, i, j, k)] = B0;
        // the layer at y = +0.5 for y > 0, its mirror image below
        const double s = d->y[j] > 0 ? 1.0 : -1.0;
        const double yl = d->y[j] - s * 0.5;
        const double t = std::tanh(yl / a);
        P[ID(0, i, j, k)] = d->y[j] > 0 ? rho0 + rho1 * t : rho0 - rho1 * t;
        P[ID(1, i, j, k)] = d->y[j] > 0 ? vShear * t : -vShear * t;
        const double bump = A0 * vShear * std::sin(2 * pi * d->x[i]) * (std::exp(-std::pow(yl, 2) / (sig * sig)));
        P[ID(2, i, j, k)] = d->y[j] > 0 ? bump : -bump;
        if (d->Ncons > 9) {
          P[ID(8, i, j, k)] = -(P[ID(2, i, j, k)] * P[ID(7, i, j, k)] - P[ID(3, i, j, k)] * P[ID(6, i, j, k)]);
          P[ID(9, i, j, k)] = -(P[ID(3, i, j, k)] * P[ID(5, i, j, k)] - P[ID(1, i, j, k)] * P[ID(7, i, j, k)]);
          P[ID(10, i, j, k)] = -(P[ID(1, i, j, k)] * P[ID(6, i, j, k)] - P[ID(2, i, j, k)] * P[ID(5, i, j, k)]);
        }
      }
}

namespace {
const double kPi = 3.14159265358979323846;
}

OTVortexSingleFluid::OTVortexSingleFluid(Data *data) : InitialFunc(data) {
  Data *d = data;
  if (d->xmin != 0.0 || d->xmax != 1.0 || d->ymin != 0.0 || d->ymax != 1.0)
    throw std::invalid_argument("Boundaries are not as expected for OTVortex, uasually (x,y) E [0, 1]");
  if (d->Nprims > 15)
    throw std::invalid_argument(
        "Trying to implement a single fluid initial state on multifluid model.\nModel has too many primitive "
        "variables to be single fluid.");
  if (d->ny == 0)
    throw std::invalid_argument(
        "System must be at least two dimensional for this initial set up. \nSet ny > 0 and try again.\n");
  const double rB = std::sqrt(4.0 * kPi);
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        double *P = d->prims;
        P[ID(0, i, j, k)] = 25.0 / 36.0 / kPi;
        P[ID(4, i, j, k)] = 5.0 / 12.0 / kPi;
        P[ID(5, i, j, k)] = -std::sin(2.0 * kPi * d->y[j]) / rB;
        P[ID(6, i, j, k)] = std::sin(4.0 * kPi * d->x[i]) / rB;
        if (d->dims == 2) {
          P[ID(1, i, j, k)] = -0.5 * std::sin(2.0 * kPi * d->y[j]);
          P[ID(2, i, j, k)] = 0.5 * std::sin(2.0 * kPi * d->x[i]);
        } else {
          // the reference modulates with z taken at the j index (initFunc.cc:261-266)
          const double mz = std::sin(2.0 * kPi * d->z[j]);
          P[ID(1, i, j, k)] = -0.5 * std::sin(2.0 * kPi * d->y[j]) * (1 + 0.2 * mz);
          P[ID(2, i, j, k)] = 0.5 * std::sin(2.0 * kPi * d->x[i]) * (1 + 0.2 * mz);
          P[ID(3, i, j, k)] = 0.2 * mz;
        }
      }
}

CurrentSheetSingleFluid::CurrentSheetSingleFluid(Data *data, int direction) : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims > 15)
    throw std::invalid_argument(
        "Trying to implement a single fluid initial state on incorrect model.\n\tModel has wrong number of primitive "
        "variables to be single fluid model.");
  if (d->xmin != -3.0 || d->xmax != 3.0) throw std::invalid_argument("Domain has incorrect values. Expected x E [-3, 3]\n");
  if (d->gamma != 2.0) throw std::invalid_argument("Expected the index gamma = 2\n");
  const int comp = direction == 0 ? 6 : (direction == 1 ? 7 : (direction == 2 ? 5 : -1));
  const double root = std::sqrt(d->sigma);
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        double *P = d->prims;
        P[ID(0, i, j, k)] = 1.0;
        P[ID(4, i, j, k)] = 50.0;
        if (comp < 0) continue;
        const double s = direction == 0 ? d->x[i] : (direction == 1 ? d->y[j] : d->z[k]);
        P[ID(comp, i, j, k)] = std::erf(0.5 * s * root);
      }
}

BrioWuSingleFluid::BrioWuSingleFluid(Data *data, int dir) : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims >= 16)
    throw std::invalid_argument(
        "Trying to implement a single fluid initial state on incorrect model.\n\tModel has wrong number of primitive "
        "variables to be single fluid model.");
  if (d->nx % 2 || d->ny % 2 || d->nz % 2)
    throw std::invalid_argument("Please ensure even number of cells in each direction for Brio Wu initial data.\n");
  // the axis split in two, and the transverse field component that jumps
  const int fac[3] = {dir == 0 ? 2 : 1, dir == 1 ? 2 : 1, dir != 0 && dir != 1 ? 2 : 1};
  const int bcomp = dir == 0 ? 6 : (dir == 1 ? 7 : 5);
  const double lo[3] = {(d->xmax - d->xmin) / fac[0] + d->xmin, (d->ymax - d->ymin) / fac[1] + d->ymin,
                        (d->zmax - d->zmin) / fac[2] + d->zmin};
  const double hi[3] = {d->xmax - (d->xmax - d->xmin) / fac[0], d->ymax - (d->ymax - d->ymin) / fac[1],
                        d->zmax - (d->zmax - d->zmin) / fac[2]};
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        // axes the grid does not have take no part in the comparison
        const bool left = d->x[i] < lo[0] && (d->dims < 2 || d->y[j] < lo[1]) && (d->dims < 3 || d->z[k] < lo[2]);
        const bool right = d->x[i] > hi[0] && (d->dims < 2 || d->y[j] > hi[1]) && (d->dims < 3 || d->z[k] > hi[2]);
        double *P = d->prims;
        if (left) {
          P[ID(0, i, j, k)] = 1;
          P[ID(4, i, j, k)] = 1;
          P[ID(5, i, j, k)] = P[ID(6, i, j, k)] = P[ID(7, i, j, k)] = 0;
          P[ID(bcomp, i, j, k)] = 0.5;
        }
        if (right) {
          P[ID(0, i, j, k)] = 0.125;
          P[ID(4, i, j, k)] = 0.1;
          P[ID(5, i, j, k)] = P[ID(6, i, j, k)] = P[ID(7, i, j, k)] = 0;
          P[ID(bcomp, i, j, k)] = -0.5;
        }
      }
}

KHRandomInstabilitySingleFluid::KHRandomInstabilitySingleFluid(Data *data, int mag, int seed) : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims > 15)
    throw std::invalid_argument(
        "Trying to implement a single fluid initial state on incorrect model.\n\tModel has wrong number of primitive "
        "variables to be single fluid model.");
  if (d->gamma != 4.0 / 3.0) throw std::invalid_argument("Expected the index gamma = 4/3\n");
  if (d->xmin != 0.0 || d->xmax != 1.0) throw std::invalid_argument("Domain has incorrect values. Expected x E [0.0, 1.0]\n");
  if (d->ymin != 0.0 || d->ymax != 1.0) throw std::invalid_argument("Domain has incorrect values. Expected y E [0.0, 1.0]\n");
  const double vShear = 0.5, rho0 = 1.0, rho1 = 0.1, epsilon = 0.01;
  const int modes = 10;
  double aLo[modes], aHi[modes], bLo[modes], bHi[modes];
  double sumLo = 0.0, sumHi = 0.0;
  std::mt19937 gen(seed);
  std::uniform_real_distribution<> dis(0.0, 1.0);
  for (int n = 0; n < modes; n++) {  // draw order: a_lower, a_upper, b_lower, b_upper
    aLo[n] = dis(gen);
    sumLo += aLo[n];
    aHi[n] = dis(gen);
    sumHi += aHi[n];
    bLo[n] = -kPi + 2.0 * kPi * dis(gen);
    bHi[n] = -kPi + 2.0 * kPi * dis(gen);
  }
  for (int n = 0; n < modes; n++) {
    aLo[n] /= sumLo;
    aHi[n] /= sumHi;
    std::printf("Random coeffs, %d: %g, %g, %g, %g\n", n, aLo[n], aHi[n], bLo[n], bHi[n]);
  }
  for (int i = 0; i < d->Nx; i++) {
    double yLo = 0.25, yHi = 0.75;
    for (int n = 0; n < modes; n++) {
      yLo += epsilon * aLo[n] * std::cos(bLo[n] + 2 * n * kPi * d->x[i]);
      yHi += epsilon * aHi[n] * std::cos(bHi[n] + 2 * n * kPi * d->x[i]);
    }
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        double *P = d->prims;
        P[ID(4, i, j, k)] = 1.0;
        if (mag) P[ID(7, i, j, k)] = 0.1;
        const bool outside = d->y[j] < yLo || d->y[j] > yHi;
        P[ID(0, i, j, k)] = outside ? rho0 : rho1;
        P[ID(1, i, j, k)] = outside ? vShear : -vShear;
      }
  }
}

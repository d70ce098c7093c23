// Data and SerialEnv of the host API (simData.cc:11-127, serialEnv.cc:11-21).
#include "simData.h"

#include <cmath>
#include <stdexcept>

#include "gpu_link.h"
#include "serialEnv.h"

Data::Data(int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
           double endTime, PlatformEnv *env, double cfl, int Ng, double gamma, double sigma, double cp, double mu1,
           double mu2, int frameSkip, int reportItersPeriod, bool functionalSigma, double gam)
    : nx(nx), ny(ny), nz(nz), xmin(xmin), xmax(xmax), ymin(ymin), ymax(ymax), zmin(zmin), zmax(zmax),
      endTime(endTime), cfl(cfl), Ng(Ng), gamma(gamma), sigma(sigma), memSet(0), bcsSet(0), Ncons(0), Nprims(0),
      Naux(0), cp(cp), mu1(mu1), mu2(mu2), frameSkip(frameSkip), reportItersPeriod(reportItersPeriod),
      functionalSigma(functionalSigma), gam(gam), cons(nullptr), prims(nullptr), aux(nullptr), f(nullptr),
      fnet(nullptr), source(nullptr), sourceExtension(nullptr), x(nullptr), y(nullptr), z(nullptr), t(0) {
  initData(env);
}

Data::~Data() { method_host::release(this); }

// Data::initData (simData.cc:61-112): one rank per axis on a single GPU.
void Data::initData(PlatformEnv *env, int nOpt, std::vector<double> opt, std::vector<std::string> optNames) {
  Nx = nx + 2 * Ng;
  Ny = ny + 2 * Ng;
  Nz = nz + 2 * Ng;
  Ntot = Nx * Ny * Nz;
  dims = 3;
  if (nz == 0) {
    Nz = 1;
    zmin = -1e20;
    zmax = 1e20;
    dims = 2;
  }
  if (ny == 0) {
    Nz = Ny = 1;
    zmin = ymin = -1e20;
    zmax = ymax = 1e20;
    dims = 1;
  }
  Ntot = Nx * Ny * Nz;
  is = Ng;
  ie = Nx - Ng;
  js = Ng;
  je = Ny - Ng;
  ks = Ng;
  ke = Nz - Ng;
  if (dims < 3) {
    ks = 0;
    ke = 1;
  }
  if (dims < 2) {
    js = 0;
    je = 1;
  }
  if (sigma < 0.0) throw std::invalid_argument("Conductivity must be non-negative, sigma >= 0.\n");
  if (mu1 > 0.0 or mu2 < 0.0)
    throw std::invalid_argument(
        "Species 1 must have negative charge, mu1 < 0, and species 2 must have positive charge, mu2 > 0.\n");
  nOptionalSimArgs = nOpt;
  optionalSimArgs = opt;
  optionalSimArgNames = optNames;
  (void)env;
}

// Data::sigmaFunc (simData.cc:114-127).
double Data::sigmaFunc(double *cons, double *prims, double *aux, int i, int j, int k) {
  (void)prims;
  (void)aux;
  if (!functionalSigma) return sigma;
  if (i < 0 || j < 0 || k < 0) return sigma * std::pow(cons[0], gam);
  return sigma * std::pow(cons[id(0, i, j, k)], gam);
}

SerialEnv::SerialEnv(int *argcP, char **argvP[], int nxRanks, int nyRanks, int nzRanks, int testing)
    : PlatformEnv(testing) {
  (void)argcP;
  (void)argvP;
  (void)nxRanks;
  (void)nyRanks;
  (void)nzRanks;
  this->nProc = 1;
  this->rank = 0;
  this->nxRanks = this->nyRanks = this->nzRanks = 1;
  this->xRankId = this->yRankId = this->zRankId = 0;
  this->leftXNeighbourRank = this->rightXNeighbourRank = 0;
  this->leftYNeighbourRank = this->rightYNeighbourRank = 0;
  this->leftZNeighbourRank = this->rightZNeighbourRank = 0;
}

SerialEnv::~SerialEnv() {}

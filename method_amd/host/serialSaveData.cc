// Text output with the reference's file layout (serialSaveData.cc:8-360).
#include "serialSaveData.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

namespace {

// mkdir -p of path's directory part.
void make_dirs(const std::string &path) {
  for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1)) {
    const std::string sub = path.substr(0, p);
    if (sub.empty() || sub == "." || sub == "..") continue;
    mkdir(sub.c_str(), 0755);
  }
}

FILE *open_or_exit(const std::string &path, const char *what) {
  make_dirs(path);
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::printf("Error: could not open '%s' for writing.\n", what);
    std::exit(1);
  }
  return f;
}

// "<name> = a, b, c" header followed by the interior values: one row per
// (variable, i, j) holding the k values of that pencil in 3-D, one value per
// row in 2-D and 1-D.
void write_field(FILE *f, const char *name, const std::vector<std::string> &labels, Data *d, const double *a,
                 int nvar) {
  std::fprintf(f, "%s = ", name);
  for (int v = 0; v < nvar - 1; ++v) std::fprintf(f, "%s, ", labels[v].c_str());
  std::fprintf(f, "%s\n", labels[nvar - 1].c_str());
  const int nx = d->Nx - 2 * d->Ng;
  const int ny = d->dims > 1 ? d->Ny - 2 * d->Ng : 1;
  const int nz = d->dims > 2 ? d->Nz - 2 * d->Ng : 1;
  const int oy = d->dims > 1 ? d->Ng : 0, oz = d->dims > 2 ? d->Ng : 0;
  for (int v = 0; v < nvar; ++v)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        for (int k = 0; k < nz; ++k) std::fprintf(f, "%.16f ", a[d->id(v, i + d->Ng, j + oy, k + oz)]);
        std::fprintf(f, "\n");
      }
}

}  // namespace

void SerialSaveData::saveAll(bool timeSeries) {
  dir_ = root() + (timeSeries ? "Data/TimeSeries" : "Data/Final");
  app_ = timeSeries ? std::to_string(Nouts++) : std::string();
  saveCons();
  savePrims();
  saveAux();
  saveDomain();
  saveConsts();
}

void SerialSaveData::saveCons() {
  FILE *f = open_or_exit(dir_ + "/Conserved/cons" + app_ + ".dat", "cons.dat");
  write_field(f, "cons", d->consLabels, d, d->cons, d->Ncons);
  std::fclose(f);
}

void SerialSaveData::savePrims() {
  FILE *f = open_or_exit(dir_ + "/Primitive/prims" + app_ + ".dat", "prims.dat");
  write_field(f, "prims", d->primsLabels, d, d->prims, d->Nprims);
  std::fclose(f);
}

void SerialSaveData::saveAux() {
  FILE *f = open_or_exit(dir_ + "/Auxiliary/aux" + app_ + ".dat", "aux.dat");
  write_field(f, "aux", d->auxLabels, d, d->aux, d->Naux);
  std::fclose(f);
}

void SerialSaveData::saveDomain() {
  FILE *f = open_or_exit(dir_ + "/Domain/domain" + app_ + ".dat", "domain.dat");
  const double *axes[3] = {d->x, d->y, d->z};
  const int n[3] = {d->Nx, d->Ny, d->Nz};
  for (int a = 0; a < 3; ++a) {
    for (int i = 0; i < n[a]; ++i) std::fprintf(f, "%.16f ", axes[a][i]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}

void SerialSaveData::saveConsts() {
  FILE *f = open_or_exit(dir_ + "/Constants/constants" + app_ + ".dat", "constants.dat");
  std::fprintf(f,
               "constants = nx, ny, nz, Nx, Ny, Nz, xmin, xmax, ymin, ymax, zmin, zmax, endTime, cfl, Ng, gamma, "
               "sigma, Ncons, Nprims, Naux, cp, dt, t, dx, dy, dz\n");
  std::fprintf(f,
               "%d %d %d %d %d %d %.16f %.16f %.16f %.16f %.16f %.16f %.16f %.16f %d %.16f %.16f %d %d %d %.16f "
               "%.16f %.16f %.16f %.16f %.16f\n",
               d->nx, d->ny, d->nz, d->Nx, d->Ny, d->Nz, d->xmin, d->xmax, d->ymin, d->ymax, d->zmin, d->zmax,
               d->endTime, d->cfl, d->Ng, d->gamma, d->sigma, d->Ncons, d->Nprims, d->Naux, d->cp, d->dt, d->t,
               d->dx, d->dy, d->dz);
  std::fclose(f);
}

void SerialSaveData::saveVar(std::string variable, int num) {
  int which = 0, nvar = 0;  // 1 cons, 2 prims, 3 aux
  const std::vector<std::string> *lists[3] = {&d->consLabels, &d->primsLabels, &d->auxLabels};
  for (int l = 0; l < 3 && !which; ++l)
    for (size_t v = 0; v < lists[l]->size(); ++v)
      if ((*lists[l])[v] == variable) {
        which = l + 1;
        nvar = (int)v;
        break;
      }
  if (!which) {
    std::printf("Error: Could not find user specified variable '%s'\n", variable.c_str());
    std::exit(1);
  }
  const std::string base = root() + "Data/TimeSeries/UserDef/";
  FILE *f = open_or_exit(base + variable + std::to_string(Nouts) + ".dat", "user-defined file");
  const double *a = which == 1 ? d->cons : (which == 2 ? d->prims : d->aux);
  std::fprintf(f, "var = %s, t = %18.16f\n", variable.c_str(), d->t);
  for (int i = 0; i < d->Nx; ++i)
    for (int j = 0; j < d->Ny; ++j) {
      for (int k = 0; k < d->Nz; ++k) std::fprintf(f, "%.16f ", a[d->id(nvar, i, j, k)]);
      std::fprintf(f, "\n");
    }
  std::fclose(f);
  // the first output of a run lists the variables saved in it
  if (Nouts == 0) {
    std::ofstream info(base + "info", Ncount == 0 ? std::ios::out : std::ios::app);
    info << variable << std::endl;
  }
  if (++Ncount == num) {
    Ncount = 0;
    ++Nouts;
  }
}

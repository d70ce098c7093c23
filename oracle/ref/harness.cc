// Test-infrastructure only: a driver that links the reference METHOD serial CPU
// sources, compiled in place from the reference tree by oracle/ref/Makefile, and
// runs them on initial data read from a binary file. Nothing in the product
// (method_amd/, include/) links or calls this program; tests and the fixture
// generator under tests/golden/ use its outputs as the reference answer.
//
// Modes (key=value arguments):
//   mode=run   : Simulation::set then `steps` calls of Simulation::updateTime
//                (simulation.cc:80-133), dumping cons/prims/aux after the steps
//                listed in dump=..., plus one text line per step with t, dt.
//   mode=fnet  : Simulation::set then one FVS::F (fluxVectorSplitting.cc:58-131)
//                on the initial state; dumps fnet.
//   mode=regime: Simulation::set then one REGIME::sourceExtension
//                (REGIME.cc:83-138) of the initial state; dumps it.
//   mode=frecon: Simulation::set then, per axis, Model::fluxVector and
//                FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56) of the
//                set-up state; dumps f and frecon of each axis.
//   mode=c2p   : reads cons, prims and aux (the C2P guess) and calls
//                Model::getPrimitiveVars once (srmhd.cc:390-562,
//                srrmhd.cc:314-356); dumps prims and aux.
//
// Options: integrator=rk2 | ssp2 | rksplit | rk2b | rk3 | rk4 | rk4_10 (RK2,
// SSP2, RKSplit and the RKPlus family, RKPlus.h:35-515); init=file
// (prims_in=...) or one of the reference's initial states: kh
// (KHInstabilitySingleFluid, mag=...), khrandom (KHRandomInstabilitySingleFluid,
// mag=..., seed=...), ot (OTVortexSingleFluid), briowu (BrioWuSingleFluid,
// dir=...), currentsheet (CurrentSheetSingleFluid, dir=...) (initFunc.cc);
// regime=1 hands a REGIME model extension to RKSplit / RKPlus; save=1 calls SerialSaveData::saveAll
// (serialSaveData.cc:8-31) after the steps, writing under ./Data/Final.
//
// Boundary conditions: bcx/bcy/bcz = periodic | outflow. When the triple matches
// one of the reference classes it is used as is (Periodic, Outflow, Flow,
// boundaryConds.cc:5-145,449-730); other triples (periodic x and z with outflow
// y) use AxisBcs below, which applies the same per-axis copies in the same
// x -> y -> z order.
//
// Binary files are raw little-endian float64 in the reference layout
// [var][i][j][k] over the ghost-padded grid (simData.h:10).

#include "simData.h"
#include "simulation.h"
#include "initFunc.h"
#include "srmhd.h"
#include "srrmhd.h"
#include "boundaryConds.h"
#include "RK2.h"
#include "SSP2.h"
#include "rkSplit.h"
#include "RKPlus.h"
#include "REGIME.h"
#include "serialSaveData.h"
#include "fluxVectorSplitting.h"
#include "serialEnv.h"
#include "weno.h"
#include "cminpack.h"

#include <chrono>
#include <sys/stat.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

int SRMHDresidual(void *p, int n, const double *x, double *fvec, int iflag);

namespace {

std::map<std::string, std::string> g_args;

std::string arg(const char *k, const char *def) {
  auto it = g_args.find(k);
  return it == g_args.end() ? std::string(def) : it->second;
}
double argd(const char *k, double def) {
  auto it = g_args.find(k);
  return it == g_args.end() ? def : strtod(it->second.c_str(), nullptr);
}
int argi(const char *k, int def) {
  auto it = g_args.find(k);
  return it == g_args.end() ? def : atoi(it->second.c_str());
}

void read_doubles(const std::string &path, double *dst, size_t n) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  size_t got = fread(dst, sizeof(double), n, f);
  fclose(f);
  if (got != n) throw std::runtime_error("short read from " + path);
}

void write_doubles(FILE *f, const double *src, size_t n) {
  if (fwrite(src, sizeof(double), n, f) != n) throw std::runtime_error("short write");
}

// Per-axis periodic/outflow ghost fill in the reference's x -> y -> z order.
class AxisBcs : public Bcs {
 public:
  int periodic[3];
  AxisBcs(Data *d, int px, int py, int pz) : Bcs(d) {
    periodic[0] = px; periodic[1] = py; periodic[2] = pz;
  }
  void fill(double *a, int nvar) {
    Data *d(this->data);
    for (int var = 0; var < nvar; var++)
      for (int i = 0; i < d->Ng; i++)
        for (int j = 0; j < d->Ny; j++)
          for (int k = 0; k < d->Nz; k++) {
            if (periodic[0]) {
              a[ID(var, i, j, k)] = a[ID(var, d->nx + i, j, k)];
              a[ID(var, d->nx + d->Ng + i, j, k)] = a[ID(var, d->Ng + i, j, k)];
            } else {
              a[ID(var, i, j, k)] = a[ID(var, d->Ng, j, k)];
              a[ID(var, d->nx + d->Ng + i, j, k)] = a[ID(var, d->nx + d->Ng - 1, j, k)];
            }
          }
    if (d->Ny > 1)
      for (int var = 0; var < nvar; var++)
        for (int i = 0; i < d->Nx; i++)
          for (int j = 0; j < d->Ng; j++)
            for (int k = 0; k < d->Nz; k++) {
              if (periodic[1]) {
                a[ID(var, i, j, k)] = a[ID(var, i, d->ny + j, k)];
                a[ID(var, i, d->ny + d->Ng + j, k)] = a[ID(var, i, d->Ng + j, k)];
              } else {
                a[ID(var, i, j, k)] = a[ID(var, i, d->Ng, k)];
                a[ID(var, i, d->ny + d->Ng + j, k)] = a[ID(var, i, d->ny + d->Ng - 1, k)];
              }
            }
    if (d->Nz > 1)
      for (int var = 0; var < nvar; var++)
        for (int i = 0; i < d->Nx; i++)
          for (int j = 0; j < d->Ny; j++)
            for (int k = 0; k < d->Ng; k++) {
              if (periodic[2]) {
                a[ID(var, i, j, k)] = a[ID(var, i, j, d->nz + k)];
                a[ID(var, i, j, d->nz + d->Ng + k)] = a[ID(var, i, j, d->Ng + k)];
              } else {
                a[ID(var, i, j, k)] = a[ID(var, i, j, d->Ng)];
                a[ID(var, i, j, d->nz + d->Ng + k)] = a[ID(var, i, j, d->nz + d->Ng - 1)];
              }
            }
  }
  void apply(double *cons, double *prims = NULL, double *aux = NULL) {
    Data *d(this->data);
    fill(cons, d->Ncons);
    if (prims) fill(prims, d->Nprims);
    if (aux) fill(aux, d->Naux);
  }
};

// Initial primitives read from a file over the whole padded grid.
class FromFile : public InitialFunc {
 public:
  FromFile(Data *d, const std::string &path) : InitialFunc(d) {
    read_doubles(path, d->prims, (size_t)d->Nprims * d->Nx * d->Ny * d->Nz);
  }
};

long g_c2p_calls = 0, g_c2p_cells = 0, g_c2p_fail = 0;

// SRMHD that also counts how many interior cells hybrd1 fails on before the
// reference's own getPrimitiveVars runs (which then takes its smart-guess path).
class CountingSRMHD : public SRMHD {
 public:
  CountingSRMHD(Data *d) : SRMHD(d) {}
  void getPrimitiveVars(double *cons, double *prims, double *aux) {
    Data *d(this->data);
    g_c2p_calls++;
    for (int i = d->is; i < d->ie; i++)
      for (int j = d->js; j < d->je; j++)
        for (int k = d->ks; k < d->ke; k++) {
          Args a;
          a.D = cons[ID(0, i, j, k)];
          a.g = d->gamma;
          a.BS = cons[ID(5, i, j, k)] * cons[ID(1, i, j, k)] + cons[ID(6, i, j, k)] * cons[ID(2, i, j, k)] +
                 cons[ID(7, i, j, k)] * cons[ID(3, i, j, k)];
          a.Bsq = cons[ID(5, i, j, k)] * cons[ID(5, i, j, k)] + cons[ID(6, i, j, k)] * cons[ID(6, i, j, k)] +
                  cons[ID(7, i, j, k)] * cons[ID(7, i, j, k)];
          a.Ssq = cons[ID(1, i, j, k)] * cons[ID(1, i, j, k)] + cons[ID(2, i, j, k)] * cons[ID(2, i, j, k)] +
                  cons[ID(3, i, j, k)] * cons[ID(3, i, j, k)];
          a.tau = cons[ID(4, i, j, k)];
          double sol[2], res[2], wa[19];
          sol[0] = prims[ID(1, i, j, k)] * prims[ID(1, i, j, k)] + prims[ID(2, i, j, k)] * prims[ID(2, i, j, k)] +
                   prims[ID(3, i, j, k)] * prims[ID(3, i, j, k)];
          sol[1] = prims[ID(0, i, j, k)] * aux[ID(0, i, j, k)] / (1 - sol[0]);
          int info = __cminpack_func__(hybrd1)(&SRMHDresidual, &a, 2, sol, res, 1.49011612e-7, wa, 19);
          g_c2p_cells++;
          if (info != 1) g_c2p_fail++;
        }
    SRMHD::getPrimitiveVars(cons, prims, aux);
  }
};

std::set<int> parse_steps(const std::string &s) {
  std::set<int> out;
  size_t p = 0;
  while (p < s.size()) {
    size_t q = s.find(',', p);
    if (q == std::string::npos) q = s.size();
    if (q > p) out.insert(atoi(s.substr(p, q - p).c_str()));
    p = q + 1;
  }
  return out;
}

void dump_state(const std::string &path, Data *d) {
  size_t ntot = (size_t)d->Nx * d->Ny * d->Nz;
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  write_doubles(f, d->cons, ntot * d->Ncons);
  write_doubles(f, d->prims, ntot * d->Nprims);
  write_doubles(f, d->aux, ntot * d->Naux);
  fclose(f);
}

}  // namespace

int main(int argc, char *argv[]) {
  for (int a = 1; a < argc; a++) {
    const char *eq = strchr(argv[a], '=');
    if (!eq) {
      fprintf(stderr, "bad argument %s\n", argv[a]);
      return 2;
    }
    g_args[std::string(argv[a], eq - argv[a])] = std::string(eq + 1);
  }
  std::string mode = arg("mode", "run");
  std::string model_name = arg("model", "srmhd");
  std::string out = arg("out", "ref_out");

  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data data(argi("nx", 64), argi("ny", 64), argi("nz", 0), argd("xmin", -0.5), argd("xmax", 0.5),
            argd("ymin", -1.0), argd("ymax", 1.0), argd("zmin", 0.0), argd("zmax", 1.0), argd("endTime", 1e10),
            &env, argd("cfl", 0.4), argi("Ng", 4), argd("gamma", 4.0 / 3.0), argd("sigma", 0.0),
            argd("cp", 0.1), -1.0e4, 1.0e4, 10, 1000000000);

  Model *model = nullptr;
  if (model_name == "srmhd")
    model = new CountingSRMHD(&data);
  else if (model_name == "srrmhd")
    model = new SRRMHD(&data);
  else
    throw std::invalid_argument("model must be srmhd or srrmhd");

  size_t ntot = (size_t)data.Nx * data.Ny * data.Nz;

  if (mode == "c2p") {
    // getPrimitiveVars on given cons with the given prims/aux as the guess.
    std::vector<double> cons(ntot * data.Ncons), prims(ntot * data.Nprims), aux(ntot * data.Naux);
    read_doubles(arg("cons_in", ""), cons.data(), cons.size());
    read_doubles(arg("prims_in", ""), prims.data(), prims.size());
    read_doubles(arg("aux_in", ""), aux.data(), aux.size());
    model->getPrimitiveVars(cons.data(), prims.data(), aux.data());
    FILE *f = fopen((out + "_c2p.bin").c_str(), "wb");
    write_doubles(f, prims.data(), prims.size());
    write_doubles(f, aux.data(), aux.size());
    fclose(f);
    printf("c2p cells=%ld fail=%ld\n", g_c2p_cells, g_c2p_fail);
    return 0;
  }

  WenoBase *weno = nullptr;
  int order = argi("weno", 3);
  if (order == 3)
    weno = new Weno3(&data);
  else if (order == 5)
    weno = new Weno5(&data);
  else
    throw std::invalid_argument("weno must be 3 or 5");
  FVS fvs(&data, weno, model);

  int px = arg("bcx", "periodic") == "periodic";
  int py = arg("bcy", "outflow") == "periodic";
  int pz = arg("bcz", "periodic") == "periodic";
  Bcs *bcs = nullptr;
  bool is2d = argi("nz", 0) == 0;
  if (px && py && (pz || is2d))
    bcs = new Periodic(&data);
  else if (!px && !py && (!pz || is2d))
    bcs = new Outflow(&data);
  else if (px && !py && (!pz || is2d))
    bcs = new Flow(&data);
  else
    bcs = new AxisBcs(&data, px, py, pz);

  Simulation sim(&data, &env);
  InitialFunc *init = nullptr;
  const std::string init_name = arg("init", "file");
  if (init_name == "kh")
    init = new KHInstabilitySingleFluid(&data, argi("mag", 1));
  else if (init_name == "khrandom")
    init = new KHRandomInstabilitySingleFluid(&data, argi("mag", 1), argi("seed", 10));
  else if (init_name == "ot")
    init = new OTVortexSingleFluid(&data);
  else if (init_name == "briowu")
    init = new BrioWuSingleFluid(&data, argi("dir", 0));
  else if (init_name == "currentsheet")
    init = new CurrentSheetSingleFluid(&data, argi("dir", 0));
  else
    init = new FromFile(&data, arg("prims_in", ""));

  // regime=1: the REGIME model extension (REGIME.h:81) handed to the integrator
  REGIME *ext = argi("regime", 0) ? new REGIME(&data, &fvs) : nullptr;
  TimeIntegrator *ti = nullptr;
  const std::string integ = arg("integrator", "rk2");
  if (integ == "rk2")
    ti = new RK2(&data, model, bcs, &fvs);
  else if (integ == "rksplit")
    ti = new RKSplit(&data, model, bcs, &fvs, ext);
  else if (integ == "rk2b")
    ti = new RK2B(&data, model, bcs, &fvs, ext);
  else if (integ == "rk3")
    ti = new RK3(&data, model, bcs, &fvs, ext);
  else if (integ == "rk4")
    ti = new RK4(&data, model, bcs, &fvs, ext);
  else if (integ == "rk4_10")
    ti = new RK4_10(&data, model, bcs, &fvs, ext);
  else if (integ == "ssp2")
    ti = new SSP2(&data, model, bcs, &fvs);
  else
    throw std::invalid_argument("unknown integrator " + integ);

  sim.set(init, model, ti, bcs, &fvs);

  if (mode == "regime") {
    // REGIME::sourceExtension of the set-up state (REGIME.cc:83-138)
    REGIME *r = ext ? ext : new REGIME(&data, &fvs);
    r->sourceExtension(data.cons, data.prims, data.aux, data.sourceExtension);
    FILE *f = fopen((out + "_regime.bin").c_str(), "wb");
    write_doubles(f, data.sourceExtension, ntot * data.Ncons);
    fclose(f);
    dump_state(out + "_step0.bin", &data);
    return 0;
  }

  if (mode == "frecon") {
    // the calls of test_fvs.cc:56-170: fluxVector then fluxReconstruction
    std::vector<double> frecon(ntot * data.Ncons, 0.0);
    for (int dir = 0; dir < 3; dir++) {
      model->fluxVector(data.cons, data.prims, data.aux, data.f, dir);
      fvs.fluxReconstruction(data.cons, data.prims, data.aux, data.f, frecon.data(), dir);
      FILE *f = fopen((out + "_frecon" + std::to_string(dir) + ".bin").c_str(), "wb");
      write_doubles(f, data.f, ntot * data.Ncons);
      write_doubles(f, frecon.data(), frecon.size());
      fclose(f);
    }
    dump_state(out + "_step0.bin", &data);
    return 0;
  }

  if (mode == "fnet") {
    std::vector<double> fnet(ntot * data.Ncons, 0.0);
    fvs.F(data.cons, data.prims, data.aux, data.f, fnet.data());
    FILE *f = fopen((out + "_fnet.bin").c_str(), "wb");
    write_doubles(f, fnet.data(), fnet.size());
    fclose(f);
    dump_state(out + "_step0.bin", &data);
    return 0;
  }

  std::set<int> dumps = parse_steps(arg("dump", "0"));
  int steps = argi("steps", 10);
  FILE *log = fopen((out + "_log.txt").c_str(), "w");
  if (dumps.count(0)) dump_state(out + "_step0.bin", &data);
  double step_seconds = 0.0;  // time inside updateTime only (no dumps, no logging)
  for (int s = 1; s <= steps; s++) {
    const auto t0 = std::chrono::steady_clock::now();
    sim.updateTime();
    step_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(log, "step %d t %.17g dt %.17g\n", s, data.t, data.dt);
    if (dumps.count(s)) dump_state(out + "_step" + std::to_string(s) + ".bin", &data);
  }
  fprintf(log, "c2p_calls %ld c2p_cells %ld c2p_fail %ld\n", g_c2p_calls, g_c2p_cells, g_c2p_fail);
  fclose(log);
  if (argi("save", 0)) {
    const char *dirs[] = {"Data", "Data/Final", "Data/Final/Conserved", "Data/Final/Primitive",
                          "Data/Final/Auxiliary", "Data/Final/Domain", "Data/Final/Constants"};
    for (const char *dname : dirs) mkdir(dname, 0755);
    SerialSaveData save(&data, &env);
    save.saveAll();
  }
  printf("done steps=%d c2p_fail=%ld step_seconds=%.6f\n", steps, g_c2p_fail, step_seconds);
  return 0;
}

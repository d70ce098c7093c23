// Output interface of the reference (Project/CPU/Include/saveData.h), kept so
// Simulation::set has the reference signature; no writer is provided here.
#ifndef SAVEDATA_H
#define SAVEDATA_H

#include <string>

#include "simData.h"

class SaveData {
 public:
  Data *d;
  int Nouts, Ncount, test;
  SaveData(Data *data, int test = 0) : d(data), Nouts(0), Ncount(0), test(test) {}
  virtual ~SaveData() {}
  virtual void saveAll(bool timeSeries = false) = 0;
  virtual void saveVar(std::string variable, int num = 1) = 0;
};

#endif

// Text output of the reference (Project/CPU/Include/serialSaveData.h,
// Src/serialSaveData.cc:8-360): the same directories, file names, header
// lines and "%.16f" number format, so the reference's own plotting and
// comparison scripts read the files. Missing output directories are created.
#ifndef SERIALSAVEDATA_H
#define SERIALSAVEDATA_H

#include <string>

#include "platformEnv.h"
#include "saveData.h"

class SerialSaveData : public SaveData {
 public:
  PlatformEnv *env;
  SerialSaveData(Data *data, PlatformEnv *env, int test = 0) : SaveData(data, test), env(env) {}
  virtual ~SerialSaveData() {}

  void saveCons();
  void savePrims();
  void saveAux();
  void saveDomain();
  void saveConsts();
  // Cons, prims, aux, domain and constants under Data/Final, or under
  // Data/TimeSeries with the output counter appended (serialSaveData.cc:8-31).
  void saveAll(bool timeSeries = false) override;
  // One named variable over the whole padded grid into
  // Data/TimeSeries/UserDef/<name><n>.dat (serialSaveData.cc:253-360).
  void saveVar(std::string variable, int num = 1) override;

 private:
  std::string dir_, app_;
  std::string root() const { return test ? "../../" : ""; }
};

#endif

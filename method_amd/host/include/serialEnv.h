// Single-process environment (Project/CPU/Include/serialEnv.h:24-27, serialEnv.cc:11-21).
#ifndef SERIAL_ENV_H
#define SERIAL_ENV_H

#include "platformEnv.h"

class SerialEnv : public PlatformEnv {
 public:
  SerialEnv(int *argcP, char **argvP[], int nxRanks, int nyRanks, int nzRanks, int testing = 0);
  virtual ~SerialEnv();
};

#endif

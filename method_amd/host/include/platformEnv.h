// Platform description of the reference API (Project/CPU/Include/platformEnv.h:16-44).
// The engine runs on one GPU, so every run is one rank of a 1x1x1 grid.
#ifndef PLATFORM_ENV_H
#define PLATFORM_ENV_H

class PlatformEnv {
 public:
  int nProc, rank, nxRanks, nyRanks, nzRanks, xRankId, yRankId, zRankId, leftXNeighbourRank, rightXNeighbourRank,
      leftYNeighbourRank, rightYNeighbourRank, leftZNeighbourRank, rightZNeighbourRank, testing;
  PlatformEnv(int testing = 0) : testing(testing) {}
  virtual ~PlatformEnv() {}
};

#endif

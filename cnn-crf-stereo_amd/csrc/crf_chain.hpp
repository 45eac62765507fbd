// Interface between the SlackProp host code and the chain-sweep kernels.
#pragma once

#include "common.hpp"

namespace ccs {

constexpr float kGamma = 0.25f;
constexpr double kGammaD = 0.25;

enum ChainMode { kRowsMinorant = 0, kColsMinorant = 1, kRowsDecode = 2 };

struct ChainArgs {
  const float* cost;   // C (row and column sweeps)
  float* A;            // state: row unaries before a row sweep, column unaries C - A before a column sweep
  const float* A_in;   // where a row sweep reads the unaries (A, or cost for the first sweep)
  const float* w;      // edge weight of the edge (node i, node i+1), indexed by pixel (P1 with w2)
  const float* w2;     // per-edge penalties: P2 of the edge (then w holds P1); nullptr: w L1, w L2
  int wmul = 1;        // elements per pixel in the arrays of w / w2 (2: interleaved x, y)
  float* msg;          // backward messages, [chain][node][D]
  int* labels;         // decode output, indexed by pixel
  float* gap;          // decode: second best - best min-marginal (optional)
  double* lam;         // per chain: minimum chain energy
  double* kept;        // per chain: sum of kept constants
  int* nan_flags;      // row minorant (optional): per chain, 1 if a unary read was NaN, else 0
  long pcs, pns;       // pixel index = chain * pcs + node * pns
  int n_chains, n, D, delta;
  float L1, L2;
};

// Launches one sweep of `mode` over all chains of `a` on `st`.
int launch_chain(ChainMode mode, ChainArgs& a, hipStream_t st);

}  // namespace ccs

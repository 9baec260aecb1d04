// Sigma for full-CI spaces whose strings hold at most one particle (dense pair GEMMs on MFMA).
#pragma once

#include "pqg_internal.hpp"

namespace pqg {

struct TensorPair {
  int d1 = 0, d2 = -1;  // active dimensions coupled (d2 < 0: one-body only)
  DevBuf<double> W;     // (n_d1 n_d2) x (n_d1 n_d2), row-major
  DevBuf<float> W32;
};

struct TensorPath {
  bool enabled = false;
  bool f32_ready = false;
  std::vector<int> dims;  // string sets holding one particle
  std::vector<TensorPair> pairs;
};

}  // namespace pqg

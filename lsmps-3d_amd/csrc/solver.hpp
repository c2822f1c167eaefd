// PPE Krylov solver workspaces (see solver.hip).
#pragma once

#include "common.hpp"

namespace lsmps {

constexpr int kMaxShadow = 16;           // largest IDR(s) shadow dimension
constexpr int kPartialStride = kMaxShadow + 4;
constexpr int kMaxReduceBlocks = 2048;

// Small device-resident state of one Krylov solve; kernels read and write it
// directly so no host round trip happens per iteration.
struct SolverScalars {
  double M[kMaxShadow * kMaxShadow];  // IDR: M(i, k) at M[i * kMaxShadow + k]
  double f[kMaxShadow];
  double c[kMaxShadow];
  double alpha[kMaxShadow];
  double beta;
  double om;
  double rr;        // squared norm of the current residual
  double rz;        // CG: r . z
  double cg_alpha;  // CG step
  double cg_beta;
  int32_t breakdown;
  int32_t pad;
};

class PpeSolver {
 public:
  DeviceArray<double> dinv64;
  DeviceArray<float> dinv32;
  DeviceArray<float> vals32;
  DeviceArray<double> vec64;   // FP64 Krylov vectors
  DeviceArray<float> vec32;    // FP32 Krylov vectors (mixed inner solve)
  DeviceArray<double> outer64; // mixed: FP64 residual
  DeviceArray<double> part_a;
  DeviceArray<double> part_b;
  DeviceArray<SolverScalars> scalars;
  double* host = nullptr;      // pinned: [0] rr, [1] bnorm2, ...

  PpeSolver() = default;
  ~PpeSolver();
  PpeSolver(const PpeSolver&) = delete;
  PpeSolver& operator=(const PpeSolver&) = delete;
  size_t bytes() const {
    return dinv64.bytes() + dinv32.bytes() + vals32.bytes() + vec64.bytes() + vec32.bytes() +
           outer64.bytes() + part_a.bytes() + part_b.bytes() + scalars.bytes();
  }
};

}  // namespace lsmps

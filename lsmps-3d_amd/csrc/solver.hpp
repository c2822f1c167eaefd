// PPE Krylov solver workspaces (see solver.hip).
#pragma once

#include "common.hpp"
#include "schwarz.hpp"

namespace lsmps {

constexpr int kMaxShadow = 16;           // largest IDR(s) shadow dimension
constexpr int kPartialStride = kMaxShadow;  // doubles per block partial record
constexpr int kMaxReduceBlocks = 4096;

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

// Sliced ELLPACK with 64-row chunks (one wavefront per chunk): entry k of the
// chunk's row r sits at chunk_start[c] + k * 64 + r, padded to the chunk's
// widest row with value 0 and the row's own index as column.
constexpr int kSellChunk = 64;
struct SellMatrix {
  size_t n = 0;
  size_t chunks = 0;
  size_t entries = 0;  // padded
  size_t nnz = 0;      // stored CSR nonzeros
  DeviceArray<int32_t> chunk_start;  // padded entry count < 2^31 is checked at build
  DeviceArray<int32_t> chunk_width;
  DeviceArray<int32_t> chunk_scaled;
  DeviceArray<int32_t> cols;
  DeviceArray<double> vals64;
  DeviceArray<float> vals32;
  size_t bytes() const {
    return chunk_start.bytes() + chunk_width.bytes() + cols.bytes() + vals64.bytes() +
           vals32.bytes();
  }
};

class PpeSolver {
 public:
  SellMatrix sell;
  SchwarzPrec schwarz;
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
    return sell.bytes() + schwarz.bytes() + dinv64.bytes() + dinv32.bytes() + vals32.bytes() + vec64.bytes() +
           vec32.bytes() + outer64.bytes() + part_a.bytes() + part_b.bytes() + scalars.bytes();
  }
};

}  // namespace lsmps

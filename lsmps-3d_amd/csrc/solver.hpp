// PPE Krylov solver workspaces (see solver.hip).
#pragma once

#include "common.hpp"
#include "schwarz.hpp"

namespace lsmps {

constexpr int kMaxShadow = 16;           // largest IDR(s) shadow dimension
constexpr int kPartialStride = kMaxShadow;  // doubles per block partial record
constexpr int kMaxReduceBlocks = 2048;  // partial records per reduction (SpMV grid cap)
constexpr int kMaxVectorBlocks = 1024;  // grid cap of the streaming vector kernels

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
// widest row with value 0.
//
// Column indices are stored in 16 bits where the chunk allows it: the chunk's
// columns are covered by at most kSellWindows windows [base_j, base_j + 4096),
// each two aligned 2048-column buckets wide and laid greedily over the buckets
// the chunk touches, and an entry holds (j << 12) | (column - base_j).
// Particles in lattice or cell order need about 7-14 windows (one or two per z
// layer of the 7x7x7 neighbourhood).  A chunk needing more windows, or with
// columns more than 64 buckets from its first row, keeps 32-bit columns
// (nwin = -1) in the same array, two 16-bit slots per entry.  Padding entries
// point at base_0 with value 0.
constexpr int kSellChunk = 64;
constexpr int kSellWindows = 16;
constexpr int kSellWindowBits = 12;
struct SellMatrix {
  size_t n = 0;
  size_t chunks = 0;
  size_t entries = 0;    // padded value slots
  size_t col_units = 0;  // 16-bit column slots (entries + entries of 32-bit chunks)
  size_t nnz = 0;        // stored CSR nonzeros
  DeviceArray<int32_t> chunk_start;  // value slot offset; padded entry count < 2^31 is checked
  DeviceArray<int32_t> chunk_width;
  DeviceArray<int32_t> chunk_scaled;
  DeviceArray<int32_t> col_rows;     // per chunk: width * (1 or 2), units of 64 slots
  DeviceArray<int32_t> col_start;    // scan of col_rows
  DeviceArray<int32_t> nwin;         // windows used, or -1 for 32-bit columns
  DeviceArray<int32_t> windows;      // chunks x kSellWindows bases
  DeviceArray<uint16_t> cols;
  DeviceArray<double> vals64;
  DeviceArray<float> vals32;
  DeviceArray<int32_t> sorted_cols;  // the CSR with each row's columns ascending
  DeviceArray<double> sorted_vals;
  size_t bytes() const {
    return chunk_start.bytes() + chunk_width.bytes() + chunk_scaled.bytes() + col_rows.bytes() +
           col_start.bytes() + nwin.bytes() + windows.bytes() + cols.bytes() + vals64.bytes() +
           vals32.bytes() + sorted_cols.bytes() + sorted_vals.bytes();
  }
};

class PpeSolver {
 public:
  SellMatrix sell;
  SchwarzPrec schwarz;
  DeviceArray<double> dinv64;  // Jacobi: 1 / diagonal
  DeviceArray<double> vec64;   // residual + Krylov basis (FP64)
  DeviceArray<double> part_a;
  DeviceArray<double> part_b;
  DeviceArray<SolverScalars> scalars;
  double* host = nullptr;      // pinned: [0] rr, [1] bnorm2, ...

  PpeSolver() = default;
  ~PpeSolver();
  PpeSolver(const PpeSolver&) = delete;
  PpeSolver& operator=(const PpeSolver&) = delete;
  size_t bytes() const {
    return sell.bytes() + schwarz.bytes() + dinv64.bytes() + vec64.bytes() + part_a.bytes() +
           part_b.bytes() + scalars.bytes();
  }
};

}  // namespace lsmps

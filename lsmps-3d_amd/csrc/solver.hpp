// PPE Krylov solver workspaces (see solver.hip).
#pragma once

#include <string>
#include <vector>

#include "common.hpp"
#include "coarse.hpp"
#include "schwarz.hpp"

namespace lsmps {

constexpr int kMaxShadow = 16;           // largest IDR(s) shadow dimension
constexpr int kPartialStride = kMaxShadow;  // doubles per block partial record
constexpr int kMaxReduceBlocks = 2048;  // partial records per reduction (SpMV grid cap)
constexpr int kMaxVectorBlocks = 2048;  // grid cap of the streaming vector kernels (<= kMaxReduceBlocks)
static_assert(kMaxVectorBlocks <= kMaxReduceBlocks, "vector kernels write one partial record per block");

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
  int32_t done;     // batched solves: this segment has converged and is frozen
};

// Device side of the IDR loop's per-cycle decisions, for cycles replayed from
// a captured graph: the same tests the host loop makes after each cycle
// (non-finite residual, reliable update, restart after a breakdown, tolerance
// and iteration limit), made by one-thread kernels at the end of the cycle.
// `halt` is the skip word of every captured launch; skip_replace and
// skip_reinit gate the true-residual and restart launches of the cycle.
struct IdrControl {
  int32_t halt;
  int32_t skip_replace;
  int32_t skip_reinit;
  int32_t iterations;
  int32_t replacements;
  int32_t max_iterations;
  int32_t s;
  int32_t pad;
  double rr;           // current squared residual norm (true after a replacement)
  double rr_replaced;  // at the last replacement
  double target2;
  double factor2;
};

// PPE matrix for the Krylov SpMVs: a hybrid of a diagonal (DIA) part over
// 256-row groups and a sliced-ELLPACK remainder over 64-row chunks (four chunks
// to a group).  One 256-thread workgroup owns a group: in the DIA part each lane
// computes four consecutive rows, in the remainder one row.
//
// DIA part: the offsets d = column - row that at least kDiaMinRows of the
// group's rows have (particles in lattice or cell order share the ~123 offsets
// of the 7x7x7 neighbourhood) are listed per group, ascending (for lookups).
// Columns are implicit.  The sorted offsets are cut into runs of consecutive
// offsets d, d + 1, ..., d + m - 1 (m <= kMaxRun): the lane owning rows
// i .. i + 3 reads the m + 3 values x[i + d .. i + d + m + 2] as one window of
// 16-byte loads and takes all 4m products from it.  The plan lists the runs by
// length, longest first (class m), then the offsets whose window would read
// partly outside [0, n) for some row of the group (class 0, gathered one value
// at a time); dia_runs holds the count of each class and dia_slot maps a sorted
// offset to its storage slot (plan order, one slot per offset).  Slot k of the
// group's row r sits at dia_start[g] + k * 256 + r, so a lane's four rows are
// one 16-byte quad in FP32.
//
// Remainder: every other entry, in row-sorted order, padded to the chunk's
// widest remainder row rounded up to a multiple of 4 (value 0), stored in
// per-lane quads so that a lane reads 4 entries of its row with one 16-byte load
// (8 bytes for 16-bit columns): entry k of the chunk's row r sits at
// chunk_start[c] + (k / 4) * 256 + r * 4 + k % 4, with the same slot numbering
// for the columns.  Its column indices are stored in 16 bits where
// the chunk allows it: the chunk's columns are covered by at most kSellWindows
// windows [base_j, base_j + 4096), each two aligned 2048-column buckets wide
// and laid greedily over the buckets the chunk touches, and an entry holds
// (j << 12) | (column - base_j).  A chunk needing more windows, or with columns
// more than 64 buckets from its first row, keeps 32-bit columns (nwin = -1) in
// the same array, two 16-bit slots per entry.  Padding points at base_0.
constexpr int kSellChunk = 64;
constexpr int kDiaGroup = 4;                       // chunks per DIA group
constexpr int kGroupRows = kDiaGroup * kSellChunk;  // 256
constexpr int kSellWindows = 16;
constexpr int kSellWindowBits = 12;
constexpr int kMaxDia = 256;
constexpr int kMaxRun = 8;
constexpr int kRunClasses = kMaxRun + 1;  // class 0: single offsets gathered value by value
constexpr int kDiaMinRows = 64;           // of the group's 256 rows
struct SellMatrix {
  size_t n = 0;
  size_t chunks = 0;
  size_t entries = 0;     // padded remainder value slots
  size_t dia_entries = 0; // DIA value slots
  size_t col_units = 0;   // 16-bit column slots (entries + entries of 32-bit chunks)
  size_t nnz = 0;         // stored CSR nonzeros
  size_t rem_nnz = 0;     // of which in the remainder
  size_t groups = 0;
  DeviceArray<int32_t> dia_count;    // diagonals per group (sorted list length)
  DeviceArray<int32_t> dia_offsets;  // groups x kMaxDia, ascending
  DeviceArray<int32_t> dia_slot;     // groups x kMaxDia: storage slot of sorted entry i
  DeviceArray<int32_t> dia_plan;     // groups x kMaxDia: run start offsets by class, then class 0
  DeviceArray<int32_t> dia_runs;     // groups x kRunClasses: runs per class (class 0: offsets)
  DeviceArray<int32_t> dia_slots;    // stored slots * 256 (scan input)
  DeviceArray<int32_t> dia_start;    // scan of dia_slots
  DeviceArray<int32_t> chunk_start;  // remainder value slot offset (< 2^31 is checked)
  DeviceArray<int32_t> chunk_width;  // remainder width
  DeviceArray<int32_t> chunk_scaled;
  DeviceArray<int32_t> col_rows;     // per chunk: width * (1 or 2), units of 64 slots
  DeviceArray<int32_t> col_start;    // scan of col_rows
  DeviceArray<int32_t> nwin;         // windows used, or -1 for 32-bit columns
  DeviceArray<int32_t> windows;      // chunks x kSellWindows bases
  DeviceArray<unsigned long long> rem_counter;
  DeviceArray<uint16_t> cols;
  DeviceArray<double> vals64, dia64;
  DeviceArray<float> vals32, dia32;
  DeviceArray<int32_t> sorted_cols;  // the CSR with each row's columns ascending
  DeviceArray<double> sorted_vals;
  size_t bytes() const {
    return dia_count.bytes() + dia_offsets.bytes() + dia_slot.bytes() + dia_plan.bytes() +
           dia_runs.bytes() + dia_slots.bytes() + dia_start.bytes() +
           chunk_start.bytes() + chunk_width.bytes() + chunk_scaled.bytes() + col_rows.bytes() +
           col_start.bytes() + nwin.bytes() + windows.bytes() + rem_counter.bytes() +
           cols.bytes() + vals64.bytes() + dia64.bytes() + vals32.bytes() + dia32.bytes() +
           sorted_cols.bytes() + sorted_vals.bytes();
  }
};

// Batched solve of independent systems (tanks) stacked in one matrix: segment k
// owns rows [start[k], start[k+1]); row_seg[i] is row i's segment.  Every segment
// runs its own IDR(s) recurrences (own scalars, dot products and convergence
// test) through shared kernel launches.
struct SolverSegments {
  int count = 0;
  const int32_t* start = nullptr;    // device, count + 1
  const int32_t* row_seg = nullptr;  // device, n
};

class PpeSolver {
 public:
  SolverSegments segments;         // set by the caller before ppe_solve (count 0: one system)
  DeviceArray<double> seg_status;  // per segment: rr, breakdown (packed for one copy)
  DeviceArray<int32_t> seg_flags;   // per segment: 1 converged, 2 restart
  double* seg_host = nullptr;      // pinned copy of seg_status
  size_t seg_host_count = 0;
  SellMatrix sell;
  SchwarzPrec schwarz;
  // set by a caller solving A^T after a solve with A: the blocks are taken
  // transposed from that solve's preconditioner instead of built from A^T
  const SchwarzPrec* schwarz_transpose_source = nullptr;
  CoarseSpace coarse;          // aggregation set by the caller (count 0: none)
  DeviceArray<double> dinv64;  // Jacobi: 1 / diagonal
  DeviceArray<double> diag64;  // row diagonal
  DeviceArray<double> offabs;  // row sum of |off-diagonal|
  DeviceArray<double> vec64;   // residual + Krylov basis (FP64)
  DeviceArray<double> part_a;
  DeviceArray<double> part_b;
  DeviceArray<SolverScalars> scalars;
  double* host = nullptr;      // pinned: [0] rr, [1] bnorm2, ...
  // captured IDR cycles (LSMPS_FLAG_GRAPH): two executable graphs of the same
  // cycle, replayed alternately so that each in-flight cycle has its own
  // control record slot; the signature lists every pointer and size baked in
  DeviceArray<IdrControl> control;
  IdrControl* host_control = nullptr;  // pinned, one slot per graph
  hipGraphExec_t cycle_exec[2] = {nullptr, nullptr};
  hipEvent_t cycle_done[2] = {nullptr, nullptr};
  hipStream_t capture_stream = nullptr;
  std::vector<uintptr_t> cycle_signature;
  bool graph_requested = false;  // set per solve from the context's LSMPS_FLAG_GRAPH
  int last_graph_cycles = 0;     // cycles the last solve replayed from the graphs
  std::string graph_error;       // why the last capture failed (empty: none)
  void release_graphs() noexcept;
  PpeSolver() = default;
  ~PpeSolver();
  PpeSolver(const PpeSolver&) = delete;
  PpeSolver& operator=(const PpeSolver&) = delete;
  size_t bytes() const {
    return sell.bytes() + schwarz.bytes() + coarse.bytes() + dinv64.bytes() + diag64.bytes() +
           offabs.bytes() + vec64.bytes() +
           part_a.bytes() + part_b.bytes() + scalars.bytes();
  }
};

}  // namespace lsmps

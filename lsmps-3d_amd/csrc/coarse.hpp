// Coarse-space correction of the PPE preconditioner (no reference counterpart:
// the reference hands the system to AMGX).
//
// Rows are grouped into aggregates (blocks of f x f x f cells of the fluid cell
// grid; Dirichlet rows -- Surface and Splash particles -- belong to none).  With
// P the piecewise-constant prolongation over the aggregates, the Galerkin coarse
// matrix Ac = P^T A P is formed once per solve and inverted explicitly
// (Gauss-Jordan with partial pivoting), and each preconditioner application
// adds P Ac^-1 P^T w to the fine-level (Schwarz + Jacobi) result: an additive
// two-level method.  It removes the smooth error components the local
// preconditioner cannot reach, which is what dominates the Krylov iteration
// count on large tanks.
//
// Determinism: Ac is accumulated in 64-bit fixed point (integer additions are
// associative), and every per-aggregate or per-row sum runs in a fixed order, so
// a solve is bit-reproducible run to run.
//
// Batched tanks (independent systems stacked in one matrix): each tank is
// aggregated over its own cell grid and its aggregates are numbered after the
// previous tank's, so Ac is block diagonal.  Each tank's block is stored
// densely (packed one after the other), inverted by its own workgroup, and
// applied by its own rows only: a tank's coarse correction depends on its rows
// alone.
#pragma once

#include <vector>

#include "common.hpp"
#include "schwarz.hpp"

namespace lsmps {

constexpr int kMaxCoarse = 4096;      // LDS row of the coarse-matrix build (int64)
constexpr int kMaxTankCoarse = 256;   // aggregates per tank of a batched coarse space

// One tank's cell grid: dims and the index of its first cell in the batched
// cell list (cells of a tank are numbered (z, y, x) from `base`).
struct CoarseTankGrid {
  int32_t nx, ny, nz, base;
};
// A tank's coarse grid on the device: fine dims, coarse dims, factor, first
// fine cell and first coarse cell of the tank.
struct TankCoarseGrid {
  int32_t nx, ny, nz, cx, cy, cz, f, base, cc_base;
};

struct CoarseSpace {
  size_t n = 0;                      // fine rows the aggregation was built for
  int count = 0;                     // aggregates (0: no correction)
  int factor = 0;                    // cells per aggregate side actually used
  DeviceArray<int32_t> agg_of_row;   // n: aggregate of a row, -1 outside
  DeviceArray<int32_t> agg_start;    // count + 1
  DeviceArray<int32_t> agg_rows;     // members, in cell order then index order
  // setup scratch (per coarse cell of the grid)
  DeviceArray<int32_t> cell_count, cell_start, cell_flag, cell_id;
  // per solve
  DeviceArray<double> ainv;          // count x count, row-major, Ac^-1
  DeviceArray<double> work;          // count x count, Gauss-Jordan in place
  DeviceArray<double> colk;          // Gauss-Jordan panel columns f (b x count) and 1 / pivot (b)
  DeviceArray<int32_t> perm;         // count: pivot row of step k
  DeviceArray<int32_t> status;       // [0]: singular pivot met
  DeviceArray<double> wfac;          // n: w = v * wfac (1 / pdiag)
  DeviceArray<double> rowabs;        // n: setup scratch (row bounds)
  DeviceArray<double> rc, yc;        // count
  bool ready = false;                // ainv valid for the current solve
  // batched tanks (tanks > 0): ainv and work hold one dense block per tank,
  // block k (c_k = tank_agg[k + 1] - tank_agg[k] rows, row-major) at entry
  // block_offset[k]; status holds one singular-pivot flag per tank
  int tanks = 0;
  size_t block_entries = 0;                // sum of c_k^2
  std::vector<int32_t> tank_agg_host;      // tanks + 2: tank_agg's host copy
  std::vector<TankCoarseGrid> grids_host;  // sources of the uploads (kept alive)
  std::vector<int32_t> tank_cc_host;
  std::vector<int64_t> offsets_host;
  std::vector<int32_t> status_host;
  DeviceArray<TankCoarseGrid> tank_grids;  // tanks
  DeviceArray<int32_t> tank_cc;            // tanks + 1: first coarse cell of each tank
  DeviceArray<int32_t> tank_agg;           // tanks + 2: first aggregate of each tank, members
  DeviceArray<int64_t> block_offset;       // tanks + 1
  DeviceArray<int32_t> agg_tank;           // count: tank of an aggregate
  size_t bytes() const {
    return agg_of_row.bytes() + agg_start.bytes() + agg_rows.bytes() + cell_count.bytes() +
           cell_start.bytes() + cell_flag.bytes() + cell_id.bytes() + ainv.bytes() +
           work.bytes() + colk.bytes() + perm.bytes() + status.bytes() + wfac.bytes() + rowabs.bytes() +
           rc.bytes() + yc.bytes() + tank_grids.bytes() + tank_cc.bytes() + tank_agg.bytes() +
           block_offset.bytes() + agg_tank.bytes();
  }
};

// Aggregation from a fluid cell list: cells (cx, cy, cz) of a dims grid grouped
// as (cx / f, cy / f, cz / f); rows whose label is Inner or NearSurface join
// their group.  Members are listed cell by cell in (z, y, x) order, each cell's
// particles ascending (the cell list's sorted order).  When f gives more than
// kMaxCoarse non-empty aggregates, f grows by one until the count fits
// (C.factor records the value used), so a large tank keeps a coarser space
// instead of none.
void coarse_aggregate(Context& ctx, CoarseSpace& C, size_t n, const int32_t* labels,
                      const int32_t* cell_sorted, const int32_t* cell_begin,
                      const int32_t* cell_end, const int32_t dims[3], int factor);

// Batched tanks: tank k's rows are aggregated as above over its own grid
// (grids[k]; its cells start at grids[k].base in the batched cell list), with
// one factor f for every tank, grown by one while some tank has more than
// kMaxTankCoarse aggregates.  Aggregates are numbered tank by tank.
void coarse_aggregate_tanks(Context& ctx, CoarseSpace& C, size_t n, const int32_t* labels,
                            const int32_t* cell_sorted, const int32_t* cell_begin,
                            const int32_t* cell_end, const std::vector<CoarseTankGrid>& grids,
                            int factor);

// Reserves every per-solve buffer of coarse_setup / coarse_set_weights for the
// current aggregation (the only device allocations of the coarse solve).
void coarse_reserve(CoarseSpace& C);
// Ac = P^T A P and its inverse, launched on stream st (no host wait, no device
// allocation: coarse_reserve runs first): diag and offabs are A's row
// diagonals and off-diagonal magnitude sums (they bound the fixed-point rows).
// Leaves C.ready false when the aggregation is empty; a zero pivot sets
// C.status, after which every coarse product is 0.
void coarse_setup(Context& ctx, CoarseSpace& C, const lsmps_csr& A, const double* diag,
                  const double* offabs, hipStream_t st);
// wfac = 1 / pdiag (the fine preconditioner's diagonal part), on ctx's stream.
void coarse_set_weights(Context& ctx, CoarseSpace& C, const double* pdiag);
// After the setup has completed: whether the inverse is valid (host wait).  For
// batched tanks: whether it is valid for at least one tank (a tank whose block
// met a zero pivot gets no correction).
bool coarse_usable(Context& ctx, CoarseSpace& C);

// inverse = matrix^-1 (n x n, row-major, device memory) by the coarse matrix's
// Gauss-Jordan kernels; raises LSMPS_ERR_NUMERICAL on a zero pivot.
void dense_inverse(Context& ctx, int n, const double* matrix, double* inverse);

// yc = Ac^-1 P^T (wfac * w) (wfac may be skipped: then yc = Ac^-1 P^T w).
template <typename W>
void coarse_solve(Context& ctx, CoarseSpace& C, const W* w, bool scale_by_wfac,
                  const int32_t* skip = nullptr);

// yc = Ac^-1 rc (the product half of coarse_solve, rc already restricted).
void coarse_product(Context& ctx, CoarseSpace& C, const int32_t* skip = nullptr);

// The Schwarz block products of wa (into P.ybuf, as schwarz_products) and the
// restriction rc = P^T (wfac * wr) (wfac skipped unless scale_by_wfac) in one
// launch; P must have blocks and C an aggregation.
template <typename T>
void products_and_restrict(Context& ctx, const SchwarzPrec& P, CoarseSpace& C, const T* wa,
                           const T* wr, bool scale_by_wfac, const int32_t* skip = nullptr);

// v[i] += yc[agg_of_row[i]] on the aggregated rows.
template <typename W>
void coarse_prolong_add(Context& ctx, const CoarseSpace& C, W* v, const int32_t* skip = nullptr);

// schwarz_complete followed by coarse_prolong_add, in one pass over the rows:
// covered rows take their blocks' mean, then aggregated rows add their yc.
template <typename T>
void schwarz_complete_prolong_add(Context& ctx, const SchwarzPrec& P, const CoarseSpace& C, T* v,
                                  const int32_t* skip = nullptr);

}  // namespace lsmps

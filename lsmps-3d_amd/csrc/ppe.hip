// PPE assembly on gfx950.
//
// Replaces assemble_ppe_matrix_and_rhs (src/ppe/ppe_matrix.cu:517-551):
// build_ppe_row_offsets_kernel (:133-146) and assemble_ppe_system_kernel
// (:243-370).  Row layout and values follow the reference exactly: row i
// starts at i + fluid_offsets[i], the diagonal comes first, then one entry per
// fluid neighbor in CSR order (value 0 when the weight vanishes), Dirichlet rows
// (Surface / Splash) are e_i with their off-diagonal columns kept at value 0.
// The pressure Laplacian row (rows 3+4+5 of the pressure Type-A inverse) and
// rows 0..2 of the velocity inverse come from the operator rows.
#include "common.hpp"
#include "device_math.hpp"
#include "stages.hpp"
#include "tune.hpp"

namespace lsmps {
namespace {

struct PpeArgs {
  size_t n;
  const double* x;
  const double* y;
  const double* z;
  const int32_t* surface_type;
  const double* tvx;
  const double* tvy;
  const double* tvz;
  const double* wx;
  const double* wy;
  const double* wz;
  const double* wnx;
  const double* wny;
  const double* wnz;
  const double* twx;
  const double* twy;
  const double* twz;
  const int32_t* f_off;
  const int32_t* f_idx;
  const int32_t* w_off;
  const int32_t* w_idx;
  const double* vel_ops;
  const double* pre_ops;
  double re;
  double density;
  double dt;
  double wall_scale;
  double g[3];
  int32_t kernel;
  int32_t* row_offsets;
  int32_t* cols;
  double* vals;
  double* rhs;
  double* divergence;
  double* wall_term;
};

__global__ void row_offsets_kernel(size_t n, const int32_t* f_off, int32_t* row_offsets) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > n) return;
  row_offsets[i] = static_cast<int32_t>(i) + f_off[i];
}

// L adjacent lanes per row take its neighbours round-robin, so the index loads
// and the CSR writes of a wavefront are contiguous runs and the neighbour
// gathers of a row are in flight together; the per-row sums follow the 8-slot
// tree of SlotSums (the same result for L = 4 and 8).  Blocks are numbered
// XCD by XCD (xcd_slab_block), so each XCD writes one contiguous slab of the
// CSR: 1.93 -> 1.79 ms at 1M particles; the gather-bound stage kernels showed
// no gain from it.
constexpr int kAsmSlots = 8;

// per-slot sums: velocity rhs x (0..8), y (9..17), z (18..26), diagonal (27),
// wall Neumann term (28)
constexpr int kAsmSums = 29;

template <int L>
__global__ void __launch_bounds__(kBlock) assemble_kernel(PpeArgs a) {
  using Sums = SlotSums<L, kAsmSlots>;
  const size_t t = xcd_slab_block() * blockDim.x + threadIdx.x;
  const size_t i = t / L;
  const int sub = static_cast<int>(t % L);
  if (i >= a.n) return;  // whole groups
  const int32_t row_begin = static_cast<int32_t>(i) + a.f_off[i];
  const int32_t nb = a.f_off[i], ne = a.f_off[i + 1];
  const int32_t st = a.surface_type ? a.surface_type[i] : LSMPS_SURFACE_INNER;
  if (st == LSMPS_SURFACE_SURFACE || st == LSMPS_SURFACE_SPLASH) {
    for (int32_t c = nb + sub; c < ne; c += L) {
      const int32_t out = row_begin + 1 + c - nb;
      a.cols[out] = a.f_idx[c];
      a.vals[out] = 0.0;
    }
    if (sub == 0) {
      a.cols[row_begin] = static_cast<int32_t>(i);
      a.vals[row_begin] = 1.0;
      a.rhs[i] = 0.0;
      a.divergence[i] = 0.0;
      if (a.wall_term) a.wall_term[i] = 0.0;
    }
    return;
  }
  double Lap[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Lap[k] = a.pre_ops[static_cast<size_t>(kOpLap + k) * a.n + i];
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  const double ui = a.tvx[i], vi = a.tvy[i], wi = a.tvz[i];
  const double inv_re = 1.0 / a.re;
  const double coef = 2.0 / (a.re * a.re * a.density);
  double acc[Sums::M][kAsmSums];
#pragma unroll
  for (int m = 0; m < Sums::M; ++m)
#pragma unroll
    for (int k = 0; k < kAsmSums; ++k) acc[m][k] = 0.0;
  struct Nb {
    int32_t j;
    double x, y, z, u, v, w;
  };
  Sums::template for_each_batched<kNbBatch>(
      nb, ne, sub,
      [&](int32_t c) {
        const int32_t j = a.f_idx[c];
        return Nb{j, a.x[j], a.y[j], a.z[j], a.tvx[j], a.tvy[j], a.tvz[j]};
      },
      [&](int32_t c, int m, const Nb& q) {
    const int32_t j = q.j;
    const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
    const double w = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    double off = 0.0;
    if (w > 0.0) {
      double b[9];
      type_a_basis(dx, dy, dz, inv_re, b);
      double lb = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) lb += Lap[k] * b[k];
      off = coef * w * lb;
      acc[m][27] -= off;
      const double du = q.u - ui, dv = q.v - vi, dw = q.w - wi;
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const double wb = w * b[k];
        acc[m][k] += wb * du;
        acc[m][9 + k] += wb * dv;
        acc[m][18 + k] += wb * dw;
      }
    }
    const int32_t out = row_begin + 1 + c - nb;
    a.cols[out] = j;
    a.vals[out] = off;
  });
  Sums::for_each(a.w_off[i], a.w_off[i + 1], sub, [&](int32_t c, int m) {
    const int32_t j = a.w_idx[c];
    const double dx = a.wx[j] - xi, dy = a.wy[j] - yi, dz = a.wz[j] - zi;
    const double bw = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (bw <= 0.0) return;
    double b[9];
    type_a_basis(dx, dy, dz, inv_re, b);
    const double w = a.wall_scale * bw;
    const double du = a.twx[j] - ui, dv = a.twy[j] - vi, dw = a.twz[j] - wi;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double wb = w * b[k];
      acc[m][k] += wb * du;
      acc[m][9 + k] += wb * dv;
      acc[m][18 + k] += wb * dw;
    }
    const double nx = a.wnx[j], ny = a.wny[j], nz = a.wnz[j];
    const double ng = a.g[0] * nx + a.g[1] * ny + a.g[2] * nz;
    double q[9];
    wall_neumann_q(dx, dy, dz, nx, ny, nz, inv_re, q);
    double lq = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) lq += Lap[k] * q[k];
    acc[m][28] += 2.0 * a.wall_scale * bw * ng * lq / a.re;
  });
  Sums::reduce(acc);
  if (sub != 0) return;
  // divergence from velocity rows 0..2 (first_coefficient_from_inverse_rhs, ppe_matrix.cu:233-241)
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    d0 += a.vel_ops[static_cast<size_t>(k) * a.n + i] * acc[0][k];
    d1 += a.vel_ops[static_cast<size_t>(9 + k) * a.n + i] * acc[0][9 + k];
    d2 += a.vel_ops[static_cast<size_t>(18 + k) * a.n + i] * acc[0][18 + k];
  }
  const double divergence = (d0 + d1 + d2) / a.re;
  const double diagonal = acc[0][27], wall_rhs = acc[0][28];
  a.cols[row_begin] = static_cast<int32_t>(i);
  a.vals[row_begin] = diagonal;
  a.divergence[i] = divergence;
  a.rhs[i] = divergence / a.dt - wall_rhs;
  if (a.wall_term) a.wall_term[i] = wall_rhs;
}

}  // namespace

void ppe_assemble(Context& ctx, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                  const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                  const lsmps_neighbor_list& wn, const lsmps_fluid_soa& tmp,
                  const lsmps_wall_soa& tmpw, const double* vel_ops, const double* pre_ops,
                  const lsmps_ppe_workspace& ws) {
  // validate_matrix_workspace / validate_rhs_workspace (ppe_matrix.cu:62-106)
  require(ws.matrix.rows >= fluid.count && ws.matrix.cols >= fluid.count &&
              ws.matrix.nnz >= fluid.count + fn.neighbor_count,
          "PPE matrix workspace is smaller than the requested fluid system");
  if (fluid.count == 0) return;
  require(ws.matrix.row_offsets && ws.matrix.col_indices && ws.matrix.values && ws.rhs &&
              ws.pressure && ws.divergence,
          "PPE workspace requires matrix, rhs, pressure, and diagnostic buffers");
  require(fn.particle_count >= fluid.count && fn.offsets,
          "PPE assembly requires a valid fluid-neighbor CSR table");
  require(tmp.vx && tmp.vy && tmp.vz,
          "PPE RHS assembly requires temporary velocity, divergence and rhs buffers");
  require(tmpw.count == 0 || (tmpw.vx && tmpw.vy && tmpw.vz),
          "PPE RHS assembly requires temporary wall velocity buffers");
  row_offsets_kernel<<<blocks_for(fluid.count + 1), kBlock, 0, ctx.stream>>>(
      fluid.count, fn.offsets, ws.matrix.row_offsets);
  LSMPS_LAUNCH_CHECK();
  PpeArgs a;
  a.n = fluid.count;
  a.x = fluid.x;
  a.y = fluid.y;
  a.z = fluid.z;
  a.surface_type = fluid.surface_type;
  a.tvx = tmp.vx;
  a.tvy = tmp.vy;
  a.tvz = tmp.vz;
  a.wx = walls.x;
  a.wy = walls.y;
  a.wz = walls.z;
  a.wnx = walls.normal_x;
  a.wny = walls.normal_y;
  a.wnz = walls.normal_z;
  a.twx = tmpw.vx;
  a.twy = tmpw.vy;
  a.twz = tmpw.vz;
  a.f_off = fn.offsets;
  a.f_idx = fn.indices;
  a.w_off = wn.offsets;
  a.w_idx = wn.indices;
  a.vel_ops = vel_ops;
  a.pre_ops = pre_ops;
  a.re = cfg.support_radius;
  a.density = cfg.density;
  a.dt = cfg.time_step;
  a.wall_scale = cfg.lsmps_wall_weight_scale;
  for (int k = 0; k < 3; ++k) a.g[k] = cfg.gravity[k];
  a.kernel = cfg.weight_kernel;
  a.row_offsets = ws.matrix.row_offsets;
  a.cols = ws.matrix.col_indices;
  a.vals = ws.matrix.values;
  a.rhs = ws.rhs;
  a.divergence = ws.divergence;
  a.wall_term = ws.pressure_laplacian;
  const size_t n = fluid.count;
  launch_tuned(ctx, kTuneAssembly, {4, 8}, 8, n, fn.offsets, [&](int lanes) {
    if (lanes == 4) {
      assemble_kernel<4><<<blocks_for(n * 4), kBlock, 0, ctx.stream>>>(a);
    } else {
      assemble_kernel<8><<<blocks_for(n * 8), kBlock, 0, ctx.stream>>>(a);
    }
  });
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps

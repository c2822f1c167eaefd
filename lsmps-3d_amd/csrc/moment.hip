// LSMPS moment matrices on gfx950.
//
// Replaces build_matrix_cache_kernel + cusolverDn{S,D}potrf/potrsBatched
// (src/moment_matrix/moment_matrix.cu:254-309, 398-476).  One thread owns one
// particle's matrix for its whole life: the lower triangle (45 or 55 entries)
// is accumulated in registers over the particle's CSR rows (the reference
// accumulated n^2 entries with read-modify-writes to global memory), the
// regularisation lambda*tr/n*I is added (moment_matrix.cu:297-306), the matrix is
// factored by Cholesky-Banachiewicz in registers and solved in registers.
//
// Two outputs:
//  * operator rows (the step's path): rows 0, 1, 2 and the row sum 3+4+5 of
//    M^-1, 36 values per particle in SoA layout.  By symmetry of M, row r of
//    M^-1 is the solution of M y = e_r, so four solves replace the n identity
//    solves of the reference.
//  * the full explicit inverse, column-major per particle (MomentMatrixView
//    parity, moment_matrix.cuh:20-34), plus the assembled matrix, trace,
//    regularisation and factorisation info.
//
// Mixed precision: the matrix is always accumulated in FP64 (the pressure is
// sensitive to O(1e-7) perturbations of the operator, which FP32 accumulation
// would introduce).  In mixed mode the factorisation and the four solves run in
// FP32 and one FP64 iterative-refinement step (residual e - M y in FP64, FP32
// correction solve) restores FP64-level accuracy; rows whose FP32 factorisation
// breaks down fall back to the FP64 factorisation in the same thread.
#include <cmath>

#include "common.hpp"
#include "device_math.hpp"
#include "stages.hpp"

namespace lsmps {
namespace {

template <int N>
struct Packed {
  static constexpr int kSize = N * (N + 1) / 2;
  __device__ static constexpr int at(int r, int c) { return r * (r + 1) / 2 + c; }  // c <= r
};

struct MomentArgs {
  size_t n;
  const double* x;
  const double* y;
  const double* z;
  const double* wx;
  const double* wy;
  const double* wz;
  const double* wnx;
  const double* wny;
  const double* wnz;
  const int32_t* f_off;
  const int32_t* f_idx;
  const int32_t* w_off;
  const int32_t* w_idx;
  double re;
  double lambda;
  double wall_scale;
  int32_t kernel;
  int32_t boundary;  // 0 none, 1 wall Dirichlet samples, 2 wall pressure Neumann
  int32_t mixed;
  // outputs
  double* ops;              // operator rows (SoA, kOpRows x n) or null
  double* inverse;          // full inverse (n x N x N) or null
  double* matrices;         // assembled matrices (n x N x N) or null
  double* trace;            // or null
  double* regularization;   // or null
  int32_t* info;            // or null
};

template <int N>
__device__ __forceinline__ void basis_full(double dx, double dy, double dz, double inv_re,
                                           double b[N]) {
  const double sx = dx * inv_re, sy = dy * inv_re, sz = dz * inv_re;
  int o = 0;
  if (N == 10) b[o++] = 1.0;
  b[o++] = sx;
  b[o++] = sy;
  b[o++] = sz;
  b[o++] = sx * sx;
  b[o++] = sy * sy;
  b[o++] = sz * sz;
  b[o++] = sx * sy;
  b[o++] = sy * sz;
  b[o++] = sz * sx;
}

template <int N>
__device__ __forceinline__ void neumann_full(double dx, double dy, double dz, double nx,
                                             double ny, double nz, double inv_re, double q[N]) {
  int o = 0;
  if (N == 10) q[o++] = 0.0;
  q[o++] = nx;
  q[o++] = ny;
  q[o++] = nz;
  q[o++] = 2.0 * dx * nx * inv_re;
  q[o++] = 2.0 * dy * ny * inv_re;
  q[o++] = 2.0 * dz * nz * inv_re;
  q[o++] = (dy * nx + dx * ny) * inv_re;
  q[o++] = (dz * ny + dy * nz) * inv_re;
  q[o++] = (dx * nz + dz * nx) * inv_re;
}

template <int N>
__device__ __forceinline__ void add_outer(double m[Packed<N>::kSize], const double v[N], double w) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const double wr = w * v[r];
#pragma unroll
    for (int c = 0; c <= r; ++c) m[Packed<N>::at(r, c)] += wr * v[c];
  }
}

// In-place Cholesky of the packed lower triangle; returns 0 or the 1-based
// order of the first non-positive pivot.
template <int N, typename T>
__device__ __forceinline__ int cholesky(T l[Packed<N>::kSize]) {
  int info = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    T s = l[Packed<N>::at(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= l[Packed<N>::at(j, k)] * l[Packed<N>::at(j, k)];
    if (!(s > T(0)) && info == 0) info = j + 1;
    const T d = sqrt(s > T(0) ? s : T(1));
    l[Packed<N>::at(j, j)] = d;
    const T inv = T(1) / d;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      T t = l[Packed<N>::at(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= l[Packed<N>::at(i, k)] * l[Packed<N>::at(j, k)];
      l[Packed<N>::at(i, j)] = t * inv;
    }
  }
  return info;
}

// Solves L L^T y = rhs in place.
template <int N, typename T>
__device__ __forceinline__ void chol_solve(const T l[Packed<N>::kSize], T y[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T t = y[i];
#pragma unroll
    for (int k = 0; k < i; ++k) t -= l[Packed<N>::at(i, k)] * y[k];
    y[i] = t / l[Packed<N>::at(i, i)];
  }
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    T t = y[i];
#pragma unroll
    for (int k = i + 1; k < N; ++k) t -= l[Packed<N>::at(k, i)] * y[k];
    y[i] = t / l[Packed<N>::at(i, i)];
  }
}

// Symmetric product y = M v with M given by its packed lower triangle.
template <int N>
__device__ __forceinline__ void sym_matvec(const double m[Packed<N>::kSize], const double v[N],
                                           double y[N]) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < N; ++c) s += (c <= r ? m[Packed<N>::at(r, c)] : m[Packed<N>::at(c, r)]) * v[c];
    y[r] = s;
  }
}

template <int N>
__device__ __forceinline__ void rhs_vector(int which, double e[N]) {
  // selects, not e[which]: `which` may differ between lanes
#pragma unroll
  for (int k = 0; k < N; ++k)
    e[k] = which < 3 ? (k == which ? 1.0 : 0.0) : (k >= 3 && k <= 5 ? 1.0 : 0.0);
}

// kMomLanes adjacent lanes per particle: lane l accumulates neighbours l,
// l + kMomLanes, ... (contiguous index loads per wavefront, a particle's gathers
// in flight together); the moment sums meet in a xor butterfly, which leaves
// the same total in every lane of the group (each pairwise sum is commutative),
// so every lane factors the same matrix and the lanes share the solves.
constexpr int kMomLanes = 4;
constexpr int kMomBatch = 4;  // neighbours per lane loaded together

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = kMomLanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int N, bool MIXED>
__global__ void __launch_bounds__(kBlock) moment_kernel(MomentArgs a) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = t / kMomLanes;
  const int sub = static_cast<int>(t % kMomLanes);
  if (i >= a.n) return;  // whole groups
  constexpr int P = Packed<N>::kSize;
  double m[P];
#pragma unroll
  for (int k = 0; k < P; ++k) m[k] = 0.0;
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  const double inv_re = 1.0 / a.re;
  double v[N];
  // fluid samples: add_sample_to_matrix (moment_matrix.cu:211-229).  A lane
  // takes its neighbours in batches of kMomBatch: the batch's indices, then its
  // positions, are loaded before the first is used (the sums keep the order
  // of a one-at-a-time loop).
  auto fluid_sample = [&](double px, double py, double pz) {
    const double dx = px - xi, dy = py - yi, dz = pz - zi;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 <= 0.0) return;
    const double w = weight_of(sqrt(d2), a.re, a.kernel);
    if (w <= 0.0) return;
    basis_full<N>(dx, dy, dz, inv_re, v);
    add_outer<N>(m, v, w);
  };
  {
    const int32_t c1 = a.f_off[i + 1];
    int32_t c = a.f_off[i] + sub;
    for (; c + (kMomBatch - 1) * kMomLanes < c1; c += kMomBatch * kMomLanes) {
      int32_t j[kMomBatch];
#pragma unroll
      for (int b = 0; b < kMomBatch; ++b) j[b] = a.f_idx[c + b * kMomLanes];
      double px[kMomBatch], py[kMomBatch], pz[kMomBatch];
#pragma unroll
      for (int b = 0; b < kMomBatch; ++b) {
        px[b] = a.x[j[b]];
        py[b] = a.y[j[b]];
        pz[b] = a.z[j[b]];
      }
#pragma unroll
      for (int b = 0; b < kMomBatch; ++b) fluid_sample(px[b], py[b], pz[b]);
    }
    for (; c < c1; c += kMomLanes) {
      const int32_t j = a.f_idx[c];
      fluid_sample(a.x[j], a.y[j], a.z[j]);
    }
  }
  if (a.boundary != 0) {
    for (int32_t c = a.w_off[i] + sub; c < a.w_off[i + 1]; c += kMomLanes) {
      const int32_t j = a.w_idx[c];
      const double dx = a.wx[j] - xi, dy = a.wy[j] - yi, dz = a.wz[j] - zi;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= 0.0) continue;
      const double w = a.wall_scale * weight_of(sqrt(d2), a.re, a.kernel);
      if (w <= 0.0) continue;
      if (a.boundary == 2) {
        // add_wall_neumann_to_matrix (moment_matrix.cu:231-252)
        neumann_full<N>(dx, dy, dz, a.wnx[j], a.wny[j], a.wnz[j], inv_re, v);
      } else {
        basis_full<N>(dx, dy, dz, inv_re, v);
      }
      add_outer<N>(m, v, w);
    }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) m[k] = group_sum(m[k]);
  const bool lead = sub == 0;
  double trace = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) trace += m[Packed<N>::at(r, r)];
  const double reg = a.lambda * (trace > 0.0 ? trace / static_cast<double>(N) : 1.0);
#pragma unroll
  for (int r = 0; r < N; ++r) m[Packed<N>::at(r, r)] += reg;
  if (lead && a.trace) a.trace[i] = trace;
  if (lead && a.regularization) a.regularization[i] = reg;
  if (lead && a.matrices) {
    double* out = a.matrices + i * N * N;
    for (int r = 0; r < N; ++r)
      for (int c = 0; c < N; ++c) out[r + c * N] = c <= r ? m[Packed<N>::at(r, c)] : m[Packed<N>::at(c, r)];
  }

  // FP64 mode factors m in place.  Mixed mode keeps m (FP64) for the refinement
  // residual and factors an FP32 copy; a row whose FP32 factorisation breaks down
  // is factored in FP64 in place instead.
  int info = 0;
  bool fp32_factor = false;
  float lf[MIXED ? P : 1];
  if (MIXED) {
#pragma unroll
    for (int k = 0; k < P; ++k) lf[k] = static_cast<float>(m[k]);
    fp32_factor = cholesky<N, float>(lf) == 0;
  }
  if (!fp32_factor) info = cholesky<N, double>(m);
  if (lead && a.info) a.info[i] = info;
  const double nan = __longlong_as_double(0x7ff8000000000000LL);

  auto solve = [&](double e[N], double y[N]) {
    if (!fp32_factor) {
#pragma unroll
      for (int k = 0; k < N; ++k) y[k] = e[k];
      chol_solve<N, double>(m, y);
    } else {
      float yf[N];
#pragma unroll
      for (int k = 0; k < N; ++k) yf[k] = static_cast<float>(e[k]);
      chol_solve<N, float>(lf, yf);
#pragma unroll
      for (int k = 0; k < N; ++k) y[k] = static_cast<double>(yf[k]);
      double my[N];
      sym_matvec<N>(m, y, my);
#pragma unroll
      for (int k = 0; k < N; ++k) yf[k] = static_cast<float>(e[k] - my[k]);
      chol_solve<N, float>(lf, yf);
#pragma unroll
      for (int k = 0; k < N; ++k) y[k] += static_cast<double>(yf[k]);
    }
    if (info != 0) {
#pragma unroll
      for (int k = 0; k < N; ++k) y[k] = nan;
    }
  };

  static_assert(kMomLanes == 4, "one operator row per lane");
  if (a.ops) {
    // N == 9 here (Type-A kinds only feed the operators); lane `sub` solves row `sub`
    const int which = sub;
    double e[N], y[N];
    rhs_vector<N>(which, e);
    solve(e, y);
#pragma unroll
    for (int k = 0; k < N; ++k) a.ops[static_cast<size_t>(which * 9 + k) * a.n + i] = y[k];
  }
  if (a.inverse) {
    double* out = a.inverse + i * N * N;
    for (int col = sub; col < N; col += kMomLanes) {
      double e[N], y[N];
#pragma unroll
      for (int k = 0; k < N; ++k) e[k] = (k == col) ? 1.0 : 0.0;
      solve(e, y);
#pragma unroll
      for (int k = 0; k < N; ++k) out[k + col * N] = y[k];
    }
  }
}

// Adjoint of the operator rows of a Type-A kind with respect to the fluid
// positions.  Given Rbar (the gradient of a loss with respect to the four
// stored rows R_k = M^-1 e_k), dR_k = -M^-1 dM R_k gives
// Mbar = -sum_k (M^-1 Rbar_k) R_k^T; with the regularisation lambda tr(M)/n I
// the raw moment gradient is S = sym(Mbar) + lambda/n tr(Mbar) I, and each
// sample w(|d|) v(d) v(d)^T contributes w' d/|d| (v.S v) + 2 w J_v^T S v to the
// gradient with respect to its pair vector d = x_j - x_i.  The matrix is
// re-accumulated and factored in FP64 (nothing of the forward factor is kept).
struct MomentAdjArgs {
  MomentArgs m;
  const double* ops_bar;  // kOpRows x n
  double* f;              // pair gradients, 3 x fcap
  size_t fcap;
  double* xbar;  // 3 x n
};

__global__ void __launch_bounds__(kBlock) moment_adjoint_kernel(MomentAdjArgs A) {
  const MomentArgs& a = A.m;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  constexpr int N = 9;
  constexpr int P = Packed<N>::kSize;
  double m[P];
#pragma unroll
  for (int k = 0; k < P; ++k) m[k] = 0.0;
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  const double inv_re = 1.0 / a.re;
  double v[N];
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 <= 0.0) continue;
    const double w = weight_of(sqrt(d2), a.re, a.kernel);
    if (w <= 0.0) continue;
    basis_full<N>(dx, dy, dz, inv_re, v);
    add_outer<N>(m, v, w);
  }
  if (a.boundary != 0) {
    for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
      const int32_t j = a.w_idx[c];
      const double dx = a.wx[j] - xi, dy = a.wy[j] - yi, dz = a.wz[j] - zi;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= 0.0) continue;
      const double w = a.wall_scale * weight_of(sqrt(d2), a.re, a.kernel);
      if (w <= 0.0) continue;
      if (a.boundary == 2) neumann_full<N>(dx, dy, dz, a.wnx[j], a.wny[j], a.wnz[j], inv_re, v);
      else basis_full<N>(dx, dy, dz, inv_re, v);
      add_outer<N>(m, v, w);
    }
  }
  double trace = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) trace += m[Packed<N>::at(r, r)];
  const double reg = a.lambda * (trace > 0.0 ? trace / static_cast<double>(N) : 1.0);
#pragma unroll
  for (int r = 0; r < N; ++r) m[Packed<N>::at(r, r)] += reg;
  if (cholesky<N, double>(m) != 0) return;  // the step's operators were NaN already
  double S[P];
#pragma unroll
  for (int k = 0; k < P; ++k) S[k] = 0.0;
#pragma unroll
  for (int which = 0; which < 4; ++which) {
    double R[N], y[N];
    rhs_vector<N>(which, R);
    chol_solve<N, double>(m, R);
#pragma unroll
    for (int k = 0; k < N; ++k) y[k] = A.ops_bar[static_cast<size_t>(which * 9 + k) * a.n + i];
    chol_solve<N, double>(m, y);
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = 0; c <= r; ++c) S[Packed<N>::at(r, c)] -= 0.5 * (y[r] * R[c] + R[r] * y[c]);
  }
  if (trace > 0.0) {
    double ts = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) ts += S[Packed<N>::at(r, r)];
    const double add = a.lambda / static_cast<double>(N) * ts;
#pragma unroll
    for (int r = 0; r < N; ++r) S[Packed<N>::at(r, r)] += add;
  }
  double sv[N], jt[3];
  const size_t fc = A.fcap;
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 <= 0.0) continue;
    const double r = sqrt(d2);
    const double w = weight_of(r, a.re, a.kernel);
    if (w <= 0.0) continue;
    const double dwr = weight_slope(r, a.re, a.kernel) / r;
    basis_full<N>(dx, dy, dz, inv_re, v);
    sym_matvec<N>(S, v, sv);
    double vsv = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) vsv += v[k] * sv[k];
    type_a_basis_jt(dx, dy, dz, inv_re, sv, jt);
    A.f[c] += dwr * vsv * dx + 2.0 * w * jt[0];
    A.f[fc + c] += dwr * vsv * dy + 2.0 * w * jt[1];
    A.f[2 * fc + c] += dwr * vsv * dz + 2.0 * w * jt[2];
  }
  if (a.boundary != 0) {
    double xl[3] = {0.0, 0.0, 0.0};
    for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
      const int32_t j = a.w_idx[c];
      const double dx = a.wx[j] - xi, dy = a.wy[j] - yi, dz = a.wz[j] - zi;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= 0.0) continue;
      const double r = sqrt(d2);
      const double bw = weight_of(r, a.re, a.kernel);
      const double w = a.wall_scale * bw;
      if (w <= 0.0) continue;
      const double dwr = a.wall_scale * weight_slope(r, a.re, a.kernel) / r;
      if (a.boundary == 2) {
        neumann_full<N>(dx, dy, dz, a.wnx[j], a.wny[j], a.wnz[j], inv_re, v);
        sym_matvec<N>(S, v, sv);
        wall_neumann_q_jt(a.wnx[j], a.wny[j], a.wnz[j], inv_re, sv, jt);
      } else {
        basis_full<N>(dx, dy, dz, inv_re, v);
        sym_matvec<N>(S, v, sv);
        type_a_basis_jt(dx, dy, dz, inv_re, sv, jt);
      }
      double vsv = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) vsv += v[k] * sv[k];
      // d = x_wall - x_i
      xl[0] -= dwr * vsv * dx + 2.0 * w * jt[0];
      xl[1] -= dwr * vsv * dy + 2.0 * w * jt[1];
      xl[2] -= dwr * vsv * dz + 2.0 * w * jt[2];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) A.xbar[k * a.n + i] += xl[k];
  }
}

}  // namespace

// make_matrix_request (moment_matrix.cu:494-538): boundary handling and size per kind.
int moment_size(int kind) { return kind == LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_B ? 10 : 9; }
static int moment_boundary(int kind) {
  switch (kind) {
    case LSMPS_MOMENT_VELOCITY_WALL_DIRICHLET_A: return 1;
    case LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A:
    case LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_B: return 2;
    default: return 0;
  }
}

static MomentArgs moment_args(Context& ctx, int kind, const lsmps_config& cfg,
                              const lsmps_fluid_soa& fluid, const lsmps_wall_soa& walls,
                              const lsmps_neighbor_list& fn, const lsmps_neighbor_list& wn) {
  MomentArgs a{};
  a.n = fluid.count;
  a.x = fluid.x;
  a.y = fluid.y;
  a.z = fluid.z;
  a.wx = walls.x;
  a.wy = walls.y;
  a.wz = walls.z;
  a.wnx = walls.normal_x;
  a.wny = walls.normal_y;
  a.wnz = walls.normal_z;
  a.f_off = fn.offsets;
  a.f_idx = fn.indices;
  a.w_off = wn.offsets;
  a.w_idx = wn.indices;
  a.re = cfg.support_radius;
  a.lambda = cfg.lsmps_regularization;
  a.wall_scale = cfg.lsmps_wall_weight_scale;
  a.kernel = cfg.weight_kernel;
  a.boundary = moment_boundary(kind);
  a.mixed = ctx.precision == LSMPS_PRECISION_MIXED ? 1 : 0;
  return a;
}

void compute_moments(Context& ctx, int kind, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                     const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                     const lsmps_neighbor_list& wn, const MomentOutputs& out) {
  // validate_matrix_request (moment_matrix.cu:102-115)
  require(cfg.support_radius > 0.0, "Moment support radius must be positive");
  require(cfg.lsmps_regularization >= 0.0, "Moment regularization must be non-negative");
  require(cfg.lsmps_wall_weight_scale >= 0.0, "Moment wall weight scale must be non-negative");
  require(fluid.count <= fn.particle_count && fluid.count <= wn.particle_count,
          "Moment matrix neighbor lists are smaller than fluid particle count");
  if (fluid.count == 0) return;
  const int size = moment_size(kind);
  require(out.ops == nullptr || size == 9, "Operator rows exist for Type-A kinds only");
  MomentArgs a = moment_args(ctx, kind, cfg, fluid, walls, fn, wn);
  a.ops = out.ops;
  a.inverse = out.inverse;
  a.matrices = out.matrices;
  a.trace = out.trace;
  a.regularization = out.regularization;
  a.info = out.info;
  const unsigned grid = blocks_for(fluid.count * kMomLanes);
  if (size == 10) {
    if (a.mixed) moment_kernel<10, true><<<grid, kBlock, 0, ctx.stream>>>(a);
    else moment_kernel<10, false><<<grid, kBlock, 0, ctx.stream>>>(a);
  } else {
    if (a.mixed) moment_kernel<9, true><<<grid, kBlock, 0, ctx.stream>>>(a);
    else moment_kernel<9, false><<<grid, kBlock, 0, ctx.stream>>>(a);
  }
  LSMPS_LAUNCH_CHECK();
}

void moment_adjoint(Context& ctx, int kind, const lsmps_config& cfg, const lsmps_fluid_soa& fluid,
                    const lsmps_wall_soa& walls, const lsmps_neighbor_list& fn,
                    const lsmps_neighbor_list& wn, const double* ops_bar, double* pair_grad,
                    size_t pair_stride, double* xbar) {
  require(moment_size(kind) == 9, "Operator-row adjoints exist for Type-A kinds only");
  if (fluid.count == 0) return;
  MomentAdjArgs A;
  A.m = moment_args(ctx, kind, cfg, fluid, walls, fn, wn);
  A.ops_bar = ops_bar;
  A.f = pair_grad;
  A.fcap = pair_stride;
  A.xbar = xbar;
  moment_adjoint_kernel<<<blocks_for(fluid.count), kBlock, 0, ctx.stream>>>(A);
  LSMPS_LAUNCH_CHECK();
}


namespace {
// Operator rows from an explicit inverse (MomentMatrixView::inverse_matrices,
// column-major per particle): rows 0, 1, 2 and the row sum 3 + 4 + 5, in the
// order compute_pressure_laplacian_row adds them (ppe_matrix.cu:203-215).
__global__ void __launch_bounds__(kBlock) view_rows_kernel(size_t n, int size,
                                                           const double* __restrict__ inverse,
                                                           double* __restrict__ ops) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* m = inverse + i * static_cast<size_t>(size) * size;
  for (int k = 0; k < 9; ++k) {
    const double* col = m + static_cast<size_t>(k) * size;
    ops[static_cast<size_t>(k) * n + i] = col[0];
    ops[static_cast<size_t>(9 + k) * n + i] = col[1];
    ops[static_cast<size_t>(18 + k) * n + i] = col[2];
    ops[static_cast<size_t>(kOpLap + k) * n + i] = col[3] + col[4] + col[5];
  }
}
}  // namespace

void operator_rows_from_view(Context& ctx, const lsmps_moment_view& view, size_t n,
                             double* ops) {
  if (n == 0) return;
  view_rows_kernel<<<blocks_for(n), kBlock, 0, ctx.stream>>>(n, view.matrix_size,
                                                              view.inverse_matrices, ops);
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps

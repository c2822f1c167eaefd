// Adjoint (reverse-mode derivative) of one step on gfx950.
//
// No reference counterpart: the reference steps forward only.  Given the
// upstream gradients dL/dv_new (and optionally dL/dx_new) of the stepper's last
// step, the kernels below return dL/dx and dL/dv at the start of the step,
// dL/dp for the pressure the correction consumed and dL/dg for the body force.
// The neighbor lists and surface labels of the step are frozen (they are
// piecewise constant in the positions); every per-pair weight is recomputed
// from positions rather than stored.  The chain, last stage first:
//
//   smoothing (weights at the corrected positions)          smooth_adjoint_kernel
//   anti-penetration (sequence of wall projections)         smooth_adjoint_kernel
//   velocity/position update                                smooth_adjoint_kernel
//   particle shifting                                       shift_adjoint_kernel
//   pressure gradient (operator rows 0..2, pressure kind)   gradient_adjoint_kernel
//   clamp + PPE solve: lambda = A^-T dL/dp                  ppe_solve on A^T
//   PPE assembly (A(x), rhs(x, u*, g))                      assembly_adjoint_kernel
//   provision u* = v + dt (nu lap(v) + g)                   provision_adjoint_kernel
//   moment operators M^-1 rows (both kinds)                 moment_adjoint (moment.hip)
//
// Pair terms are written as a gradient f_c with respect to d_c = x_j - x_i for
// CSR entry c = (i, j) (so x_j gains f_c and x_i loses it); one gather at the
// end adds f at the mirrored entry (the fluid lists are symmetric: the radius
// test is symmetric in the pair), so there are no atomics and the result is
// deterministic.  Terms of stages whose per-row state is small (smoothing,
// pressure, u*, velocity) are gathered directly from the neighbour's row.
#include <algorithm>
#include <cmath>

#include "coarse.hpp"
#include "objects.hpp"

namespace lsmps {
namespace {

constexpr int kWallWords = 8;  // anti-penetration history: 256 wall neighbours per particle

struct AdjArgs {
  size_t n;
  size_t fcap;  // stride of the three pair-gradient components
  const double* x[3];
  const double* v[3];
  const int32_t* st;
  const double* wp[3];
  const double* wn[3];
  const double* wv[3];
  const double* tw[3];  // u* of the walls
  const double* tv[3];  // u* of the fluid
  const double* grad[3];
  const double* ps[3];
  const double* p_raw;
  const double* vel_ops;
  const double* pre_ops;
  const int32_t* f_off;
  const int32_t* f_idx;
  const int32_t* w_off;
  const int32_t* w_idx;
  int32_t* rev;
  double re, inv_re, ws, rho, nu, dt, g[3], clearance, alpha, ps_scale, ps_min, ps_max;
  int32_t kernel;
  int32_t smooth_on;
  const double* vbar_out[3];
  const double* xbar_out[3];
  double* xpp;      // 3n corrected positions (after anti-penetration)
  double* vpp;      // 3n velocities before smoothing
  double* sw;       // n smoothing weight sums (0: the particle is not smoothed)
  double* sm;       // 3n smoothing means
  double* xbar;     // 3n
  double* vbar;     // 3n
  double* ubar;     // 3n dL/du*
  double* gbar;     // 3n per-particle dL/dg
  double* gradbar;  // 3n dL/d(grad p)
  double* psbar;    // 3n
  double* rhsbar;   // 9n
  double* rbar;     // 27n
  double* sbar;     // 3n
  double* opbar_v;  // 36n, operator-row layout
  double* opbar_p;  // 36n
  double* pbar;     // n
  double* psolve;   // n
  const double* lam;
  double* f;  // 3 fcap
  int32_t* error;
};

__device__ __forceinline__ bool splash_at(const int32_t* st, size_t i) {
  return st && st[i] == LSMPS_SURFACE_SPLASH;
}
__device__ __forceinline__ bool dirichlet_at(const int32_t* st, size_t i) {
  if (!st) return false;
  const int32_t s = st[i];
  return s == LSMPS_SURFACE_SURFACE || s == LSMPS_SURFACE_SPLASH;
}
__device__ __forceinline__ double dot9(const double* a, const double* b) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) s += a[k] * b[k];
  return s;
}
__device__ __forceinline__ void load9(const double* ops, int row, size_t n, size_t i,
                                      double out[9]) {
#pragma unroll
  for (int k = 0; k < 9; ++k) out[k] = ops[static_cast<size_t>(row * 9 + k) * n + i];
}

// mirrored CSR position of every fluid entry
// One wavefront per row i, one lane per entry (i, j): each lane searches row j
// for i, so a row's searches run side by side.
__global__ void __launch_bounds__(kBlock) rev_kernel(AdjArgs a) {
  const size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
  const int lane = static_cast<int>(threadIdx.x & 63);
  if (i >= a.n) return;  // whole wavefronts
  for (int32_t c = a.f_off[i] + lane; c < a.f_off[i + 1]; c += 64) {
    const int32_t j = a.f_idx[c];
    int32_t found = -1;
    for (int32_t e = a.f_off[j]; e < a.f_off[j + 1]; ++e) {
      if (a.f_idx[e] == static_cast<int32_t>(i)) {
        found = e;
        break;
      }
    }
    a.rev[c] = found < 0 ? c : found;
    if (found < 0) a.error[0] = 1;
  }
}

// Recomputes the corrected state of the step: update_kernel then antipen_kernel
// (correction.hip) with the same operations, so xpp equals the step's output
// positions bit for bit.
__device__ __forceinline__ void corrected_start(const AdjArgs& a, size_t i, double xp[3],
                                                double nv[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) nv[k] = a.tv[k][i];
  if (!splash_at(a.st, i)) {
    const double s = a.dt / a.rho;
#pragma unroll
    for (int k = 0; k < 3; ++k) nv[k] -= s * a.grad[k][i];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) xp[k] = a.x[k][i] + (0.5 * a.dt * (a.v[k][i] + nv[k]) + a.ps[k][i]);
}

__global__ void __launch_bounds__(kBlock) forward_kernel(AdjArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double p[3], u[3];
  corrected_start(a, i, p, u);
  if (a.clearance > 0.0) {
    for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
      const int32_t j = a.w_idx[c];
      const double nx = a.wn[0][j], ny = a.wn[1][j], nz = a.wn[2][j];
      const double sd = (p[0] - a.wp[0][j]) * nx + (p[1] - a.wp[1][j]) * ny + (p[2] - a.wp[2][j]) * nz;
      if (sd >= a.clearance) continue;
      const double corr = a.clearance - sd;
      p[0] += corr * nx;
      p[1] += corr * ny;
      p[2] += corr * nz;
      const double vn = u[0] * nx + u[1] * ny + u[2] * nz;
      if (vn < 0.0) {
        u[0] -= vn * nx;
        u[1] -= vn * ny;
        u[2] -= vn * nz;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.xpp[k * a.n + i] = p[k];
    a.vpp[k * a.n + i] = u[k];
  }
}

// smooth_kernel's sums: W_i = sum w_ij, m_i = sum w_ij v_j / W_i
__global__ void __launch_bounds__(kBlock) smooth_stats_kernel(AdjArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double ws = 0.0, m[3] = {0.0, 0.0, 0.0};
  if (a.smooth_on && !splash_at(a.st, i)) {
    const double xi = a.xpp[i], yi = a.xpp[a.n + i], zi = a.xpp[2 * a.n + i];
    for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
      const int32_t j = a.f_idx[c];
      if (splash_at(a.st, j)) continue;
      const double dx = a.xpp[j] - xi, dy = a.xpp[a.n + j] - yi, dz = a.xpp[2 * a.n + j] - zi;
      const double w = correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
      if (w <= 0.0) continue;
      ws += w;
#pragma unroll
      for (int k = 0; k < 3; ++k) m[k] += w * a.vpp[k * a.n + j];
    }
  }
  if (!(ws > 0.0)) ws = 0.0;
  a.sw[i] = ws;
#pragma unroll
  for (int k = 0; k < 3; ++k) a.sm[k * a.n + i] = ws > 0.0 ? m[k] / ws : 0.0;
}

// Smoothing, anti-penetration and update adjoints of particle k.
__global__ void __launch_bounds__(kBlock) smooth_adjoint_kernel(AdjArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const size_t n = a.n;
  double vb[3], xb[3], vbk[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    vbk[c] = a.vbar_out[c][k];
    xb[c] = a.xbar_out[c] ? a.xbar_out[c][k] : 0.0;
  }
  const double swk = a.sw[k];
  const bool elig = swk > 0.0;
  const bool splash_k = splash_at(a.st, k);
  const double alpha = a.alpha;
#pragma unroll
  for (int c = 0; c < 3; ++c) vb[c] = elig ? (1.0 - alpha) * vbk[c] : vbk[c];
  const double xk = a.xpp[k], yk = a.xpp[n + k], zk = a.xpp[2 * n + k];
  const double vk[3] = {a.vpp[k], a.vpp[n + k], a.vpp[2 * n + k]};
  const double mk[3] = {a.sm[k], a.sm[n + k], a.sm[2 * n + k]};
  if (a.smooth_on) {
    for (int32_t e = a.f_off[k]; e < a.f_off[k + 1]; ++e) {
      const int32_t j = a.f_idx[e];
      const double d[3] = {a.xpp[j] - xk, a.xpp[n + j] - yk, a.xpp[2 * n + j] - zk};
      const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double w = correction_weight_of(r, a.re, a.kernel);
      if (w <= 0.0) continue;
      const double dwr = correction_weight_slope(r, a.re, a.kernel) / r;
      const bool splash_j = splash_at(a.st, j);
      if (elig && !splash_j) {
        // row k: d out_k / d w_kj = alpha (v_j - m_k) / W_k
        double coeff = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) coeff += (a.vpp[c * n + j] - mk[c]) * vbk[c];
        coeff *= alpha / swk;
#pragma unroll
        for (int c = 0; c < 3; ++c) xb[c] -= coeff * dwr * d[c];
      }
      const double swj = a.sw[j];
      if (swj > 0.0 && !splash_k) {
        // row j reads v_k with weight w_jk / W_j
        const double vbj[3] = {a.vbar_out[0][j], a.vbar_out[1][j], a.vbar_out[2][j]};
        const double s = alpha * w / swj;
        double coeff = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          vb[c] += s * vbj[c];
          coeff += (vk[c] - a.sm[c * n + j]) * vbj[c];
        }
        coeff *= alpha / swj;
#pragma unroll
        for (int c = 0; c < 3; ++c) xb[c] -= coeff * dwr * d[c];
      }
    }
  }
  // anti-penetration: replay the forward sequence, remember which wall
  // neighbours projected the position (sd < clearance) and the velocity
  // (also vn < 0), then apply the projections I - n n^T in reverse order.
  double p[3], u[3];
  corrected_start(a, k, p, u);
  double xpb[3] = {xb[0], xb[1], xb[2]};
  double nvb[3] = {vb[0], vb[1], vb[2]};
  const int32_t wb = a.w_off[k], we = a.w_off[k + 1];
  if (a.clearance > 0.0 && we > wb) {
    if (we - wb > 32 * kWallWords) a.error[0] = 2;
    uint32_t pos_bits[kWallWords], vel_bits[kWallWords];
#pragma unroll
    for (int q = 0; q < kWallWords; ++q) {
      pos_bits[q] = 0u;
      vel_bits[q] = 0u;
      for (int b = 0; b < 32; ++b) {
        const int32_t c = wb + 32 * q + b;
        if (c >= we) break;
        const int32_t j = a.w_idx[c];
        const double nx = a.wn[0][j], ny = a.wn[1][j], nz = a.wn[2][j];
        const double sd =
            (p[0] - a.wp[0][j]) * nx + (p[1] - a.wp[1][j]) * ny + (p[2] - a.wp[2][j]) * nz;
        if (sd >= a.clearance) continue;
        const double corr = a.clearance - sd;
        p[0] += corr * nx;
        p[1] += corr * ny;
        p[2] += corr * nz;
        pos_bits[q] |= 1u << b;
        const double vn = u[0] * nx + u[1] * ny + u[2] * nz;
        if (vn < 0.0) {
          u[0] -= vn * nx;
          u[1] -= vn * ny;
          u[2] -= vn * nz;
          vel_bits[q] |= 1u << b;
        }
      }
    }
#pragma unroll
    for (int q = kWallWords - 1; q >= 0; --q) {
      for (int b = 31; b >= 0; --b) {
        if (!((pos_bits[q] >> b) & 1u)) continue;
        const int32_t j = a.w_idx[wb + 32 * q + b];
        const double nx = a.wn[0][j], ny = a.wn[1][j], nz = a.wn[2][j];
        const double px = xpb[0] * nx + xpb[1] * ny + xpb[2] * nz;
        xpb[0] -= px * nx;
        xpb[1] -= px * ny;
        xpb[2] -= px * nz;
        if ((vel_bits[q] >> b) & 1u) {
          const double pv = nvb[0] * nx + nvb[1] * ny + nvb[2] * nz;
          nvb[0] -= pv * nx;
          nvb[1] -= pv * ny;
          nvb[2] -= pv * nz;
        }
      }
    }
  }
  // update: x' = x + dt/2 (v + nv) + ps, nv = u* - dt/rho grad (non-splash)
  const double s = splash_k ? 0.0 : -a.dt / a.rho;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    nvb[c] += 0.5 * a.dt * xpb[c];
    a.ubar[c * n + k] = nvb[c];
    a.gradbar[c * n + k] = s * nvb[c];
    a.psbar[c * n + k] = xpb[c];
    a.xbar[c * n + k] = xpb[c];
    a.vbar[c * n + k] = 0.5 * a.dt * xpb[c];
  }
}

// shifting_kernel (correction.hip) reversed; initialises f for row i.
__global__ void __launch_bounds__(kBlock) shift_adjoint_kernel(AdjArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const size_t n = a.n, fc = a.fcap;
  const int32_t nb = a.f_off[i], ne = a.f_off[i + 1];
  const int32_t st = a.st ? a.st[i] : LSMPS_SURFACE_INNER;
  if (a.ps_scale <= 0.0 || st == LSMPS_SURFACE_SPLASH) {
    for (int32_t c = nb; c < ne; ++c) {
      a.f[c] = 0.0;
      a.f[fc + c] = 0.0;
      a.f[2 * fc + c] = 0.0;
    }
    return;
  }
  const double m = a.ps_min, scale = a.ps_scale;
  const double xi[3] = {a.x[0][i], a.x[1][i], a.x[2][i]};
  double sh[3] = {0.0, 0.0, 0.0}, na[3] = {0.0, 0.0, 0.0};
  for (int32_t c = nb; c < ne; ++c) {
    const int32_t j = a.f_idx[c];
    const double d[3] = {xi[0] - a.x[0][j], xi[1] - a.x[1][j], xi[2] - a.x[2][j]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r <= 0.0 || r >= m) continue;
    const double push = scale * (m - r) / r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sh[k] += push * d[k];
      na[k] += d[k] / r;
    }
  }
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double d[3] = {xi[0] - a.wp[0][j], xi[1] - a.wp[1][j], xi[2] - a.wp[2][j]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r <= 0.0 || r >= m) continue;
    const double push = scale * (m - r);
#pragma unroll
    for (int k = 0; k < 3; ++k) sh[k] += push * a.wn[k][j];
  }
  double y[3] = {sh[0], sh[1], sh[2]}, nh[3] = {0.0, 0.0, 0.0};
  double nn = 0.0, comp = 0.0;
  bool proj = false;
  if (st == LSMPS_SURFACE_SURFACE) {
    nn = sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
    if (nn > 0.0) {
      proj = true;
#pragma unroll
      for (int k = 0; k < 3; ++k) nh[k] = na[k] / nn;
      comp = sh[0] * nh[0] + sh[1] * nh[1] + sh[2] * nh[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) y[k] = sh[k] - comp * nh[k];
    }
  }
  const double norm = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  double yb[3] = {a.psbar[i], a.psbar[n + i], a.psbar[2 * n + i]};
  if (norm > a.ps_max && norm > 0.0) {
    // y_out = max y / |y|
    const double yu = (yb[0] * y[0] + yb[1] * y[1] + yb[2] * y[2]) / (norm * norm);
#pragma unroll
    for (int k = 0; k < 3; ++k) yb[k] = a.ps_max / norm * (yb[k] - yu * y[k]);
  }
  double shb[3] = {yb[0], yb[1], yb[2]}, nab[3] = {0.0, 0.0, 0.0};
  if (proj) {
    const double ybn = yb[0] * nh[0] + yb[1] * nh[1] + yb[2] * nh[2];
    double nhb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      shb[k] = yb[k] - ybn * nh[k];
      nhb[k] = -ybn * sh[k] - comp * yb[k];
    }
    const double t = nhb[0] * nh[0] + nhb[1] * nh[1] + nhb[2] * nh[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) nab[k] = (nhb[k] - t * nh[k]) / nn;
  }
  for (int32_t c = nb; c < ne; ++c) {
    const int32_t j = a.f_idx[c];
    const double d[3] = {xi[0] - a.x[0][j], xi[1] - a.x[1][j], xi[2] - a.x[2][j]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double gd[3] = {0.0, 0.0, 0.0};
    if (r > 0.0 && r < m) {
      const double ds = d[0] * shb[0] + d[1] * shb[1] + d[2] * shb[2];
      const double dn = d[0] * nab[0] + d[1] * nab[1] + d[2] * nab[2];
      const double r2 = r * r;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        gd[k] = scale * ((m / r - 1.0) * shb[k] - m * ds * d[k] / (r2 * r)) +
                (nab[k] - dn * d[k] / r2) / r;
    }
    // d = x_i - x_j = -(x_j - x_i)
    a.f[c] = -gd[0];
    a.f[fc + c] = -gd[1];
    a.f[2 * fc + c] = -gd[2];
  }
  double xl[3] = {0.0, 0.0, 0.0};
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double d[3] = {xi[0] - a.wp[0][j], xi[1] - a.wp[1][j], xi[2] - a.wp[2][j]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r <= 0.0 || r >= m) continue;
    const double t = -scale * (a.wn[0][j] * shb[0] + a.wn[1][j] * shb[1] + a.wn[2][j] * shb[2]) / r;
#pragma unroll
    for (int k = 0; k < 3; ++k) xl[k] += t * d[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) a.xbar[k * n + i] += xl[k];
}

// gradient_kernel (correction.hip) reversed.
__global__ void __launch_bounds__(kBlock) gradient_adjoint_kernel(AdjArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const size_t n = a.n, fc = a.fcap;
  if (splash_at(a.st, i)) {
    for (int k = 0; k < 27; ++k) a.opbar_p[static_cast<size_t>(k) * n + i] = 0.0;
    for (int k = 0; k < 9; ++k) a.rhsbar[static_cast<size_t>(k) * n + i] = 0.0;
    for (int k = 0; k < 3; ++k) a.gbar[k * n + i] = 0.0;
    return;
  }
  const double inv_re = a.inv_re;
  const double xi = a.x[0][i], yi = a.x[1][i], zi = a.x[2][i];
  const double pi = a.p_raw[i] < 0.0 ? 0.0 : a.p_raw[i];
  const double gb[3] = {a.gradbar[i], a.gradbar[n + i], a.gradbar[2 * n + i]};
  double rhs[9], b[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) rhs[k] = 0.0;
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double dx = a.x[0][j] - xi, dy = a.x[1][j] - yi, dz = a.x[2][j] - zi;
    const double w = correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) continue;
    type_a_basis(dx, dy, dz, inv_re, b);
    const double pj = a.p_raw[j] < 0.0 ? 0.0 : a.p_raw[j];
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[k] += w * b[k] * (pj - pi);
  }
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double dx = a.wp[0][j] - xi, dy = a.wp[1][j] - yi, dz = a.wp[2][j] - zi;
    const double w = a.ws * correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) continue;
    const double nx = a.wn[0][j], ny = a.wn[1][j], nz = a.wn[2][j];
    wall_neumann_q(dx, dy, dz, nx, ny, nz, inv_re, b);
    const double wr = a.re * a.rho * (a.g[0] * nx + a.g[1] * ny + a.g[2] * nz);
#pragma unroll
    for (int k = 0; k < 9; ++k) rhs[k] += w * b[k] * wr;
  }
  double rb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) rb[k] = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double G[9];
    load9(a.pre_ops, r, n, i, G);
    const double s = gb[r] * inv_re;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      a.opbar_p[static_cast<size_t>(r * 9 + k) * n + i] = s * rhs[k];
      rb[k] += s * G[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) a.rhsbar[static_cast<size_t>(k) * n + i] = rb[k];
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double dx = a.x[0][j] - xi, dy = a.x[1][j] - yi, dz = a.x[2][j] - zi;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double w = correction_weight_of(r, a.re, a.kernel);
    if (w <= 0.0) continue;
    const double dwr = correction_weight_slope(r, a.re, a.kernel) / r;
    type_a_basis(dx, dy, dz, inv_re, b);
    const double pj = a.p_raw[j] < 0.0 ? 0.0 : a.p_raw[j];
    const double dp = pj - pi;
    const double ub = dot9(b, rb);
    double jt[3];
    type_a_basis_jt(dx, dy, dz, inv_re, rb, jt);
    a.f[c] += dp * (dwr * ub * dx + w * jt[0]);
    a.f[fc + c] += dp * (dwr * ub * dy + w * jt[1]);
    a.f[2 * fc + c] += dp * (dwr * ub * dz + w * jt[2]);
  }
  double xl[3] = {0.0, 0.0, 0.0}, gl[3] = {0.0, 0.0, 0.0};
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double dx = a.wp[0][j] - xi, dy = a.wp[1][j] - yi, dz = a.wp[2][j] - zi;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double wc = correction_weight_of(r, a.re, a.kernel);
    if (a.ws * wc <= 0.0) continue;
    const double dwr = correction_weight_slope(r, a.re, a.kernel) / r;
    const double nx = a.wn[0][j], ny = a.wn[1][j], nz = a.wn[2][j];
    wall_neumann_q(dx, dy, dz, nx, ny, nz, inv_re, b);
    const double uq = dot9(b, rb);
    const double wr = a.re * a.rho * (a.g[0] * nx + a.g[1] * ny + a.g[2] * nz);
    double jt[3];
    wall_neumann_q_jt(nx, ny, nz, inv_re, rb, jt);
    xl[0] -= a.ws * wr * (dwr * uq * dx + wc * jt[0]);
    xl[1] -= a.ws * wr * (dwr * uq * dy + wc * jt[1]);
    xl[2] -= a.ws * wr * (dwr * uq * dz + wc * jt[2]);
    const double t = a.ws * wc * uq * a.re * a.rho;
    gl[0] += t * nx;
    gl[1] += t * ny;
    gl[2] += t * nz;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.xbar[k * n + i] += xl[k];
    a.gbar[k * n + i] = gl[k];
  }
}

// dL/dp (pressure read by the correction) and the transpose-solve rhs.
__global__ void __launch_bounds__(kBlock) pressure_adjoint_kernel(AdjArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const size_t n = a.n;
  const bool splash_k = splash_at(a.st, k);
  double rk[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) rk[q] = a.rhsbar[static_cast<size_t>(q) * n + k];
  const double xk = a.x[0][k], yk = a.x[1][k], zk = a.x[2][k];
  double pb = 0.0, b[9];
  for (int32_t e = a.f_off[k]; e < a.f_off[k + 1]; ++e) {
    const int32_t j = a.f_idx[e];
    const double dx = a.x[0][j] - xk, dy = a.x[1][j] - yk, dz = a.x[2][j] - zk;
    const double w = correction_weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) continue;
    if (!splash_k) {
      type_a_basis(dx, dy, dz, a.inv_re, b);
      pb -= w * dot9(b, rk);
    }
    if (!splash_at(a.st, j)) {
      type_a_basis(-dx, -dy, -dz, a.inv_re, b);
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) s += b[q] * a.rhsbar[static_cast<size_t>(q) * n + j];
      pb += w * s;
    }
  }
  a.pbar[k] = pb;
  a.psolve[k] = a.p_raw[k] > 0.0 ? pb : 0.0;
}

// A^T on A's pattern (symmetric lists): entry c = (i, j) of row i takes A_ji.
__global__ void __launch_bounds__(kBlock) transpose_kernel(size_t n, const int32_t* f_off,
                                                           const int32_t* f_idx,
                                                           const int32_t* rev, const double* av,
                                                           double* atv) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t nb = f_off[i], ne = f_off[i + 1];
  atv[i + nb] = av[i + nb];
  for (int32_t c = nb; c < ne; ++c) {
    const int32_t j = f_idx[c];
    atv[i + 1 + c] = av[static_cast<size_t>(j) + 1 + rev[c]];
  }
}

// assemble_kernel (ppe.hip) reversed for the row's b_i = lambda_i and
// A_ij-bar = -lambda_i p_j.
__global__ void __launch_bounds__(kBlock) assembly_adjoint_kernel(AdjArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const size_t n = a.n, fc = a.fcap;
  if (dirichlet_at(a.st, i)) {
    for (int k = 0; k < 9; ++k) a.opbar_p[static_cast<size_t>(kOpLap + k) * n + i] = 0.0;
    for (int k = 0; k < 27; ++k) {
      a.opbar_v[static_cast<size_t>(k) * n + i] = 0.0;
      a.rbar[static_cast<size_t>(k) * n + i] = 0.0;
    }
    return;
  }
  const double inv_re = a.inv_re;
  const double lam = a.lam[i];
  const double divb = lam / a.dt, wallb = -lam;
  const double coef = 2.0 / (a.re * a.re * a.rho);
  const double xi = a.x[0][i], yi = a.x[1][i], zi = a.x[2][i];
  const double ui[3] = {a.tv[0][i], a.tv[1][i], a.tv[2][i]};
  const double pi = a.p_raw[i];
  double L[9], b[9];
  load9(a.pre_ops, 3, n, i, L);  // rows 27..35
  double rr[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) rr[k] = 0.0;
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double dx = a.x[0][j] - xi, dy = a.x[1][j] - yi, dz = a.x[2][j] - zi;
    const double w = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (!(w > 0.0)) continue;
    type_a_basis(dx, dy, dz, inv_re, b);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double du = a.tv[q][j] - ui[q];
#pragma unroll
      for (int k = 0; k < 9; ++k) rr[q * 9 + k] += w * b[k] * du;
    }
  }
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double dx = a.wp[0][j] - xi, dy = a.wp[1][j] - yi, dz = a.wp[2][j] - zi;
    const double bw = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (bw <= 0.0) continue;
    type_a_basis(dx, dy, dz, inv_re, b);
    const double w = a.ws * bw;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double du = a.tw[q][j] - ui[q];
#pragma unroll
      for (int k = 0; k < 9; ++k) rr[q * 9 + k] += w * b[k] * du;
    }
  }
  // divergence = (V_0 . rx + V_1 . ry + V_2 . rz) / re
  double rb[27];
  const double sdiv = divb * inv_re;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double V[9];
    load9(a.vel_ops, q, n, i, V);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      a.opbar_v[static_cast<size_t>(q * 9 + k) * n + i] = sdiv * rr[q * 9 + k];
      rb[q * 9 + k] = sdiv * V[k];
      a.rbar[static_cast<size_t>(q * 9 + k) * n + i] = rb[q * 9 + k];
    }
  }
  double Lb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Lb[k] = 0.0;
  double ub[3] = {0.0, 0.0, 0.0};
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double d[3] = {a.x[0][j] - xi, a.x[1][j] - yi, a.x[2][j] - zi};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double w = weight_of(r, a.re, a.kernel);
    if (!(w > 0.0)) continue;
    const double dwr = r > 0.0 ? weight_slope(r, a.re, a.kernel) / r : 0.0;
    type_a_basis(d[0], d[1], d[2], inv_re, b);
    const double lb = dot9(L, b);
    const double offb = -lam * (a.p_raw[j] - pi) * coef;
    double jt[3];
    type_a_basis_jt(d[0], d[1], d[2], inv_re, L, jt);
    double gr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) gr[k] = offb * (dwr * lb * d[k] + w * jt[k]);
#pragma unroll
    for (int k = 0; k < 9; ++k) Lb[k] += offb * w * b[k];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double du = a.tv[q][j] - ui[q];
      const double bq = dot9(b, rb + 9 * q);
      type_a_basis_jt(d[0], d[1], d[2], inv_re, rb + 9 * q, jt);
#pragma unroll
      for (int k = 0; k < 3; ++k) gr[k] += du * (dwr * bq * d[k] + w * jt[k]);
      ub[q] -= w * bq;
    }
    a.f[c] += gr[0];
    a.f[fc + c] += gr[1];
    a.f[2 * fc + c] += gr[2];
  }
  double xl[3] = {0.0, 0.0, 0.0}, gl[3] = {0.0, 0.0, 0.0};
  const double cw = 2.0 * a.ws / a.re;
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double d[3] = {a.wp[0][j] - xi, a.wp[1][j] - yi, a.wp[2][j] - zi};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double bw = weight_of(r, a.re, a.kernel);
    if (bw <= 0.0) continue;
    const double dwr = r > 0.0 ? weight_slope(r, a.re, a.kernel) / r : 0.0;
    const double w = a.ws * bw;
    type_a_basis(d[0], d[1], d[2], inv_re, b);
    double jt[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double du = a.tw[q][j] - ui[q];
      const double bq = dot9(b, rb + 9 * q);
      type_a_basis_jt(d[0], d[1], d[2], inv_re, rb + 9 * q, jt);
#pragma unroll
      for (int k = 0; k < 3; ++k) xl[k] -= du * a.ws * (dwr * bq * d[k] + bw * jt[k]);
      ub[q] -= w * bq;
      gl[q] += a.dt * w * bq;  // u*_wall = v_wall + dt g
    }
    const double nx = a.wn[0][j], ny = a.wn[1][j], nz = a.wn[2][j];
    const double ng = a.g[0] * nx + a.g[1] * ny + a.g[2] * nz;
    double q9[9];
    wall_neumann_q(d[0], d[1], d[2], nx, ny, nz, inv_re, q9);
    const double lq = dot9(L, q9);
    const double t = wallb * cw * bw * lq;
    gl[0] += t * nx;
    gl[1] += t * ny;
    gl[2] += t * nz;
#pragma unroll
    for (int k = 0; k < 9; ++k) Lb[k] += wallb * cw * bw * ng * q9[k];
    wall_neumann_q_jt(nx, ny, nz, inv_re, L, jt);
#pragma unroll
    for (int k = 0; k < 3; ++k) xl[k] -= wallb * cw * ng * (dwr * lq * d[k] + bw * jt[k]);
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) a.opbar_p[static_cast<size_t>(kOpLap + k) * n + i] = Lb[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.xbar[k * n + i] += xl[k];
    a.gbar[k * n + i] += gl[k];
    a.ubar[k * n + i] += ub[k];
  }
}

// u*_k gains sum over rows j that read u*_k: w_jk (b_jk . rbar_j)
__global__ void __launch_bounds__(kBlock) ustar_gather_kernel(AdjArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const size_t n = a.n;
  const double xk = a.x[0][k], yk = a.x[1][k], zk = a.x[2][k];
  double acc[3] = {0.0, 0.0, 0.0}, b[9];
  for (int32_t e = a.f_off[k]; e < a.f_off[k + 1]; ++e) {
    const int32_t j = a.f_idx[e];
    if (dirichlet_at(a.st, j)) continue;
    const double dx = xk - a.x[0][j], dy = yk - a.x[1][j], dz = zk - a.x[2][j];
    const double w = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (!(w > 0.0)) continue;
    type_a_basis(dx, dy, dz, a.inv_re, b);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < 9; ++t) s += b[t] * a.rbar[static_cast<size_t>(q * 9 + t) * n + j];
      acc[q] += w * s;
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) a.ubar[q * n + k] += acc[q];
}

// provision_kernel (provision.hip) reversed.
__global__ void __launch_bounds__(kBlock) provision_adjoint_kernel(AdjArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const size_t n = a.n, fc = a.fcap;
  const double inv_re = a.inv_re;
  const double ub[3] = {a.ubar[i], a.ubar[n + i], a.ubar[2 * n + i]};
  const double nu = splash_at(a.st, i) ? 0.0 : a.nu;
  const double scale = 2.0 / (a.re * a.re) * a.dt * nu;
  const double sb[3] = {scale * ub[0], scale * ub[1], scale * ub[2]};
  const double xi = a.x[0][i], yi = a.x[1][i], zi = a.x[2][i];
  const double vi[3] = {a.v[0][i], a.v[1][i], a.v[2][i]};
  double L[9], b[9], Lb[9], jt[3];
  load9(a.vel_ops, 3, n, i, L);
#pragma unroll
  for (int k = 0; k < 9; ++k) Lb[k] = 0.0;
  double vs[3] = {0.0, 0.0, 0.0};
  for (int32_t c = a.f_off[i]; c < a.f_off[i + 1]; ++c) {
    const int32_t j = a.f_idx[c];
    const double d[3] = {a.x[0][j] - xi, a.x[1][j] - yi, a.x[2][j] - zi};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double w = weight_of(r, a.re, a.kernel);
    if (w <= 0.0) continue;
    const double dwr = r > 0.0 ? weight_slope(r, a.re, a.kernel) / r : 0.0;
    type_a_basis(d[0], d[1], d[2], inv_re, b);
    const double lb = dot9(L, b);
    const double cc = (a.v[0][j] - vi[0]) * sb[0] + (a.v[1][j] - vi[1]) * sb[1] +
                      (a.v[2][j] - vi[2]) * sb[2];
#pragma unroll
    for (int k = 0; k < 9; ++k) Lb[k] += w * cc * b[k];
    type_a_basis_jt(d[0], d[1], d[2], inv_re, L, jt);
    a.f[c] += cc * (dwr * lb * d[0] + w * jt[0]);
    a.f[fc + c] += cc * (dwr * lb * d[1] + w * jt[1]);
    a.f[2 * fc + c] += cc * (dwr * lb * d[2] + w * jt[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) vs[k] -= w * lb * sb[k];
  }
  double xl[3] = {0.0, 0.0, 0.0};
  for (int32_t c = a.w_off[i]; c < a.w_off[i + 1]; ++c) {
    const int32_t j = a.w_idx[c];
    const double d[3] = {a.wp[0][j] - xi, a.wp[1][j] - yi, a.wp[2][j] - zi};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double bw = weight_of(r, a.re, a.kernel);
    const double w = a.ws * bw;
    if (w <= 0.0) continue;
    const double dwr = r > 0.0 ? weight_slope(r, a.re, a.kernel) / r : 0.0;
    type_a_basis(d[0], d[1], d[2], inv_re, b);
    const double lb = dot9(L, b);
    const double cc = (a.wv[0][j] - vi[0]) * sb[0] + (a.wv[1][j] - vi[1]) * sb[1] +
                      (a.wv[2][j] - vi[2]) * sb[2];
#pragma unroll
    for (int k = 0; k < 9; ++k) Lb[k] += w * cc * b[k];
    type_a_basis_jt(d[0], d[1], d[2], inv_re, L, jt);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      xl[k] -= cc * a.ws * (dwr * lb * d[k] + bw * jt[k]);
      vs[k] -= w * lb * sb[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) a.opbar_v[static_cast<size_t>(kOpLap + k) * n + i] = Lb[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.sbar[k * n + i] = sb[k];
    a.xbar[k * n + i] += xl[k];
    a.vbar[k * n + i] += ub[k] + vs[k];
    a.gbar[k * n + i] += a.dt * ub[k];
  }
}

// v_k gains sum over rows j that read v_k: w_jk (L_j . b_jk) sbar_j
__global__ void __launch_bounds__(kBlock) velocity_gather_kernel(AdjArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const size_t n = a.n;
  const double xk = a.x[0][k], yk = a.x[1][k], zk = a.x[2][k];
  double acc[3] = {0.0, 0.0, 0.0}, b[9], L[9];
  for (int32_t e = a.f_off[k]; e < a.f_off[k + 1]; ++e) {
    const int32_t j = a.f_idx[e];
    const double dx = xk - a.x[0][j], dy = yk - a.x[1][j], dz = zk - a.x[2][j];
    const double w = weight_of(sqrt(dx * dx + dy * dy + dz * dz), a.re, a.kernel);
    if (w <= 0.0) continue;
    type_a_basis(dx, dy, dz, a.inv_re, b);
    load9(a.vel_ops, 3, n, j, L);
    const double t = w * dot9(L, b);
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] += t * a.sbar[q * n + j];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) a.vbar[q * n + k] += acc[q];
}

// dL/dx_k = local terms + sum over entries of row k of (f at the mirrored entry - f)
__global__ void __launch_bounds__(kBlock) final_gather_kernel(AdjArgs a, double* ox, double* oy,
                                                              double* oz, double* vx, double* vy,
                                                              double* vz) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const size_t n = a.n, fc = a.fcap;
  double acc[3] = {a.xbar[k], a.xbar[n + k], a.xbar[2 * n + k]};
  for (int32_t c = a.f_off[k]; c < a.f_off[k + 1]; ++c) {
    const int32_t m = a.rev[c];
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] += a.f[q * fc + m] - a.f[q * fc + c];
  }
  if (ox) {
    ox[k] = acc[0];
    oy[k] = acc[1];
    oz[k] = acc[2];
  }
  if (vx) {
    vx[k] = a.vbar[k];
    vy[k] = a.vbar[n + k];
    vz[k] = a.vbar[2 * n + k];
  }
}

// Deterministic sum of the per-particle dL/dg terms: one workgroup, fixed
// per-thread strides and a fixed LDS tree.
constexpr int kSumThreads = 1024;
__global__ void __launch_bounds__(kSumThreads) sum3_kernel(size_t n, const double* v,
                                                           double* out) {
  __shared__ double lds[3][kSumThreads];
  const int t = threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  for (size_t i = t; i < n; i += kSumThreads) {
#pragma unroll
    for (int q = 0; q < 3; ++q) s[q] += v[q * n + i];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) lds[q][t] = s[q];
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < 3; ++q) lds[q][t] += lds[q][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) out[q] = lds[q][0];
  }
}

}  // namespace

void adjoint_step(Context& c, lsmps_stepper& S, const lsmps_fluid_soa& fluid,
                  const lsmps_wall_soa& walls, const lsmps_vector3& vbar_out,
                  const lsmps_vector3* xbar_out, lsmps_adjoint_grads& out) {
  require(S.recorded, "adjoint needs a step taken with LSMPS_STEP_RECORD", LSMPS_ERR_NOT_PREPARED);
  const size_t n = S.recorded_fluid;
  require(fluid.count == n && walls.count == S.recorded_walls,
          "adjoint particle counts differ from the recorded step");
  require(vbar_out.count >= n && (n == 0 || (vbar_out.x && vbar_out.y && vbar_out.z)),
          "adjoint requires dL/dv_new for every fluid particle");
  require(!xbar_out || xbar_out->count >= n, "dL/dx_new is smaller than the fluid count");
  require(walls.count == 0 || (walls.vx && walls.vy && walls.vz && walls.normal_x &&
                               walls.normal_y && walls.normal_z),
          "adjoint requires wall velocities and normals");
  out.transpose_report = lsmps_solve_report{};
  for (int k = 0; k < 3; ++k) out.gravity[k] = 0.0;
  if (n == 0) return;
  const lsmps_config& cfg = S.config;
  auto& W = S.adjoint;
  const size_t fcap = S.fluid_capacity * S.max_neighbors;
  const size_t ncap = std::max<size_t>(S.fluid_capacity, 1);
  W.rev.reserve(std::max<size_t>(fcap, 1));
  W.pair.reserve(3 * std::max<size_t>(fcap, 1));
  W.work.reserve(static_cast<size_t>(kAdjointWork) * ncap);
  W.at_vals.reserve(n + fcap);
  W.flags.reserve(4);
  W.sums.reserve(4);
  double* w = W.work.data();
  auto take = [&](size_t rows) {
    double* p = w;
    w += rows * n;
    return p;
  };
  const size_t nf1 = ncap;
  const double* start = S.start_state.data();
  AdjArgs a{};
  a.n = n;
  a.fcap = fcap;
  for (int k = 0; k < 3; ++k) {
    a.x[k] = start + k * nf1;
    a.v[k] = start + (3 + k) * nf1;
    a.tv[k] = S.tmp_fluid.data() + k * nf1;
    a.tw[k] = S.tmp_wall.data() + k * std::max<size_t>(S.wall_capacity, 1);
    a.grad[k] = S.corr.data() + k * nf1;
    a.ps[k] = S.corr.data() + (3 + k) * nf1;
    a.g[k] = cfg.gravity[k];
  }
  a.st = fluid.surface_type;
  a.wp[0] = walls.x;
  a.wp[1] = walls.y;
  a.wp[2] = walls.z;
  a.wn[0] = walls.normal_x;
  a.wn[1] = walls.normal_y;
  a.wn[2] = walls.normal_z;
  a.wv[0] = walls.vx;
  a.wv[1] = walls.vy;
  a.wv[2] = walls.vz;
  a.p_raw = S.raw_pressure.data();
  a.vel_ops = S.moments.velocity_ops();
  a.pre_ops = S.moments.pressure_ops();
  a.f_off = S.f_off.data();
  a.f_idx = S.f_idx.data();
  a.w_off = S.w_off.data();
  a.w_idx = S.w_idx.data();
  a.rev = W.rev.data();
  a.re = cfg.support_radius;
  a.inv_re = 1.0 / cfg.support_radius;
  a.ws = cfg.lsmps_wall_weight_scale;
  a.rho = cfg.density;
  a.nu = cfg.kinematic_viscosity;
  a.dt = cfg.time_step;
  a.clearance = cfg.wall_clearance_ratio * cfg.particle_spacing;
  const double strength = cfg.velocity_smoothing_strength;
  a.smooth_on = strength > 0.0 ? 1 : 0;
  a.alpha = strength < 0.0 ? 0.0 : (strength > 1.0 ? 1.0 : strength);
  a.ps_scale = cfg.ps_displacement_scale;
  a.ps_min = cfg.ps_min_distance_ratio * cfg.particle_spacing;
  a.ps_max = cfg.ps_max_displacement_ratio * cfg.particle_spacing;
  a.kernel = cfg.weight_kernel;
  a.vbar_out[0] = vbar_out.x;
  a.vbar_out[1] = vbar_out.y;
  a.vbar_out[2] = vbar_out.z;
  for (int k = 0; k < 3; ++k) a.xbar_out[k] = nullptr;
  if (xbar_out && xbar_out->x) {
    a.xbar_out[0] = xbar_out->x;
    a.xbar_out[1] = xbar_out->y;
    a.xbar_out[2] = xbar_out->z;
  }
  a.xpp = take(3);
  a.vpp = take(3);
  a.sw = take(1);
  a.sm = take(3);
  a.xbar = take(3);
  a.vbar = take(3);
  a.ubar = take(3);
  a.gbar = take(3);
  a.gradbar = take(3);
  a.psbar = take(3);
  a.rhsbar = take(9);
  a.rbar = take(27);
  a.sbar = take(3);
  a.opbar_v = take(kOpRows);
  a.opbar_p = take(kOpRows);
  a.pbar = take(1);
  a.psolve = take(1);
  double* lam = take(1);
  a.lam = lam;
  a.f = W.pair.data();
  a.error = W.flags.data();
  fill_i32(c.stream, W.flags.data(), 1, 0);
  const unsigned grid = blocks_for(n);
  const hipStream_t s = c.stream;
  rev_kernel<<<blocks_for(n * 64), kBlock, 0, s>>>(a);
  forward_kernel<<<grid, kBlock, 0, s>>>(a);
  smooth_stats_kernel<<<grid, kBlock, 0, s>>>(a);
  smooth_adjoint_kernel<<<grid, kBlock, 0, s>>>(a);
  shift_adjoint_kernel<<<grid, kBlock, 0, s>>>(a);
  gradient_adjoint_kernel<<<grid, kBlock, 0, s>>>(a);
  pressure_adjoint_kernel<<<grid, kBlock, 0, s>>>(a);
  LSMPS_LAUNCH_CHECK();
  LSMPS_HIP(hipMemcpyAsync(c.host_i32, W.flags.data(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  c.sync();
  const int32_t flag = c.host_i32[0];
  if (flag == 1) raise(LSMPS_ERR_INVALID_ARGUMENT, "adjoint: fluid neighbor lists are not symmetric");
  if (flag == 2) raise(LSMPS_ERR_CAPACITY, "adjoint: more than 256 wall neighbors on a particle");
  // lambda = A^-T dL/dp on the unclamped rows
  transpose_kernel<<<grid, kBlock, 0, s>>>(n, a.f_off, a.f_idx, a.rev, S.ppe_vals.data(),
                                            W.at_vals.data());
  LSMPS_LAUNCH_CHECK();
  lsmps_csr at{n, n, n + fcap, S.ppe_rows.data(), S.ppe_cols.data(), W.at_vals.data()};
  lsmps_solver_options opt = S.solver_options;
  opt.clamp_negative = 0;
  W.solver.segments = S.solver.segments;
  // The coarse space of the forward solve, for A^T: with the same aggregation P
  // the Galerkin operator of A^T is P^T A^T P = Ac^T, so the transpose solve
  // gets the two-level preconditioner of the forward one (the aggregation needs
  // the recorded step's labels and fluid cell lists, which are still current).
  W.solver.coarse.count = 0;
  // and its Schwarz blocks, transposed: (A^T)_b = (A_b)^T
  W.solver.schwarz_transpose_source =
      S.solver.schwarz.n == n && S.solver.schwarz.blocks > 0 ? &S.solver.schwarz : nullptr;
  if (opt.preconditioner == LSMPS_PRECOND_SCHWARZ_COARSE && n > 0) {
    const int f = opt.coarse_cells > 0 ? opt.coarse_cells : 4;
    if (S.segment_count > 0) {
      coarse_aggregate_tanks(c, W.solver.coarse, n, fluid.surface_type, S.fc_sorted.data(),
                             S.fc_begin.data(), S.fc_end.data(), S.coarse_grids, f);
    } else {
      coarse_aggregate(c, W.solver.coarse, n, fluid.surface_type, S.fc_sorted.data(),
                       S.fc_begin.data(), S.fc_end.data(), cfg.cell_dims, f);
    }
  }
  const uint32_t saved_flags = c.flags;
  c.flags |= LSMPS_FLAG_ASYNC;
  bool converged = true;
  try {
    ppe_solve(c, W.solver, at, a.psolve, lam, opt, &out.transpose_report);
    W.solver.schwarz_transpose_source = nullptr;
  } catch (const Error& e) {
    W.solver.schwarz_transpose_source = nullptr;
    c.flags = saved_flags;
    if (e.code != LSMPS_ERR_NOT_CONVERGED) throw;
    converged = false;
  } catch (...) {
    W.solver.schwarz_transpose_source = nullptr;
    c.flags = saved_flags;
    throw;
  }
  c.flags = saved_flags;
  assembly_adjoint_kernel<<<grid, kBlock, 0, s>>>(a);
  ustar_gather_kernel<<<grid, kBlock, 0, s>>>(a);
  provision_adjoint_kernel<<<grid, kBlock, 0, s>>>(a);
  velocity_gather_kernel<<<grid, kBlock, 0, s>>>(a);
  LSMPS_LAUNCH_CHECK();
  // moment operators of both kinds at the step's start positions
  lsmps_fluid_soa f0 = fluid;
  f0.x = const_cast<double*>(a.x[0]);
  f0.y = const_cast<double*>(a.x[1]);
  f0.z = const_cast<double*>(a.x[2]);
  const lsmps_neighbor_list fn = S.fluid_neighbors();
  const lsmps_neighbor_list wn = S.wall_neighbors();
  moment_adjoint(c, LSMPS_MOMENT_VELOCITY_WALL_DIRICHLET_A, cfg, f0, walls, fn, wn, a.opbar_v,
                 a.f, fcap, a.xbar);
  moment_adjoint(c, LSMPS_MOMENT_PRESSURE_WALL_NEUMANN_A, cfg, f0, walls, fn, wn, a.opbar_p,
                 a.f, fcap, a.xbar);
  final_gather_kernel<<<grid, kBlock, 0, s>>>(a, out.position.x, out.position.y, out.position.z,
                                               out.velocity.x, out.velocity.y, out.velocity.z);
  sum3_kernel<<<1, kSumThreads, 0, s>>>(n, a.gbar, W.sums.data());
  LSMPS_LAUNCH_CHECK();
  if (out.pressure) copy_f64(s, out.pressure, a.pbar, n);
  LSMPS_HIP(hipMemcpyAsync(c.host_f64, W.sums.data(), 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  c.sync();
  for (int k = 0; k < 3; ++k) out.gravity[k] = c.host_f64[k];
  if (!converged)
    raise(LSMPS_ERR_NOT_CONVERGED,
          "adjoint transpose PPE solve did not reach the tolerance (relative residual " +
              std::to_string(out.transpose_report.relative_residual) + ")");
}

}  // namespace lsmps

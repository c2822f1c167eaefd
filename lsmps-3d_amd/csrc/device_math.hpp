// Device-side LSMPS building blocks shared by the stage kernels: weight
// kernels, the Type-A basis and the wall Neumann vector.  Each mirrors the
// formula of the reference function cited beside it, in the same operation
// order, so FP64 results track the oracle to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lsmps3d_amd.h"

namespace lsmps {

// particle_weight (surface_detection.cu:26-31) / moment_weight (moment_matrix.cu:145-150) /
// provision_weight (explicit_update.cu:121-126): zero at r >= re.
template <typename T>
__device__ __forceinline__ T weight_of(T distance, T re, int kernel) {
  if (distance >= re) return T(0);
  const T t = T(1) - distance / re;
  return kernel == LSMPS_WEIGHT_QUADRATIC ? t * t : t;
}

// correction_weight (pressure_correction.cu:127-132): also zero at r <= 0.
template <typename T>
__device__ __forceinline__ T correction_weight_of(T distance, T re, int kernel) {
  if (distance >= re || distance <= T(0)) return T(0);
  const T t = T(1) - distance / re;
  return kernel == LSMPS_WEIGHT_QUADRATIC ? t * t : t;
}

// Type-A basis with s = d * (1/re): provision_type_a_basis (explicit_update.cu:128-146),
// ppe_type_a_basis (ppe_matrix.cu:155-173), correction_type_a_basis (pressure_correction.cu:138-156).
template <typename T>
__device__ __forceinline__ void type_a_basis(T dx, T dy, T dz, T inv_re, T b[9]) {
  const T sx = dx * inv_re, sy = dy * inv_re, sz = dz * inv_re;
  b[0] = sx;
  b[1] = sy;
  b[2] = sz;
  b[3] = sx * sx;
  b[4] = sy * sy;
  b[5] = sz * sz;
  b[6] = sx * sy;
  b[7] = sy * sz;
  b[8] = sz * sx;
}

// Wall Neumann vector q of Type-A (wall_neumann_vector, moment_matrix.cu:177-201;
// ppe_wall_neumann_vector, ppe_matrix.cu:175-193).
template <typename T>
__device__ __forceinline__ void wall_neumann_q(T dx, T dy, T dz, T nx, T ny, T nz, T inv_re,
                                               T q[9]) {
  q[0] = nx;
  q[1] = ny;
  q[2] = nz;
  q[3] = T(2) * dx * nx * inv_re;
  q[4] = T(2) * dy * ny * inv_re;
  q[5] = T(2) * dz * nz * inv_re;
  q[6] = (dy * nx + dx * ny) * inv_re;
  q[7] = (dz * ny + dy * nz) * inv_re;
  q[8] = (dx * nz + dz * nx) * inv_re;
}

// ---- derivatives used by the adjoint of the step ----------------------------

// dw/dr of weight_of inside its support (0 for r >= re).  At r = 0 the pair
// vector has no direction and callers drop the dw/dr term.
template <typename T>
__device__ __forceinline__ T weight_slope(T distance, T re, int kernel) {
  if (distance >= re) return T(0);
  return kernel == LSMPS_WEIGHT_QUADRATIC ? T(-2) * (T(1) - distance / re) / re : T(-1) / re;
}

// dw/dr of correction_weight_of (0 outside (0, re)).
template <typename T>
__device__ __forceinline__ T correction_weight_slope(T distance, T re, int kernel) {
  if (distance <= T(0)) return T(0);
  return weight_slope(distance, re, kernel);
}

// J_b^T u for the Type-A basis b(d) (s = d / re): out_c = sum_k u_k db_k/dd_c.
template <typename T>
__device__ __forceinline__ void type_a_basis_jt(T dx, T dy, T dz, T inv_re, const T u[9],
                                                T out[3]) {
  const T sx = dx * inv_re, sy = dy * inv_re, sz = dz * inv_re;
  out[0] = (u[0] + T(2) * sx * u[3] + sy * u[6] + sz * u[8]) * inv_re;
  out[1] = (u[1] + T(2) * sy * u[4] + sx * u[6] + sz * u[7]) * inv_re;
  out[2] = (u[2] + T(2) * sz * u[5] + sy * u[7] + sx * u[8]) * inv_re;
}

// J_q^T u for the wall Neumann vector q(d, n) (linear in d).
template <typename T>
__device__ __forceinline__ void wall_neumann_q_jt(T nx, T ny, T nz, T inv_re, const T u[9],
                                                  T out[3]) {
  out[0] = (T(2) * nx * u[3] + ny * u[6] + nz * u[8]) * inv_re;
  out[1] = (T(2) * ny * u[4] + nx * u[6] + nz * u[7]) * inv_re;
  out[2] = (T(2) * nz * u[5] + ny * u[7] + nx * u[8]) * inv_re;
}

// Operator rows kept per particle and moment kind (SoA, index k*N + i):
//   k = 0..8   row 0 of M^-1 (x-gradient coefficients)
//   k = 9..17  row 1 (y-gradient)
//   k = 18..26 row 2 (z-gradient)
//   k = 27..35 rows 3+4+5 summed (Laplacian coefficients)
// These are the only parts of the inverse the consumers read:
// explicit_update.cu:164-176 (rows 3..5), ppe_matrix.cu:203-215 (3+4+5) and
// :355-361 (rows 0..2), pressure_correction.cu:268-275 (rows 0..2).
constexpr int kOpRows = 36;
constexpr int kOpLap = 27;

// ---- lane-count-invariant neighbour sums ----------------------------------
//
// A particle's neighbour list is dealt round-robin onto kSlots virtual slots
// (entry e of a list goes to slot e % kSlots); each slot sums its entries in
// list order (the fluid list, then the wall list), and the slots meet in an xor
// butterfly of strides kSlots / 2, ..., 2, 1.  A group of L adjacent lanes
// (L divides kSlots and 64) runs one particle: lane `sub` holds the kSlots / L
// slots sub, sub + L, ..., and takes list entries sub, sub + L, ... in turn.
// The butterfly strides of at least L combine slots inside a lane and the
// smaller ones go across lanes, so every L yields the same sums bit for bit;
// L only changes speed, which is what the autotuner picks it by.
constexpr int kNbBatch = 4;  // entries per slot whose gathers for_each_batched issues together
template <int L, int kSlots>
struct SlotSums {
  static_assert(kSlots % L == 0 && 64 % L == 0, "lanes per particle must divide the slots");
  static constexpr int M = kSlots / L;  // slots per lane

  // body(entry, slot) for this lane's entries of [begin, end); slot in [0, M).
  template <typename F>
  __device__ __forceinline__ static void for_each(int32_t begin, int32_t end, int sub, F body) {
    for (int32_t c0 = begin + sub; c0 < end; c0 += L * M) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int32_t c = c0 + m * L;
        if (c < end) body(c, m);
      }
    }
  }

  // for_each with this lane's entries taken in batches of about B (whole
  // rounds of its M slots): load(c) returns what the body needs of entry c (its
  // neighbour index and gathered fields), and the loads of a batch are all
  // issued before its first body runs, so a lane waits for one round of gathers
  // per batch instead of one per entry; body(c, slot, loaded) runs in
  // for_each's order.
  template <int B, typename Load, typename F>
  __device__ __forceinline__ static void for_each_batched(int32_t begin, int32_t end, int sub,
                                                          Load load, F body) {
    using Loaded = decltype(load(int32_t{0}));
    constexpr int Q = (B / M > 0 ? B / M : 1) * M;  // entries per batch
    int32_t c0 = begin + sub;
    for (; c0 + (Q - 1) * L < end; c0 += Q * L) {
      Loaded v[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) v[q] = load(c0 + q * L);
#pragma unroll
      for (int q = 0; q < Q; ++q) body(c0 + q * L, q % M, v[q]);
    }
    for (; c0 < end; c0 += L * M) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int32_t c = c0 + m * L;
        if (c < end) body(c, m, load(c));
      }
    }
  }

  // Butterfly over the slots of v[m][k] (k < W): on return v[0][k] holds the
  // total in every lane of the group.
  template <int W>
  __device__ __forceinline__ static void reduce(double (&v)[M][W]) {
#pragma unroll
    for (int q = M / 2; q > 0; q >>= 1) {  // stride q * L: inside the lane
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if ((m & q) == 0) {
#pragma unroll
          for (int k = 0; k < W; ++k) {
            const double t = v[m][k] + v[m | q][k];
            v[m][k] = t;
            v[m | q][k] = t;
          }
        }
      }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {  // strides below L: across lanes
#pragma unroll
      for (int k = 0; k < W; ++k) v[0][k] += __shfl_xor(v[0][k], off, 64);
    }
  }
};

}  // namespace lsmps

// Device helpers shared by the sweep kernels (sweep.hip: level and persistent
// sweeps; sweep_tile.hip: the 2D tile sweep).  Everything here works in an
// octant's canonical frame: local node indices are reflected along each axis
// where Omega_a < 0, so the upwind face of every cell is node 0 and the
// downwind face node p along each axis.
#pragma once

#include <hip/hip_runtime.h>

#include "plan.hpp"

namespace ou {
namespace {

// Read-only data that every lane of a wave indexes identically is loaded
// through the constant address space, which lets the compiler use scalar
// loads (one SGPR operand per FMA instead of a vector register).
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> as_const(const void *p) {
  return (cptr<T>)(p);
}

// Natural (lexicographic, x fastest) index of canonical local node i in
// octant o.
template <int DIM, int P>
__device__ __forceinline__ int natural_index(int i, int o) {
  constexpr int P1 = P + 1;
  int ix = i % P1, iy = (i / P1) % P1, iz = i / (P1 * P1);
  if (o & 1) ix = P - ix;
  if (o & 2) iy = P - iy;
  if (DIM == 3 && (o & 4)) iz = P - iz;
  return ix + P1 * (iy + P1 * iz);
}

// Canonical local index of face node fi on the face normal to axis A at
// node SIDE (0: upwind, P: downwind).
template <int DIM, int P, int A, int SIDE>
__device__ constexpr int face_dof(int fi) {
  constexpr int P1 = P + 1;
  const int u = fi % P1, v = fi / P1;
  if (DIM == 2) return A == 0 ? SIDE + P1 * u : u + P1 * SIDE;
  return A == 0 ? SIDE + P1 * (u + P1 * v)
                : (A == 1 ? u + P1 * (SIDE + P1 * v) : u + P1 * (v + P1 * SIDE));
}

// Upwind face term for axis A (3D: t indexed u + P1 v over the face):
//   b[face(u,v)] += c (m1 (x) m1 t)[u,v],  c = |Omega_A| x face area,
// 2D: b[face(u)] += c (m1 t)[u].  m1 is the 1D reference mass matrix.
template <typename T, int DIM, int P, int A>
__device__ __forceinline__ void add_upwind(const T *m1, T c, const T *t, T *b) {
  constexpr int P1 = P + 1;
  if constexpr (DIM == 2) {
#pragma unroll
    for (int u = 0; u < P1; ++u) {
      T s = T(0);
#pragma unroll
      for (int q = 0; q < P1; ++q) s += m1[u * P1 + q] * t[q];
      b[face_dof<DIM, P, A, 0>(u)] += c * s;
    }
  } else {
    T t1[P1 * P1];
#pragma unroll
    for (int v = 0; v < P1; ++v)
#pragma unroll
      for (int u = 0; u < P1; ++u) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < P1; ++q) s += m1[u * P1 + q] * t[q + P1 * v];
        t1[u + P1 * v] = s;
      }
#pragma unroll
    for (int v = 0; v < P1; ++v)
#pragma unroll
      for (int u = 0; u < P1; ++u) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < P1; ++q) s += m1[v * P1 + q] * t1[u + P1 * q];
        b[face_dof<DIM, P, A, 0>(u + P1 * v)] += c * s;
      }
  }
}

// Traces are stored [dir][slot][F]: the F face values of a slot are
// contiguous, so the runs of consecutive slots a wave touches are whole
// cache lines.
template <typename T, int F>
__device__ __forceinline__ void load_trace(const T *dir_base, unsigned slot, T *t) {
  const T *p = dir_base + slot * F;
#pragma unroll
  for (int f = 0; f < F; ++f) t[f] = p[f];
}

template <typename T, int DIM, int P, int A>
__device__ __forceinline__ void store_trace(T *dir_base, unsigned slot, const T *psi) {
  constexpr int P1 = P + 1, F = DIM == 2 ? P1 : P1 * P1;
  T *p = dir_base + slot * F;
#pragma unroll
  for (int f = 0; f < F; ++f) p[f] = psi[face_dof<DIM, P, A, P>(f)];
}

// Orders the LDS accesses of one wave (its lanes exchange data through LDS).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Workgroup barrier ordering LDS accesses only.  __syncthreads() is a
// release/acquire fence on all memory, which waits for every outstanding
// global load and store of the wave (the prefetched traces and inverse, the
// downwind trace stores) at each barrier; the waves of a workgroup only
// exchange data through LDS here.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
__device__ __forceinline__ const T *m1_of(const SweepView &v);
template <>
__device__ __forceinline__ const float *m1_of<float>(const SweepView &v) { return v.m1f; }
template <>
__device__ __forceinline__ const double *m1_of<double>(const SweepView &v) { return v.m1d; }

}  // namespace
}  // namespace ou

// Shared device helpers for the FWI kernels (gfx950).
//
// All arithmetic is written in the evaluation order of the reference kernels (src/parallelcalls.cu) and
// the library is compiled with -ffp-contract=off, so every stored value is the IEEE result of the written
// expression.  Stencil sums use S = double for the reference's mixed precision (its coefficients are
// double literals, parallelcalls.cu:57,127) or S = float for the FP32 path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace fwi {

// Arithmetic of the stencils.  float / double: sums in that type in the reference's expression order
// (parallelcalls.cu:57,127).  F32D: float, with the second derivative written in difference form
//   D2 u = sum_k c_k ((u_{+k} - u_0) + (u_{-k} - u_0)) / h^2
// (exact because the weights sum to zero), which removes the cancellation between the O(|u|) terms of
// the reference form and keeps the FP32 error of a 4000-step propagation near 3e-6 instead of 1e-4.
struct F32D {};
template <class S>
struct SumType {
  using type = S;
};
template <>
struct SumType<F32D> {
  using type = float;
};

constexpr int kRim = 4;     // stencil radius; the rim of 4 cells is never updated
constexpr int kLayer = 15;  // rows iz <= 15 are excluded from model updates (include/parallelcalls.h:6)

// 8th-order second derivative from the 9 samples u[-4..4]; pair sums in float, weighted sum in S.
template <class S>
__device__ __forceinline__ float d2s(float m4, float m3, float m2, float m1, float c, float p1, float p2,
                                     float p3, float p4, float h2) {
  if constexpr (std::is_same<S, F32D>::value) {
    const float c1 = (float)(8.0 / 5.0), c2 = (float)(1.0 / 5.0), c3 = (float)(8.0 / 315.0),
                c4 = (float)(1.0 / 560.0);
    const float r = c1 * ((p1 - c) + (m1 - c)) - c2 * ((p2 - c) + (m2 - c)) + c3 * ((p3 - c) + (m3 - c)) -
                    c4 * ((p4 - c) + (m4 - c));
    return r / h2;
  } else {
    const S c0 = (S)(205.0 / 72.0), c1 = (S)(8.0 / 5.0), c2 = (S)(1.0 / 5.0), c3 = (S)(8.0 / 315.0),
            c4 = (S)(1.0 / 560.0);
    S r = -c0 * (S)c + c1 * (S)(p1 + m1) - c2 * (S)(p2 + m2) + c3 * (S)(p3 + m3) - c4 * (S)(p4 + m4);
    return (float)(r / (S)h2);
  }
}

// 8th-order first derivative from u[-4..4] (the centre sample is not used).
template <class S0>
__device__ __forceinline__ float d1s(float m4, float m3, float m2, float m1, float p1, float p2, float p3,
                                     float p4, float h) {
  using S = typename SumType<S0>::type;
  const S c1 = (S)(4.0 / 5.0), c2 = (S)(1.0 / 5.0), c3 = (S)(4.0 / 105.0), c4 = (S)(1.0 / 280.0);
  S r = c1 * (S)(p1 - m1) - c2 * (S)(p2 - m2) + c3 * (S)(p3 - m3) - c4 * (S)(p4 - m4);
  return (float)(r / (S)h);
}

// The same stencils on FP64 samples (the near-surface band): pair sums and the weighted sum in double.
__device__ __forceinline__ double d2d(double m4, double m3, double m2, double m1, double c, double p1, double p2,
                                      double p3, double p4, double h2) {
  const double c0 = 205.0 / 72.0, c1 = 8.0 / 5.0, c2 = 1.0 / 5.0, c3 = 8.0 / 315.0, c4 = 1.0 / 560.0;
  const double r = -c0 * c + c1 * (p1 + m1) - c2 * (p2 + m2) + c3 * (p3 + m3) - c4 * (p4 + m4);
  return r / h2;
}
__device__ __forceinline__ double d1d(double m4, double m3, double m2, double m1, double p1, double p2, double p3,
                                      double p4, double h) {
  const double c1 = 4.0 / 5.0, c2 = 1.0 / 5.0, c3 = 4.0 / 105.0, c4 = 1.0 / 280.0;
  const double r = c1 * (p1 - m1) - c2 * (p2 - m2) + c3 * (p3 - m3) - c4 * (p4 - m4);
  return r / h;
}

// Strided loads of the 9-point stencil around index i.
template <class S>
__device__ __forceinline__ float d2_at(const float* __restrict__ u, long i, long st, float h2) {
  return d2s<S>(u[i - 4 * st], u[i - 3 * st], u[i - 2 * st], u[i - st], u[i], u[i + st], u[i + 2 * st],
                u[i + 3 * st], u[i + 4 * st], h2);
}
template <class S>
__device__ __forceinline__ float d1_at(const float* __restrict__ u, long i, long st, float h) {
  return d1s<S>(u[i - 4 * st], u[i - 3 * st], u[i - 2 * st], u[i - st], u[i + st], u[i + 2 * st],
                u[i + 3 * st], u[i + 4 * st], h);
}

// C-PML coefficient of cell i on an axis of n cells from the per-distance table tab[e], e = distance from
// the edge (0 = outermost).  Both ends of x carry C-PML; along z only the bottom (the top is a free surface).
__device__ __forceinline__ float coef_x(const float* __restrict__ tab, int i, int n, int cpml) {
  if (i < cpml) return tab[i];
  if (i > n - cpml - 1) return tab[n - 1 - i];
  return 0.f;
}
__device__ __forceinline__ float coef_z(const float* __restrict__ tab, int i, int n, int cpml) {
  return (i > n - cpml - 1) ? tab[n - 1 - i] : 0.f;
}
// The same for a wave-uniform row i, read through the scalar cache (constant address space), so that the
// value lands in a scalar register and its wait does not involve the vector memory counter.
__device__ __forceinline__ float coef_z_uniform(const float* __restrict__ tab, int i, int n, int cpml) {
  const __attribute__((address_space(4))) float* t = (const __attribute__((address_space(4))) float*)tab;
  return (i > n - cpml - 1) ? t[n - 1 - i] : 0.f;
}

// Wave-private row exchange through LDS.  Lane l holds the value v of band column l; lanes 0..3 also hold
// the halo value h of column l-4, lanes 4..7 that of column 64+(l-4).  On return u[k] is the value of
// column l-4+k.  buf is this wave's 72-float LDS slice.  LDS operations of one wavefront execute in
// order, so the wavefront-scope fences (which only stop the compiler from moving the accesses) are the
// whole synchronisation; no workgroup barrier is involved.
__device__ __forceinline__ void xrow9(float* buf, int lane, float v, float h, float (&u)[9]) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  buf[4 + lane] = v;
  if (lane < 4) buf[lane] = h;
  else if (lane < 8) buf[64 + lane] = h;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < 9; ++k) u[k] = buf[lane + k];
}

// Band column of the halo value a lane supplies to xrow9: lanes 0..3 -> -4..-1, lanes 4..7 -> 64..67.
__device__ __forceinline__ int halo_col(int lane) { return lane < 4 ? lane - 4 : 60 + lane; }

// xrow9 without lane-dependent branches: every lane writes its value and its halo value, lanes >= 8 (whose halo
// value is their own) to their own slot again.  hidx = xrow9_halo_slot(lane).
__device__ __forceinline__ int xrow9_halo_slot(int lane) { return lane < 4 ? lane : lane < 8 ? 64 + lane : 4 + lane; }
__device__ __forceinline__ void xrow9u(float* buf, int lane, int hidx, float v, float h, float (&u)[9]) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  buf[4 + lane] = v;
  buf[hidx] = h;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < 9; ++k) u[k] = buf[lane + k];
}

// Load/store through a wave-uniform base pointer and a 32-bit element index: the address is
// base + zext(4 * idx), which the compiler emits as one global access with the base in scalar registers
// and a 32-bit vector offset (no per-lane 64-bit address arithmetic).  Per-shot fields are far below
// 2^30 elements.
__device__ __forceinline__ float ldu(const float* base, unsigned idx) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (idx << 2));
}
__device__ __forceinline__ void stu(float* base, unsigned idx, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (idx << 2)) = v;
}
__device__ __forceinline__ double ldud(const double* base, unsigned idx) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + (idx << 3));
}
__device__ __forceinline__ void stud(double* base, unsigned idx, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + (idx << 3)) = v;
}
// Streaming (non-temporal) variants for values touched once per launch (centre-only state, imaging terms,
// gradient slots), so the L2 keeps the rows that neighbouring work items read again (stencil halos, model).
__device__ __forceinline__ float ldn(const float* base, unsigned idx) {
  return __builtin_nontemporal_load(reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (idx << 2)));
}
__device__ __forceinline__ void stn(float* base, unsigned idx, float v) {
  __builtin_nontemporal_store(v, reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (idx << 2)));
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Raw buffer access (bounds-checked by the hardware against the resource's byte range): a load at an offset
// past the range returns 0 and a store there is dropped.  Lane conditions become offsets instead of branches,
// so a loop body has one path and the compiler's memory-counter waits need not assume a shorter one.
// The resource is wave-uniform (scalar registers); word 3 is the gfx9 raw-buffer format.
using brsrc = __amdgpu_buffer_rsrc_t;
constexpr unsigned kOutOfRange = 0x80000000u;  // byte offset past every resource range used here (< 2 GiB)
constexpr int kBufNT = 2;                      // cache policy bit: non-temporal
__device__ __forceinline__ brsrc make_rsrc(const float* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)bytes, 0x00020000);
}
template <int POL = 0>
__device__ __forceinline__ float bld(brsrc r, unsigned byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, POL));
}
template <int POL = 0>
__device__ __forceinline__ void bst(brsrc r, unsigned byte_off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)byte_off, 0, POL);
}
// Wave index within the workgroup as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Map a linear workgroup id onto a logical id so that the workgroups one XCD receives (hardware deals
// them round-robin over 8 XCDs) form one contiguous block of logical work.  Bijective for any n.
__device__ __forceinline__ int xcd_logical_block(int bid, int n) {
  const int nx = 8;
  int q = n / nx, r = n % nx;
  int xcd = bid % nx, k = bid / nx;
  return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

}  // namespace fwi

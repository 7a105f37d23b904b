// Finite-difference operators and small device helpers shared by the wave
// stage kernels (wave_stage_kernels.h) and the split-node fault kernels
// (kernels.hip).  Everything here is header-only device code.
#pragma once
#include "cgfdm_internal.h"
#include <stdint.h>

namespace cgfdm {

// Workgroups are dispatched round-robin over the 8 XCDs (block L runs on
// XCD L % 8).  Give XCD x the contiguous work range [x*per, ...), so blocks
// resident on one XCD at the same time are neighbours in the work order.
__device__ __forceinline__ int xcd_remap(int L, int n) {
  const int per = (n + 7) >> 3;
  const int x = L & 7, s = L >> 3;
  const int full = n - 8 * (per - 1);  // XCDs that take 'per' items; the rest take per-1
  return x < full ? x * per + s : full * per + (x - full) * (per - 1) + s;
}

namespace fd {

// DRP/MacCormack coefficients (macdrp.h:23-40).
constexpr float C1 = 0.29164704779069994f, C2 = 0.66674514216519998f, C3 = 1.2501177132452999f;
constexpr float C4 = 0.33341180882999999f, C5 = 0.041686285540599999f;
constexpr float A0P = -1.250039237746600f, A0M = 1.250039237746600f;
constexpr float A1 = 1.541764761036000f, A2 = -0.333411808829999f, A3 = 0.0416862855405999f;
constexpr float C24_1 = 7.0f / 6.0f, C24_2 = 4.0f / 3.0f, C24_3 = 1.0f / 6.0f;

// One-sided 5-point operator with sign FLAG on values f[0..4] = points
// (-1..3) for FLAG=+1 or (-3..1) for FLAG=-1 (macdrp.h:60-75).
template <int FLAG>
__device__ __forceinline__ float drp5(float f0, float f1, float f2, float f3, float f4) {
  if (FLAG > 0) return -C1 * f0 - C2 * f1 + C3 * f2 - C4 * f3 + C5 * f4;
  return -C5 * f0 + C4 * f1 - C3 * f2 + C2 * f3 + C1 * f4;
}
__device__ __forceinline__ float drpF(const float* f, long s) { return -C1 * f[-s] - C2 * f[0] + C3 * f[s] - C4 * f[2 * s] + C5 * f[3 * s]; }
__device__ __forceinline__ float drpB(const float* f, long s) { return -C5 * f[-3 * s] + C4 * f[-2 * s] - C3 * f[-s] + C2 * f[0] + C1 * f[s]; }
__device__ __forceinline__ float drp(const float* f, long s, int flag) { return flag > 0 ? drpF(f, s) : drpB(f, s); }
__device__ __forceinline__ float d22(const float* f, long s, int flag) { return flag > 0 ? f[s] - f[0] : f[0] - f[-s]; }
__device__ __forceinline__ float d24F(const float* f, long s) { return -C24_1 * f[0] + C24_2 * f[s] - C24_3 * f[2 * s]; }
__device__ __forceinline__ float d24B(const float* f, long s) { return C24_3 * f[-2 * s] - C24_2 * f[-s] + C24_1 * f[0]; }
__device__ __forceinline__ float d24(const float* f, long s, int flag) { return flag > 0 ? d24F(f, s) : d24B(f, s); }

// Directional RHS: momentum and Hooke's law for derivatives d (Vx Vy Vz Txx
// Tyy Tzz Txy Txz Tyz) along one axis with metric row e (wave.cu:133-141).
__device__ __forceinline__ void axis_rhs(const float* d, float e0, float e1, float e2, float rrho, float lam, float mu,
                                         float* o) {
  float l2m = lam + 2.0f * mu;
  o[0] = (d[3] * e0 + d[6] * e1 + d[7] * e2) * rrho;
  o[1] = (d[6] * e0 + d[4] * e1 + d[8] * e2) * rrho;
  o[2] = (d[7] * e0 + d[8] * e1 + d[5] * e2) * rrho;
  o[3] = d[0] * e0 * l2m + d[1] * e1 * lam + d[2] * e2 * lam;
  o[4] = d[0] * e0 * lam + d[1] * e1 * l2m + d[2] * e2 * lam;
  o[5] = d[0] * e0 * lam + d[1] * e1 * lam + d[2] * e2 * l2m;
  o[6] = (d[0] * e1 + d[1] * e0) * mu;
  o[7] = (d[0] * e2 + d[2] * e0) * mu;
  o[8] = (d[1] * e2 + d[2] * e1) * mu;
}

// Component c of axis_rhs, for callers that consume one component at a time
// (c is a constant after unrolling).
__device__ __forceinline__ float axis_rhs_c(int c, const float* d, float e0, float e1, float e2, float rrho, float lam,
                                            float l2m, float mu) {
  if (c == 0) return (d[3] * e0 + d[6] * e1 + d[7] * e2) * rrho;
  if (c == 1) return (d[6] * e0 + d[4] * e1 + d[8] * e2) * rrho;
  if (c == 2) return (d[7] * e0 + d[8] * e1 + d[5] * e2) * rrho;
  if (c == 3) return d[0] * e0 * l2m + d[1] * e1 * lam + d[2] * e2 * lam;
  if (c == 4) return d[0] * e0 * lam + d[1] * e1 * l2m + d[2] * e2 * lam;
  if (c == 5) return d[0] * e0 * lam + d[1] * e1 * lam + d[2] * e2 * l2m;
  if (c == 6) return (d[0] * e1 + d[1] * e0) * mu;
  if (c == 7) return (d[0] * e2 + d[2] * e0) * mu;
  return (d[1] * e2 + d[2] * e1) * mu;
}

// Adds the RHS terms of one axis, derivative by derivative: D(c) is the
// derivative of field c, (e0,e1,e2) the metric row of the axis (already
// multiplied by 1/beta).  vel=false skips the velocity derivatives.
template <class F>
__device__ __forceinline__ void accum_axis(float* h, F D, float e0, float e1, float e2, float rrho, float lam, float l2m,
                                           float mu, bool vel) {
  if (vel) {
    float d = D(0);  // Vx
    h[3] += l2m * e0 * d; h[4] += lam * e0 * d; h[5] += lam * e0 * d; h[6] += mu * e1 * d; h[7] += mu * e2 * d;
    d = D(1);  // Vy
    h[3] += lam * e1 * d; h[4] += l2m * e1 * d; h[5] += lam * e1 * d; h[6] += mu * e0 * d; h[8] += mu * e2 * d;
    d = D(2);  // Vz
    h[3] += lam * e2 * d; h[4] += lam * e2 * d; h[5] += l2m * e2 * d; h[7] += mu * e0 * d; h[8] += mu * e1 * d;
  }
  const float r0 = rrho * e0, r1 = rrho * e1, r2 = rrho * e2;
  float d = D(3);  // Txx
  h[0] += r0 * d;
  d = D(4);  // Tyy
  h[1] += r1 * d;
  d = D(5);  // Tzz
  h[2] += r2 * d;
  d = D(6);  // Txy
  h[0] += r1 * d; h[1] += r0 * d;
  d = D(7);  // Txz
  h[0] += r2 * d; h[2] += r0 * d;
  d = D(8);  // Tyz
  h[1] += r2 * d; h[2] += r1 * d;
}

// RK_Syn stage combine (rk4.cu:20-33) of one value: irk 0 starts the
// accumulator T and returns w + rka h; irk 1,2 add to T and return
// S + rka h; irk 3 returns T + rkb h.
__device__ __forceinline__ float rk_combine(int irk, float rka, float rkb, float w, float h, const float* S, float* T, long x) {
  if (irk == 0) {
    T[x] = w + rkb * h;
    return w + rka * h;
  }
  if (irk == 3) return T[x] + rkb * h;
  T[x] = T[x] + rkb * h;
  return S[x] + rka * h;
}

// Loads and stores at a 32-bit byte offset from a wave-uniform base, so the
// compiler can address with one 32-bit VGPR against a scalar base.
template <class T>
__device__ __forceinline__ T ldo(const T* base, uint32_t bo) { return *(const T*)((const char*)base + bo); }
template <class T>
__device__ __forceinline__ void sto(T* base, uint32_t bo, T v) { *(T*)((char*)base + bo) = v; }

// Same combine at byte offset bo of the field bases S, T.
__device__ __forceinline__ float rk_combine_o(int irk, float rka, float rkb, float w, float h, const float* S, float* T,
                                              uint32_t bo) {
  if (irk == 0) {
    sto(T, bo, w + rkb * h);
    return w + rka * h;
  }
  if (irk == 3) return ldo(T, bo) + rkb * h;
  sto(T, bo, ldo(T, bo) + rkb * h);
  return ldo(S, bo) + rka * h;
}

}  // namespace fd

// Pointer tables used off the hot path (PML, Cerjan, free-surface maps) live
// in device memory so that runtime face indices do not force the kernel
// argument block into scratch.
struct WaveCold {
  const float* pA[3];
  const float* pB[3];
  const float* pD[3];
  float* auxY[6];
  float* auxS[6];
  float* auxT[6];
  const float* cj[3];
  const float* mVx;
  const float* mVy;
  int on[6];
  int N;
};

}  // namespace cgfdm

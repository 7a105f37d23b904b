// Learned initial guess of the rate-and-state Newton solve (extension; no
// counterpart in the reference).  A 4-64-64-1 perceptron with ReLU hidden
// layers maps four features of a fault node's trial to log10 of its slip
// rate.  The solve is f(x) = 2 V0 e^(-psi/a) sinh x + coef x - Vt = 0 with
// coef = c a |sigma_n| and Vt = c |T_s| (c: velocity per traction of the
// split node).  With e = ln(2 V0 / coef) - psi/a it reads e^e sinh x + x =
// Vt/coef, and V = coef e^e sinh x.  The features are
//   f0 = T(s), s = e + Vt/coef        (state and shear traction)
//   f1 = Vt / coef                    (shear traction over a |sigma_n|)
//   f2 = log10 coef                   (normal traction, scaled by c a)
//   f3 = log10 V_prev                 (slip rate at the step start, the plain guess)
// with T(s) = s for s <= 0 and ln(1 + s) above, so that log10 V =
// f2 + (s - f1 + ln sinh x*) / ln 10 exactly.  For x* >= 20 (friction >= 20 a)
// u = x* + e solves e^u / 2 + u = s up to e^(2e-u): u ~ s for s << 0 and
// ~ ln(2 s) for s >> 0, close to linear in f0 on both sides, so log10 V is
// nearly f2 + (u(f0) - ln 2) / ln 10, which the network fits.  Layer 1
// (4 -> 64) runs in FP32, layer 2 (64 -> 64, 94% of the arithmetic) in bf16
// matrix instructions with FP32 accumulation, the output layer in FP32.
// The Newton solve (FP64, unchanged) starts from the better of this guess and
// the plain one, so convergence and tolerance are those of the plain solve.
//
// Parameter blob (float32, the layout of weights files and the C-ABI):
//   xmean[4] xscale[4] ymean yscale W1[64][4] b1[64] W2[64][64] b2[64] w3[64] b3
// features are standardised as (f - xmean) * xscale, the output is
// y * yscale + ymean.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cgfdm {

constexpr int kMlpIn = 4, kMlpH = 64;
constexpr int kMlpParams = 4 + 4 + 2 + kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;  // 4555

// Device copy used by the kernels: W2 in bf16, everything else FP32.
struct MlpDev {
  float xmean[4], xscale[4], ymean, yscale;
  float W1[kMlpH][kMlpIn];
  uint16_t W2[kMlpH][kMlpH];
  float b1[kMlpH], b2[kMlpH], w3[kMlpH], b3;
};

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  union { float f; uint32_t u; } v;
  v.f = f;
  const uint32_t r = v.u + 0x7fffu + ((v.u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  union { float f; uint32_t u; } v;
  v.u = (uint32_t)h << 16;
  return v.f;
}

// Pack a parameter blob into the device layout (host side).
inline void mlp_pack(const float* blob, MlpDev* m) {
  const float* p = blob;
  for (int i = 0; i < 4; i++) m->xmean[i] = *p++;
  for (int i = 0; i < 4; i++) m->xscale[i] = *p++;
  m->ymean = *p++;
  m->yscale = *p++;
  for (int h = 0; h < kMlpH; h++)
    for (int i = 0; i < kMlpIn; i++) m->W1[h][i] = p[h * kMlpIn + i];
  p += kMlpH * kMlpIn;
  for (int h = 0; h < kMlpH; h++) m->b1[h] = *p++;
  for (int h = 0; h < kMlpH; h++)
    for (int i = 0; i < kMlpH; i++) m->W2[h][i] = f2bf(p[h * kMlpH + i]);
  p += kMlpH * kMlpH;
  for (int h = 0; h < kMlpH; h++) m->b2[h] = *p++;
  for (int h = 0; h < kMlpH; h++) m->w3[h] = *p++;
  m->b3 = *p++;
}

// Layer 1 for one hidden unit (FP32; the activation is then rounded to
// bf16, the operand format of layer 2).
__device__ __forceinline__ float mlp_h1(const MlpDev* __restrict__ m, int h, const float* x) {
  float s = m->b1[h];
  for (int i = 0; i < kMlpIn; i++) s = fmaf(m->W1[h][i], x[i], s);
  return bf2f(f2bf(fmaxf(s, 0.0f)));
}

// One node, scalar, with the roundings of the wave version.
__device__ inline float mlp_forward_scalar(const MlpDev* __restrict__ m, const float* feat) {
  float x[4];
  for (int i = 0; i < 4; i++) x[i] = (feat[i] - m->xmean[i]) * m->xscale[i];
  float h1[kMlpH];
  for (int h = 0; h < kMlpH; h++) h1[h] = mlp_h1(m, h, x);
  float y = 0.0f;
  for (int h = 0; h < kMlpH; h++) {
    float s = 0.0f;
    for (int i = 0; i < kMlpH; i++) s += bf2f(m->W2[h][i]) * h1[i];
    y += m->w3[h] * fmaxf(s + m->b2[h], 0.0f);
  }
  return (y + m->b3) * m->yscale + m->ymean;
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// The whole wavefront evaluates the network for its 64 lanes' nodes (lane l
// holds node l), layer 2 with 16x16x16 bf16 matrix instructions: hidden
// units are the M/K dimension, nodes the N dimension, 16 nodes per tile.
// Operand layouts of v_mfma_f32_16x16x16bf16_1k: A[m][k] in lane m%16 +
// 16*(k/4), element k%4; B[k][n] in lane n%16 + 16*(k/4), element k%4;
// C[m][n] in lane n%16 + 16*(m/4), element m%4.  Each lane computes the
// layer-1 units its B operand holds.  All 64 lanes must be active.
__device__ inline float mlp_guess_wave(const MlpDev* __restrict__ m, const float* feat) {
  const int lane = __lane_id();
  const int g = lane >> 4, c = lane & 15;
  float xs[4];
  for (int i = 0; i < 4; i++) xs[i] = (feat[i] - m->xmean[i]) * m->xscale[i];
  // W2 tiles (mt, kt) as A operands
  bf16x4 w2[4][4];
  for (int mt = 0; mt < 4; mt++) {
    for (int kt = 0; kt < 4; kt++) {
      const uint16_t* r2 = &m->W2[16 * mt + c][16 * kt + 4 * g];
      w2[mt][kt] = bf16x4{(short)r2[0], (short)r2[1], (short)r2[2], (short)r2[3]};
    }
  }
  float part[4];  // output-layer partial sums of this lane for node 16*nt + c
  for (int nt = 0; nt < 4; nt++) {
    // standardised features of node 16*nt + c
    float v[4];
    for (int i = 0; i < 4; i++) v[i] = __shfl(xs[i], 16 * nt + c);
    bf16x4 h1[4];  // B operand tiles kt: hidden units 16*kt + 4*g + r of that node
    for (int kt = 0; kt < 4; kt++) {
      short hv[4];
      for (int r = 0; r < 4; r++) hv[r] = (short)f2bf(mlp_h1(m, 16 * kt + 4 * g + r, v));
      h1[kt] = bf16x4{hv[0], hv[1], hv[2], hv[3]};
    }
    float s = 0.0f;
    for (int mt = 0; mt < 4; mt++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int kt = 0; kt < 4; kt++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[mt][kt], h1[kt], acc, 0, 0, 0);
      for (int r = 0; r < 4; r++) {
        const int h = 16 * mt + 4 * g + r;
        s += m->w3[h] * fmaxf(acc[r] + m->b2[h], 0.0f);
      }
    }
    // sum the four lane groups holding node 16*nt + c
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    part[nt] = s;
  }
  const float y = part[0] * (g == 0) + part[1] * (g == 1) + part[2] * (g == 2) + part[3] * (g == 3);
  return (y + m->b3) * m->yscale + m->ymean;
}

// Start of the Newton solve: the learned guess (log10 slip rate lv) or the
// plain one (residual f_pln), whichever has the smaller residual |f| of
// f(x) = 2 V0 e^(-psi/a) sinh x + coef x - Vt.  x of the learned guess is
// asinh(V / (2 V0 e^(-psi/a))); for x > 20 that is ln V + psi/a - ln V0 to
// a relative 1e-17 and its sinh term is V itself.
__device__ inline double choose_guess(float lv, double x_plain, double f_pln, double st, double a, double Vt,
                                      double coef) {
  const double V0 = 1e-6, ln10 = 2.302585092994046;
  const double lvc = (double)fminf(fmaxf(lv, -40.0f), 3.0f);
  const double v = exp(lvc * ln10);
  double x_mlp = lvc * ln10 + st / a - log(V0);
  double f_mlp;
  if (x_mlp > 20.0) {
    f_mlp = v + coef * x_mlp - Vt;
  } else {
    const double ea = 2.0 * V0 * exp(-st / a);
    x_mlp = asinh(v / ea);
    f_mlp = ea * sinh(x_mlp) + coef * x_mlp - Vt;
  }
  return fabs(f_mlp) < fabs(f_pln) ? x_mlp : x_plain;
}

}  // namespace cgfdm

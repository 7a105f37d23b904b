// Split-node fault kernels of the CGFDM time step for CDNA4 (gfx950), 2-D
// over the fault plane with one thread per (j,k) node:
//   fault_pre   = state_rk + trial_sw | trial_rs + fault2wave
//   fault_deriv = fault_dvelo + fault_dstrs_{f,b}
//   fault_post  = slab RK + fault_rk + fault_rk_stage2 + fault2wave (+Cerjan)
//   rup_sensor  = rupture-front sensor (two passes)
// The volume stencil that precedes them in each stage is in wave_stage_kernels.h.
#include "fd_ops.h"
#include "mlp.h"
#include <hip/hip_bf16.h>

namespace cgfdm {
using namespace fd;

namespace {

// =============================================================================
// Fault kernels (2-D over the fault plane, one thread per node).
// =============================================================================
struct FArgs {
  FaultDev F;
  const float* __restrict__ W;   // stage input wavefield (read)
  float* __restrict__ Wi0;       // wavefield receiving fault2wave
  const float* __restrict__ M;   // FP32 metric and media words (the volume, or the copy around the fault)
  const float* mVx; const float* mVy;
  float* slabH;
  const float* S; float* T;      // for the slab RK
  const float* Win;              // stage input for irk 0 mW
  const float* cj[3];
  int ni, nj, nk, nx, ny, nz, i0;
  long nyz;        // words of a 2-D fault array (ny*nz)
  long pyz;        // words of one field of one plane of a volume array (py*nz, padded rows)
  long wps, mps;   // plane strides of the wave (9 pyz) and metric (13 pyz) volume arrays
  int py, jpad;    // row pitch and front pad of volume rows
  int NY, NZ;
  float DH, rDH, DT;
  double DTd;
  int it, irk, FX, FY, FZ;
  int freenode, friction, stress_type, cerjan;
  int G[4];        // Fault_grid
  float Dc0;       // scalar Dc of the JSON (smooth.cu:318)
  const MlpDev* mlp;  // learned Newton guess, or null
  float mu_d, viscosity, smooth_load_T, Vini;
  double f0, V0, fw;
  int newton_mode, newton_fallback;
  double newton_tol, mlp_gate;
};

// Metric and media reads of the fault kernels: FP32 words, from the metric
// volume with FP32 storage, or from the FP32 copy of the planes around the
// fault that the bfloat16 storage forms keep (a.M is then that copy's base
// moved back by its first plane, so full-volume plane indices address it).
__device__ __forceinline__ float LM(const FArgs& a, long i) { return a.M[i]; }
// Metric word c (0-9) at plane offset p.
__device__ __forceinline__ float LMF(const FArgs& a, int c, long p) { return a.M[c * a.pyz + p]; }
// Media word c (10 lambda, 11 mu, 12 rho) of padded plane x at column pc.
__device__ __forceinline__ float LMEDIA(const FArgs& a, int c, long x, long pc) {
  return a.M[c * a.pyz + pc + x * a.mps];
}
__device__ float Gt_func(float t, float T) {  // device.cu:134-143
  if (t < T) {
    float t1 = t - T, t2 = t - 2.0f * T;
    return expf(t1 * t1 / t / t2);
  }
  return 1.0f;
}
__device__ float Fr_func(float r, float R) { return r < R ? expf(r * r / (r * r - R * R)) : 0.0f; }

// Volume arrays are plane-interleaved with padded rows: field c of padded
// plane i at volume column jk = jpad + j + k*py sits at i*wps + c*pyz + jk
// (wavefields) or i*mps + c*pyz + jk (metric).  2-D fault arrays keep the
// reference's [nz][ny] rows and are indexed with pf = j + k*ny.
// Conservative traction J*(grad xi . sigma) at plane i, volume column jk along metric row 'row'.
__device__ __forceinline__ void trac(const FArgs& a, const float* W, int i, long jk, int row, float* o) {
  const long nyz = a.pyz, pm = jk + i * a.mps, pw = jk + i * a.wps;
  float e0 = LMF(a, 3 * row, pm), e1 = LMF(a, 3 * row + 1, pm), e2 = LMF(a, 3 * row + 2, pm);
  float J = LMF(a, 9, pm);
  float txx = W[3 * nyz + pw], tyy = W[4 * nyz + pw], tzz = W[5 * nyz + pw];
  float txy = W[6 * nyz + pw], txz = W[7 * nyz + pw], tyz = W[8 * nyz + pw];
  o[0] = J * (e0 * txx + e1 * txy + e2 * txz);
  o[1] = J * (e0 * txy + e1 * tyy + e2 * tyz);
  o[2] = J * (e0 * txz + e1 * tyz + e2 * tzz);
}

// NRsearch (trial_ratestate.cu:20-57) in FP64, with the optional residual
// stopping rule and plain-guess fallback of the learned-guess extension.
__device__ int nr_search(double x, double psi, double A, double Vhat, double coef, int mode, double tol, int fb_iter,
                         double x_plain, double* xnew, int* fell_back) {
  const double V0 = 1e-6;
  double ea = 2.0 * exp(-psi / A) * V0;
  int iter = 0;
  *fell_back = 0;
  if (mode == 0) {
    *xnew = x;
    while (iter < 100) {
      double f = ea * sinh(x) + coef * x - Vhat;
      double df = ea * cosh(x) + coef;
      double d = -f / (df + 1e-100);
      *xnew = x + d;
      double err = fabs(d) / (fabs(x) + 1e-100);
      x = *xnew;
      iter++;
      if (err < 1e-12) break;
    }
    return iter;
  }
  // residual rule: |f| <= tol * max(|Vhat|, coef*|x|, tiny); fall back to the
  // plain guess after fb_iter iterations without convergence.
  for (int attempt = 0; attempt < 2; attempt++) {
    int lim = attempt == 0 ? fb_iter : 100;
    int n = 0;
    bool conv = false;
    while (n < lim) {
      double sh = sinh(x);
      double f = ea * sh + coef * x - Vhat;
      double scale = fmax(fabs(Vhat), fmax(fabs(coef * x), 1e-300));
      if (fabs(f) <= tol * scale) { conv = true; break; }
      double df = ea * cosh(x) + coef;
      x = x - f / (df + 1e-100);
      n++;
    }
    iter += n;
    if (conv || attempt == 1) { *xnew = x; return iter; }
    *fell_back = 1;
    x = x_plain;
  }
  *xnew = x;
  return iter;
}

// Shared end of the friction trial (trial_slipweakening.cu:640-709,
// trial_ratestate.cu:800-880): T1 on the fault plane with the optional
// viscosity term, the stage traction averages, rupture time, then
// fault2wave (transform.cu:224-403) into the stage input at i0.
__device__ __forceinline__ void fault_tail(const FArgs& a, const FaultDev& F, long q, long pf, long pc,
                                           const float* Tau, float Tau_n, bool ifchange, float ratecheck,
                                           const float* Trial, float jacvec, const float* s1, const float* s2) {
  const long nyz = a.nyz, pyz = a.pyz, wps = a.wps, mps = a.mps;
  const int i0 = a.i0, irk = a.irk;
  const float DT = a.DT;
  long p1 = 3 * nyz + pf;
  float T1[3];
  if (ifchange) {
    T1[0] = (Tau[0] - F.str_init_x[q]) * jacvec;
    T1[1] = (Tau[1] - F.str_init_y[q]) * jacvec;
    T1[2] = (Tau[2] - F.str_init_z[q]) * jacvec;
  } else {
    T1[0] = Trial[0]; T1[1] = Trial[1]; T1[2] = Trial[2];
  }
  float vis = a.viscosity * DT;
  float h1 = (T1[0] - F.mT11[pf]) / DT, h2 = (T1[1] - F.mT12[pf]) / DT, h3 = (T1[2] - F.mT13[pf]) / DT;
  if (irk == 3) { F.hT11[pf] = h1; F.hT12[pf] = h2; F.hT13[pf] = h3; }
  T1[0] += vis * h1; T1[1] += vis * h2; T1[2] += vis * h3;
  F.T11[p1] = T1[0]; F.T12[p1] = T1[1]; F.T13[p1] = T1[2];
  F.tTs1[q] = Tau[0] * s1[0] + Tau[1] * s1[1] + Tau[2] * s1[2];
  F.tTs2[q] = Tau[0] * s2[0] + Tau[1] * s2[1] + Tau[2] * s2[2];
  F.tTn[q] = Tau_n;
  if (!F.init_t0_flag[q] && ratecheck > 1e-3f) {
    F.init_t0[q] = a.it * DT;
    F.init_t0_flag[q] = 1;
    F.flag_rup[q] = 1;
  }

  // ---- fault2wave into the stage input at i0
  {
    const long p = pc + i0 * mps, pw = pc + i0 * wps;
    float mt[3][3] = {{LMF(a, 0, p), LMF(a, 3, p), LMF(a, 6, p)},
                      {LMF(a, 1, p), LMF(a, 4, p), LMF(a, 7, p)},
                      {LMF(a, 2, p), LMF(a, 5, p), LMF(a, 8, p)}};
    float rj = 1.0f / LMF(a, 9, p);
    float inv[3][3];
    inv[0][0] = mt[1][1] * mt[2][2] - mt[2][1] * mt[1][2];
    inv[0][1] = mt[2][1] * mt[0][2] - mt[0][1] * mt[2][2];
    inv[0][2] = mt[0][1] * mt[1][2] - mt[0][2] * mt[1][1];
    inv[1][0] = mt[1][2] * mt[2][0] - mt[1][0] * mt[2][2];
    inv[1][1] = mt[0][0] * mt[2][2] - mt[2][0] * mt[0][2];
    inv[1][2] = mt[1][0] * mt[0][2] - mt[0][0] * mt[1][2];
    inv[2][0] = mt[1][0] * mt[2][1] - mt[1][1] * mt[2][0];
    inv[2][1] = mt[2][0] * mt[0][1] - mt[0][0] * mt[2][1];
    inv[2][2] = mt[0][0] * mt[1][1] - mt[0][1] * mt[1][0];
    float det = 1.0f / (inv[0][0] * mt[0][0] + inv[0][1] * mt[1][0] + inv[0][2] * mt[2][0]);
    float tr[3][3];
    for (int c = 0; c < 3; c++) {
      tr[c][0] = T1[c];
      tr[c][1] = (F.W[(3 + c) * 2 * nyz + pf] + F.W[(3 + c) * 2 * nyz + nyz + pf]) * 0.5f;
      tr[c][2] = (F.W[(6 + c) * 2 * nyz + pf] + F.W[(6 + c) * 2 * nyz + nyz + pf]) * 0.5f;
    }
    float sg[3][3];
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++)
        sg[r][cc] = (tr[r][0] * inv[0][cc] + tr[r][1] * inv[1][cc] + tr[r][2] * inv[2][cc]) * det * rj;
    float* W = a.Wi0;
    for (int c = 0; c < 3; c++) W[c * pyz + pw] = (F.W[c * 2 * nyz + pf] + F.W[c * 2 * nyz + nyz + pf]) * 0.5f;
    W[3 * pyz + pw] = sg[0][0];
    W[4 * pyz + pw] = sg[1][1];
    W[5 * pyz + pw] = sg[2][2];
    W[6 * pyz + pw] = (sg[0][1] + sg[1][0]) * 0.5f;
    W[7 * pyz + pw] = (sg[0][2] + sg[2][0]) * 0.5f;
    W[8 * pyz + pw] = (sg[1][2] + sg[2][1]) * 0.5f;
  }
}

// fault_pre: state_rk (rk4.cu:791-938) + trial_sw (trial_slipweakening.cu:23-732)
// or trial_rs (trial_ratestate.cu:59-893) + fault2wave (transform.cu:224-403).
__global__ void __launch_bounds__(256) fault_pre_kernel(FArgs a) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.nj || k >= a.nk) return;
  const FaultDev& F = a.F;
  long q = j + (long)k * a.nj;
  if (F.united[q]) return;
  const int j1 = j + 3, k1 = k + 3, ny = a.ny;
  const long nyz = a.nyz, pyz = a.pyz, wps = a.wps, mps = a.mps, pf = j1 + (long)k1 * ny, pv = pf * 3;
  const long pc = a.jpad + j1 + (long)k1 * a.py;  // volume column
  const int i0 = a.i0;
  const float DH = a.DH, rDH = a.rDH, DT = a.DT, DH2 = DH * DH;
  const double RA[4] = {0.5, 0.5, 1.0, 1e10}, RB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const int irk = a.irk;
  const double rka = RA[irk] * a.DTd, rkb = RB[irk] * a.DTd;
  float s1[3] = {F.vec_s1[pv], F.vec_s1[pv + 1], F.vec_s1[pv + 2]};
  float s2[3] = {F.vec_s2[pv], F.vec_s2[pv + 1], F.vec_s2[pv + 2]};

  // ---- state_rk: slip (and state) RK4 with the current split-node rate
  {
    float dV[3];
    for (int c = 0; c < 3; c++) dV[c] = F.W[c * 2 * nyz + nyz + pf] - F.W[c * 2 * nyz + pf];
    float vs1 = dV[0] * s1[0] + dV[1] * s1[1] + dV[2] * s1[2];
    float vs2 = dV[0] * s2[0] + dV[1] * s2[1] + dV[2] * s2[2];
    float hs = sqrtf(vs2 * vs2 + vs1 * vs1);
    F.hslip[q] = hs;
    double V = fmax((double)hs, 1e-30);
    double st = F.State[q];
    double hState = 0.0;
    if (a.friction == 1) {
      double b = F.b[q], L = F.L[q];
      hState = b * a.V0 / L * (exp((a.f0 - st) / b) - V / a.V0);
    } else if (a.friction >= 2) {
      double A = F.a[q], b = F.b[q], L = F.L[q], Vw = F.Vw[q];
      double flv = a.f0 - (b - A) * log(V / a.V0);
      double fss = a.fw + (flv - a.fw) / pow(1. + pow(V / Vw, 8), 0.125);
      double psiss = A * (log(sinh(fss / A)) + log(2. * (a.V0 / V)));
      hState = -V / L * (st - psiss);
    }
    if (irk == 0) {
      F.mState[q] = st; F.tState[q] = st + rkb * hState;
      F.mslip[q] = F.slip[q]; F.tslip[q] = F.slip[q] + rkb * hs;
    } else if (irk < 3) {
      F.tState[q] += rkb * hState;
      F.tslip[q] += rkb * hs;
    }
    if (irk == 3) { F.State[q] = F.tState[q] + rkb * hState; F.slip[q] = F.tslip[q] + rkb * hs; }
    else { F.State[q] = F.mState[q] + rka * hState; F.slip[q] = F.mslip[q] + rka * hs; }
  }

  // ---- trial: conservative traction on planes i0-+1..3, R and lumped mass
  if (irk == 0) { F.mT11[pf] = F.T11[3 * nyz + pf]; F.mT12[pf] = F.T12[3 * nyz + pf]; F.mT13[pf] = F.T13[3 * nyz + pf]; }
  const int km = a.nk - k;
  float R[2][3], Mr[2], jac_l = 1.f, rho_l = 1.f;
  for (int m = 0; m < 2; m++) {
    float rho = F.rho_f[m * nyz + pf];
    float jac = 1.f;
    for (int l = 1; l <= 3; l++) {
      const int ii = i0 + (2 * m - 1) * l;
      jac = LMF(a, 9, pc + ii * mps);
      float o[3];
      trac(a, a.W, ii, pc, 0, o);
      long pl = (3 + (2 * m - 1) * l) * nyz + pf;
      F.T11[pl] = o[0]; F.T12[pl] = o[1]; F.T13[pl] = o[2];
    }
    float DyT[3], DzT[3];
    for (int c = 0; c < 3; c++) {
      const float* t2 = &F.W[(3 + c) * 2 * nyz + m * nyz + pf];
      DyT[c] = (F.rup_index_y[q] % 7 ? d22(t2, 1, a.FY) : drp(t2, 1, a.FY)) * rDH;
      float v[7];
      for (int l = -3; l <= 3; l++) v[l + 3] = F.W[(6 + c) * 2 * nyz + m * nyz + pf + (long)l * ny];
      if (a.freenode && km <= 3) {
        v[km + 2] = 0.0f;
        for (int l = km + 3; l < 7; l++) v[l] = -v[2 * (km + 2) - l];
      }
      DzT[c] = (F.rup_index_z[q] % 7 ? d22(&v[3], 1, a.FZ) : drp(&v[3], 1, a.FZ)) * rDH;
    }
    const float* TT[3] = {F.T11, F.T12, F.T13};
    for (int c = 0; c < 3; c++) {
      const float* T = TT[c];
      float r = m == 0 ? A1 * T[2 * nyz + pf] + A2 * T[1 * nyz + pf] + A3 * T[0 * nyz + pf]
                       : A1 * T[4 * nyz + pf] + A2 * T[5 * nyz + pf] + A3 * T[6 * nyz + pf];
      R[m][c] = 0.5f * ((2 * m - 1) * r + (DyT[c] + DzT[c]) * DH) * DH2;
    }
    Mr[m] = 0.5f * jac * rho * DH * DH2;  // J of i0-+3, as the reference's inner loop leaves it
    jac_l = jac; rho_l = rho;
  }
  float t = irk == 0 ? a.it * DT : (irk == 3 ? (a.it + 1) * DT : (a.it + 0.5f) * DT);
  long pm = pc + i0 * mps;
  float n[3] = {LMF(a, 0, pm), LMF(a, 1, pm), LMF(a, 2, pm)};
  float n0 = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  float jacvec = LMF(a, 9, pm) * n0;
  for (int c = 0; c < 3; c++) n[c] /= n0;
  float dV[3];
  for (int c = 0; c < 3; c++) dV[c] = F.mW[c * 2 * nyz + nyz + pf] - F.mW[c * 2 * nyz + pf];
  float Tau[3], Tau_n, ratecheck;
  bool ifchange;
  float Trial[3];
  if (a.friction == 0) {
    if (a.stress_type == 1) {
      float Gt = Gt_func(t, a.smooth_load_T);
      F.str_init_x[q] = F.T0x[q] + Gt * F.dT0x[q];
      F.str_init_y[q] = F.T0y[q] + Gt * F.dT0y[q];
      F.str_init_z[q] = F.T0z[q] + Gt * F.dT0z[q];
    }
    float den = DH2 * (Mr[0] * A0P - A0M * Mr[1]);
    for (int c = 0; c < 3; c++) Trial[c] = -(Mr[0] * Mr[1] * dV[c] / DT + Mr[0] * R[1][c] - Mr[1] * R[0][c]) / den * 2.0f;
    float TL[3] = {Trial[0] / jacvec + F.str_init_x[q], Trial[1] / jacvec + F.str_init_y[q], Trial[2] / jacvec + F.str_init_z[q]};
    float Tn_ = TL[0] * n[0] + TL[1] * n[1] + TL[2] * n[2];
    float Ts[3] = {TL[0] - n[0] * Tn_, TL[1] - n[1] * Tn_, TL[2] - n[2] * Tn_};
    float Ts0 = sqrtf(Ts[0] * Ts[0] + Ts[1] * Ts[1] + Ts[2] * Ts[2]);
    ifchange = false;
    if (Tn_ >= 0.0f) { Tau_n = 0.0f; ifchange = true; } else Tau_n = Tn_;
    float mu1 = F.str_peak[q], sl = (float)F.slip[q], Dc = F.Dc[q];
    float fr = sl <= Dc ? mu1 - (mu1 - a.mu_d) * sl / Dc : a.mu_d;
    float Tau_c = -fr * Tau_n + F.C0[q];
    float Tau_s;
    if (Ts0 >= Tau_c) { Tau_s = Tau_c; F.flag_rup[q] = 1; ifchange = true; }
    else { Tau_s = Ts0; F.flag_rup[q] = 0; }
    if (ifchange) {
      if (fabsf(Ts0) < 1e-1f) for (int c = 0; c < 3; c++) Tau[c] = Tau_n * n[c];
      else for (int c = 0; c < 3; c++) Tau[c] = Tau_n * n[c] + (Tau_s / Ts0) * Ts[c];
    } else {
      for (int c = 0; c < 3; c++) Tau[c] = TL[c];
    }
    ratecheck = F.hslip[q];
  } else {
    // rate-and-state with the hard-coded nucleation unless stresses are read
    float y1 = (float)(j - a.NY / 2);
    float z1 = (float)((a.NZ - 1 - k) - (int)(7.5e3f / DH));
    float r = sqrtf(y1 * y1 + z1 * z1) * DH;
    float Gt = Gt_func(t, a.smooth_load_T);
    float Fr = Fr_func(r, 3.0e3f);
    F.str_init_x[q] = -120e6f;
    F.str_init_y[q] = -25e6f * Fr * Gt + -75e6f;
    F.str_init_z[q] = 0.0f;
    if (a.stress_type == 1) {
      F.str_init_x[q] = F.T0x[q] + Gt * F.dT0x[q];
      F.str_init_y[q] = F.T0y[q] + Gt * F.dT0y[q];
      F.str_init_z[q] = F.T0z[q] + Gt * F.dT0z[q];
    }
    float T0[3] = {F.str_init_x[q], F.str_init_y[q], F.str_init_z[q]};
    double dn[3] = {n[0], n[1], n[2]}, ds1[3] = {s1[0], s1[1], s1[2]}, ds2[3] = {s2[0], s2[1], s2[2]};
    double T0l[3] = {dn[0] * T0[0] + dn[1] * T0[1] + dn[2] * T0[2], ds1[0] * T0[0] + ds1[1] * T0[1] + ds1[2] * T0[2],
                     ds2[0] * T0[0] + ds2[1] * T0[1] + ds2[2] * T0[2]};
    double R1[2], R2[2], R3[2];
    for (int m = 0; m < 2; m++) {
      R1[m] = n[0] * R[m][0] + n[1] * R[m][1] + n[2] * R[m][2];
      R2[m] = s1[0] * R[m][0] + s1[1] * R[m][1] + s1[2] * R[m][2];
      R3[m] = s2[0] * R[m][0] + s2[1] * R[m][1] + s2[2] * R[m][2];
    }
    double dVd[3] = {dV[0], dV[1], dV[2]};
    double Rl[3] = {dn[0] * dVd[0] + dn[1] * dVd[1] + dn[2] * dVd[2], ds1[0] * dVd[0] + ds1[1] * dVd[1] + ds1[2] * dVd[2],
                    ds2[0] * dVd[0] + ds2[1] * dVd[1] + ds2[2] * dVd[2]};
    Rl[1] -= a.Vini;
    double dt1 = a.DTd;
    double M0 = Mr[0], M1 = Mr[1];
    double den = (double)DH2 * (M0 * A0P - A0M * M1) * jacvec;
    double Tt[3];
    Tt[0] = T0l[0] - 2.0 * (M0 * M1 * Rl[0] / dt1 + M0 * R1[1] - M1 * R1[0]) / den;
    Tt[1] = T0l[1] - 2.0 * (M0 * M1 * Rl[1] / dt1 + M0 * R2[1] - M1 * R2[0]) / den;
    Tt[2] = T0l[2] - 2.0 * (M0 * M1 * Rl[2] / dt1 + M0 * R3[1] - M1 * R3[0]) / den;
    double cc = -((double)A0P - (double)A0M) * dt1 * jacvec / ((double)jac_l * rho_l * DH);
    double Vt1 = cc * Tt[1], Vt2 = cc * Tt[2];
    double st = F.State[q], RS_a = F.a[q];
    double Tau_nd = Tt[0] >= -1e-1 ? -1e-1 : Tt[0];
    // thermal pressurization: the fault's pore pressure makes the normal
    // stress effective for the solve and the strength (trial_ratestate.cu:636-644)
    if (F.TP_P) Tau_nd += F.TP_P[q];
    double coef = cc * RS_a * fabs(Tau_nd);
    double Vtn = sqrt(Vt1 * Vt1 + Vt2 * Vt2);
    double Ttn = sqrt(Tt[1] * Tt[1] + Tt[2] * Tt[2]);
    double Rn = sqrt(Rl[1] * Rl[1] + Rl[2] * Rl[2]);
    const double RS_V0 = 1e-6;
    const double x_plain = asinh(exp(st / RS_a) * Rn / (2.0 * RS_V0));
    // The Newton solve and the rest of the trial run in fault_rs_kernel:
    // hand over its inputs, what the trial needs after it, and the features
    // of the learned initial guess (Newton unknown x = friction / a).
    const long n2 = (long)a.nj * a.nk;
    double* C = F.rs_carry;
    C[q] = st; C[n2 + q] = RS_a; C[2 * n2 + q] = Vtn; C[3 * n2 + q] = coef; C[4 * n2 + q] = x_plain;
    C[5 * n2 + q] = Tau_nd; C[6 * n2 + q] = Tt[1]; C[7 * n2 + q] = Tt[2]; C[8 * n2 + q] = Ttn;
    C[9 * n2 + q] = Rn;
    float* Cf = F.rs_carryf;
    Cf[q] = n[0]; Cf[n2 + q] = n[1]; Cf[2 * n2 + q] = n[2]; Cf[3 * n2 + q] = jacvec;
    // x* of f(x) = 2 V0 e^(-psi/a) sinh x + coef x - Vt = 0 depends on the
    // two numbers e = ln(2 V0 / coef) - psi/a and Vt/coef (features
    // T(e + Vt/coef) and Vt/coef, T(s) = s for s <= 0, ln(1 + s) above);
    // log10 V* adds log10 coef.  The fourth is the log previous slip rate.
    const double s_rs = log(2.0 * RS_V0 / coef) - st / RS_a + Vtn / coef;
    Cf[4 * n2 + q] = (float)(s_rs <= 0.0 ? s_rs : log1p(s_rs));
    Cf[5 * n2 + q] = (float)(Vtn / coef);
    Cf[6 * n2 + q] = (float)log10(coef);
    Cf[7 * n2 + q] = (float)log10(fmax(Rn, 1e-30));
    return;
  }
  fault_tail(a, F, q, pf, pc, Tau, Tau_n, ifchange, ratecheck, Trial, jacvec, s1, s2);
}

// Rate-and-state trial after the hand-over from fault_pre_kernel
// (trial_ratestate.cu:580-800 and NRsearch, trial_ratestate.cu:20-57): FP64
// Newton for x = friction / a from the plain guess asinh(e^(psi/a) V/(2 V0))
// (or the learned guess when one is loaded), then shear traction along the
// trial direction, Tau[0] = str_init_x (trial_ratestate.cu:797) and the
// shared tail.  Every lane of a wavefront stays active up to the guess, which
// evaluates the network with matrix-core instructions for the whole wave.
__global__ void __launch_bounds__(256) fault_rs_kernel(FArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  const bool inrange = j < a.nj && k < a.nk;
  const int jc = min(j, a.nj - 1), kc = min(k, a.nk - 1);
  const FaultDev& F = a.F;
  const long q = jc + (long)kc * a.nj, n2 = (long)a.nj * a.nk;
  const bool valid = inrange && !F.united[q];
  const double* C = F.rs_carry;
  const float* Cf = F.rs_carryf;
  const double st = C[q], RS_a = C[n2 + q], Vtn = C[2 * n2 + q], coef = C[3 * n2 + q], x_plain = C[4 * n2 + q];
  double x0 = x_plain;
  bool consulted = false;
  if (a.mlp) {
    // 2 V0 e^(-psi/a) sinh(x_plain) is the previous slip rate Rn, so the
    // plain guess's residual costs no transcendental.  The network is
    // consulted by the waves in which some node's plain guess is poor
    // (relative residual above the gate, NewtonMLPGate); the vote is wave-uniform, so all
    // 64 lanes enter the matrix instructions together.
    const double Rn = C[9 * n2 + q];
    const double f_pln = Rn + coef * x_plain - Vtn;
    const bool poor = valid && fabs(f_pln) > a.mlp_gate * fmax(Vtn, 1e-300);
    if (__any(poor)) {
      consulted = true;
      float feat[4];
      for (int f = 0; f < 4; f++) feat[f] = valid ? Cf[(4 + f) * n2 + q] : 0.0f;
      const float lv = mlp_guess_wave(a.mlp, feat);  // log10 of the predicted slip rate
      if (poor) x0 = choose_guess(lv, x_plain, f_pln, st, RS_a, Vtn, coef);
    }
  }
  // Newton on the valid lanes; the wave's largest iteration count goes to a
  // histogram (one atomic per wave, lane 0) before the other lanes leave
  double xnew = x0;
  int fb = 0, iters = 0;
  if (valid) iters = nr_search(x0, st, RS_a, Vtn, coef, a.newton_mode, a.newton_tol, a.newton_fallback, x_plain, &xnew, &fb);
  int wmax = iters;
  for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off));
  if (__any(valid) && __lane_id() == 0) atomicAdd(&F.newton_stats[5 * n2 + min(wmax, kNewtonHistBins - 1)], 1u);
  if (!valid) return;
  // C[5] holds the effective normal stress; the traction vector takes the
  // total one (trial_ratestate.cu:785-788)
  const double Tau_eff = C[5 * n2 + q], Tt1 = C[6 * n2 + q], Tt2 = C[7 * n2 + q];
  const double Tau_nd = F.TP_P ? Tau_eff - F.TP_P[q] : Tau_eff;
  double Ttn = C[8 * n2 + q];
  const float n[3] = {Cf[q], Cf[n2 + q], Cf[2 * n2 + q]};
  const float jacvec = Cf[3 * n2 + q];
  const int j1 = j + 3, k1 = k + 3;
  const long pf = j1 + (long)k1 * a.ny, pv = pf * 3, pc = a.jpad + j1 + (long)k1 * a.py;
  const float s1[3] = {F.vec_s1[pv], F.vec_s1[pv + 1], F.vec_s1[pv + 2]};
  const float s2[3] = {F.vec_s2[pv], F.vec_s2[pv + 1], F.vec_s2[pv + 2]};
  F.newton_stats[q] += 1u;  // per-node counters: no contention between nodes
  F.newton_stats[n2 + q] += (unsigned)iters;
  F.newton_stats[2 * n2 + q] += fb ? 1u : 0u;
  F.newton_stats[3 * n2 + q] += x0 != x_plain ? 1u : 0u;  // started from the learned guess
  F.newton_stats[4 * n2 + q] += consulted ? 1u : 0u;      // its wave evaluated the network
  const double RS_V0 = 1e-6;
  float V = (float)(2.0 * RS_V0 * exp(-st / RS_a) * sinh(xnew));
  const double Tm = RS_a * xnew * fabs(Tau_eff);
  F.friction[q] = (float)(RS_a * xnew);
  Ttn = fmax(Ttn, 1e-8);
  const double Ts1 = Tt1 / Ttn * Tm, Ts2 = Tt2 / Ttn * Tm;
  V = fmaxf(V, 1e-30f);
  float Tau[3];
  for (int c = 0; c < 3; c++) Tau[c] = (float)(Tau_nd * (double)n[c] + Ts1 * (double)s1[c] + Ts2 * (double)s2[c]);
  Tau[0] = F.str_init_x[q];  // trial_ratestate.cu:797
  const float Tau_n = Tau[0] * n[0] + Tau[1] * n[1] + Tau[2] * n[2];
  const float Trial[3] = {0.0f, 0.0f, 0.0f};
  fault_tail(a, F, q, pf, pc, Tau, Tau_n, true, V, Trial, jacvec, s1, s2);
}

__device__ __forceinline__ void mv3(const float* A, const float* x, float* y) {
  for (int r = 0; r < 3; r++) y[r] = A[3 * r] * x[0] + A[3 * r + 1] * x[1] + A[3 * r + 2] * x[2];
}

// x-derivative of velocity component c on off-fault plane i0 + l of the
// slab (nn = l + 3 for l < 0, l + 4 for l > 0), as fault_dstrs_f/_b form it
// (fault_dstrs_f.cu:180-330): one-sided DRP, the 2-4 operator and the
// two-point difference next to the fault, reaching the split node's own
// side (F.W of side 0 or 1) instead of the plane i0.
__device__ __forceinline__ float dxv_off(const FArgs& a, int nn, int c, long pc, long pf) {
  const FaultDev& F = a.F;
  const long wps = a.wps;
  const float* W = a.W;
  const int i0 = a.i0;
  auto wv = [&](int i) -> const float* { return &W[c * a.pyz + pc + i * wps]; };
  const float f0 = F.W[c * 2 * a.nyz + pf], f1 = F.W[c * 2 * a.nyz + a.nyz + pf];
  float d;
  if (a.FX > 0) {
    if (nn == 7) d = drpF(wv(i0 + 3), wps);
    else if (nn == 6) d = d24F(wv(i0 + 2), wps);
    else if (nn == 5) d = wv(i0 + 1)[wps] - wv(i0 + 1)[0];
    else if (nn == 0) {
      const float v5[5] = {wv(i0 - 4)[0], wv(i0 - 3)[0], wv(i0 - 2)[0], wv(i0 - 1)[0], f0};
      d = drpF(&v5[1], 1);
    } else if (nn == 1) {
      const float v3[3] = {wv(i0 - 2)[0], wv(i0 - 1)[0], f0};
      d = d24F(&v3[0], 1);
    } else {
      d = f0 - wv(i0 - 1)[0];
    }
  } else {
    if (nn == 0) d = drpB(wv(i0 - 3), wps);
    else if (nn == 1) d = d24B(wv(i0 - 2), wps);
    else if (nn == 2) d = wv(i0 - 1)[0] - wv(i0 - 1)[-wps];
    else if (nn == 7) {
      const float v5[5] = {f1, wv(i0 + 1)[0], wv(i0 + 2)[0], wv(i0 + 3)[0], wv(i0 + 4)[0]};
      d = drpB(&v5[3], 1);
    } else if (nn == 6) {
      const float v3[3] = {f1, wv(i0 + 1)[0], wv(i0 + 2)[0]};
      d = d24B(&v3[2], 1);
    } else {
      d = wv(i0 + 1)[0] - f1;
    }
  }
  return d * a.rDH;
}

// fault_deriv: fault_dvelo_cu (fault_dvelo.cu:17-389) + fault_dstrs_f/b_cu
// (fault_dstrs_f.cu:17-1254, fault_dstrs_b.cu) writing the slab RHS.  The
// parts that are independent per slab plane run in parallel, one per block:
// 0-5 the velocity rates of planes i0-3..-1, +1..+3,
// 6-11 the stress rates of the same planes, 12-13 the split-node velocity
// rates of sides 0 and 1, 14-15 their stress rates.  The parts of one
// 64 x 4 tile of fault nodes are consecutive work items, kept on one XCD:
// they read the same slab planes around the tile, which then come from that
// XCD's L2 instead of being fetched once per part.  Each group of parts
// (G 0: 0-5, 1: 6-11, 2: 12-15) is a kernel of its own, so that the
// registers of the split-node stress part do not set the occupancy of the
// others.
constexpr int kDerivGroupBase[3] = {0, 6, 12}, kDerivGroupParts[3] = {6, 6, 4};
template <int G>
__global__ void __launch_bounds__(256) fault_deriv_kernel(FArgs a, int tiles_j, int items) {
  const int item = xcd_remap((int)blockIdx.x, items);
  const int part = kDerivGroupBase[G] + item % kDerivGroupParts[G], tile = item / kDerivGroupParts[G];
  int j = (tile % tiles_j) * 64 + threadIdx.x;
  int k = (tile / tiles_j) * 4 + threadIdx.y;
  if (j >= a.nj || k >= a.nk) return;
  const FaultDev& F = a.F;
  long q = j + (long)k * a.nj;
  if (F.united[q]) return;
  const int j1 = j + 3, k1 = k + 3, ny = a.ny, i0 = a.i0;
  const long nyz = a.nyz, pyz = a.pyz, wps = a.wps, mps = a.mps, pf = j1 + (long)k1 * ny;
  const long pc = a.jpad + j1 + (long)k1 * a.py;  // volume column
  const int py = a.py;
  const float rDH = a.rDH;
  const int km = a.nk - k;
  const int FX = a.FX, FY = a.FY, FZ = a.FZ;
  const float* W = a.W;
  float* slab = a.slabH;
  auto slabx = [&](int i, int c) -> float& { return slab[(long)(i - (i0 - 3)) * 9 * pyz + c * pyz + pc]; };

  if (G == 0) {  // velocity rates at i0-3..-1, +1..+3 (fault_dvelo.cu:153-255)
    const int i = i0 + (part < 3 ? part - 3 : part - 2);
    {
      // the three axes one after another, each array indexed with constants
      // only (a runtime index would put it in scratch memory)
      const int nn = i0 - i;
      float dxt[3], dyt[3], dzt[3];
      {  // across the fault: the traction on the fault plane replaces the point at i0
        float v[3][7];
        for (int l = -3; l <= 3; l++) {
          float o[3];
          trac(a, W, i + l, pc, 0, o);
          for (int c = 0; c < 3; c++) v[c][l + 3] = o[c];
        }
        const float tf[3] = {F.T11[3 * nyz + pf], F.T12[3 * nyz + pf], F.T13[3 * nyz + pf]};
        for (int c = 0; c < 3; c++) {
#pragma unroll
          for (int l = 0; l < 7; l++)
            if (l == nn + 3) v[c][l] = tf[c];
          if (nn == 2) v[c][6] = 2.0f * v[c][5] - v[c][4];
          if (nn == 1) { v[c][5] = 2.0f * v[c][4] - v[c][3]; v[c][6] = 2.0f * v[c][4] - v[c][2]; }
          if (nn == -1) { v[c][0] = 2.0f * v[c][2] - v[c][4]; v[c][1] = 2.0f * v[c][2] - v[c][3]; }
          if (nn == -2) v[c][0] = 2.0f * v[c][1] - v[c][2];
          dxt[c] = drp(&v[c][3], 1, FX) * rDH;
        }
      }
      {  // along strike
        float v[3][7];
        for (int l = -3; l <= 3; l++) {
          float o[3];
          trac(a, W, i, pc + l, 1, o);
          for (int c = 0; c < 3; c++) v[c][l + 3] = o[c];
        }
        for (int c = 0; c < 3; c++) dyt[c] = drp(&v[c][3], 1, FY) * rDH;
      }
      {  // along dip, with the odd image above the free surface
        float v[3][7];
        for (int l = -3; l <= 3; l++) {
          float o[3];
          trac(a, W, i, pc + (long)l * py, 2, o);
          for (int c = 0; c < 3; c++) v[c][l + 3] = o[c];
        }
        if (a.freenode && km <= 3)
          for (int c = 0; c < 3; c++) {
            if (km == 1) { v[c][3] = 0.0f; v[c][4] = -v[c][2]; v[c][5] = -v[c][1]; v[c][6] = -v[c][0]; }
            else if (km == 2) { v[c][4] = 0.0f; v[c][5] = -v[c][3]; v[c][6] = -v[c][2]; }
            else { v[c][5] = 0.0f; v[c][6] = -v[c][4]; }
          }
        for (int c = 0; c < 3; c++) dzt[c] = drp(&v[c][3], 1, FZ) * rDH;
      }
      const long p = pc + i * mps;
      float rrj = 1.0f / (LMEDIA(a, 12, i, pc) * LMF(a, 9, p));
      for (int c = 0; c < 3; c++) slabx(i, c) = (dxt[c] + dyt[c] + dzt[c]) * rrj;
    }
    return;
  }
  if (G == 1) {  // off-fault stress rates (fault_dstrs_f.cu:1009-1250)
    const int nn = part < 9 ? part - 6 : part - 4;
    {
      const int l = nn < 4 ? nn - 3 : nn - 4;
      const int i = i0 + l;
      const long p = pc + i * mps, pw = pc + i * wps;
      float dy[3], dz[3];
      float dx[3];
      for (int c = 0; c < 3; c++) dx[c] = dxv_off(a, nn, c, pc, pf);
      for (int c = 0; c < 3; c++) dy[c] = drp(&W[c * pyz + pw], 1, FY) * rDH;
      long ij = (j1 + (long)i * ny) * 9;
      for (int c = 0; c < 3; c++) {
        if (a.freenode && km == 1)
          dz[c] = a.mVx[ij + 3 * c] * dx[0] + a.mVx[ij + 3 * c + 1] * dx[1] + a.mVx[ij + 3 * c + 2] * dx[2]
                + a.mVy[ij + 3 * c] * dy[0] + a.mVy[ij + 3 * c + 1] * dy[1] + a.mVy[ij + 3 * c + 2] * dy[2];
        else if (a.freenode && km == 2) dz[c] = d22(&W[c * pyz + pw], py, FZ) * rDH;
        else if (a.freenode && km == 3) dz[c] = d24(&W[c * pyz + pw], py, FZ) * rDH;
        else dz[c] = drp(&W[c * pyz + pw], py, FZ) * rDH;
      }
      float lam = LMEDIA(a, 10, i, pc), mu = LMEDIA(a, 11, i, pc), l2m = lam + 2.0f * mu;
      float xi[3] = {LMF(a, 0, p), LMF(a, 1, p), LMF(a, 2, p)};
      float et[3] = {LMF(a, 3, p), LMF(a, 4, p), LMF(a, 5, p)};
      float zt[3] = {LMF(a, 6, p), LMF(a, 7, p), LMF(a, 8, p)};
      slabx(i, 3) = l2m * xi[0] * dx[0] + lam * xi[1] * dx[1] + lam * xi[2] * dx[2] + l2m * et[0] * dy[0] + lam * et[1] * dy[1] +
                    lam * et[2] * dy[2] + l2m * zt[0] * dz[0] + lam * zt[1] * dz[1] + lam * zt[2] * dz[2];
      slabx(i, 4) = lam * xi[0] * dx[0] + l2m * xi[1] * dx[1] + lam * xi[2] * dx[2] + lam * et[0] * dy[0] + l2m * et[1] * dy[1] +
                    lam * et[2] * dy[2] + lam * zt[0] * dz[0] + l2m * zt[1] * dz[1] + lam * zt[2] * dz[2];
      slabx(i, 5) = lam * xi[0] * dx[0] + lam * xi[1] * dx[1] + l2m * xi[2] * dx[2] + lam * et[0] * dy[0] + lam * et[1] * dy[1] +
                    l2m * et[2] * dy[2] + lam * zt[0] * dz[0] + lam * zt[1] * dz[1] + l2m * zt[2] * dz[2];
      slabx(i, 6) = (xi[1] * dx[0] + xi[0] * dx[1] + et[1] * dy[0] + et[0] * dy[1] + zt[1] * dz[0] + zt[0] * dz[1]) * mu;
      slabx(i, 7) = (xi[2] * dx[0] + xi[0] * dx[2] + et[2] * dy[0] + et[0] * dy[2] + zt[2] * dz[0] + zt[0] * dz[2]) * mu;
      slabx(i, 8) = (xi[2] * dx[1] + xi[1] * dx[2] + et[2] * dy[1] + et[1] * dy[2] + zt[2] * dz[1] + zt[1] * dz[2]) * mu;
    }
    return;
  }
  if (part < 14) {  // G 2: split-node velocity rates of side part-12 (fault_dvelo.cu:260-385)
    const int m = part - 12;
    const float* TT[3] = {F.T11, F.T12, F.T13};
    float rrj = 1.0f / (F.rho_f[m * nyz + pf] * LMF(a, 9, pc + i0 * mps));
    for (int c = 0; c < 3; c++) {
      const float* T = TT[c];
      float dxt = m == 0 ? rDH * (A0M * T[3 * nyz + pf] - A1 * T[2 * nyz + pf] - A2 * T[1 * nyz + pf] - A3 * T[0 * nyz + pf])
                         : rDH * (A0P * T[3 * nyz + pf] + A1 * T[4 * nyz + pf] + A2 * T[5 * nyz + pf] + A3 * T[6 * nyz + pf]);
      float v2[7], v3[7];
      for (int l = -3; l <= 3; l++) {
        v2[l + 3] = F.W[(3 + c) * 2 * nyz + m * nyz + pf + l];
        v3[l + 3] = F.W[(6 + c) * 2 * nyz + m * nyz + pf + (long)l * ny];
      }
      float dzt;
      if (a.freenode && km <= 3) {
        v3[km + 2] = 0.0f;
        for (int l = km + 3; l < 7; l++) v3[l] = -v3[2 * (km + 2) - l];
        dzt = drp(&v3[3], 1, FZ) * rDH;
      } else {
        dzt = (F.rup_index_z[q] % 7 ? d22(&v3[3], 1, FZ) : drp(&v3[3], 1, FZ)) * rDH;
      }
      float dyt = (F.rup_index_y[q] % 7 ? d22(&v2[3], 1, FY) : drp(&v2[3], 1, FY)) * rDH;
      F.hW[c * 2 * nyz + m * nyz + pf] = (dxt + dyt + dzt) * rrj;
    }
    return;
  }
  // split-node stress rates of side part-14; the x-derivatives of the two
  // sides are coupled on the surface row, so each side forms both
  const int ms = part - 14;

  // ---- stress rates (fault_dstrs_f/b)
  auto wv = [&](int c, int i) -> const float* { return &W[c * pyz + pc + i * wps]; };
  auto fv = [&](int c, int m) -> const float* { return &F.W[c * 2 * nyz + m * nyz + pf]; };
  float DxV[8][3], DyV[8][3], DzV[8][3];  // entries 3 and 4: the two sides of the split node
  for (int c = 0; c < 3; c++) {
    if (FX > 0) {
      DxV[4][c] = (wv(c, i0 + 1)[0] - fv(c, 1)[0]) * rDH;
      DxV[3][c] = 0.0f;
    } else {
      DxV[3][c] = (fv(c, 0)[0] - wv(c, i0 - 1)[0]) * rDH;
      DxV[4][c] = 0.0f;
    }
    for (int m = 0; m < 2; m++) DyV[3 + m][c] = (F.rup_index_y[q] % 7 ? d22(fv(c, m), 1, FY) : drp(fv(c, m), 1, FY)) * rDH;
  }
  if (a.freenode && km == 1) {
    const float* VX1 = &F.matVx2Vz_f[j1 * 9];
    const float* VY1 = &F.matVy2Vz_f[j1 * 9];
    const float* VX2 = &F.matVx2Vz_f[(ny + j1) * 9];
    const float* VY2 = &F.matVy2Vz_f[(ny + j1) * 9];
    float o1[3], o2[3], o3[3];
    if (FX > 0) {
      mv3(&F.matP2Mf[(0 * ny + j1) * 9], DxV[4], o1); mv3(&F.matP2Mf[(1 * ny + j1) * 9], DyV[4], o2); mv3(&F.matP2Mf[(2 * ny + j1) * 9], DyV[3], o3);
      for (int c = 0; c < 3; c++) DxV[3][c] = o1[c] + o2[c] - o3[c];
    } else {
      mv3(&F.matM2Pf[(0 * ny + j1) * 9], DxV[3], o1); mv3(&F.matM2Pf[(1 * ny + j1) * 9], DyV[3], o2); mv3(&F.matM2Pf[(2 * ny + j1) * 9], DyV[4], o3);
      for (int c = 0; c < 3; c++) DxV[4][c] = o1[c] + o2[c] - o3[c];
    }
    mv3(VX1, DxV[3], o1); mv3(VY1, DyV[3], o2);
    for (int c = 0; c < 3; c++) DzV[3][c] = o1[c] + o2[c];
    mv3(VX2, DxV[4], o1); mv3(VY2, DyV[4], o2);
    for (int c = 0; c < 3; c++) DzV[4][c] = o1[c] + o2[c];
  } else {
    for (int c = 0; c < 3; c++) {  // off the surface row each side needs only its own
      const float* f = fv(c, ms);
      float dz;
      if (a.freenode && km == 2) dz = d22(f, ny, FZ) * rDH;
      else if (a.freenode && km == 3) dz = d24(f, ny, FZ) * rDH;
      else dz = (F.rup_index_z[q] % 7 ? d22(f, ny, FZ) : drp(f, ny, FZ)) * rDH;
      if (ms == 0) DzV[3][c] = dz;
      else DzV[4][c] = dz;
    }
  }
  // this side's derivatives, selected with constant indices (a runtime
  // index into DxV/DyV/DzV would put them in scratch memory)
  float dx[3], dy[3], dz[3];
  for (int c = 0; c < 3; c++) {
    dx[c] = ms == 0 ? DxV[3][c] : DxV[4][c];
    dy[c] = ms == 0 ? DyV[3][c] : DyV[4][c];
    dz[c] = ms == 0 ? DzV[3][c] : DzV[4][c];
  }
  if (!(a.freenode && km == 1)) {
    // DxV_T1: D11^-1 (dT1/dt - D12 DyV - D13 DzV) of this side
    float dtT1[3] = {0.f, 0.f, 0.f};
    if (F.faultgrid[q]) { dtT1[0] = F.hT11[pf]; dtT1[1] = F.hT12[pf]; dtT1[2] = F.hT13[pf]; }
    float o1[3], o2[3], o3[3];
    mv3(&F.matT1toVx[(ms * nyz + pf) * 9], dtT1, o1);
    mv3(&F.matVytoVx[(ms * nyz + pf) * 9], dy, o2);
    mv3(&F.matVztoVx[(ms * nyz + pf) * 9], dz, o3);
    for (int c = 0; c < 3; c++) dx[c] = o1[c] - o2[c] - o3[c];
  }
  // split-node hT2*, hT3* (fault_dstrs_f.cu:874-1007)
  {
    const int m = ms;
    for (int row = 1; row <= 2; row++) {
      float o1[3], o2[3], o3[3];
      mv3(&F.Dmat[((long)(m * 9 + row * 3 + 0) * nyz + pf) * 9], dx, o1);
      mv3(&F.Dmat[((long)(m * 9 + row * 3 + 1) * nyz + pf) * 9], dy, o2);
      mv3(&F.Dmat[((long)(m * 9 + row * 3 + 2) * nyz + pf) * 9], dz, o3);
      for (int c = 0; c < 3; c++) F.hW[(3 * row + c) * 2 * nyz + m * nyz + pf] = o1[c] + o2[c] + o3[c];
    }
  }
}

// fault_post: RK of the fault-slab planes + fault_rk_cu + fault_rk_stage2_cu
// (rk4.cu:466-789) + fault2wave (transform.cu:224-403) + Cerjan.
__global__ void __launch_bounds__(256) fault_post_kernel(FArgs a) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.nj || k >= a.nk) return;
  const FaultDev& F = a.F;
  long q = j + (long)k * a.nj;
  const int j1 = j + 3, k1 = k + 3, ny = a.ny, i0 = a.i0;
  const long nyz = a.nyz, pyz = a.pyz, wps = a.wps, mps = a.mps, pf = j1 + (long)k1 * ny, pv = pf * 3;
  const long pc = a.jpad + j1 + (long)k1 * a.py;  // volume column
  const double RA[4] = {0.5, 0.5, 1.0, 1e10}, RB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const int irk = a.irk;
  const float rka = (float)(RA[irk] * a.DTd), rkb = (float)(RB[irk] * a.DTd);
  float dampj = a.cerjan ? fminf(a.cj[1][j1], a.cj[2][k1]) : 1.0f;
  // volume RK of the slab planes, one plane per block layer (blockIdx.z
  // 0-5); layer 6 does the split nodes and fault2wave
  if (blockIdx.z < 6) {
    const int pl = blockIdx.z < 3 ? blockIdx.z : blockIdx.z + 1;
    int ip = i0 - 3 + pl;
    const long p = pc + ip * wps;
    float damp = a.cerjan ? fminf(a.cj[0][ip], dampj) : 1.0f;
    for (int c = 0; c < 9; c++) {
      long x = c * pyz + p;
      float h = a.slabH[(long)pl * 9 * pyz + c * pyz + pc];
      float w = rk_combine(irk, rka, rkb, a.Win[x], h, a.S, a.T, x);
      a.Wi0[x] = w * damp;
    }
    return;
  }
  float s1[3] = {F.vec_s1[pv], F.vec_s1[pv + 1], F.vec_s1[pv + 2]};
  float s2[3] = {F.vec_s2[pv], F.vec_s2[pv + 1], F.vec_s2[pv + 2]};
  bool uni = F.united[q] != 0;
  if (!uni) {
    float vis = a.viscosity * (float)a.DTd;
    for (int m = 0; m < 2; m++)
      for (int c = 0; c < 9; c++) {
        long x = c * 2 * nyz + m * nyz + pf;
        float h = F.hW[x];
        float w0 = F.W[x];
        if (irk == 0) F.mW[x] = w0;
        float w = rk_combine(irk, rka, rkb, w0, h, F.mW, F.tW, x);
        if (c >= 3) w += vis * h;
        F.W[x] = w;
      }
  }
  float dV[3];
  for (int c = 0; c < 3; c++) dV[c] = F.W[c * 2 * nyz + nyz + pf] - F.W[c * 2 * nyz + pf];
  float vs1 = dV[0] * s1[0] + dV[1] * s1[1] + dV[2] * s1[2];
  float vs2 = dV[0] * s2[0] + dV[1] * s2[1] + dV[2] * s2[2];
  F.hslip[q] = sqrtf(vs2 * vs2 + vs1 * vs1);
  F.Vs1[q] = vs1;
  F.Vs2[q] = vs2;
  float rkb1 = (float)(RB[irk]);
  if (irk == 0) { F.Ts1[q] = rkb1 * F.tTs1[q]; F.Ts2[q] = rkb1 * F.tTs2[q]; F.Tn[q] = rkb1 * F.tTn[q]; }
  else { F.Ts1[q] += rkb1 * F.tTs1[q]; F.Ts2[q] += rkb1 * F.tTs2[q]; F.Tn[q] += rkb1 * F.tTn[q]; }
  if (uni) return;
  // fault2wave into the stage output at i0
  const long p = pc + i0 * mps, pw = pc + i0 * wps;
  float mt[3][3] = {{LMF(a, 0, p), LMF(a, 3, p), LMF(a, 6, p)},
                    {LMF(a, 1, p), LMF(a, 4, p), LMF(a, 7, p)},
                    {LMF(a, 2, p), LMF(a, 5, p), LMF(a, 8, p)}};
  float rj = 1.0f / LMF(a, 9, p);
  float inv[3][3];
  inv[0][0] = mt[1][1] * mt[2][2] - mt[2][1] * mt[1][2];
  inv[0][1] = mt[2][1] * mt[0][2] - mt[0][1] * mt[2][2];
  inv[0][2] = mt[0][1] * mt[1][2] - mt[0][2] * mt[1][1];
  inv[1][0] = mt[1][2] * mt[2][0] - mt[1][0] * mt[2][2];
  inv[1][1] = mt[0][0] * mt[2][2] - mt[2][0] * mt[0][2];
  inv[1][2] = mt[1][0] * mt[0][2] - mt[0][0] * mt[1][2];
  inv[2][0] = mt[1][0] * mt[2][1] - mt[1][1] * mt[2][0];
  inv[2][1] = mt[2][0] * mt[0][1] - mt[0][0] * mt[2][1];
  inv[2][2] = mt[0][0] * mt[1][1] - mt[0][1] * mt[1][0];
  float det = 1.0f / (inv[0][0] * mt[0][0] + inv[0][1] * mt[1][0] + inv[0][2] * mt[2][0]);
  float tr[3][3];
  tr[0][0] = F.T11[3 * nyz + pf]; tr[1][0] = F.T12[3 * nyz + pf]; tr[2][0] = F.T13[3 * nyz + pf];
  for (int c = 0; c < 3; c++) {
    tr[c][1] = (F.W[(3 + c) * 2 * nyz + pf] + F.W[(3 + c) * 2 * nyz + nyz + pf]) * 0.5f;
    tr[c][2] = (F.W[(6 + c) * 2 * nyz + pf] + F.W[(6 + c) * 2 * nyz + nyz + pf]) * 0.5f;
  }
  float sg[3][3];
  for (int r = 0; r < 3; r++)
    for (int cc = 0; cc < 3; cc++) sg[r][cc] = (tr[r][0] * inv[0][cc] + tr[r][1] * inv[1][cc] + tr[r][2] * inv[2][cc]) * det * rj;
  float damp = a.cerjan ? fminf(a.cj[0][i0], dampj) : 1.0f;
  float* W = a.Wi0;
  for (int c = 0; c < 3; c++) W[c * pyz + pw] = (F.W[c * 2 * nyz + pf] + F.W[c * 2 * nyz + nyz + pf]) * 0.5f * damp;
  W[3 * pyz + pw] = sg[0][0] * damp;
  W[4 * pyz + pw] = sg[1][1] * damp;
  W[5 * pyz + pw] = sg[2][2] * damp;
  W[6 * pyz + pw] = (sg[0][1] + sg[1][0]) * 0.5f * damp;
  W[7 * pyz + pw] = (sg[0][2] + sg[2][0]) * 0.5f * damp;
  W[8 * pyz + pw] = (sg[1][2] + sg[2][1]) * 0.5f * damp;
}

// cal_rup_sensor_D_cu / cal_rup_sensor_magn_cu (rup_sensor.cu:7-86).
// Thermal pressurization (thermpress_cu, thermpress.cu:22-233), once per
// step before the stages: per fault node, the temperature and pore-pressure
// rises on TP_n points x = i dh normal to the fault (mirror symmetric at
// both ends), driven by the frictional heating tau V spread as a Gaussian of
// half-width h, in explicit Euler substeps over DT; tau takes the last
// step's friction coefficient and total normal stress plus the fault's pore
// pressure.  FP64, updated in place level by level: the old values of the
// levels either side ride in registers.  With TP_n = 1 dh is infinite and
// the diffusion terms vanish.
__global__ void __launch_bounds__(256) thermpress_kernel(FaultDev F, int nj, int nk, double DT) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= nj || k >= nk) return;
  const long q = j + (long)k * nj, n2 = (long)nj * nk;
  if (F.united[q]) return;
  const double Lamb = 0.1e6, rho_c = 2.7e6, h = 20e-3, alpha_th = 1.0e-6, SQRT2PI = 2.506628274631000;
  const double h2 = h * h;
  const int n = F.TP_n;
  const double v1 = F.Vs1[q], v2 = F.Vs2[q];
  const double rate = sqrt(v1 * v1 + v2 * v2);
  const double alpha_hy = F.TP_hy[q];
  const double alpha_max = fmax(alpha_th, alpha_hy);
  const double w = sqrt(4.0 * alpha_max * 25.0);
  const double dh = n > 1 ? w / (double)(n - 1) : (double)INFINITY;
  const double dh2 = dh * dh;
  const int nsub = (int)(DT / (0.1 * dh2 / alpha_max)) + 1;
  const double dt = DT / (double)nsub;
  const double tn = F.Tn[q], fr = F.friction[q];
  double* T = F.TP_T;
  double* P = F.TP_P;
  for (int s = 0; s < nsub; s++) {
    const double tauV = -fr * (tn + P[q]) * rate;
    const double cw = tauV / (rho_c * h * SQRT2PI);
    double To = T[q], Po = P[q];                                  // level i
    double Tl = 0.0, Pl = 0.0;                                    // level i-1 (old)
    double Tr = n > 1 ? T[n2 + q] : 0.0, Pr = n > 1 ? P[n2 + q] : 0.0;  // level i+1 (old)
    for (int i = 0; i < n; i++) {
      const double Tm = i == 0 ? Tr : Tl, Pm = i == 0 ? Pr : Pl;
      const double Tp = i == n - 1 ? Tl : Tr, Pp = i == n - 1 ? Pl : Pr;
      const double x = n > 1 ? i * dh : 0.0;
      const double om = exp(-0.5 * x * x / h2) * cw;
      const double dTv = alpha_th * ((Tp - 2.0 * To + Tm) / dh2) + om;
      const double dPv = alpha_hy * ((Pp - 2.0 * Po + Pm) / dh2) + Lamb * dTv;
      T[i * n2 + q] = To + dt * dTv;
      P[i * n2 + q] = Po + dt * dPv;
      Tl = To; Pl = Po; To = Tr; Po = Pr;
      if (i + 2 < n) { Tr = T[(i + 2) * n2 + q]; Pr = P[(i + 2) * n2 + q]; }
    }
  }
}

__global__ void rup_sensor_d_kernel(FaultDev F, int nj, int nk, int g0, int g1, int g2, int g3) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= nj || k >= nk) return;
  int gj1 = j + 1, gk1 = k + 1;
  if (!(gj1 >= g0 + 3 && gj1 <= g1 - 3 && gk1 >= g2 + 3 && gk1 <= g3 - 3)) return;
  long q = j + (long)nj * k;
  if (F.united[q]) return;
  const float* p = F.Ts1;
  F.rsD_y[q] = 0.25f * (-p[q + 1] + 2.0f * p[q] - p[q - 1]);
  F.rsD_z[q] = 0.25f * (-p[q + nj] + 2.0f * p[q] - p[q - nj]);
}
__global__ void rup_sensor_m_kernel(FaultDev F, int nj, int nk, int g0, int g1, int g2, int g3) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= nj || k >= nk) return;
  int gj1 = j + 1, gk1 = k + 1;
  if (!(gj1 >= g0 + 3 && gj1 <= g1 - 3 && gk1 >= g2 + 3 && gk1 <= g3 - 3)) return;
  long q = j + (long)nj * k;
  const float* Dy = F.rsD_y;
  const float* Dz = F.rsD_z;
  float D1 = Dy[q] - Dy[q + 1], D2 = Dy[q] - Dy[q - 1];
  float mg = 0.5f * (D1 * D1 + D2 * D2);
  D1 = Dz[q] - Dy[q + nj];  // rup_sensor.cu:75-76 pairs Dz with Dy
  D2 = Dz[q] - Dy[q - nj];
  mg += 0.5f * (D1 * D1 + D2 * D2);
  float p0 = F.Ts1[q] * F.Ts1[q];
  F.rup_sensor[q] = mg / (p0 + 1e-16f) + 1e-16f;
  if (F.united[q]) F.rup_sensor[q] = 0;
}

// smooth_gauss_volume0 (smooth.cu:15-62): mirror the split nodes, the
// fault-plane T1 and the wavefield planes i0-3..i0+3 across the free
// surface into the three ghost rows above it (one thread per j).
__global__ void smooth_ext_kernel(FArgs a, float* __restrict__ Wv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.nj) return;
  const FaultDev& F = a.F;
  const long nyz = a.nyz, ny = a.ny;
  const int j1 = j + 3, k1 = a.nk - 1 + 3;
  for (int l = 0; l < 9; l++)
    for (int m = 0; m < 2; m++) {
      float* f = F.W + (long)l * 2 * nyz + m * nyz;
      for (int n = 1; n <= 3; n++) f[(k1 + n) * ny + j1] = 2.0f * f[k1 * ny + j1] - f[(k1 - n) * ny + j1];
    }
  float* T[3] = {F.T11 + 3 * nyz, F.T12 + 3 * nyz, F.T13 + 3 * nyz};
  for (int t = 0; t < 3; t++)
    for (int n = 1; n <= 3; n++) T[t][(k1 + n) * ny + j1] = 2.0f * T[t][k1 * ny + j1] - T[t][(k1 - n) * ny + j1];
  for (int l = 0; l < 9; l++)
    for (int i1 = a.i0 - 3; i1 <= a.i0 + 3; i1++) {
      float* w = Wv + i1 * a.wps + (long)l * a.pyz + a.jpad + j1;
      for (int n = 1; n <= 3; n++) w[(long)(k1 + n) * a.py] = 2.0f * w[(long)k1 * a.py] - w[(long)(k1 - n) * a.py];
    }
}

// smooth_gauss_volume1 (smooth.cu:64-296): T2*/T3* of each split node as a
// half-volume Gaussian average of the node's own plane and the conservative
// tractions J (sigma . grad eta), J (sigma . grad zeta) of the three volume
// planes on its side.
__global__ void __launch_bounds__(256) smooth_avg_kernel(FArgs a, const float* __restrict__ Wv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.nj || k >= a.nk) return;
  const FaultDev& F = a.F;
  const long nyz = a.nyz, ny = a.ny, pyz = a.pyz;
  const int j1 = j + 3, k1 = k + 3;
  for (int m = 0; m < 2; m++) {
    const float* wp = F.smooth_w + m * 4 * 49;
    float Ts[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int aa = 0; aa < 7; aa++)
      for (int b = 0; b < 7; b++) {
        const long q = (long)(k1 + aa - 3) * ny + (j1 + b - 3);
        const float w = wp[aa * 7 + b];
        for (int c = 0; c < 6; c++) Ts[c] += F.W[(3 + c) * 2 * nyz + m * nyz + q] * w;
      }
    for (int l = 1; l <= 3; l++) {
      const int i = m == 0 ? a.i0 - l : a.i0 + l;
      for (int aa = 0; aa < 7; aa++)
        for (int b = 0; b < 7; b++) {
          const long pc = a.jpad + (j1 + b - 3) + (long)(k1 + aa - 3) * a.py;
          const long pm = pc + i * a.mps, pw = pc + i * a.wps;
          float E[3][3];  // columns grad xi, grad eta, grad zeta
          for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) E[r][cc] = LMF(a, 3 * cc + r, pm);
          const float sxx = Wv[3 * pyz + pw], syy = Wv[4 * pyz + pw], szz = Wv[5 * pyz + pw];
          const float sxy = Wv[6 * pyz + pw], sxz = Wv[7 * pyz + pw], syz = Wv[8 * pyz + pw];
          const float S[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}};
          const float w = LMF(a, 9, pm) * wp[(l * 7 + aa) * 7 + b];
          for (int r = 0; r < 3; r++) {
            const float t1 = S[r][0] * E[0][1] + S[r][1] * E[1][1] + S[r][2] * E[2][1];
            const float t2 = S[r][0] * E[0][2] + S[r][1] * E[1][2] + S[r][2] * E[2][2];
            Ts[r] += t1 * w;
            Ts[3 + r] += t2 * w;
          }
        }
    }
    for (int c = 0; c < 6; c++) F.Ws[(3 + c) * 2 * nyz + m * nyz + (long)k1 * ny + j1] = Ts[c];
  }
}

// smooth_gauss_volume2 (smooth.cu:298-345): blend the smoothed tractions in
// by min(slip/Dc, 1) inside Fault_grid; outside it the smoothed value wins.
__global__ void smooth_blend_kernel(FArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.nj || k >= a.nk) return;
  const FaultDev& F = a.F;
  const int gj = j + 1, gk = k + 1;
  float beta = 1.0f;
  if (gj >= a.G[0] && gj <= a.G[1] && gk >= a.G[2] && gk <= a.G[3])
    beta = fminf((float)F.slip[(long)k * a.nj + j] / a.Dc0, 1.0f);
  const long pf = (long)(k + 3) * a.ny + (j + 3);
  for (int l = 3; l < 9; l++)
    for (int m = 0; m < 2; m++) {
      const long x = (long)l * 2 * a.nyz + m * a.nyz + pf;
      F.W[x] = beta * F.Ws[x] + (1.0f - beta) * F.W[x];
    }
}

__global__ void fill_kernel(float* p, long n, float v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

FArgs make_fargs(Ctx& c) {
  FArgs a{};
  a.F = c.f;
  a.M = c.p.MetricBF16 ? c.met_fault_base : (const float*)c.met;
  a.mVx = c.matVx2Vz; a.mVy = c.matVy2Vz;
  a.slabH = c.slabH;
  for (int x = 0; x < 3; x++) a.cj[x] = c.cj[x];
  a.ni = c.d.ni; a.nj = c.d.nj; a.nk = c.d.nk; a.nx = c.d.nx; a.ny = c.d.ny; a.nz = c.d.nz; a.i0 = c.d.i0;
  a.nyz = c.d.nyz; a.pyz = c.d.pyz; a.wps = 9 * c.d.pyz; a.mps = 13 * c.d.pyz;
  a.py = c.d.py; a.jpad = c.d.jpad;
  for (int q = 0; q < 4; q++) a.G[q] = c.p.Fault_grid[q];
  a.mlp = c.p.NewtonMode == 2 ? (const MlpDev*)c.mlp_dev : nullptr;
  a.Dc0 = (float)c.p.Dc;
  a.NY = c.p.NY; a.NZ = c.p.NZ;
  a.DH = (float)c.p.DH; a.rDH = (float)(1.0 / c.p.DH); a.DT = (float)c.p.DT; a.DTd = c.p.DT;
  a.freenode = c.p.FreeSurface ? 1 : 0;
  a.friction = c.p.Friction_type;
  a.stress_type = c.p.INPORT_STRESS_TYPE;
  a.cerjan = c.p.useCerjan ? 1 : 0;
  a.mu_d = (float)c.p.mu_d; a.viscosity = (float)c.p.viscosity; a.smooth_load_T = (float)c.p.smooth_load_T;
  a.Vini = (float)c.p.Vini;
  a.f0 = c.p.f0; a.V0 = c.p.V0; a.fw = c.p.fw;
  a.newton_mode = c.p.NewtonMode; a.newton_tol = c.p.NewtonTol; a.newton_fallback = c.p.NewtonFallbackIter;
  a.mlp_gate = c.p.NewtonMLPGate;
  return a;
}

dim3 fgrid(const Ctx& c) { return dim3((c.d.nj + 63) / 64, (c.d.nk + 3) / 4, 1); }

}  // namespace

// Arguments of the trial kernels of one stage: the stage input Win (read, and
// written at plane i0 by fault2wave), the step and stage, the FD flags.
static FArgs trial_args(Ctx& c, const float* Win, int it, int irk, int fx, int fy, int fz) {
  FArgs a = make_fargs(c);
  a.W = Win; a.Wi0 = const_cast<float*>(Win);
  a.it = it; a.irk = irk; a.FX = fx; a.FY = fy; a.FZ = fz;
  return a;
}

void launch_fault_pre(Ctx& c, const float* Win, int it, int irk, int fx, int fy, int fz) {
  FArgs a = trial_args(c, Win, it, irk, fx, fy, fz);
  c.prof.begin(c.stream, PK_FAULT_PRE);
  hipLaunchKernelGGL(fault_pre_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a);
  c.prof.end(c.stream);
}

void launch_fault_rs(Ctx& c, const float* Win, int it, int irk, int fx, int fy, int fz) {
  FArgs a = trial_args(c, Win, it, irk, fx, fy, fz);
  c.prof.begin(c.stream, PK_FAULT_RS);
  hipLaunchKernelGGL(fault_rs_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a);
  c.prof.end(c.stream);
}

void launch_fault_deriv(Ctx& c, const float* Win, int fx, int fy, int fz) {
  FArgs a = make_fargs(c);
  a.W = Win;
  a.FX = fx; a.FY = fy; a.FZ = fz;
  c.prof.begin(c.stream, PK_FAULT_DERIV);
  const dim3 g = fgrid(c);
  const int tiles = (int)(g.x * g.y);
  // per-plane velocity parts, per-plane stress parts, split-node parts of each tile
  hipLaunchKernelGGL(fault_deriv_kernel<0>, dim3(tiles * kDerivGroupParts[0]), dim3(64, 4, 1), 0, c.stream, a, (int)g.x,
                     tiles * kDerivGroupParts[0]);
  hipLaunchKernelGGL(fault_deriv_kernel<1>, dim3(tiles * kDerivGroupParts[1]), dim3(64, 4, 1), 0, c.stream, a, (int)g.x,
                     tiles * kDerivGroupParts[1]);
  hipLaunchKernelGGL(fault_deriv_kernel<2>, dim3(tiles * kDerivGroupParts[2]), dim3(64, 4, 1), 0, c.stream, a, (int)g.x,
                     tiles * kDerivGroupParts[2]);
  c.prof.end(c.stream);
}

void launch_fault_post(Ctx& c, const StageBufs& b, int irk) {
  FArgs a = make_fargs(c);
  a.Win = b.Win; a.Wi0 = b.Wout; a.S = b.S; a.T = b.T;
  a.irk = irk;
  c.prof.begin(c.stream, PK_FAULT_POST);
  dim3 g = fgrid(c);
  g.z = 7;  // six slab planes and the split-node part
  hipLaunchKernelGGL(fault_post_kernel, g, dim3(64, 4, 1), 0, c.stream, a);
  c.prof.end(c.stream);
}

void launch_thermpress(Ctx& c) {
  dim3 g((c.d.nj + 63) / 64, (c.d.nk + 3) / 4, 1);
  hipLaunchKernelGGL(thermpress_kernel, g, dim3(64, 4, 1), 0, c.stream, c.f, c.d.nj, c.d.nk, c.p.DT);
}

void launch_rup_sensor(Ctx& c) {
  dim3 g((c.d.nj + 63) / 64, (c.d.nk + 3) / 4, 1);
  const int* G = c.p.Fault_grid;
  c.prof.begin(c.stream, PK_RUP_SENSOR);
  hipLaunchKernelGGL(rup_sensor_d_kernel, g, dim3(64, 4, 1), 0, c.stream, c.f, c.d.nj, c.d.nk, G[0], G[1], G[2], G[3]);
  hipLaunchKernelGGL(rup_sensor_m_kernel, g, dim3(64, 4, 1), 0, c.stream, c.f, c.d.nj, c.d.nk, G[0], G[1], G[2], G[3]);
  c.prof.end(c.stream);
}

void launch_fault_smooth(Ctx& c) {
  FArgs a = make_fargs(c);
  if (c.p.FreeSurface)
    hipLaunchKernelGGL(smooth_ext_kernel, dim3((c.d.nj + 255) / 256), dim3(256), 0, c.stream, a, c.cur);
  hipLaunchKernelGGL(smooth_avg_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a, (const float*)c.cur);
  hipLaunchKernelGGL(smooth_blend_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a);
}

void launch_fill(Ctx& c, float* p, long n, float v) {
  long blocks = std::min<long>((n + 255) / 256, 65536);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, c.stream, p, n, v);
}

}  // namespace cgfdm

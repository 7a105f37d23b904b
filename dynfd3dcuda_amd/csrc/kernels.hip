// HIP kernels of the CGFDM time step for CDNA4 (gfx950).
//
// Wave stage (the HBM-bound hot path): one fused kernel computes the
// velocity-stress RHS on the curvilinear grid (wave_deriv_cu, wave.cu:8-494),
// adds the ADE-PML terms and advances their auxiliary fields
// (abs_deriv_{x,y,z}_cu, pml{x,y,z}.cu), applies the RK4 stage combine
// (wave_rk_cu/RK_Syn, rk4.cu:20-33,154-257) and the Cerjan sponge
// (apply_cerjan_cu, cerjan.cu:166-200), reading each wavefield word once from
// HBM.  Blocks stream along i (the fault-normal axis) over a TJ x TK tile of
// (j,k) columns: the five planes the x-stencil needs sit in a register queue,
// the current plane is staged in LDS for the y- and z-stencils.  Points of
// the fault slab (i0-3..i0+3 without i0) keep their RHS in a slab buffer so
// the split-node kernels can overwrite it (fault_dvelo/fault_dstrs) before
// the slab is advanced by fault_post.
//
// Fault (2-D, one thread per (j,k) node): fault_pre = state_rk + trial_sw |
// trial_rs + fault2wave; fault_deriv = fault_dvelo + fault_dstrs_{f,b};
// fault_post = slab RK + fault_rk + fault_rk_stage2 + fault2wave (+Cerjan).
#include "cgfdm_internal.h"
#include <hip/hip_bf16.h>

namespace cgfdm {

namespace {

// DRP/MacCormack coefficients (macdrp.h:23-40).
constexpr float C1 = 0.29164704779069994f, C2 = 0.66674514216519998f, C3 = 1.2501177132452999f;
constexpr float C4 = 0.33341180882999999f, C5 = 0.041686285540599999f;
constexpr float A0P = -1.250039237746600f, A0M = 1.250039237746600f;
constexpr float A1 = 1.541764761036000f, A2 = -0.333411808829999f, A3 = 0.0416862855405999f;
constexpr float C24_1 = 7.0f / 6.0f, C24_2 = 4.0f / 3.0f, C24_3 = 1.0f / 6.0f;

constexpr int TJ = 64;  // one wavefront per k-row of the tile
constexpr int TK = 8;

__device__ __forceinline__ float ldm(const float* p, long i) { return p[i]; }
__device__ __forceinline__ float ldm(const uint16_t* p, long i) {
  return __uint_as_float(((unsigned)p[i]) << 16);
}

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

template <class MT>
struct WaveArgs {
  const float* __restrict__ Win;
  float* __restrict__ Wout;
  const float* __restrict__ S;
  float* __restrict__ T;
  const MT* __restrict__ met;
  float* __restrict__ slabH;
  float* __restrict__ rhs;
  const WaveCold* __restrict__ cold;
  int ni, nj, nk, ny, ii0;
  long nxyz, nyz;
  float rDH, rka, rkb;
  int irk, freenode, cerjan, fault, rhs_only, ichunk, pml;
};

// Traction-image hV for the top three layers (wave.cu:398-487): the
// conservative traction vectors along each axis, zero at the surface and
// odd-mirrored above it.  Direct (cached) loads: only 3 of nk layers.
template <int FX, int FY, int FZ, class MT>
__device__ void traction_image(const WaveArgs<MT>& a, long p, int k, float rb1, float rb2, float rb3, float rrho, float* h) {
  const float* W = a.Win;
  const MT* M = a.met;
  const long nxyz = a.nxyz;
  int ik = a.nk - 1 - k;
  float rrhojac = rrho / ldm(M, 9 * nxyz + p);
  const long st[3] = {a.nyz, 1, (long)a.ny};
  const int fl[3] = {FX, FY, FZ};
  float d[3][3];  // d[axis][component]
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    float v[3][7];
#pragma unroll
    for (int l = 0; l < 7; l++) {
      long q = p + (l - 3) * st[ax];
      float e0 = ldm(M, (3 * ax + 0) * nxyz + q), e1 = ldm(M, (3 * ax + 1) * nxyz + q), e2 = ldm(M, (3 * ax + 2) * nxyz + q);
      float J = ldm(M, 9 * nxyz + q);
      float txx = W[3 * nxyz + q], tyy = W[4 * nxyz + q], tzz = W[5 * nxyz + q];
      float txy = W[6 * nxyz + q], txz = W[7 * nxyz + q], tyz = W[8 * nxyz + q];
      v[0][l] = (e0 * txx + e1 * txy + e2 * txz) * J;
      v[1][l] = (e0 * txy + e1 * tyy + e2 * tyz) * J;
      v[2][l] = (e0 * txz + e1 * tyz + e2 * tzz) * J;
    }
    if (ax == 2) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (ik == 0) { v[c][3] = 0.0f; v[c][6] = -v[c][0]; v[c][5] = -v[c][1]; v[c][4] = -v[c][2]; }
        else if (ik == 1) { v[c][4] = 0.0f; v[c][6] = -v[c][2]; v[c][5] = -v[c][3]; }
        else { v[c][5] = 0.0f; v[c][6] = -v[c][4]; }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
      d[ax][c] = a.rDH * (fl[ax] > 0 ? drp5<1>(v[c][2], v[c][3], v[c][4], v[c][5], v[c][6])
                                     : drp5<-1>(v[c][0], v[c][1], v[c][2], v[c][3], v[c][4]));
  }
#pragma unroll
  for (int c = 0; c < 3; c++) h[c] = (d[0][c] * rb1 + d[1][c] * rb2 + d[2][c] * rb3) * rrhojac;
}

// RK_Syn stage combine (rk4.cu:20-33) of one value.
__device__ __forceinline__ float rk_combine(int irk, float rka, float rkb, float w, float h, const float* S, float* T, long x) {
  if (irk == 0) {
    T[x] = w + rkb * h;
    return w + rka * h;
  }
  if (irk == 3) return T[x] + rkb * h;
  T[x] = T[x] + rkb * h;
  return S[x] + rka * h;
}

template <int FX, int FY, int FZ, class MT>
__global__ void __launch_bounds__(TJ* TK) wave_stage_kernel(WaveArgs<MT> a) {
  constexpr int HJL = FY > 0 ? 1 : 3;
  constexpr int HKL = FZ > 0 ? 1 : 3;
  constexpr int LJ = TJ + 4, LK = TK + 4;
  constexpr int QLO = FX > 0 ? -1 : -3;  // queue holds planes i+QLO .. i+QLO+4
  constexpr int CTR = -QLO;
  __shared__ float lds[9][LK][LJ];

  const int tj = threadIdx.x, tk = threadIdx.y;
  const int tid = tj + tk * TJ;
  const int j0 = blockIdx.x * TJ, k0 = blockIdx.y * TK;
  const int j = j0 + tj, k = k0 + tk;
  const int ni = a.ni, nj = a.nj, nk = a.nk, ny = a.ny;
  const long nxyz = a.nxyz, nyz = a.nyz;
  const bool act = j < nj && k < nk;
  const int ibeg = blockIdx.z * a.ichunk;
  const int iend = min(ni, ibeg + a.ichunk);
  if (ibeg >= iend) return;  // uniform per block
  const long col = (j + 3) + (long)(k + 3) * ny;
  const float* __restrict__ Win = a.Win;

  float q[5][9];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    long p = col + (long)(ibeg + QLO + m + 3) * nyz;
#pragma unroll
    for (int c = 0; c < 9; c++) q[m][c] = act ? Win[c * nxyz + p] : 0.0f;
  }

  for (int i = ibeg; i < iend; ++i) {
    float nq[9];
    const bool more = i + 1 < iend;
    {
      long p = col + (long)(i + QLO + 5 + 3) * nyz;
#pragma unroll
      for (int c = 0; c < 9; c++) nq[c] = (act && more) ? Win[c * nxyz + p] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 9; c++) lds[c][tk + HKL][tj + HJL] = q[CTR][c];
    // halo: 4 rows of LJ cells, then 4 side cells of each of the TK rows
    constexpr int NH = 4 * LJ + 4 * TK;
    for (int hh = tid; hh < NH; hh += TJ * TK) {
      int lk, lj;
      if (hh < 4 * LJ) {
        int r = hh / LJ;
        lk = r < HKL ? r : r + TK;
        lj = hh % LJ;
      } else {
        int h2 = hh - 4 * LJ;
        lk = HKL + h2 / 4;
        int cc = h2 % 4;
        lj = cc < HJL ? cc : cc + TJ;
      }
      int gj = j0 + lj - HJL, gk = k0 + lk - HKL;
      bool ok = gj >= -3 && gj < nj + 3 && gk >= -3 && gk < nk + 3;
      long p = (gj + 3) + (long)(gk + 3) * ny + (long)(i + 3) * nyz;
#pragma unroll
      for (int c = 0; c < 9; c++) lds[c][lk][lj] = ok ? Win[c * nxyz + p] : 0.0f;
    }
    __syncthreads();

    if (act) {
      const long p = col + (long)(i + 3) * nyz;
      const MT* __restrict__ M = a.met;
      const float xix = ldm(M, p), xiy = ldm(M, nxyz + p), xiz = ldm(M, 2 * nxyz + p);
      const float etx = ldm(M, 3 * nxyz + p), ety = ldm(M, 4 * nxyz + p), etz = ldm(M, 5 * nxyz + p);
      const float ztx = ldm(M, 6 * nxyz + p), zty = ldm(M, 7 * nxyz + p), ztz = ldm(M, 8 * nxyz + p);
      const float lam = ldm(M, 10 * nxyz + p), mu = ldm(M, 11 * nxyz + p), rrho = 1.0f / ldm(M, 12 * nxyz + p);
      const float l2m = lam + 2.0f * mu;
      const float rDH = a.rDH;
      const WaveCold* __restrict__ cold = a.cold;
      const float rb1 = a.pml ? 1.0f / cold->pB[0][i] : 1.0f;
      const float rb2 = a.pml ? 1.0f / cold->pB[1][j] : 1.0f;
      const float rb3 = a.pml ? 1.0f / cold->pB[2][k] : 1.0f;
      const int lj = tj + HJL, lk = tk + HKL;
      // one-sided DRP derivative of field c along each axis (macdrp.h:91-95)
      auto Dx = [&](int c) { return rDH * drp5<FX>(q[0][c], q[1][c], q[2][c], q[3][c], q[4][c]); };
      auto Dy = [&](int c) {
        return FY > 0 ? rDH * drp5<1>(lds[c][lk][lj - 1], lds[c][lk][lj], lds[c][lk][lj + 1], lds[c][lk][lj + 2], lds[c][lk][lj + 3])
                      : rDH * drp5<-1>(lds[c][lk][lj - 3], lds[c][lk][lj - 2], lds[c][lk][lj - 1], lds[c][lk][lj], lds[c][lk][lj + 1]);
      };
      auto Dz = [&](int c) {
        return FZ > 0 ? rDH * drp5<1>(lds[c][lk - 1][lj], lds[c][lk][lj], lds[c][lk + 1][lj], lds[c][lk + 2][lj], lds[c][lk + 3][lj])
                      : rDH * drp5<-1>(lds[c][lk - 3][lj], lds[c][lk - 2][lj], lds[c][lk - 1][lj], lds[c][lk][lj], lds[c][lk + 1][lj]);
      };
      const bool top = a.freenode && k >= nk - 3;
      // RHS accumulated derivative by derivative (wave.cu:133-276): each
      // field's derivative adds its momentum / Hooke terms, metric rows
      // pre-scaled by the PML 1/beta of their axis.
      float h[9];
#pragma unroll
      for (int c = 0; c < 9; c++) h[c] = 0.0f;
      accum_axis(h, Dx, xix * rb1, xiy * rb1, xiz * rb1, rrho, lam, l2m, mu, true);
      accum_axis(h, Dy, etx * rb2, ety * rb2, etz * rb2, rrho, lam, l2m, mu, true);
      accum_axis(h, Dz, ztx * rb3, zty * rb3, ztz * rb3, rrho, lam, l2m, mu, !top);
      if (top) {
        float dzv[3];
        if (k == nk - 1) {  // DzV from the traction-free condition (wave.cu:207-232)
          const long ij = ((j + 3) + (long)(i + 3) * ny) * 9;
          const float* mVx = cold->mVx;
          const float* mVy = cold->mVy;
          const float dxv[3] = {Dx(0), Dx(1), Dx(2)}, dyv[3] = {Dy(0), Dy(1), Dy(2)};
#pragma unroll
          for (int r = 0; r < 3; r++)
            dzv[r] = mVx[ij + 3 * r] * dxv[0] + mVx[ij + 3 * r + 1] * dxv[1] + mVx[ij + 3 * r + 2] * dxv[2]
                   + mVy[ij + 3 * r] * dyv[0] + mVy[ij + 3 * r + 1] * dyv[1] + mVy[ij + 3 * r + 2] * dyv[2];
        } else if (k == nk - 2) {  // L22 (wave.cu:233-238)
#pragma unroll
          for (int c = 0; c < 3; c++) dzv[c] = rDH * (FZ > 0 ? lds[c][lk + 1][lj] - lds[c][lk][lj] : lds[c][lk][lj] - lds[c][lk - 1][lj]);
        } else {  // L24 (wave.cu:239-245)
#pragma unroll
          for (int c = 0; c < 3; c++)
            dzv[c] = rDH * (FZ > 0 ? -C24_1 * lds[c][lk][lj] + C24_2 * lds[c][lk + 1][lj] - C24_3 * lds[c][lk + 2][lj]
                                   : C24_3 * lds[c][lk - 2][lj] - C24_2 * lds[c][lk - 1][lj] + C24_1 * lds[c][lk][lj]);
        }
        const float e0 = ztx * rb3, e1 = zty * rb3, e2 = ztz * rb3;
        h[3] += (l2m * e0 * dzv[0] + lam * e1 * dzv[1] + lam * e2 * dzv[2]);
        h[4] += (lam * e0 * dzv[0] + l2m * e1 * dzv[1] + lam * e2 * dzv[2]);
        h[5] += (lam * e0 * dzv[0] + lam * e1 * dzv[1] + l2m * e2 * dzv[2]);
        h[6] += (e1 * dzv[0] + e0 * dzv[1]) * mu;
        h[7] += (e2 * dzv[0] + e0 * dzv[2]) * mu;
        h[8] += (e2 * dzv[1] + e1 * dzv[2]) * mu;
        traction_image<FX, FY, FZ, MT>(a, p, k, rb1, rb2, rb3, rrho, h);
      }

      // ADE-PML (pmlx.cu:126-210 and the y/z variants): -aux/beta in the
      // RHS and the auxiliary ODE from the plain derivative along the axis.
      if (a.pml) {
        const int N = cold->N;
        const int n3[3] = {ni, nj, nk};
        const int pos3[3] = {i, j, k};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
          const int li = pos3[ax];
#pragma unroll
          for (int side = 0; side < 2; side++) {
            const int f = 2 * ax + side;
            const int off = side ? n3[ax] - N : 0;
            if (!cold->on[f] || li < off || li >= off + N) continue;
            float dd[9], Dr[9];
#pragma unroll
            for (int c = 0; c < 9; c++) dd[c] = ax == 0 ? Dx(c) : ax == 1 ? Dy(c) : Dz(c);
            if (ax == 0) axis_rhs(dd, xix, xiy, xiz, rrho, lam, mu, Dr);
            else if (ax == 1) axis_rhs(dd, etx, ety, etz, rrho, lam, mu, Dr);
            else axis_rhs(dd, ztx, zty, ztz, rrho, lam, mu, Dr);
            const int pi = ax == 0 ? i - off : i, pj = ax == 1 ? j - off : j, pk = ax == 2 ? k - off : k;
            const int pnj = ax == 1 ? N : nj, pnk = ax == 2 ? N : nk, pni = ax == 0 ? N : ni;
            const long ps = (long)pni * pnj * pnk;
            const long pa = pj + (long)pk * pnj + (long)pi * pnj * pnk;
            const float b1 = cold->pB[ax][li], d1 = cold->pD[ax][li], ad1 = -(cold->pA[ax][li] + d1);
            const float rbb = 1.0f / b1;
            float* Y = cold->auxY[f];
            float ah[9];
#pragma unroll
            for (int c = 0; c < 9; c++) {
              float y = Y[c * ps + pa];
              h[c] += -y * rbb;
              ah[c] = ad1 * y + d1 * Dr[c];
            }
            if (ax < 2 && a.freenode && k == nk - 1) {  // pmlx.cu:185-210
              const float* mat = ax == 0 ? cold->mVx : cold->mVy;
              const long ij = ((j + 3) + (long)(i + 3) * ny) * 9;
              float z1[3];
#pragma unroll
              for (int r = 0; r < 3; r++) z1[r] = mat[ij + 3 * r] * dd[0] + mat[ij + 3 * r + 1] * dd[1] + mat[ij + 3 * r + 2] * dd[2];
              const float c0 = d1 * b1;
              ah[3] += c0 * (l2m * ztx * z1[0] + lam * zty * z1[1] + lam * ztz * z1[2]);
              ah[4] += c0 * (lam * ztx * z1[0] + l2m * zty * z1[1] + lam * ztz * z1[2]);
              ah[5] += c0 * (lam * ztx * z1[0] + lam * zty * z1[1] + l2m * ztz * z1[2]);
              ah[6] += c0 * mu * (zty * z1[0] + ztx * z1[1]);
              ah[7] += c0 * mu * (ztz * z1[0] + ztx * z1[2]);
              ah[8] += c0 * mu * (ztz * z1[1] + zty * z1[2]);
            }
            if (!a.rhs_only) {
              float* S = cold->auxS[f];
              float* T = cold->auxT[f];
#pragma unroll
              for (int c = 0; c < 9; c++) {
                const long x = c * ps + pa;
                const float y = Y[x];
                if (a.irk == 0) S[x] = y;
                Y[x] = rk_combine(a.irk, a.rka, a.rkb, y, ah[c], S, T, x);
              }
            }
          }
        }
      }

      if (a.rhs_only) {
#pragma unroll
        for (int c = 0; c < 9; c++) a.rhs[c * nxyz + p] = h[c];
      } else if (a.fault && i >= a.ii0 - 3 && i <= a.ii0 + 3 && i != a.ii0) {
        long sp = (long)(i - (a.ii0 - 3)) * 9 * nyz + col;
#pragma unroll
        for (int c = 0; c < 9; c++) a.slabH[sp + c * nyz] = h[c];
      } else {
        float damp = 1.0f;
        if (a.cerjan) damp = fminf(cold->cj[0][i + 3], fminf(cold->cj[1][j + 3], cold->cj[2][k + 3]));
#pragma unroll
        for (int c = 0; c < 9; c++) {
          long x = c * nxyz + p;
          float w = rk_combine(a.irk, a.rka, a.rkb, q[CTR][c], h[c], a.S, a.T, x);
          a.Wout[x] = w * damp;
        }
      }
    }

    // advance the register queue by one plane
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int c = 0; c < 9; c++) q[m][c] = q[m + 1][c];
#pragma unroll
    for (int c = 0; c < 9; c++) q[4][c] = nq[c];
  }
}

template <int FX, int FY, int FZ, class MT>
void launch_wave_t(const WaveArgs<MT>& a, int nj, int nk, int ni, hipStream_t s) {
  dim3 block(TJ, TK, 1);
  dim3 grid((nj + TJ - 1) / TJ, (nk + TK - 1) / TK, (ni + a.ichunk - 1) / a.ichunk);
  hipLaunchKernelGGL((wave_stage_kernel<FX, FY, FZ, MT>), grid, block, 0, s, a);
}

template <class MT>
void launch_wave_dispatch(const WaveArgs<MT>& a, int fx, int fy, int fz, int nj, int nk, int ni, hipStream_t s) {
  int key = (fx > 0) * 4 + (fy > 0) * 2 + (fz > 0);
  switch (key) {
    case 0: launch_wave_t<-1, -1, -1, MT>(a, nj, nk, ni, s); break;
    case 1: launch_wave_t<-1, -1, 1, MT>(a, nj, nk, ni, s); break;
    case 2: launch_wave_t<-1, 1, -1, MT>(a, nj, nk, ni, s); break;
    case 3: launch_wave_t<-1, 1, 1, MT>(a, nj, nk, ni, s); break;
    case 4: launch_wave_t<1, -1, -1, MT>(a, nj, nk, ni, s); break;
    case 5: launch_wave_t<1, -1, 1, MT>(a, nj, nk, ni, s); break;
    case 6: launch_wave_t<1, 1, -1, MT>(a, nj, nk, ni, s); break;
    default: launch_wave_t<1, 1, 1, MT>(a, nj, nk, ni, s); break;
  }
}

template <class MT>
WaveArgs<MT> make_args(Ctx& c, const StageBufs& b, int irk, bool rhs_only, float* rhs_out) {
  WaveArgs<MT> a{};
  a.Win = b.Win; a.Wout = b.Wout; a.S = b.S; a.T = b.T;
  a.met = (const MT*)c.met;
  a.slabH = c.slabH;
  a.rhs = rhs_out;
  a.cold = (const WaveCold*)c.wave_cold;
  bool any = false;
  for (int f = 0; f < 6; f++) any |= c.pml.on[f] != 0;
  a.pml = any ? 1 : 0;
  a.ni = c.d.ni; a.nj = c.d.nj; a.nk = c.d.nk; a.ny = c.d.ny; a.ii0 = c.d.i0 - 3;
  a.nxyz = c.d.nxyz; a.nyz = c.d.nyz;
  a.rDH = (float)(1.0 / c.p.DH);
  const double RA[4] = {0.5, 0.5, 1.0, 1e10}, RB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  a.rka = (float)(RA[irk] * c.p.DT);
  a.rkb = (float)(RB[irk] * c.p.DT);
  a.irk = irk;
  a.freenode = c.p.FreeSurface ? 1 : 0;
  a.cerjan = c.p.useCerjan ? 1 : 0;
  a.fault = c.p.Rupture ? 1 : 0;
  a.rhs_only = rhs_only ? 1 : 0;
  // planes per block: enough blocks to fill 256 CUs several times over
  int tiles = ((c.d.nj + TJ - 1) / TJ) * ((c.d.nk + TK - 1) / TK);
  int chunks = std::max(1, std::min(c.d.ni, (2048 + tiles - 1) / tiles));
  a.ichunk = (c.d.ni + chunks - 1) / chunks;
  if (a.ichunk < 8) a.ichunk = std::min(8, c.d.ni);
  return a;
}

// =============================================================================
// Fault kernels (2-D over the fault plane, one thread per node).
// =============================================================================
struct FArgs {
  FaultDev F;
  const float* __restrict__ W;   // stage input wavefield (read)
  float* __restrict__ Wi0;       // wavefield receiving fault2wave
  const float* __restrict__ M;   // metric (fp32 path; bf16 handled via Mb)
  const uint16_t* __restrict__ Mb;
  int bf16;
  const float* mVx; const float* mVy;
  float* slabH;
  const float* S; float* T;      // for the slab RK
  const float* Win;              // stage input for irk 0 mW
  const float* cj[3];
  int ni, nj, nk, nx, ny, nz, i0;
  long nxyz, nyz;
  int NY, NZ;
  float DH, rDH, DT;
  double DTd;
  int it, irk, FX, FY, FZ;
  int freenode, friction, stress_type, cerjan;
  float mu_d, viscosity, smooth_load_T, Vini;
  double f0, V0, fw;
  int newton_mode, newton_fallback;
  double newton_tol;
};

__device__ __forceinline__ float LM(const FArgs& a, long i) {
  return a.bf16 ? __uint_as_float(((unsigned)a.Mb[i]) << 16) : a.M[i];
}

__device__ float Gt_func(float t, float T) {  // device.cu:134-143
  if (t < T) {
    float t1 = t - T, t2 = t - 2.0f * T;
    return expf(t1 * t1 / t / t2);
  }
  return 1.0f;
}
__device__ float Fr_func(float r, float R) { return r < R ? expf(r * r / (r * r - R * R)) : 0.0f; }

// Conservative traction J*(grad xi . sigma) at volume index p along metric row 'row'.
__device__ __forceinline__ void trac(const FArgs& a, const float* W, long p, int row, float* o) {
  float e0 = LM(a, (3 * row) * a.nxyz + p), e1 = LM(a, (3 * row + 1) * a.nxyz + p), e2 = LM(a, (3 * row + 2) * a.nxyz + p);
  float J = LM(a, 9 * a.nxyz + p);
  float txx = W[3 * a.nxyz + p], tyy = W[4 * a.nxyz + p], tzz = W[5 * a.nxyz + p];
  float txy = W[6 * a.nxyz + p], txz = W[7 * a.nxyz + p], tyz = W[8 * a.nxyz + p];
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
  const long nyz = a.nyz, nxyz = a.nxyz, pf = j1 + (long)k1 * ny, pv = pf * 3;
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
      long p = pf + (long)(i0 + (2 * m - 1) * l) * nyz;
      jac = LM(a, 9 * nxyz + p);
      float o[3];
      trac(a, a.W, p, 0, o);
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
  long pm = pf + (long)i0 * nyz;
  float n[3] = {LM(a, pm), LM(a, nxyz + pm), LM(a, 2 * nxyz + pm)};
  float n0 = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  float jacvec = LM(a, 9 * nxyz + pm) * n0;
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
    double coef = cc * RS_a * fabs(Tau_nd);
    double Vtn = sqrt(Vt1 * Vt1 + Vt2 * Vt2);
    double Ttn = sqrt(Tt[1] * Tt[1] + Tt[2] * Tt[2]);
    double Rn = sqrt(Rl[1] * Rl[1] + Rl[2] * Rl[2]);
    const double RS_V0 = 1e-6;
    double x_plain = asinh(exp(st / RS_a) * Rn / (2.0 * RS_V0));
    double x0 = x_plain;
    double xnew;
    int fb;
    int iters = nr_search(x0, st, RS_a, Vtn, coef, a.newton_mode, a.newton_tol, a.newton_fallback, x_plain, &xnew, &fb);
    atomicAdd(&F.newton_stats[0], 1ULL);
    atomicAdd(&F.newton_stats[1], (unsigned long long)iters);
    if (fb) atomicAdd(&F.newton_stats[2], 1ULL);
    float V = (float)(2.0 * RS_V0 * exp(-st / RS_a) * sinh(xnew));
    double Tm = RS_a * xnew * fabs(Tau_nd);
    F.friction[q] = (float)(RS_a * xnew);
    Ttn = fmax(Ttn, 1e-8);
    double Ts1 = Tt[1] / Ttn * Tm, Ts2 = Tt[2] / Ttn * Tm;
    V = fmaxf(V, 1e-30f);
    for (int c = 0; c < 3; c++) Tau[c] = (float)(Tau_nd * dn[c] + Ts1 * ds1[c] + Ts2 * ds2[c]);
    Tau[0] = F.str_init_x[q];  // trial_ratestate.cu:797
    Tau_n = Tau[0] * n[0] + Tau[1] * n[1] + Tau[2] * n[2];
    ifchange = true;
    ratecheck = V;
  }
  // ---- shared tail (trial_slipweakening.cu:640-709)
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
    long p = pf + (long)i0 * nyz;
    float mt[3][3] = {{LM(a, p), LM(a, 3 * nxyz + p), LM(a, 6 * nxyz + p)},
                      {LM(a, nxyz + p), LM(a, 4 * nxyz + p), LM(a, 7 * nxyz + p)},
                      {LM(a, 2 * nxyz + p), LM(a, 5 * nxyz + p), LM(a, 8 * nxyz + p)}};
    float rj = 1.0f / LM(a, 9 * nxyz + p);
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
    for (int c = 0; c < 3; c++) W[c * nxyz + p] = (F.W[c * 2 * nyz + pf] + F.W[c * 2 * nyz + nyz + pf]) * 0.5f;
    W[3 * nxyz + p] = sg[0][0];
    W[4 * nxyz + p] = sg[1][1];
    W[5 * nxyz + p] = sg[2][2];
    W[6 * nxyz + p] = (sg[0][1] + sg[1][0]) * 0.5f;
    W[7 * nxyz + p] = (sg[0][2] + sg[2][0]) * 0.5f;
    W[8 * nxyz + p] = (sg[1][2] + sg[2][1]) * 0.5f;
  }
}

__device__ __forceinline__ void mv3(const float* A, const float* x, float* y) {
  for (int r = 0; r < 3; r++) y[r] = A[3 * r] * x[0] + A[3 * r + 1] * x[1] + A[3 * r + 2] * x[2];
}

// fault_deriv: fault_dvelo_cu (fault_dvelo.cu:17-389) + fault_dstrs_f/b_cu
// (fault_dstrs_f.cu:17-1254, fault_dstrs_b.cu) writing the slab RHS.
__global__ void __launch_bounds__(256) fault_deriv_kernel(FArgs a) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.nj || k >= a.nk) return;
  const FaultDev& F = a.F;
  long q = j + (long)k * a.nj;
  if (F.united[q]) return;
  const int j1 = j + 3, k1 = k + 3, ny = a.ny, i0 = a.i0;
  const long nyz = a.nyz, nxyz = a.nxyz, pf = j1 + (long)k1 * ny;
  const float rDH = a.rDH;
  const int km = a.nk - k;
  const int FX = a.FX, FY = a.FY, FZ = a.FZ;
  const float* W = a.W;
  float* slab = a.slabH;
  auto slabx = [&](int i, int c) -> float& { return slab[(long)(i - (i0 - 3)) * 9 * nyz + c * nyz + pf]; };

  // ---- velocity rates at i0-3..-1, +1..+3 (fault_dvelo.cu:153-255)
  for (int i = i0 - 3; i <= i0 + 3; i++) {
    int nn = i0 - i;
    if (nn == 0) continue;
    float v1[3][7], v2[3][7], v3[3][7];
    for (int l = -3; l <= 3; l++) {
      float o[3];
      trac(a, W, pf + (long)(i + l) * nyz, 0, o);
      for (int c = 0; c < 3; c++) v1[c][l + 3] = o[c];
      trac(a, W, pf + l + (long)i * nyz, 1, o);
      for (int c = 0; c < 3; c++) v2[c][l + 3] = o[c];
      trac(a, W, pf + (long)l * ny + (long)i * nyz, 2, o);
      for (int c = 0; c < 3; c++) v3[c][l + 3] = o[c];
    }
    v1[0][nn + 3] = F.T11[3 * nyz + pf];
    v1[1][nn + 3] = F.T12[3 * nyz + pf];
    v1[2][nn + 3] = F.T13[3 * nyz + pf];
    for (int c = 0; c < 3; c++) {
      float* v = v1[c];
      if (nn == 2) v[6] = 2.0f * v[5] - v[4];
      if (nn == 1) { v[5] = 2.0f * v[4] - v[3]; v[6] = 2.0f * v[4] - v[2]; }
      if (nn == -1) { v[0] = 2.0f * v[2] - v[4]; v[1] = 2.0f * v[2] - v[3]; }
      if (nn == -2) v[0] = 2.0f * v[1] - v[2];
    }
    if (a.freenode && km <= 3)
      for (int c = 0; c < 3; c++) {
        v3[c][km + 2] = 0.0f;
        for (int l = km + 3; l < 7; l++) v3[c][l] = -v3[c][2 * (km + 2) - l];
      }
    long p = pf + (long)i * nyz;
    float rrj = 1.0f / (LM(a, 12 * nxyz + p) * LM(a, 9 * nxyz + p));
    for (int c = 0; c < 3; c++) {
      float dxt = drp(&v1[c][3], 1, FX) * rDH, dyt = drp(&v2[c][3], 1, FY) * rDH, dzt = drp(&v3[c][3], 1, FZ) * rDH;
      slabx(i, c) = (dxt + dyt + dzt) * rrj;
    }
  }
  // ---- split-node velocity rates (fault_dvelo.cu:260-385)
  for (int m = 0; m < 2; m++) {
    const float* TT[3] = {F.T11, F.T12, F.T13};
    float rrj = 1.0f / (F.rho_f[m * nyz + pf] * LM(a, 9 * nxyz + pf + (long)i0 * nyz));
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
  }

  // ---- stress rates (fault_dstrs_f/b)
  auto wv = [&](int c, int i) -> const float* { return &W[c * nxyz + pf + (long)i * nyz]; };
  auto fv = [&](int c, int m) -> const float* { return &F.W[c * 2 * nyz + m * nyz + pf]; };
  float DxV[8][3], DyV[8][3], DzV[8][3];
  for (int c = 0; c < 3; c++) {
    if (FX > 0) {
      DxV[7][c] = drpF(wv(c, i0 + 3), nyz) * rDH;
      DxV[6][c] = d24F(wv(c, i0 + 2), nyz) * rDH;
      DxV[5][c] = (wv(c, i0 + 1)[nyz] - wv(c, i0 + 1)[0]) * rDH;
      DxV[4][c] = (wv(c, i0 + 1)[0] - fv(c, 1)[0]) * rDH;
      float v5[5] = {wv(c, i0 - 4)[0], wv(c, i0 - 3)[0], wv(c, i0 - 2)[0], wv(c, i0 - 1)[0], fv(c, 0)[0]};
      DxV[0][c] = drpF(&v5[1], 1) * rDH;
      float v3[3] = {wv(c, i0 - 2)[0], wv(c, i0 - 1)[0], fv(c, 0)[0]};
      DxV[1][c] = d24F(&v3[0], 1) * rDH;
      DxV[2][c] = (fv(c, 0)[0] - wv(c, i0 - 1)[0]) * rDH;
      DxV[3][c] = 0.0f;
    } else {
      DxV[0][c] = drpB(wv(c, i0 - 3), nyz) * rDH;
      DxV[1][c] = d24B(wv(c, i0 - 2), nyz) * rDH;
      DxV[2][c] = (wv(c, i0 - 1)[0] - wv(c, i0 - 1)[-nyz]) * rDH;
      DxV[3][c] = (fv(c, 0)[0] - wv(c, i0 - 1)[0]) * rDH;
      float v5[5] = {fv(c, 1)[0], wv(c, i0 + 1)[0], wv(c, i0 + 2)[0], wv(c, i0 + 3)[0], wv(c, i0 + 4)[0]};
      DxV[7][c] = drpB(&v5[3], 1) * rDH;
      float v3[3] = {fv(c, 1)[0], wv(c, i0 + 1)[0], wv(c, i0 + 2)[0]};
      DxV[6][c] = d24B(&v3[2], 1) * rDH;
      DxV[5][c] = (wv(c, i0 + 1)[0] - fv(c, 1)[0]) * rDH;
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
    for (int c = 0; c < 3; c++)
      for (int m = 0; m < 2; m++) {
        const float* f = fv(c, m);
        if (a.freenode && km == 2) DzV[3 + m][c] = d22(f, ny, FZ) * rDH;
        else if (a.freenode && km == 3) DzV[3 + m][c] = d24(f, ny, FZ) * rDH;
        else DzV[3 + m][c] = (F.rup_index_z[q] % 7 ? d22(f, ny, FZ) : drp(f, ny, FZ)) * rDH;
      }
    // DxV_T1: D11^-1 (dT1/dt - D12 DyV - D13 DzV) on both sides
    float dtT1[3] = {0.f, 0.f, 0.f};
    if (F.faultgrid[q]) { dtT1[0] = F.hT11[pf]; dtT1[1] = F.hT12[pf]; dtT1[2] = F.hT13[pf]; }
    for (int m = 0; m < 2; m++) {
      float o1[3], o2[3], o3[3];
      mv3(&F.matT1toVx[(m * nyz + pf) * 9], dtT1, o1);
      mv3(&F.matVytoVx[(m * nyz + pf) * 9], DyV[3 + m], o2);
      mv3(&F.matVztoVx[(m * nyz + pf) * 9], DzV[3 + m], o3);
      for (int c = 0; c < 3; c++) DxV[3 + m][c] = o1[c] - o2[c] - o3[c];
    }
  }
  // split-node hT2*, hT3* (fault_dstrs_f.cu:874-1007)
  for (int m = 0; m < 2; m++)
    for (int row = 1; row <= 2; row++) {
      float o1[3], o2[3], o3[3];
      mv3(&F.Dmat[((long)(m * 9 + row * 3 + 0) * nyz + pf) * 9], DxV[3 + m], o1);
      mv3(&F.Dmat[((long)(m * 9 + row * 3 + 1) * nyz + pf) * 9], DyV[3 + m], o2);
      mv3(&F.Dmat[((long)(m * 9 + row * 3 + 2) * nyz + pf) * 9], DzV[3 + m], o3);
      for (int c = 0; c < 3; c++) F.hW[(3 * row + c) * 2 * nyz + m * nyz + pf] = o1[c] + o2[c] + o3[c];
    }
  // off-fault stress rates (fault_dstrs_f.cu:1009-1250)
  for (int nn = 0; nn < 8; nn++) {
    int l = nn < 4 ? nn - 3 : nn - 4;
    if (l == 0) continue;
    int i = i0 + l;
    long p = pf + (long)i * nyz;
    float dy[3], dz[3];
    const float* dx = DxV[nn];
    for (int c = 0; c < 3; c++) dy[c] = drp(&W[c * nxyz + p], 1, FY) * rDH;
    long ij = (j1 + (long)i * ny) * 9;
    for (int c = 0; c < 3; c++) {
      if (a.freenode && km == 1)
        dz[c] = a.mVx[ij + 3 * c] * dx[0] + a.mVx[ij + 3 * c + 1] * dx[1] + a.mVx[ij + 3 * c + 2] * dx[2]
              + a.mVy[ij + 3 * c] * dy[0] + a.mVy[ij + 3 * c + 1] * dy[1] + a.mVy[ij + 3 * c + 2] * dy[2];
      else if (a.freenode && km == 2) dz[c] = d22(&W[c * nxyz + p], ny, FZ) * rDH;
      else if (a.freenode && km == 3) dz[c] = d24(&W[c * nxyz + p], ny, FZ) * rDH;
      else dz[c] = drp(&W[c * nxyz + p], ny, FZ) * rDH;
    }
    float lam = LM(a, 10 * nxyz + p), mu = LM(a, 11 * nxyz + p), l2m = lam + 2.0f * mu;
    float xi[3] = {LM(a, p), LM(a, nxyz + p), LM(a, 2 * nxyz + p)};
    float et[3] = {LM(a, 3 * nxyz + p), LM(a, 4 * nxyz + p), LM(a, 5 * nxyz + p)};
    float zt[3] = {LM(a, 6 * nxyz + p), LM(a, 7 * nxyz + p), LM(a, 8 * nxyz + p)};
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
  const long nyz = a.nyz, nxyz = a.nxyz, pf = j1 + (long)k1 * ny, pv = pf * 3;
  const double RA[4] = {0.5, 0.5, 1.0, 1e10}, RB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const int irk = a.irk;
  const float rka = (float)(RA[irk] * a.DTd), rkb = (float)(RB[irk] * a.DTd);
  float dampj = a.cerjan ? fminf(a.cj[1][j1], a.cj[2][k1]) : 1.0f;
  // volume RK of the slab planes
  for (int pl = 0; pl < 7; pl++) {
    if (pl == 3) continue;
    int ip = i0 - 3 + pl;
    long p = pf + (long)ip * nyz;
    float damp = a.cerjan ? fminf(a.cj[0][ip], dampj) : 1.0f;
    for (int c = 0; c < 9; c++) {
      long x = c * nxyz + p;
      float h = a.slabH[(long)pl * 9 * nyz + c * nyz + pf];
      float w = rk_combine(irk, rka, rkb, a.Win[x], h, a.S, a.T, x);
      a.Wi0[x] = w * damp;
    }
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
  long p = pf + (long)i0 * nyz;
  float mt[3][3] = {{LM(a, p), LM(a, 3 * nxyz + p), LM(a, 6 * nxyz + p)},
                    {LM(a, nxyz + p), LM(a, 4 * nxyz + p), LM(a, 7 * nxyz + p)},
                    {LM(a, 2 * nxyz + p), LM(a, 5 * nxyz + p), LM(a, 8 * nxyz + p)}};
  float rj = 1.0f / LM(a, 9 * nxyz + p);
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
  for (int c = 0; c < 3; c++) W[c * nxyz + p] = (F.W[c * 2 * nyz + pf] + F.W[c * 2 * nyz + nyz + pf]) * 0.5f * damp;
  W[3 * nxyz + p] = sg[0][0] * damp;
  W[4 * nxyz + p] = sg[1][1] * damp;
  W[5 * nxyz + p] = sg[2][2] * damp;
  W[6 * nxyz + p] = (sg[0][1] + sg[1][0]) * 0.5f * damp;
  W[7 * nxyz + p] = (sg[0][2] + sg[2][0]) * 0.5f * damp;
  W[8 * nxyz + p] = (sg[1][2] + sg[2][1]) * 0.5f * damp;
}

// cal_rup_sensor_D_cu / cal_rup_sensor_magn_cu (rup_sensor.cu:7-86).
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

__global__ void fill_kernel(float* p, long n, float v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

FArgs make_fargs(Ctx& c) {
  FArgs a{};
  a.F = c.f;
  a.bf16 = c.p.MetricBF16 ? 1 : 0;
  a.M = a.bf16 ? nullptr : (const float*)c.met;
  a.Mb = a.bf16 ? (const uint16_t*)c.met : nullptr;
  a.mVx = c.matVx2Vz; a.mVy = c.matVy2Vz;
  a.slabH = c.slabH;
  for (int x = 0; x < 3; x++) a.cj[x] = c.cj[x];
  a.ni = c.d.ni; a.nj = c.d.nj; a.nk = c.d.nk; a.nx = c.d.nx; a.ny = c.d.ny; a.nz = c.d.nz; a.i0 = c.d.i0;
  a.nxyz = c.d.nxyz; a.nyz = c.d.nyz;
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
  return a;
}

dim3 fgrid(const Ctx& c) { return dim3((c.d.nj + 63) / 64, (c.d.nk + 3) / 4, 1); }

}  // namespace

void upload_wave_cold(Ctx& c) {
  WaveCold w{};
  for (int x = 0; x < 3; x++) { w.pA[x] = c.pml.A[x]; w.pB[x] = c.pml.B[x]; w.pD[x] = c.pml.D[x]; w.cj[x] = c.cj[x]; }
  for (int f = 0; f < 6; f++) { w.auxY[f] = c.pml.Y[f]; w.auxS[f] = c.pml.S[f]; w.auxT[f] = c.pml.T[f]; w.on[f] = c.pml.on[f]; }
  w.mVx = c.matVx2Vz; w.mVy = c.matVy2Vz;
  w.N = c.pml.N;
  if (!c.wave_cold) {
    hipMalloc(&c.wave_cold, sizeof(WaveCold));
    c.allocations.push_back(c.wave_cold);
  }
  hipMemcpy(c.wave_cold, &w, sizeof w, hipMemcpyHostToDevice);
}

void launch_wave_stage(Ctx& c, const StageBufs& b, int irk, int fx, int fy, int fz, bool rhs_only, float* rhs_out) {
  if (c.p.MetricBF16) {
    WaveArgs<uint16_t> a = make_args<uint16_t>(c, b, irk, rhs_only, rhs_out);
    launch_wave_dispatch(a, fx, fy, fz, c.d.nj, c.d.nk, c.d.ni, c.stream);
  } else {
    WaveArgs<float> a = make_args<float>(c, b, irk, rhs_only, rhs_out);
    launch_wave_dispatch(a, fx, fy, fz, c.d.nj, c.d.nk, c.d.ni, c.stream);
  }
}

void launch_fault_pre(Ctx& c, const float* Win, int it, int irk, int fx, int fy, int fz) {
  FArgs a = make_fargs(c);
  a.W = Win; a.Wi0 = const_cast<float*>(Win);
  a.it = it; a.irk = irk; a.FX = fx; a.FY = fy; a.FZ = fz;
  hipLaunchKernelGGL(fault_pre_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a);
}

void launch_fault_deriv(Ctx& c, const float* Win, int fx, int fy, int fz) {
  FArgs a = make_fargs(c);
  a.W = Win;
  a.FX = fx; a.FY = fy; a.FZ = fz;
  hipLaunchKernelGGL(fault_deriv_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a);
}

void launch_fault_post(Ctx& c, const StageBufs& b, int irk) {
  FArgs a = make_fargs(c);
  a.Win = b.Win; a.Wi0 = b.Wout; a.S = b.S; a.T = b.T;
  a.irk = irk;
  hipLaunchKernelGGL(fault_post_kernel, fgrid(c), dim3(64, 4, 1), 0, c.stream, a);
}

void launch_rup_sensor(Ctx& c) {
  dim3 g((c.d.nj + 63) / 64, (c.d.nk + 3) / 4, 1);
  const int* G = c.p.Fault_grid;
  hipLaunchKernelGGL(rup_sensor_d_kernel, g, dim3(64, 4, 1), 0, c.stream, c.f, c.d.nj, c.d.nk, G[0], G[1], G[2], G[3]);
  hipLaunchKernelGGL(rup_sensor_m_kernel, g, dim3(64, 4, 1), 0, c.stream, c.f, c.d.nj, c.d.nk, G[0], G[1], G[2], G[3]);
}

void launch_fill(Ctx& c, float* p, long n, float v) {
  long blocks = std::min<long>((n + 255) / 256, 65536);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, c.stream, p, n, v);
}

}  // namespace cgfdm

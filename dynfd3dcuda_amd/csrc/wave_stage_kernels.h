// Wave stage kernel of the CGFDM time step for CDNA4 (gfx950): the
// HBM-bound hot path.
//
// One fused kernel computes the velocity-stress RHS on the curvilinear grid
// (wave_deriv_cu, wave.cu:8-494), adds the ADE-PML terms and advances their
// auxiliary fields (abs_deriv_{x,y,z}_cu, pml{x,y,z}.cu), applies the RK4
// stage combine (wave_rk_cu/RK_Syn, rk4.cu:20-33,154-257) and the Cerjan
// sponge (apply_cerjan_cu, cerjan.cu:166-200), reading each wavefield word
// once from HBM.
//
// Work decomposition: a block owns a TJ x TK tile of (j,k) columns and
// streams a chunk of i planes (i is the fault-normal, slowest axis).  The
// five planes the x-stencil needs sit in a per-thread register queue; the
// current plane plus its y/z halo is staged in a double-buffered LDS tile,
// so each plane costs one barrier and the next plane's loads are in flight
// while the current plane is computed.  Blocks are numbered so that each
// XCD (and its private L2) receives a contiguous range of work items of one
// i-chunk, which keeps the k-halo rows of neighbouring tiles L2 hits.
//
// Addressing: every global access is a 32-bit byte offset from a
// wave-uniform field base, so one VGPR addresses all 9 wave fields and the
// metric fields of a point (the grid is limited to 2^30 points).
//
// Points of the fault slab (i0-3..i0+3 without i0) keep their RHS in a slab
// buffer so the split-node kernels can overwrite it (fault_dvelo/
// fault_dstrs) before fault_post advances the slab.
//
// This header holds the stage kernels and their launchers; each
// wave_stage_<metric storage>.hip instantiates them for one metric storage
// type (so the three compile in parallel), and wave_stage.hip builds the
// work lists and calls them.
#pragma once
#include "fd_ops.h"
#include <algorithm>
#include <type_traits>
#include <vector>

namespace cgfdm {
namespace wave {
using namespace fd;


constexpr int NTHR = 512;  // threads per block
constexpr int TJ = 64;     // default tile: one wavefront per k-row, j contiguous
constexpr int TK = NTHR / TJ;
constexpr int TJN = 32;    // narrow tile for the absorbing bands along j (32 j x 8 k)
constexpr int NTHRN = 256;  // threads per block of the narrow tiles
constexpr int TKN = NTHRN / TJN;
constexpr int TJB = 20;    // exact-width tiles of the absorbing bands along j when they are at most 20
constexpr int TKB = 12;    // columns wide (20 x 12, 240 threads; the 20-layer PML and sponge)
constexpr int TK6 = 6;     // rows of the 64 x 6 interior tiles (WaveInteriorKernel 2)
constexpr int TK4 = 4;     // rows of the 64 x 4 interior tiles (WaveInteriorKernel 3)
constexpr int kPmlBoxInts = 8;  // ints of one box of the auxiliary update


// Volume arrays are plane-interleaved with padded rows: word (field c,
// padded plane x, column jk = jpad + j + k*py) sits at (x*nf + c)*pyz + jk,
// nf = 9 for wavefields, 13 for the metric (16 in its bfloat16-deviation form); py
// is a multiple of 32 words.  A plane of an array is addressed as one buffer resource
// (base = the plane's first word, range = the plane's bytes) with the lane's
// column byte offset as voffset and the field's offset c*nyz*size as
// soffset: no per-lane 64-bit address arithmetic, a handful of scalar
// registers per plane, and one block touches few memory pages per plane.
// Word 3 of the resource is the gfx9-family default (32-bit data, no swizzle).
constexpr int kBufWord3 = 0x00020000;
typedef __amdgpu_buffer_rsrc_t Rsrc;
template <class T>
__device__ __forceinline__ Rsrc plane_rsrc(const T* a, int plane, uint32_t nf, uint32_t nyz) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(a) + (size_t)plane * nf * nyz, (short)0,
                                           (int)(nf * nyz * (uint32_t)sizeof(T)), kBufWord3);
}
// Cache policy of an access: kReuse for stage-input words that neighbouring
// tiles read again (halo rows and columns, the x-queue warm-up), kStream
// (non-temporal) for words touched once per stage, so they do not push the
// reused rows out of the XCD's L2.
constexpr int kReuse = 0, kStream = 2;
template <int POL = kReuse>
__device__ __forceinline__ float bld(Rsrc r, uint32_t voff, uint32_t soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, POL));
}
template <int POL = kStream>
__device__ __forceinline__ void bst(Rsrc r, uint32_t voff, uint32_t soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, POL);
}
template <class MT>
__device__ __forceinline__ float bldm(Rsrc r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(MT) == 2) {
    const unsigned short v = __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, kStream);
    return __uint_as_float(((unsigned)v) << 16);
  } else {
    return bld<kStream>(r, voff, soff);
  }
}
// Word of field c at padded plane x, column jk (plain loads, off the hot path).
__device__ __forceinline__ float wat(const float* W, int c, int x, uint32_t jk, uint32_t nyz) {
  return W[((size_t)x * 9 + c) * nyz + jk];
}
__device__ __forceinline__ float mat_(const float* M, int c, int x, uint32_t jk, uint32_t nyz, const float*) {
  return M[((size_t)x * 13 + c) * nyz + jk];
}
__device__ __forceinline__ float mat_(const uint16_t* M, int c, int x, uint32_t jk, uint32_t nyz, const float*) {
  return __uint_as_float(((unsigned)M[((size_t)x * 13 + c) * nyz + jk]) << 16);
}
// Metric in bfloat16 with the media words exact (18 16-bit slots per
// plane): the bfloat16 of each metric word's deviation from a reference
// value of its field (ref[c], the midrange of the field over the grid), so
// the rounding scales with the field's variation rather than its size,
// packed two to a 32-bit word (slots 2p and 2p+1 form 32-bit plane p, field
// 2p in the low half).  Slots 0-8 hold the nine metric coefficients, slot 9
// the point's entry in the media palette (lambda, mu, rho of the few
// distinct media of a layered or two-half model, see WaveArgs::pal; 0 when
// the media have no palette), slots 10-15 lambda, mu and rho as three FP32
// planes, slot 16 the Jacobian (its deviation, field 9 of ref), slot 17 0.
struct MetBf16Delta {
  uint16_t h;
};
constexpr uint32_t kMetDeltaSlots = 18;
constexpr uint32_t kMetDeltaJac = 16;  // slot of the Jacobian
template <class MT>
constexpr uint32_t met_nf() {
  return std::is_same<MT, MetBf16Delta>::value ? kMetDeltaSlots : 13u;
}
__device__ __forceinline__ float mat_(const MetBf16Delta* Ms, int c, int x, uint32_t jk, uint32_t nyz, const float* ref) {
  const uint16_t* M = &Ms->h;
  if (c < 10) {  // field c of the pair at slot c & ~1 (c < 9), or the Jacobian in the low half at its slot
    const uint32_t s = c == 9 ? kMetDeltaJac : (uint32_t)(c & ~1);
    const uint32_t u = ((const uint32_t*)(M + ((size_t)x * kMetDeltaSlots + s) * nyz))[jk];
    return __uint_as_float((c < 9 && (c & 1)) ? (u & 0xffff0000u) : (u << 16)) + ref[c];
  }
  return ((const float*)(M + ((size_t)x * kMetDeltaSlots + 10) * nyz))[(size_t)(c - 10) * nyz + jk];
}
// Field c (a metric coefficient 0..8 or a media word 10..12) of a metric
// plane resource r at byte column voff (mpb bytes a field); ref is the
// field's reference value in the deviation form.
template <class MT>
__device__ __forceinline__ float bldmc(Rsrc r, uint32_t voff, int c, uint32_t mpb, float ref) {
  if constexpr (std::is_same<MT, MetBf16Delta>::value) {
    if (c >= 10) return bld<kStream>(r, 2 * voff, (uint32_t)(10 + 2 * (c - 10)) * mpb);  // FP32 word
    const uint32_t u = __float_as_uint(bld<kStream>(r, 2 * voff, (uint32_t)(c & ~1) * mpb));  // pair of c
    return __uint_as_float((c & 1) ? (u & 0xffff0000u) : (u << 16)) + ref;
  }
  return bldm<MT>(r, voff, c * mpb);
}

// RK_Syn combine (rk4.cu:20-33) of field c of the planes rs (S) and rt (T).
// Stage 0 sets the accumulator T = S + (dt/6) h0 beside the stage value
// W1 = S + (dt/2) h0, so T = S + (W1 - S) / 3: stage 1, which reads S and
// W1 anyway, derives it (to FP32 rounding) instead of stage 0 writing it and
// stage 1 reading it, 18 of the 144 wavefield words of a step.  A point
// whose stage-0 value changes after the combine (Cerjan-damped, or on the
// fault plane, where fault2wave rewrites it) keeps the stored accumulator:
// keep_t.  The same point takes the same branch in every stage.
__device__ __forceinline__ float rk_derived_t(float s, float w1) { return s + (w1 - s) * (1.0f / 3.0f); }
__device__ __forceinline__ float rk_combine_b(int irk, float rka, float rkb, float w, float h, Rsrc rs, Rsrc rt,
                                              uint32_t voff, uint32_t soff, bool keep_t) {
  if (irk == 0) {
    if (keep_t) bst(rt, voff, soff, w + rkb * h);
    return w + rka * h;
  }
  if (irk == 3) return bld<kStream>(rt, voff, soff) + rkb * h;
  const float s = bld<kStream>(rs, voff, soff);
  const float t = irk == 1 && !keep_t ? rk_derived_t(s, w) : bld<kStream>(rt, voff, soff);
  bst(rt, voff, soff, t + rkb * h);
  return s + rka * h;
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
  float* ttop;  // traction vectors of the six top rows (free surface)
  float* itop;  // image velocity RHS of the three top rows (free surface)
  const WaveCold* __restrict__ cold;
  const int4* __restrict__ work;  // two int4 per block: (j0, k0, first plane, end plane), (j lanes, k rows, first active j lane, first active k row)
  unsigned nyz;                   // words of one field of one plane (py*nz, padded rows)
  int py, jpad;                   // row pitch and front pad of volume rows
  int nwork, ni, nj, nk, ny, ii0;
  float rDH, rka, rkb;
  float mref[10];   // metric reference values (bfloat16-deviation form)
  // bfloat16-deviation form: lambda, mu, rho (and 0) of each media palette
  // entry, or null when the media have no palette (the pipelined kernels
  // then read the FP32 media planes)
  const float4* __restrict__ pal;
  int npal;
  unsigned met_skip;  // bfloat16-deviation form: bit p set when 32-bit metric word p is +-0 on the whole grid
  int irk, freenode, cerjan, fault, rhs_only, pml;
  int aux_fused;  // BAND tiles advance their points' ADE-PML auxiliary fields themselves
};

// Free surface, first pass: the conservative traction vectors
// J (grad xi_ax . sigma) (wave.cu:398-440) of the six rows nk-6..nk-1 that
// the image stencil reaches, for every padded plane and column, from the
// stage input: ttop[x][r][3 ax + c][jj], r = row - (nk-6), jj the padded
// column.  One thread per point.
template <class MT>
__global__ void __launch_bounds__(256) traction_top_kernel(WaveArgs<MT> a, int nx) {
  const int ny = a.ny;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nx * 6 * ny) return;
  const int jj = (int)(idx % ny);
  const int r = (int)((idx / ny) % 6);
  const int x = (int)(idx / ((long)ny * 6));
  const uint32_t nyz = a.nyz;
  const uint32_t q = (uint32_t)a.jpad + (uint32_t)jj + (uint32_t)(a.nk - 6 + r + 3) * (uint32_t)a.py;
  const float* W = a.Win;
  const MT* M = a.met;
  const float J = mat_(M, 9, x, q, nyz, a.mref);
  const float txx = wat(W, 3, x, q, nyz), tyy = wat(W, 4, x, q, nyz), tzz = wat(W, 5, x, q, nyz);
  const float txy = wat(W, 6, x, q, nyz), txz = wat(W, 7, x, q, nyz), tyz = wat(W, 8, x, q, nyz);
  float* o = a.ttop + ((long)x * 6 + r) * 9 * ny + jj;
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    const float e0 = mat_(M, 3 * ax, x, q, nyz, a.mref), e1 = mat_(M, 3 * ax + 1, x, q, nyz, a.mref), e2 = mat_(M, 3 * ax + 2, x, q, nyz, a.mref);
    o[(3 * ax + 0) * ny] = (e0 * txx + e1 * txy + e2 * txz) * J;
    o[(3 * ax + 1) * ny] = (e0 * txy + e1 * tyy + e2 * tyz) * J;
    o[(3 * ax + 2) * ny] = (e0 * txz + e1 * tyz + e2 * tzz) * J;
  }
}

// Free surface, second pass: the traction-image velocity RHS of the three
// top rows (wave.cu:398-487): DRP derivatives of the traction vectors along
// each axis, the z stencil zero at the surface and odd-mirrored above it,
// with the PML 1/beta of each axis; itop[c][row][i][j], row = k - (nk-3).
template <int FX, int FY, int FZ, class MT>
__global__ void __launch_bounds__(256) image_top_kernel(WaveArgs<MT> a) {
  const int ni = a.ni, nj = a.nj, nk = a.nk, ny = a.ny;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)ni * 3 * nj) return;
  const int j = (int)(idx % nj);
  const int row = (int)((idx / nj) % 3);
  const int i = (int)(idx / ((long)nj * 3));
  const int k = nk - 3 + row, ik = 2 - row, x = i + 3;
  const uint32_t col = (uint32_t)a.jpad + (uint32_t)(j + 3) + (uint32_t)(k + 3) * (uint32_t)a.py;
  const MT* M = a.met;
  const float rrhojac = 1.0f / (mat_(M, 12, x, col, a.nyz, a.mref) * mat_(M, 9, x, col, a.nyz, a.mref));
  float rb[3] = {1.0f, 1.0f, 1.0f};
  if (a.pml) {
    const WaveCold* cold = a.cold;
    rb[0] = 1.0f / cold->pB[0][i];
    rb[1] = 1.0f / cold->pB[1][j];
    rb[2] = 1.0f / cold->pB[2][k];
  }
  const int fl[3] = {FX, FY, FZ};
  const int r0 = k - (nk - 6);  // 3, 4, 5
  float hv[3] = {0.0f, 0.0f, 0.0f};
  for (int ax = 0; ax < 3; ax++) {
    float v[3][7];
    for (int l = 0; l < 7; l++) {
      const int xq = ax == 0 ? x + l - 3 : x;
      const int r = ax == 2 ? min(r0 + l - 3, 5) : r0;  // rows above the surface: replaced below
      const int jq = j + 3 + (ax == 1 ? l - 3 : 0);
      const float* t = a.ttop + ((long)xq * 6 + r) * 9 * ny + (long)(3 * ax) * ny + jq;
      for (int c = 0; c < 3; c++) v[c][l] = t[(long)c * ny];
    }
    if (ax == 2) {
      for (int c = 0; c < 3; c++) {
        if (ik == 0) { v[c][3] = 0.0f; v[c][6] = -v[c][0]; v[c][5] = -v[c][1]; v[c][4] = -v[c][2]; }
        else if (ik == 1) { v[c][4] = 0.0f; v[c][6] = -v[c][2]; v[c][5] = -v[c][3]; }
        else { v[c][5] = 0.0f; v[c][6] = -v[c][4]; }
      }
    }
    for (int c = 0; c < 3; c++) {
      const float d = a.rDH * (fl[ax] > 0 ? drp5<1>(v[c][2], v[c][3], v[c][4], v[c][5], v[c][6])
                                          : drp5<-1>(v[c][0], v[c][1], v[c][2], v[c][3], v[c][4]));
      hv[c] += d * rb[ax];
    }
  }
  const long plane = (long)ni * nj;
  for (int c = 0; c < 3; c++) a.itop[((long)c * 3 + row) * plane + (long)i * nj + j] = hv[c] * rrhojac;
}

// Extra auxiliary-field source of the x/y PML faces on the free surface
// (pmlx.cu:185-210, pmly.cu): the z-derivative of velocity eliminated through
// the traction-free condition, component c in 3..8.
__device__ __forceinline__ float pml_fs_term(int c, const float* z1, float ztx, float zty, float ztz, float lam,
                                             float l2m, float mu, float c0) {
  if (c == 3) return c0 * (l2m * ztx * z1[0] + lam * zty * z1[1] + lam * ztz * z1[2]);
  if (c == 4) return c0 * (lam * ztx * z1[0] + l2m * zty * z1[1] + lam * ztz * z1[2]);
  if (c == 5) return c0 * (lam * ztx * z1[0] + lam * zty * z1[1] + l2m * ztz * z1[2]);
  if (c == 6) return c0 * mu * (zty * z1[0] + ztx * z1[1]);
  if (c == 7) return c0 * mu * (ztz * z1[0] + ztx * z1[2]);
  if (c == 8) return c0 * mu * (ztz * z1[1] + zty * z1[2]);
  return 0.0f;
}

// KIND selects the instantiation: 0 interior (no PML, free-surface or Cerjan
// code: the host gives it only points where every absorbing profile is
// trivial), 1 absorbing bands (PML correction, Cerjan), 2 tiles holding the
// free-surface rows (everything).
template <int FX, int FY, int FZ, class MT, int KIND, int TJ_ = TJ, int NT_ = NTHR>
__global__ void __launch_bounds__(NT_, KIND == 2 ? 2 : 4) wave_stage_kernel(WaveArgs<MT> a) {
  constexpr int TJ = TJ_, TK = NT_ / TJ_;  // tile of TJ (j) x TK (k) columns
  constexpr int LJ = TJ + 4, LK = TK + 4;   // LDS tile with halos
  constexpr int NH = 4 * LJ + 4 * TK;       // halo cells of one plane: 4 rows + 4 side cells per row
  static_assert(NH <= NT_, "one halo cell per thread");
  constexpr int HJL = FY > 0 ? 1 : 3;    // halo width below j (the stencil reaches HJL back, 4-HJL forward)
  constexpr int HKL = FZ > 0 ? 1 : 3;
  constexpr int QLO = FX > 0 ? -1 : -3;  // queue holds planes i+QLO .. i+QLO+4
  constexpr int CTR = -QLO;
  __shared__ float lds[2][9][LK][LJ];
  // the absorbing layers' uniform descriptors, staged once per block: read
  // through `cold` inside the plane loop they would be fetched again from
  // global memory after every store (a chain of dependent round trips per
  // plane and face), and registers for them would spill
  __shared__ const float* s_auxY[6];
  __shared__ const float* s_pBx;
  __shared__ const float* s_cjx;
  __shared__ const float* s_mV[2];
  __shared__ int s_on[6];
  __shared__ int s_N;

  const int item = xcd_remap((int)blockIdx.x, a.nwork);
  const int4 wk = a.work[2 * item], we = a.work[2 * item + 1];
  const int j0 = wk.x, k0 = wk.y, ibeg = wk.z, iend = wk.w;
  const int tj = threadIdx.x, tk = threadIdx.y, tid = tj + tk * TJ;
  const int ni = a.ni, nj = a.nj, nk = a.nk, ny = a.ny;
  const int j = j0 + tj, k = k0 + tk;
  const bool act = tj >= we.z && tj < we.x && tk >= we.w && tk < we.y;
  const bool pml = KIND >= 1 && a.pml, cerjan = KIND >= 1 && a.cerjan;
  const int nxp = ni + 6;  // padded planes
  const uint32_t nyz = a.nyz;
  // Every lane loads (no predicated loads): lanes past the padded array read
  // a clamped in-bounds column whose value no active point ever uses.
  const uint32_t py = a.py, jpad = a.jpad;
  const uint32_t col = jpad + (uint32_t)(min(max(j, -3), nj + 2) + 3) + (uint32_t)(min(max(k, -3), nk + 2) + 3) * py;
  const uint32_t col4 = col << 2;
  const uint32_t mcol = col * (uint32_t)sizeof(MT);
  const uint32_t pb = nyz * 4, mpb = nyz * (uint32_t)sizeof(MT);  // bytes of one field of one plane
  const float* __restrict__ Win = a.Win;
  const WaveCold* __restrict__ cold = a.cold;

  // this thread's halo cell: 4 full rows of LJ cells, then 4 side cells per row
  int hlk = 0, hlj = 0;
  uint32_t hcol = col;
  if (tid < NH) {
    if (tid < 4 * LJ) {
      const int r = tid / LJ;
      hlk = r < HKL ? r : r + TK;
      hlj = tid % LJ;
    } else {
      const int h2 = tid - 4 * LJ;
      hlk = HKL + h2 / 4;
      const int cc = h2 % 4;
      hlj = cc < HJL ? cc : cc + TJ;
    }
    const int gj = min(max(j0 + hlj - HJL, -3), nj + 2), gk = min(max(k0 + hlk - HKL, -3), nk + 2);
    hcol = jpad + (uint32_t)(gj + 3) + (uint32_t)(gk + 3) * py;
  }
  const uint32_t hcol4 = hcol << 2;

  if (KIND >= 1 && threadIdx.y == 0 && threadIdx.x < 6) {
    s_auxY[threadIdx.x] = cold->auxY[threadIdx.x];
    s_on[threadIdx.x] = cold->on[threadIdx.x];
    if (threadIdx.x == 0) { s_N = cold->N; s_pBx = cold->pB[0]; s_cjx = cold->cj[0]; }
    if (threadIdx.x < 2) s_mV[threadIdx.x] = threadIdx.x ? cold->mVy : cold->mVx;
  }
  float rb2 = 1.0f, rb3 = 1.0f, cjk = 1.0f;
  if (act && pml) { rb2 = 1.0f / cold->pB[1][j]; rb3 = 1.0f / cold->pB[2][k]; }
  if (act && cerjan) cjk = fminf(cold->cj[1][j + 3], cold->cj[2][k + 3]);

  float q[5][9];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const Rsrc r = plane_rsrc(Win, ibeg + QLO + m + 3, 9, nyz);
#pragma unroll
    for (int c = 0; c < 9; c++) q[m][c] = bld(r, col4, c * pb);
  }
  float hv[9];
  {
    const Rsrc r = plane_rsrc(Win, ibeg + 3, 9, nyz);
#pragma unroll
    for (int c = 0; c < 9; c++) hv[c] = bld(r, hcol4, c * pb);
  }
#pragma unroll
  for (int c = 0; c < 9; c++) lds[0][c][tk + HKL][tj + HJL] = q[CTR][c];
  if (tid < NH) {
#pragma unroll
    for (int c = 0; c < 9; c++) lds[0][c][hlk][hlj] = hv[c];
  }
  __syncthreads();

  for (int i = ibeg; i < iend; ++i) {
    const int b = (i - ibeg) & 1;
    const bool more = i + 1 < iend;  // uniform per block
    // next plane of the queue and the halo of plane i+1, in flight during the compute
    // (planes clamped in-bounds: past the chunk's end the values go unused)
    float nq[9];
    {
      const Rsrc r = plane_rsrc(Win, min(i + QLO + 8, nxp - 1), 9, nyz);
#pragma unroll
      for (int c = 0; c < 9; c++) nq[c] = bld(r, col4, c * pb);
    }
    {
      const Rsrc r = plane_rsrc(Win, min(i + 4, nxp - 1), 9, nyz);
#pragma unroll
      for (int c = 0; c < 9; c++) hv[c] = bld(r, hcol4, c * pb);
    }

    if (act) {
      const Rsrc rm = plane_rsrc(a.met, i + 3, met_nf<MT>(), nyz);
      auto MET = [&](int c) { return bldmc<MT>(rm, mcol, c, mpb, c < 10 ? a.mref[c] : 0.0f); };
      const float xix = MET(0), xiy = MET(1), xiz = MET(2);
      const float etx = MET(3), ety = MET(4), etz = MET(5);
      const float ztx = MET(6), zty = MET(7), ztz = MET(8);
      const float lam = MET(10), mu = MET(11);
      const float rrho = 1.0f / MET(12);
      const float l2m = lam + 2.0f * mu;
      const float rDH = a.rDH;
      const float rb1 = pml ? 1.0f / s_pBx[i] : 1.0f;
      const float(*L)[LK][LJ] = lds[b];
      const int lj = tj + HJL, lk = tk + HKL;
      // one-sided DRP difference of field c along each axis (macdrp.h:91-95),
      // without the 1/DH factor (folded into the metric rows below)
      auto Dx = [&](int c) { return drp5<FX>(q[0][c], q[1][c], q[2][c], q[3][c], q[4][c]); };
      auto Dy = [&](int c) {
        return FY > 0 ? drp5<1>(L[c][lk][lj - 1], L[c][lk][lj], L[c][lk][lj + 1], L[c][lk][lj + 2], L[c][lk][lj + 3])
                      : drp5<-1>(L[c][lk][lj - 3], L[c][lk][lj - 2], L[c][lk][lj - 1], L[c][lk][lj], L[c][lk][lj + 1]);
      };
      auto Dz = [&](int c) {
        return FZ > 0 ? drp5<1>(L[c][lk - 1][lj], L[c][lk][lj], L[c][lk + 1][lj], L[c][lk + 2][lj], L[c][lk + 3][lj])
                      : drp5<-1>(L[c][lk - 3][lj], L[c][lk - 2][lj], L[c][lk - 1][lj], L[c][lk][lj], L[c][lk + 1][lj]);
      };
      const bool top = KIND == 2 && a.freenode && k >= nk - 3;  // uniform per wavefront (one k-row)
      // RHS accumulated derivative by derivative (wave.cu:133-276): each
      // field's derivative adds its momentum / Hooke terms, metric rows
      // pre-scaled by the PML 1/beta of their axis.
      float h[9];
#pragma unroll
      for (int c = 0; c < 9; c++) h[c] = 0.0f;
      {
        const float s1 = rb1 * rDH, s2 = rb2 * rDH, s3 = rb3 * rDH;
        accum_axis(h, Dx, xix * s1, xiy * s1, xiz * s1, rrho, lam, l2m, mu, true);
        accum_axis(h, Dy, etx * s2, ety * s2, etz * s2, rrho, lam, l2m, mu, true);
        accum_axis(h, Dz, ztx * s3, zty * s3, ztz * s3, rrho, lam, l2m, mu, !top);
      }
      if (top) {
        float dzv[3];
        if (k == nk - 1) {  // DzV from the traction-free condition (wave.cu:207-232)
          const uint32_t ij = ((uint32_t)(j + 3) + (uint32_t)(i + 3) * (uint32_t)ny) * 9;
          const float* mVx = s_mV[0];
          const float* mVy = s_mV[1];
          const float dxv[3] = {rDH * Dx(0), rDH * Dx(1), rDH * Dx(2)}, dyv[3] = {rDH * Dy(0), rDH * Dy(1), rDH * Dy(2)};
#pragma unroll
          for (int r = 0; r < 3; r++)
            dzv[r] = mVx[ij + 3 * r] * dxv[0] + mVx[ij + 3 * r + 1] * dxv[1] + mVx[ij + 3 * r + 2] * dxv[2]
                   + mVy[ij + 3 * r] * dyv[0] + mVy[ij + 3 * r + 1] * dyv[1] + mVy[ij + 3 * r + 2] * dyv[2];
        } else if (k == nk - 2) {  // L22 (wave.cu:233-238)
#pragma unroll
          for (int c = 0; c < 3; c++) dzv[c] = rDH * (FZ > 0 ? L[c][lk + 1][lj] - L[c][lk][lj] : L[c][lk][lj] - L[c][lk - 1][lj]);
        } else {  // L24 (wave.cu:239-245)
#pragma unroll
          for (int c = 0; c < 3; c++)
            dzv[c] = rDH * (FZ > 0 ? -C24_1 * L[c][lk][lj] + C24_2 * L[c][lk + 1][lj] - C24_3 * L[c][lk + 2][lj]
                                   : C24_3 * L[c][lk - 2][lj] - C24_2 * L[c][lk - 1][lj] + C24_1 * L[c][lk][lj]);
        }
        const float e0 = ztx * rb3, e1 = zty * rb3, e2 = ztz * rb3;
        h[3] += (l2m * e0 * dzv[0] + lam * e1 * dzv[1] + lam * e2 * dzv[2]);
        h[4] += (lam * e0 * dzv[0] + l2m * e1 * dzv[1] + lam * e2 * dzv[2]);
        h[5] += (lam * e0 * dzv[0] + lam * e1 * dzv[1] + l2m * e2 * dzv[2]);
        h[6] += (e1 * dzv[0] + e0 * dzv[1]) * mu;
        h[7] += (e2 * dzv[0] + e0 * dzv[2]) * mu;
        h[8] += (e2 * dzv[1] + e1 * dzv[2]) * mu;
        {  // the image velocity RHS (image_top_kernel)
          const long plane = (long)ni * nj, o = ((long)(k - (nk - 3))) * plane + (long)i * nj + j;
#pragma unroll
          for (int c = 0; c < 3; c++) h[c] = a.itop[(long)c * 3 * plane + o];
        }
      }

      // ADE-PML (pmlx.cu:126-210 and the y/z variants): the RHS of a point
      // inside the band of face f gets -aux_f/beta; the auxiliary fields
      // themselves are advanced afterwards by pml_aux_kernel from the same
      // stage input.
      if (pml) {
        const int N = s_N;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
          const int li = ax == 0 ? i : ax == 1 ? j : k;
          const int nn = ax == 0 ? ni : ax == 1 ? nj : nk;
#pragma unroll
          for (int side = 0; side < 2; side++) {
            const int f = 2 * ax + side;
            const int off = side ? nn - N : 0;
            if (!s_on[f] || li < off || li >= off + N) continue;
            const uint32_t pi = ax == 0 ? i - off : i, pj = ax == 1 ? j - off : j, pk = ax == 2 ? k - off : k;
            const uint32_t pnj = ax == 1 ? N : nj, pnk = ax == 2 ? N : nk, pni = ax == 0 ? N : ni;
            const size_t ps = (size_t)pni * pnj * pnk;
            const uint32_t pa4 = (pj + pk * pnj + pi * pnj * pnk) << 2;
            const float rbb = ax == 0 ? rb1 : ax == 1 ? rb2 : rb3;  // 1 / beta at li
            const float* Y = s_auxY[f];
#pragma unroll
            for (int c = 0; c < 9; c++) h[c] += -ldo(Y + c * ps, pa4) * rbb;
          }
        }
      }

      if (a.rhs_only) {
#pragma unroll
        for (int c = 0; c < 9; c++) bst(plane_rsrc(a.rhs, i + 3, 9, nyz), col4, c * pb, h[c]);
      } else if (a.fault && i >= a.ii0 - 3 && i <= a.ii0 + 3 && i != a.ii0) {
        const Rsrc r = plane_rsrc(a.slabH, i - (a.ii0 - 3), 9, nyz);  // slab planes share the layout
#pragma unroll
        for (int c = 0; c < 9; c++) bst(r, col4, c * pb, h[c]);
      } else {
        const float damp = cerjan ? fminf(s_cjx[i + 3], cjk) : 1.0f;
        const bool keep_t = damp != 1.0f || (a.fault && i == a.ii0);
        const Rsrc rs = plane_rsrc(a.S, i + 3, 9, nyz), rt = plane_rsrc(a.T, i + 3, 9, nyz);
        const Rsrc ro = plane_rsrc(a.Wout, i + 3, 9, nyz);
#pragma unroll
        for (int c = 0; c < 9; c++) {
          const float w = rk_combine_b(a.irk, a.rka, a.rkb, q[CTR][c], h[c], rs, rt, col4, c * pb, keep_t);
          bst(ro, col4, c * pb, w * damp);
        }
      }
    }

    if (more) {  // stage plane i+1 into the other LDS buffer
#pragma unroll
      for (int c = 0; c < 9; c++) lds[b ^ 1][c][tk + HKL][tj + HJL] = q[CTR + 1][c];
      if (tid < NH) {
#pragma unroll
        for (int c = 0; c < 9; c++) lds[b ^ 1][c][hlk][hlj] = hv[c];
      }
    }
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int c = 0; c < 9; c++) q[m][c] = q[m + 1][c];
#pragma unroll
    for (int c = 0; c < 9; c++) q[4][c] = nq[c];
    __syncthreads();
  }
}


// Raw metric words of one point (MetricBF16=2: five 32-bit pairs of bfloat16
// deviations, the fifth with the palette entry in its high half, and, when
// the media have no palette, the FP32 lambda, mu, rho; FP32: the 12 words
// xix..ztz, lam, mu, rho; bfloat16: the same 12 as 16-bit words), held
// undecoded while they wait for their plane.
template <class MT>
constexpr int met_raw_n() { return std::is_same<MT, MetBf16Delta>::value ? 8 : 12; }
template <class MT>
__device__ __forceinline__ void met_raw_load(Rsrc r, uint32_t voff, uint32_t mpb, uint32_t (&w)[met_raw_n<MT>()],
                                             bool palette, unsigned skip) {
  if constexpr (std::is_same<MT, MetBf16Delta>::value) {
#pragma unroll
    for (int p = 0; p < 5; p++)  // a word that is +-0 on the whole grid (skip bit p) is not loaded
      w[p] = (skip >> p) & 1u ? 0u : __float_as_uint(bld<kStream>(r, 2 * voff, (uint32_t)(2 * p) * mpb));
    if (palette) {
#pragma unroll
      for (int m = 0; m < 3; m++) w[5 + m] = 0u;
    } else {
#pragma unroll
      for (int m = 0; m < 3; m++) w[5 + m] = __float_as_uint(bld<kStream>(r, 2 * voff, (uint32_t)(10 + 2 * m) * mpb));
    }
  } else if constexpr (sizeof(MT) == 2) {
#pragma unroll
    for (int q = 0; q < 12; q++)
      w[q] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)((q < 9 ? q : q + 1) * mpb), kStream) << 16;
  } else {
#pragma unroll
    for (int q = 0; q < 12; q++) w[q] = __float_as_uint(bld<kStream>(r, voff, (uint32_t)(q < 9 ? q : q + 1) * mpb));
  }
}
// Field c (0..8 metric, 10..12 media) of raw metric words; pal is the
// palette in LDS, or null.
template <class MT>
__device__ __forceinline__ float met_raw_get(const uint32_t (&w)[met_raw_n<MT>()], int c, const float* ref,
                                             const float4* pal) {
  if constexpr (std::is_same<MT, MetBf16Delta>::value) {
    if (c >= 10) {
      if (pal) {
        const float4 e = pal[w[4] >> 16];
        return c == 10 ? e.x : c == 11 ? e.y : e.z;
      }
      return __uint_as_float(w[5 + c - 10]);
    }
    const uint32_t u = w[c >> 1];
    return __uint_as_float((c & 1) ? (u & 0xffff0000u) : (u << 16)) + ref[c];
  } else {
    return __uint_as_float(w[c < 9 ? c : c - 1]);
  }
}

// Interior stage kernel with every load of plane i+1 issued before plane i
// is computed: the points of the interior box (no absorbing layer, no free
// surface), where the general kernel's per-plane metric and RK-register
// loads would each cost a memory round trip that the whole block waits for
// at its barrier.  Per plane the block issues, in this order, the halo and
// the new queue plane of plane i+1, the metric of plane i+1, and the S and T
// words of plane i; vector-memory loads return in issue order, so waiting
// for the S and T words at the RK combine leaves nothing outstanding of the
// earlier ones, and those were issued a whole compute phase earlier.  The
// arithmetic is the general kernel's (same operators, same order of the
// sums), so results are identical to it bit for bit.
//
// BAND: the same pipeline on the points of the absorbing layers (the general
// kernel's KIND 1 tiles; no free-surface rows): the RHS of a point inside the
// ADE-PML band of a face gets -aux/beta of that face, and the Cerjan sponge
// damps the new value.  The aux words of plane i are issued with its S and T
// words, one face per axis at most (the host uses this kernel only when the
// two bands of an axis do not overlap), and added after the three axis sums
// in the general kernel's order; 1/beta and the sponge factor of plane i+1
// are fetched during plane i.
template <int FX, int FY, int FZ, class MT, int TJ_, int NT_, int MINB, bool BAND = false, bool RING = true,
          bool TOP = false>
__global__ void __launch_bounds__(NT_, MINB) wave_interior_kernel(WaveArgs<MT> a) {
  static_assert(!TOP || (BAND && TJ_ == 64), "free-surface tiles: band terms, one k row per wavefront");
  constexpr int TJ = TJ_, TK = NT_ / TJ_;
  constexpr int LJ = TJ + 4, LK = TK + 4;
  // halo cells of one plane: the 4 k-halo rows over the tile's own columns
  // and 4 j-halo cells per row; the corners of the LDS tile are neither
  // written nor read (the y and z operators reach along one axis only), so
  // each k-halo row costs the cache lines of the tile's own columns alone
  constexpr int NH = 4 * TJ + 4 * TK;
  constexpr int NHC = (NH + NT_ - 1) / NT_;  // halo cells per thread
  static_assert(NHC <= 2, "at most two halo cells per thread");
  constexpr int HJL = FY > 0 ? 1 : 3;
  constexpr int HKL = FZ > 0 ? 1 : 3;
  constexpr int QLO = FX > 0 ? -1 : -3;
  constexpr int CTR = -QLO;
  constexpr int NMW = met_raw_n<MT>();
  // RING: an LDS ring of the planes i .. i + LEAD: plane i + LEAD enters the
  // x-queue during step i, and its y/z halo is loaded in the same step, so
  // that a tile reads its neighbours' edge lines when they load those lines
  // as their own queue plane (still in L2), not LEAD planes later.
  // Otherwise a double buffer of planes i and i + 1 (LEAD 1), whose own
  // values come from the queue: a third of the LDS, for blocks that share a CU.
  constexpr int QLEAD = QLO + 5;  // the queue's newest plane is i + QLEAD
  constexpr int LEAD = RING ? QLEAD : 1;
  constexpr int NS = LEAD + 1;
  __shared__ float lds[NS][9][LK][LJ];
  // BAND: the absorbing layers' uniform descriptors, staged once per block
  __shared__ float* s_auxY[6];
  __shared__ float* s_auxS[6];
  __shared__ float* s_auxT[6];
  __shared__ const float* s_pA[3];
  __shared__ const float* s_pB[3];
  __shared__ const float* s_pD[3];
  __shared__ const float* s_pBx;
  __shared__ const float* s_cjx;
  __shared__ int s_on[6];
  __shared__ int s_N;
  __shared__ const float* s_mV[2];  // TOP: the free-surface velocity maps
  // the media palette of the bfloat16-deviation form, staged once per block
  __shared__ float4 s_pal[std::is_same<MT, MetBf16Delta>::value ? kMaxMediaPalette : 1];

  const int item = xcd_remap((int)blockIdx.x, a.nwork);
  const int4 wk = a.work[2 * item], we = a.work[2 * item + 1];
  const int j0 = wk.x, k0 = wk.y, ibeg = wk.z, iend = wk.w;
  const int tj = threadIdx.x, tk = threadIdx.y, tid = tj + tk * TJ;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int j = j0 + tj, k = k0 + tk;
  const bool act = tj >= we.z && tj < we.x && tk >= we.w && tk < we.y;
  const int nxp = ni + 6;
  const uint32_t nyz = a.nyz;
  const uint32_t py = a.py, jpad = a.jpad;
  const uint32_t col = jpad + (uint32_t)(min(max(j, -3), nj + 2) + 3) + (uint32_t)(min(max(k, -3), nk + 2) + 3) * py;
  const uint32_t col4 = col << 2;
  const uint32_t mcol = col * (uint32_t)sizeof(MT);
  const uint32_t pb = nyz * 4, mpb = nyz * (uint32_t)sizeof(MT);
  const float* __restrict__ Win = a.Win;
  const int irk = a.irk;

  // this thread's halo cells tid and tid + NT_: the 4 k-halo rows, then the
  // 4 j-halo cells of each row
  int hlk[NHC], hlj[NHC];
  uint32_t hcol4[NHC];
#pragma unroll
  for (int e = 0; e < NHC; e++) {
    const int h = tid + e * NT_;
    hlk[e] = 0; hlj[e] = 0;
    uint32_t hcol = col;
    if (h < NH) {
      if (h < 4 * TJ) {
        const int r = h / TJ;
        hlk[e] = r < HKL ? r : r + TK;
        hlj[e] = HJL + h % TJ;
      } else {
        const int h2 = h - 4 * TJ;
        hlk[e] = HKL + h2 / 4;
        const int cc = h2 % 4;
        hlj[e] = cc < HJL ? cc : cc + TJ;
      }
      const int gj = min(max(j0 + hlj[e] - HJL, -3), nj + 2), gk = min(max(k0 + hlk[e] - HKL, -3), nk + 2);
      hcol = jpad + (uint32_t)(gj + 3) + (uint32_t)(gk + 3) * py;
    }
    hcol4[e] = hcol << 2;
  }
  auto load_halo = [&](Rsrc r, float (&v)[NHC][9]) {
#pragma unroll
    for (int e = 0; e < NHC; e++)
      if (e == 0 || tid + e * NT_ < NH) {
#pragma unroll
        for (int c = 0; c < 9; c++) v[e][c] = bld(r, hcol4[e], c * pb);
      }
  };
  auto store_halo = [&](float (*L)[LK][LJ], const float (&v)[NHC][9]) {
#pragma unroll
    for (int e = 0; e < NHC; e++)
      if (tid + e * NT_ < NH) {
#pragma unroll
        for (int c = 0; c < 9; c++) L[c][hlk[e]][hlj[e]] = v[e][c];
      }
  };

  // BAND: per lane, the axes whose absorbing band (one side) holds the
  // point's column (j, k), their 1/beta, and the sponge factor of the column
  const bool pml = BAND && a.pml, cerjan = BAND && a.cerjan;
  float rb2 = 1.0f, rb3 = 1.0f, cjk = 1.0f;
  int N = 0;
  bool inj = false, ink = false;
  uint32_t offj = 0, offk = 0;
  if constexpr (BAND) {
    if (threadIdx.y == 0 && threadIdx.x < 6) {
      s_auxY[threadIdx.x] = a.cold->auxY[threadIdx.x];
      s_auxS[threadIdx.x] = a.cold->auxS[threadIdx.x];
      s_auxT[threadIdx.x] = a.cold->auxT[threadIdx.x];
      s_on[threadIdx.x] = a.cold->on[threadIdx.x];
      if (threadIdx.x < 3) {
        s_pA[threadIdx.x] = a.cold->pA[threadIdx.x];
        s_pB[threadIdx.x] = a.cold->pB[threadIdx.x];
        s_pD[threadIdx.x] = a.cold->pD[threadIdx.x];
      }
      if (threadIdx.x == 0) { s_N = a.cold->N; s_pBx = a.cold->pB[0]; s_cjx = a.cold->cj[0]; }
      if (TOP && threadIdx.x < 2) s_mV[threadIdx.x] = threadIdx.x ? a.cold->mVy : a.cold->mVx;
    }
    if (act && pml) { rb2 = 1.0f / a.cold->pB[1][j]; rb3 = 1.0f / a.cold->pB[2][k]; }
    if (act && cerjan) cjk = fminf(a.cold->cj[1][j + 3], a.cold->cj[2][k + 3]);
    __syncthreads();
    if (pml) {
      N = s_N;
      if (s_on[2] && j < N) { inj = true; offj = 0; }
      else if (s_on[3] && j >= nj - N) { inj = true; offj = nj - N; }
      if (s_on[4] && k < N) { ink = true; offk = 0; }
      else if (s_on[5] && k >= nk - N) { ink = true; offk = nk - N; }
    }
  }
  // 1/beta of plane i along x and the sponge factor of plane i (uniform)
  auto rb_x = [&](int i) { return pml ? 1.0f / s_pBx[i] : 1.0f; };
  auto cj_x = [&](int i) { return cerjan ? s_cjx[i + 3] : 1.0f; };

  float q[5][9];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const Rsrc r = plane_rsrc(Win, ibeg + QLO + m + 3, 9, nyz);
#pragma unroll
    for (int c = 0; c < 9; c++) q[m][c] = bld(r, col4, c * pb);
  }
  // planes ibeg .. ibeg + LEAD - 1 of the ring: own values from the queue,
  // halos loaded here
  float hv[NHC][9];
#pragma unroll
  for (int m = 0; m < LEAD; m++) {
    if (ibeg + m < iend) {  // uniform per block
      load_halo(plane_rsrc(Win, ibeg + m + 3, 9, nyz), hv);
#pragma unroll
      for (int c = 0; c < 9; c++) lds[m][c][tk + HKL][tj + HJL] = q[CTR + m][c];
      store_halo(lds[m], hv);
    }
  }
  const bool palette = std::is_same<MT, MetBf16Delta>::value && a.pal != nullptr;  // uniform
  if (palette)
    for (int e = tid; e < a.npal; e += NT_) s_pal[e] = a.pal[e];
  uint32_t mw[NMW];
  const unsigned met_skip = a.met_skip;  // uniform
  met_raw_load<MT>(plane_rsrc(a.met, ibeg + 3, met_nf<MT>(), nyz), mcol, mpb, mw, palette, met_skip);
  float rb1 = 1.0f, cjx = 1.0f;
  if constexpr (BAND) { rb1 = rb_x(ibeg); cjx = cj_x(ibeg); }
  __syncthreads();

  int b = 0;  // ring slot of plane i; plane i + LEAD goes to the slot before it
  for (int i = ibeg; i < iend; ++i) {
    const bool more = i + 1 < iend;       // uniform per block
    const bool ahead = i + LEAD < iend;   // uniform per block
    const int bw = b == 0 ? NS - 1 : b - 1;
    // plane i + QLEAD: queue plane, and with RING its halo (clamped
    // in-bounds past the chunk, where the queue plane goes unused); plane
    // i + 1: the halo without RING, and the metric
    float nq[9];
    {
      const Rsrc r = plane_rsrc(Win, min(i + QLEAD + 3, nxp - 1), 9, nyz);
#pragma unroll
      for (int c = 0; c < 9; c++) nq[c] = bld(r, col4, c * pb);
      if (RING && ahead) load_halo(r, hv);
    }
    if (!RING && ahead) load_halo(plane_rsrc(Win, i + 4, 9, nyz), hv);
    uint32_t mwn[NMW];
    met_raw_load<MT>(plane_rsrc(a.met, min(i + 4, nxp - 1), met_nf<MT>(), nyz), mcol, mpb, mwn, palette, met_skip);
    float rb1n = 1.0f, cjxn = 1.0f;
    if constexpr (BAND) {
      if (more) { rb1n = rb_x(i + 1); cjxn = cj_x(i + 1); }
    }
    // plane i: the step-start (S) and accumulator (T) words of the RK
    // combine; stage 1 derives T from S and the point's stage-0 value
    // unless the point keeps it (rk_combine_b)
    const bool slab = a.fault && i >= a.ii0 - 3 && i <= a.ii0 + 3 && i != a.ii0;
    const float damp = cerjan ? fminf(cjx, cjk) : 1.0f;
    const bool keep_t = damp != 1.0f || (a.fault && i == a.ii0);
    const Rsrc rs = plane_rsrc(a.S, i + 3, 9, nyz), rt = plane_rsrc(a.T, i + 3, 9, nyz);
    float sv[9], tv[9];
#pragma unroll
    for (int c = 0; c < 9; c++) { sv[c] = 0.0f; tv[c] = 0.0f; }
    if (!a.rhs_only && !slab && irk != 0) {
      if (irk != 1 || keep_t) {
#pragma unroll
        for (int c = 0; c < 9; c++) tv[c] = bld<kStream>(rt, col4, c * pb);
      }
      if (irk != 3) {
#pragma unroll
        for (int c = 0; c < 9; c++) sv[c] = bld<kStream>(rs, col4, c * pb);
      }
    }
    // BAND: the face along each axis whose band holds the point, and the aux
    // words of plane i (pmlx.cu:126-210 and the y/z variants), all issued
    // with S and T: the first face's value in ya; for a point on that face
    // alone and the fused auxiliary update, its step-start (S) and
    // accumulator (T) words in sa and ta; for a point on two or three faces
    // (an edge of the absorbing box, whose auxiliary fields pml_aux_kernel
    // advances) the second and third faces' values in sa and ta instead
    float ya[9], sa[9], ta[9];
    bool inx = false;
    int fa = -1, fb = -1;               // first and second face axes of the point
    int nfc = 0;                        // faces holding the point
    int fo[3] = {0, 0, 0};              // face of each axis
    uint32_t pa4s[3] = {0, 0, 0};       // aux byte offset of the point in the face of each axis
    auto face_ps = [&](int ax) { return (size_t)(ax == 0 ? N : ni) * (ax == 1 ? N : nj) * (ax == 2 ? N : nk); };
    if constexpr (BAND) {
#pragma unroll
      for (int c = 0; c < 9; c++) { ya[c] = 0.0f; sa[c] = 0.0f; ta[c] = 0.0f; }
      if (pml) {
        uint32_t offi = 0;
        if (s_on[0] && i < N) { inx = true; offi = 0; }
        else if (s_on[1] && i >= ni - N) { inx = true; offi = ni - N; }
        const bool fuse = !TOP && a.aux_fused && !a.rhs_only;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
          const bool in = ax == 0 ? inx : ax == 1 ? inj : ink;
          if (!act || !in) continue;
          const uint32_t off = ax == 0 ? offi : ax == 1 ? offj : offk;
          fo[ax] = 2 * ax + (off != 0);
          const uint32_t pi = ax == 0 ? i - offi : i, pj = ax == 1 ? j - offj : j, pk = ax == 2 ? k - offk : k;
          const uint32_t pnj = ax == 1 ? N : nj, pnk = ax == 2 ? N : nk;
          pa4s[ax] = (pj + pk * pnj + pi * pnj * pnk) << 2;
          if (fa < 0) fa = ax;
          else if (fb < 0) fb = ax;
          nfc++;
        }
        if (fa >= 0) {
          const int f = fo[fa];
          const size_t ps = face_ps(fa);
          const uint32_t pa4 = pa4s[fa];
#pragma unroll
          for (int c = 0; c < 9; c++) ya[c] = ldo(s_auxY[f] + c * ps, pa4);
          if (nfc == 1 && fuse && (irk == 2 || irk == 3)) {  // stage 1 derives the accumulator
#pragma unroll
            for (int c = 0; c < 9; c++) ta[c] = ldo(s_auxT[f] + c * ps, pa4);
          }
          if (nfc == 1 && fuse && (irk == 1 || irk == 2)) {
#pragma unroll
            for (int c = 0; c < 9; c++) sa[c] = ldo(s_auxS[f] + c * ps, pa4);
          }
        }
        if (nfc >= 2) {  // the second face's value
          const size_t ps = face_ps(fb);
#pragma unroll
          for (int c = 0; c < 9; c++) sa[c] = ldo(s_auxY[fo[fb]] + c * ps, pa4s[fb]);
        }
        if (nfc == 3) {  // the third face's value (the axes are then 0, 1, 2)
          const size_t ps = face_ps(2);
#pragma unroll
          for (int c = 0; c < 9; c++) ta[c] = ldo(s_auxY[fo[2]] + c * ps, pa4s[2]);
        }
      }
    }
    // aux value c of the face along axis ax (a face holding the point)
    auto aux_y = [&](int ax, int c) { return ax == fa ? ya[c] : ax == fb ? sa[c] : ta[c]; };

    if (act) {
      auto MET = [&](int c) { return met_raw_get<MT>(mw, c, a.mref, palette ? s_pal : nullptr); };
      const float xix = MET(0), xiy = MET(1), xiz = MET(2);
      const float etx = MET(3), ety = MET(4), etz = MET(5);
      const float ztx = MET(6), zty = MET(7), ztz = MET(8);
      const float lam = MET(10), mu = MET(11);
      const float rrho = 1.0f / MET(12);
      const float l2m = lam + 2.0f * mu;
      const float rDH = a.rDH;
      const float(*L)[LK][LJ] = lds[b];
      const int lj = tj + HJL, lk = tk + HKL;
      auto Dx = [&](int c) { return drp5<FX>(q[0][c], q[1][c], q[2][c], q[3][c], q[4][c]); };
      auto Dy = [&](int c) {
        return FY > 0 ? drp5<1>(L[c][lk][lj - 1], L[c][lk][lj], L[c][lk][lj + 1], L[c][lk][lj + 2], L[c][lk][lj + 3])
                      : drp5<-1>(L[c][lk][lj - 3], L[c][lk][lj - 2], L[c][lk][lj - 1], L[c][lk][lj], L[c][lk][lj + 1]);
      };
      auto Dz = [&](int c) {
        return FZ > 0 ? drp5<1>(L[c][lk - 1][lj], L[c][lk][lj], L[c][lk + 1][lj], L[c][lk + 2][lj], L[c][lk + 3][lj])
                      : drp5<-1>(L[c][lk - 3][lj], L[c][lk - 2][lj], L[c][lk - 1][lj], L[c][lk][lj], L[c][lk + 1][lj]);
      };
      float h[9];
#pragma unroll
      for (int c = 0; c < 9; c++) h[c] = 0.0f;
      {
        // the general kernel's arithmetic (in the interior beta = 1: rb * rDH = rDH)
        const float s1 = rb1 * rDH, s2 = rb2 * rDH, s3 = rb3 * rDH;
        accum_axis(h, Dx, xix * s1, xiy * s1, xiz * s1, rrho, lam, l2m, mu, true);
        accum_axis(h, Dy, etx * s2, ety * s2, etz * s2, rrho, lam, l2m, mu, true);
        accum_axis(h, Dz, ztx * s3, zty * s3, ztz * s3, rrho, lam, l2m, mu, !(TOP && a.freenode && k >= nk - 3));
      }
      if constexpr (TOP) {
        // the top three rows under the free surface (wave.cu:207-245), as
        // the general kernel computes them: DzV of the stress rates from the
        // traction-free condition, L22 or L24, and the velocity RHS of the
        // image (image_top_kernel); one k row per wavefront
        if (a.freenode && k >= nk - 3) {
          float dzv[3];
          if (k == nk - 1) {  // traction-free condition (wave.cu:207-232)
            const uint32_t ij = ((uint32_t)(j + 3) + (uint32_t)(i + 3) * (uint32_t)a.ny) * 9;
            const float* mVx = s_mV[0];
            const float* mVy = s_mV[1];
            const float dxv[3] = {rDH * Dx(0), rDH * Dx(1), rDH * Dx(2)}, dyv[3] = {rDH * Dy(0), rDH * Dy(1), rDH * Dy(2)};
#pragma unroll
            for (int r = 0; r < 3; r++)
              dzv[r] = mVx[ij + 3 * r] * dxv[0] + mVx[ij + 3 * r + 1] * dxv[1] + mVx[ij + 3 * r + 2] * dxv[2]
                     + mVy[ij + 3 * r] * dyv[0] + mVy[ij + 3 * r + 1] * dyv[1] + mVy[ij + 3 * r + 2] * dyv[2];
          } else if (k == nk - 2) {  // L22 (wave.cu:233-238)
#pragma unroll
            for (int c = 0; c < 3; c++) dzv[c] = rDH * (FZ > 0 ? L[c][lk + 1][lj] - L[c][lk][lj] : L[c][lk][lj] - L[c][lk - 1][lj]);
          } else {  // L24 (wave.cu:239-245)
#pragma unroll
            for (int c = 0; c < 3; c++)
              dzv[c] = rDH * (FZ > 0 ? -C24_1 * L[c][lk][lj] + C24_2 * L[c][lk + 1][lj] - C24_3 * L[c][lk + 2][lj]
                                     : C24_3 * L[c][lk - 2][lj] - C24_2 * L[c][lk - 1][lj] + C24_1 * L[c][lk][lj]);
          }
          const float e0 = ztx * rb3, e1 = zty * rb3, e2 = ztz * rb3;
          h[3] += (l2m * e0 * dzv[0] + lam * e1 * dzv[1] + lam * e2 * dzv[2]);
          h[4] += (lam * e0 * dzv[0] + l2m * e1 * dzv[1] + lam * e2 * dzv[2]);
          h[5] += (lam * e0 * dzv[0] + lam * e1 * dzv[1] + l2m * e2 * dzv[2]);
          h[6] += (e1 * dzv[0] + e0 * dzv[1]) * mu;
          h[7] += (e2 * dzv[0] + e0 * dzv[2]) * mu;
          h[8] += (e2 * dzv[1] + e1 * dzv[2]) * mu;
          const long plane = (long)ni * nj, o = ((long)(k - (nk - 3))) * plane + (long)i * nj + j;
#pragma unroll
          for (int c = 0; c < 3; c++) h[c] = a.itop[(long)c * 3 * plane + o];
        }
      }
      if constexpr (BAND) {
        if (pml) {
#pragma unroll
          for (int ax = 0; ax < 3; ax++) {
            const bool in = ax == 0 ? inx : ax == 1 ? inj : ink;
            const float rbb = ax == 0 ? rb1 : ax == 1 ? rb2 : rb3;  // 1 / beta at the point
            if (in) {
#pragma unroll
              for (int c = 0; c < 9; c++) h[c] += -aux_y(ax, c) * rbb;
            }
          }
          // the fused auxiliary update (what pml_aux_kernel does for the
          // point, with the same arithmetic): aux' = -(alpha + d) aux +
          // d * RHS_axis, RHS_axis from the plain derivative along the face's
          // axis, advanced with the RK4 stage combine
          if (!TOP && a.aux_fused && !a.rhs_only && nfc == 1) {  // a point on one face only
            const float rka = a.rka, rkb = a.rkb;
#pragma unroll
            for (int ax = 0; ax < 3; ax++) {
              if (ax != fa) continue;
              float dd[9];
#pragma unroll
              for (int c = 0; c < 9; c++) dd[c] = rDH * (ax == 0 ? Dx(c) : ax == 1 ? Dy(c) : Dz(c));
              const float e0 = ax == 0 ? xix : ax == 1 ? etx : ztx;
              const float e1 = ax == 0 ? xiy : ax == 1 ? ety : zty;
              const float e2 = ax == 0 ? xiz : ax == 1 ? etz : ztz;
              const int li = ax == 0 ? i : ax == 1 ? j : k;
              const float d1 = s_pD[ax][li], ad1 = -(s_pA[ax][li] + d1);
              const int f = fo[ax];
              const size_t ps = face_ps(ax);
              const uint32_t pa4 = pa4s[ax];
              float* Y = s_auxY[f];
              float* S = s_auxS[f];
              float* T = s_auxT[f];
#pragma unroll
              for (int c = 0; c < 9; c++) {
                const float y = aux_y(ax, c);
                const float ah = ad1 * y + d1 * axis_rhs_c(c, dd, e0, e1, e2, rrho, lam, l2m, mu);
                float w;
                if (irk == 0) {  // the accumulator is derived in stage 1 (rk_derived_t)
                  sto(S + c * ps, pa4, y);
                  w = y + rka * ah;
                } else if (irk == 3) {
                  w = ta[c] + rkb * ah;
                } else {
                  const float t = irk == 1 ? rk_derived_t(sa[c], y) : ta[c];
                  sto(T + c * ps, pa4, t + rkb * ah);
                  w = sa[c] + rka * ah;
                }
                sto(Y + c * ps, pa4, w);
              }
            }
          }
        }
      }
      if (a.rhs_only) {
#pragma unroll
        for (int c = 0; c < 9; c++) bst(plane_rsrc(a.rhs, i + 3, 9, nyz), col4, c * pb, h[c]);
      } else if (slab) {
        const Rsrc r = plane_rsrc(a.slabH, i - (a.ii0 - 3), 9, nyz);
#pragma unroll
        for (int c = 0; c < 9; c++) bst(r, col4, c * pb, h[c]);
      } else {
        // RK_Syn (rk4.cu:20-33); the damping factor is 1 in the interior
        const Rsrc ro = plane_rsrc(a.Wout, i + 3, 9, nyz);
        const float rka = a.rka, rkb = a.rkb;
#pragma unroll
        for (int c = 0; c < 9; c++) {
          float w;
          if (irk == 0) {
            if (keep_t) bst(rt, col4, c * pb, q[CTR][c] + rkb * h[c]);
            w = q[CTR][c] + rka * h[c];
          } else if (irk == 3) {
            w = tv[c] + rkb * h[c];
          } else {
            const float t = irk == 1 && !keep_t ? rk_derived_t(sv[c], q[CTR][c]) : tv[c];
            bst(rt, col4, c * pb, t + rkb * h[c]);
            w = sv[c] + rka * h[c];
          }
          bst(ro, col4, c * pb, w * damp);
        }
      }
    }

    if (ahead) {
#pragma unroll
      for (int c = 0; c < 9; c++) lds[bw][c][tk + HKL][tj + HJL] = RING ? nq[c] : q[CTR + 1][c];
      store_halo(lds[bw], hv);
    }
    b = b == NS - 1 ? 0 : b + 1;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int c = 0; c < 9; c++) q[m][c] = q[m + 1][c];
#pragma unroll
    for (int c = 0; c < 9; c++) q[4][c] = nq[c];
#pragma unroll
    for (int m = 0; m < NMW; m++) mw[m] = mwn[m];
    if constexpr (BAND) { rb1 = rb1n; cjx = cjxn; }
    __syncthreads();
  }
}

// ADE-PML auxiliary fields (pmlx.cu:126-210, pmly.cu, pmlz.cu): for every
// point of the band of face f, the plain one-sided derivative along the
// face's axis drives aux' = -(alpha + d/beta) aux + (d/beta) * RHS_axis, plus
// the traction-free term on the free surface for the x/y faces, advanced
// with the same RK4 stage combine as the wavefield.  Runs after the stage
// kernel (which read the old aux values) on the unchanged stage input.  One
// thread per band point, j fastest so aux accesses are contiguous; on x
// faces the band plane index comes next, so the five planes of an
// x-derivative are read by neighbouring blocks and stay in L2.
template <class MT>
struct PmlAuxArgs {
  const float* __restrict__ Win;
  const MT* __restrict__ met;
  // every face's arrays, passed directly (kernel arguments are preloaded):
  // auxiliary value, step start and RK accumulator; per axis the PML
  // profiles; the free-surface velocity maps of the x and y faces.  Picked
  // with constant indices (a runtime index into the argument block would
  // copy it to scratch memory).
  float* Y[6];
  float* S[6];
  float* T[6];
  const float* pA[3];
  const float* pB[3];
  const float* pD[3];
  const float* mat[2];
  // boxes of points to update, kPmlBoxInts ints each: face, then the
  // face-local (start, count) of j, k and i, then the box's first block
  const int* __restrict__ boxes;
  int nbox;
  int flag[3];  // one-sided operator sign along each axis
  int N;
  int ni, nj, nk, ny, py, jpad;
  unsigned pyz;
  float rDH, rka, rkb;
  float mref[10];   // metric reference values (bfloat16-deviation form)
  int irk, freenode;
};

template <class T>
__device__ __forceinline__ T pick6(int f, const T (&v)[6]) {
  return f == 0 ? v[0] : f == 1 ? v[1] : f == 2 ? v[2] : f == 3 ? v[3] : f == 4 ? v[4] : v[5];
}
template <class T>
__device__ __forceinline__ T pick3(int x, const T (&v)[3]) { return x == 0 ? v[0] : x == 1 ? v[1] : v[2]; }

// One launch for all absorbing faces: each box of points (a face, or a part
// of one) is served by a contiguous run of blocks, numbered
// XCD-contiguously so that the blocks sharing stencil planes (neighbouring
// band planes on x faces, rows on z faces) share one L2.
template <class MT>
__global__ void __launch_bounds__(256) pml_aux_kernel(PmlAuxArgs<MT> a) {
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  // the box of this block (block-uniform loads; a handful of boxes)
  const int* B = a.boxes;
  int q = 0;
  for (int r = 1; r < a.nbox; r++)
    if (b >= B[r * kPmlBoxInts + 7]) q = r;
  const int* bx = B + q * kPmlBoxInts;
  const int f = bx[0], bj0 = bx[1], bnj = bx[2], bk0 = bx[3], bnk = bx[4], bi0 = bx[5], bni = bx[6];
  const int ax = f >> 1, N = a.N, flag = pick3(ax, a.flag);
  const int nn = ax == 0 ? a.ni : ax == 1 ? a.nj : a.nk;
  const int off = (f & 1) ? nn - N : 0;
  const int pnj = ax == 1 ? N : a.nj, pnk = ax == 2 ? N : a.nk, pni = ax == 0 ? N : a.ni;
  const long ps = (long)pni * pnj * pnk;
  const long idx = (long)(b - bx[7]) * 256L + threadIdx.x;
  if (idx >= (long)bnj * bnk * bni) return;
  const int pj = bj0 + (int)(idx % bnj);
  const long t = idx / bnj;
  // j fastest, then the band plane index on x faces (its five x-derivative
  // planes are read by neighbouring blocks), k on the others
  const int pk = bk0 + (ax == 0 ? (int)(t / bni) : (int)(t % bnk));
  const int pi = bi0 + (ax == 0 ? (int)(t % bni) : (int)(t / bnk));
  const long pa = pj + (long)pk * pnj + (long)pi * pnj * pnk;  // aux layout [c][pi][pk][pj]
  const int i = pi + (ax == 0 ? off : 0), j = pj + (ax == 1 ? off : 0), k = pk + (ax == 2 ? off : 0);
  const int li = ax == 0 ? i : ax == 1 ? j : k;
  const unsigned pyz = a.pyz;
  const long col = a.jpad + (j + 3) + (long)(k + 3) * a.py;
  const int x = i + 3;
  const float* W = a.Win;
  // neighbour offset (words) along the axis, within one field's plane or across planes
  const long st = ax == 0 ? 9L * pyz : ax == 1 ? 1L : (long)a.py;
  const long base = (long)x * 9 * pyz + col;
  float dd[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* w = W + base + (long)c * pyz;
    dd[c] = a.rDH * (flag > 0 ? drp5<1>(w[-st], w[0], w[st], w[2 * st], w[3 * st])
                                : drp5<-1>(w[-3 * st], w[-2 * st], w[-st], w[0], w[st]));
  }
  const MT* M = a.met;
  auto MET = [&](int m) { return mat_(M, m, x, (uint32_t)col, pyz, a.mref); };
  const float e0 = MET(3 * ax), e1 = MET(3 * ax + 1), e2 = MET(3 * ax + 2);
  const float lam = MET(10), mu = MET(11), rrho = 1.0f / MET(12), l2m = lam + 2.0f * mu;
  const float* pAx = pick3(ax, a.pA);
  const float* pBx = pick3(ax, a.pB);
  const float* pDx = pick3(ax, a.pD);
  const float b1 = pBx[li], d1 = pDx[li], ad1 = -(pAx[li] + d1);
  const float c0 = d1 * b1;
  const bool fsx = ax < 2 && a.freenode && k == a.nk - 1;
  float z1[3] = {0.0f, 0.0f, 0.0f}, ztx = 0.0f, zty = 0.0f, ztz = 0.0f;
  if (fsx) {
    const float* mat = ax == 0 ? a.mat[0] : a.mat[1];
    const long ij = ((long)(j + 3) + (long)(i + 3) * a.ny) * 9;
#pragma unroll
    for (int r = 0; r < 3; r++) z1[r] = mat[ij + 3 * r] * dd[0] + mat[ij + 3 * r + 1] * dd[1] + mat[ij + 3 * r + 2] * dd[2];
    ztx = MET(6); zty = MET(7); ztz = MET(8);
  }
  float* Y = pick6(f, a.Y);
  float* S = pick6(f, a.S);
  float* T = pick6(f, a.T);
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const long q = c * ps + pa;
    const float y = Y[q];
    float ah = ad1 * y + d1 * axis_rhs_c(c, dd, e0, e1, e2, rrho, lam, l2m, mu);
    if (fsx && c >= 3) ah += pml_fs_term(c, z1, ztx, zty, ztz, lam, l2m, mu, c0);
    // RK_Syn (rk4.cu:20-33) on the step start S and accumulator T of the
    // auxiliary field, the accumulator of stage 0 derived in stage 1 from S
    // and the stage-0 value (rk_derived_t), as in the band tiles
    float w;
    if (a.irk == 0) {
      S[q] = y;
      w = y + a.rka * ah;
    } else if (a.irk == 3) {
      w = T[q] + a.rkb * ah;
    } else {
      const float s = S[q];
      const float t = a.irk == 1 ? rk_derived_t(s, y) : T[q];
      T[q] = t + a.rkb * ah;
      w = s + a.rka * ah;
    }
    Y[q] = w;
  }
}

template <int FX, int FY, int FZ, class MT>
void launch_wave_t(WaveArgs<MT> a, const Ctx& c, hipStream_t s) {
  // work list order: interior, bands (64-wide tiles), bands (32-wide tiles), free-surface tiles
  const int* n = c.wave_nkind;
  const int o1 = n[0], o2 = o1 + n[1], o3 = o2 + n[2];
  const int4* w = (const int4*)c.wave_work;
  Profiler& pr = const_cast<Ctx&>(c).prof;
  // StageStreams: the interior tiles and the free-surface passes share no
  // words with the bands or with each other (disjoint points, read-only stage
  // input), so they run on side streams, fenced after the work already queued
  // on the solver's stream; the auxiliary update waits for the surface tiles
  // (which read the aux fields) and the caller's next launch for the interior.
  // While kernel classes are timed the launches stay serial, so that each
  // class's events bracket its own kernels only.  A failing event call sets
  // HIP's last error, which the caller's launch check reports.
  const bool par = c.side[0] != nullptr && !c.prof.on;
  hipStream_t si = s, ss = s;
  if (par) {
    (void)hipEventRecord(c.fence[0], s);
    si = c.side[0];
    ss = c.side[1];
    (void)hipStreamWaitEvent(si, c.fence[0], 0);
    (void)hipStreamWaitEvent(ss, c.fence[0], 0);
  }
  if (n[0] > 0) {
    a.work = w; a.nwork = n[0];
    pr.begin(si, PK_STAGE_INTERIOR);
    if (c.wave_tki == TK6)
      hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, TJ * TK6, 2, false, false>), dim3(a.nwork), dim3(TJ, TK6, 1),
                         0, si, a);
    else if (c.wave_tki == TK4)  // at least 3 waves per SIMD: three blocks per CU (two with FP32 metric words,
      // whose 12 raw words per point do not fit 168 VGPRs: 148 B per lane of
      // spills, 98-100 against 94-96 ms of interior per step at 1600 x 1000 x 500)
      hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, TJ * TK4, std::is_same<MT, float>::value ? 2 : 3, false, false>),
                         dim3(a.nwork), dim3(TJ, TK4, 1), 0, si, a);
    else if (c.wave_pipe && c.wave_ring)
      hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, NTHR, 2>), dim3(a.nwork), dim3(TJ, TK, 1), 0, si, a);
    else if (c.wave_pipe)
      hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, NTHR, 2, false, false>), dim3(a.nwork), dim3(TJ, TK, 1), 0,
                         si, a);
    else if (c.wave_tall)
      hipLaunchKernelGGL((wave_stage_kernel<FX, FY, FZ, MT, 0, TJ, 2 * NTHR>), dim3(a.nwork), dim3(TJ, 2 * TK, 1), 0, si, a);
    else
      hipLaunchKernelGGL((wave_stage_kernel<FX, FY, FZ, MT, 0>), dim3(a.nwork), dim3(TJ, TK, 1), 0, si, a);
    pr.end(si);
  }
  if (n[3] > 0) {
    a.work = w + 2 * o3; a.nwork = n[3];
    pr.begin(ss, PK_STAGE_SURFACE);
    if (a.freenode) {
      const long n1 = (long)c.d.nx * 6 * c.d.ny, n2 = (long)c.d.ni * 3 * c.d.nj;
      hipLaunchKernelGGL(traction_top_kernel<MT>, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, ss, a, c.d.nx);
      hipLaunchKernelGGL((image_top_kernel<FX, FY, FZ, MT>), dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, ss, a);
    }
    if (c.wave_surf_pipe)
      hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, NTHR, 2, true, true, true>), dim3(a.nwork), dim3(TJ, TK, 1),
                         0, ss, a);
    else
      hipLaunchKernelGGL((wave_stage_kernel<FX, FY, FZ, MT, 2>), dim3(a.nwork), dim3(TJ, TK, 1), 0, ss, a);
    pr.end(ss);
  }
  if (n[1] + n[2] > 0) {
    pr.begin(s, PK_STAGE_BAND);
    if (n[1] > 0) {
      a.work = w + 2 * o1; a.nwork = n[1];
      if (c.wave_band_pipe && c.wave_band_tk == TK4)
        hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, TJ * TK4, 2, true, false>), dim3(a.nwork), dim3(TJ, TK4, 1),
                           0, s, a);
      else if (c.wave_band_pipe)
        hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJ, NTHR, 2, true>), dim3(a.nwork), dim3(TJ, TK, 1), 0, s, a);
      else
        hipLaunchKernelGGL((wave_stage_kernel<FX, FY, FZ, MT, 1>), dim3(a.nwork), dim3(TJ, TK, 1), 0, s, a);
    }
    if (n[2] > 0) {
      a.work = w + 2 * o2; a.nwork = n[2];
      if (c.wave_band_pipe && c.wave_tjn == TJB)
        hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJB, TJB * TKB, 2, true>), dim3(a.nwork), dim3(TJB, TKB, 1), 0,
                           s, a);
      else if (c.wave_band_pipe)
        hipLaunchKernelGGL((wave_interior_kernel<FX, FY, FZ, MT, TJN, NTHRN, 2, true>), dim3(a.nwork), dim3(TJN, TKN, 1), 0,
                           s, a);
      else
        hipLaunchKernelGGL((wave_stage_kernel<FX, FY, FZ, MT, 1, TJN, NTHRN>), dim3(a.nwork), dim3(TJN, TKN, 1), 0, s, a);
    }
    pr.end(s);
  }
  if (par) {
    (void)hipEventRecord(c.fence[2], ss);
    (void)hipStreamWaitEvent(s, c.fence[2], 0);
  }
  // the auxiliary fields of the points the pipelined band tiles do not
  // advance themselves (all of them with the general band kernel; with the
  // fused update those of the free-surface rows k >= ktop and of the edges
  // of the absorbing box, on two or three faces)
  if (a.pml && !a.rhs_only) {
    PmlAuxArgs<MT> p{};
    p.Win = a.Win; p.met = a.met; p.N = c.pml.N;
    p.ni = a.ni; p.nj = a.nj; p.nk = a.nk; p.ny = a.ny; p.py = a.py; p.jpad = a.jpad; p.pyz = a.nyz;
    for (int m = 0; m < 10; m++) p.mref[m] = a.mref[m];
    p.rDH = a.rDH; p.rka = a.rka; p.rkb = a.rkb; p.irk = a.irk; p.freenode = a.freenode;
    p.flag[0] = FX; p.flag[1] = FY; p.flag[2] = FZ;
    for (int ax = 0; ax < 3; ax++) { p.pA[ax] = c.pml.A[ax]; p.pB[ax] = c.pml.B[ax]; p.pD[ax] = c.pml.D[ax]; }
    p.mat[0] = c.matVx2Vz; p.mat[1] = c.matVy2Vz;
    for (int f = 0; f < 6; f++) { p.Y[f] = c.pml.Y[f]; p.S[f] = c.pml.S[f]; p.T[f] = c.pml.T[f]; }
    // the boxes (upload_wave_cold): every face whole, or with the fused
    // update the free-surface rows and the edges of the absorbing box
    p.boxes = c.pml_boxes;
    p.nbox = c.pml_nbox;
    const int blocks = c.pml_blocks;
    if (blocks > 0) {
      pr.begin(s, PK_PML_AUX);
      hipLaunchKernelGGL(pml_aux_kernel<MT>, dim3((unsigned)blocks), dim3(256), 0, s, p);
      pr.end(s);
    }
  }
  if (par) {
    (void)hipEventRecord(c.fence[1], si);
    (void)hipStreamWaitEvent(s, c.fence[1], 0);
  }
}

template <class MT>
void launch_wave_dispatch(const WaveArgs<MT>& a, const Ctx& c, int fx, int fy, int fz, hipStream_t s) {
  int key = (fx > 0) * 4 + (fy > 0) * 2 + (fz > 0);
  switch (key) {
    case 0: launch_wave_t<-1, -1, -1, MT>(a, c, s); break;
    case 1: launch_wave_t<-1, -1, 1, MT>(a, c, s); break;
    case 2: launch_wave_t<-1, 1, -1, MT>(a, c, s); break;
    case 3: launch_wave_t<-1, 1, 1, MT>(a, c, s); break;
    case 4: launch_wave_t<1, -1, -1, MT>(a, c, s); break;
    case 5: launch_wave_t<1, -1, 1, MT>(a, c, s); break;
    case 6: launch_wave_t<1, 1, -1, MT>(a, c, s); break;
    default: launch_wave_t<1, 1, 1, MT>(a, c, s); break;
  }
}

// one instantiation per metric storage type, each in its own source file
extern template void launch_wave_dispatch<float>(const WaveArgs<float>&, const Ctx&, int, int, int, hipStream_t);
extern template void launch_wave_dispatch<uint16_t>(const WaveArgs<uint16_t>&, const Ctx&, int, int, int, hipStream_t);
extern template void launch_wave_dispatch<MetBf16Delta>(const WaveArgs<MetBf16Delta>&, const Ctx&, int, int, int,
                                                        hipStream_t);

}  // namespace wave
}  // namespace cgfdm

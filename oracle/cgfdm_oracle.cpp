// cgfdm_oracle.cpp -- FP64 CPU restatement of the CGFDM dynamic-rupture time
// step of wenqiang-geophys/dynfd3dcuda.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
// path is compared against.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it.  The product library never links
// or calls it.
//
// Every routine below restates one reference routine in double precision and
// cites the reference file:line it follows (paths relative to the reference
// repository root).  Layouts follow the reference (srcd/common.h:98-184):
//   volume  : pos = j + k*ny + i*ny*nz, 9 wave components Vx Vy Vz Txx Tyy Tzz
//             Txy Txz Tyz, 13 metric/media fields xix..ztz jac lam miu rho;
//   split   : F.W[c][m][nz][ny] (c = Vx Vy Vz T21 T22 T23 T31 T32 T33,
//             m = 0 minus side, 1 plus side);
//   planes  : T11/T12/T13[7][nz][ny], plane 3 = fault, 0..2 / 4..6 = i0-3..-1 /
//             i0+1..+3;
//   2-D out : [nk][nj], index j + k*nj.
// Compile-time feature macros of the reference are runtime flags here, with
// the reference Makefile defaults (Makefile:36-121): usePML, FreeSurface,
// Rupture, NormalRefine on; useCerjan off; TractionImg and DxV_T1 always on.
//
// Parity pins: see tests/test_oracle_*.py (closed forms, invariants, brute
// force).  A function listed as "parity unpinned" in its comment has no such
// check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

typedef double real;

namespace {

// Planted mistakes for the pin tests (tests/test_oracle_pins*.py).  0 (the
// default) runs the restatement as written; a test sets a number with
// orc_set_mutation, checks that its pin then FAILS, and sets it back to 0.
// Each MUT(n) site below names the mistake it plants.
int g_mutation = 0;
inline bool MUT(int n) { return g_mutation == n; }

// DRP/MacCormack coefficients, srcd/macdrp.h:23-40.
const real c_1 = 0.29164704779069994;
const real c_2 = 0.66674514216519998;
const real c_3 = 1.2501177132452999;
const real c_4 = 0.33341180882999999;
const real c_5 = 0.041686285540599999;
const real a_0p = -1.250039237746600;  // a_0pF == a_0pB
const real a_0m = 1.250039237746600;   // a_0mF == a_0mB
const real a_1 = 1.541764761036000;
const real a_2 = -0.333411808829999;
const real a_3 = 0.0416862855405999;
const real c24_1 = 7.0 / 6.0, c24_2 = 4.0 / 3.0, c24_3 = 1.0 / 6.0;  // macdrp.h:120-122
const real PI_ = 3.141592653589793238463;

// One-sided operators on a strided array, macdrp.h:60-75 (L), 98-104 (L22),
// 124-135 (L24).  'f' points at the evaluation point.
inline real opLF(const real* f, long s) {
  return -c_1 * f[-s] - c_2 * f[0] + c_3 * f[s] - c_4 * f[2 * s] + c_5 * f[3 * s];
}
inline real opLB(const real* f, long s) {
  return -c_5 * f[-3 * s] + c_4 * f[-2 * s] - c_3 * f[-s] + c_2 * f[0] + c_1 * f[s];
}
inline real opL(const real* f, long s, int flag) { return flag == 1 ? opLF(f, s) : opLB(f, s); }
inline real opL22(const real* f, long s, int flag) {
  return flag == 1 ? (f[s] - f[0]) : (f[0] - f[-s]);
}
inline real opL24F(const real* f, long s) { return -c24_1 * f[0] + c24_2 * f[s] - c24_3 * f[2 * s]; }
inline real opL24B(const real* f, long s) { return c24_3 * f[-2 * s] - c24_2 * f[-s] + c24_1 * f[0]; }
inline real opL24(const real* f, long s, int flag) { return flag == 1 ? opL24F(f, s) : opL24B(f, s); }

// 3x3 helpers, srcd/device.cu:12-101.
void inv3(real m[3][3]) {
  real v[3][3];
  v[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  v[0][1] = m[2][1] * m[0][2] - m[0][1] * m[2][2];
  v[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  v[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  v[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
  v[1][2] = m[1][0] * m[0][2] - m[0][0] * m[1][2];
  v[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  v[2][1] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
  v[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  real det = v[0][0] * m[0][0] + v[0][1] * m[1][0] + v[0][2] * m[2][0];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) m[a][b] = v[a][b] / det;
}
void mm3(const real A[3][3], const real B[3][3], real C[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      real s = 0;
      for (int c = 0; c < 3; c++) s += A[a][c] * B[c][b];
      C[a][b] = s;
    }
}
void mv3(const real A[3][3], const real x[3], real y[3]) {
  for (int a = 0; a < 3; a++) y[a] = A[a][0] * x[0] + A[a][1] * x[1] + A[a][2] * x[2];
}
inline real norm3(const real* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline void cross3(const real* a, const real* b, real* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
inline real dot3(const real* a, const real* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// device.cu:125-143
real Fr_func(real r, real R) { return r < R ? std::exp(r * r / (r * r - R * R)) : 0.0; }
real Gt_func(real t, real T) {
  if (t < T) {
    real t1 = t - T, t2 = t - 2.0 * T;
    return std::exp(t1 * t1 / t / t2);
  }
  return 1.0;
}

// RK4 stage weights, rk4.cu:188-198.
void rk_coef(int irk, real DT, real& rka, real& rkb) {
  const real A[4] = {0.5, 0.5, 1.0, 1e10};
  const real B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  rka = A[irk] * DT;
  rkb = B[irk] * DT;
}
// RK_Syn, rk4.cu:20-33: three-register classical RK4.
inline void rk_syn(int irk, real rka, real rkb, real& W, real& mW, real& tW, real hW) {
  if (irk == 0) {
    mW = W;
    tW = rkb * hW + mW;
  } else if (irk == 1 || irk == 2) {
    tW += rkb * hW;
  }
  if (irk == 3) W = rkb * hW + tW;
  else W = rka * hW + mW;
}

}  // namespace

struct Oracle {
  // ---- parameters (srcd/common.h:49-84, defaults set_params.cu:60-103, 264-267)
  int NX = 100, NY = 90, NZ = 80;
  real TMAX = 1.0, DT = 0.01, DH = 100.0;
  int INPORT_GRID_TYPE = 0, INPORT_STRESS_TYPE = 0, Friction_type = 0;
  int TP_n = 1;  // thermal-pressurization grid points normal to the fault (set_params.cu:103, 281-282)
  int PML_N = 20;
  real PML_velocity = 3000.0, PML_bmax = 3.0, PML_fc = 1.0;
  int DAMP_N = 20;
  real mu_s = 0.677, mu_d = 0.525, mu_0 = 0.0, Dc = 0.4, C0 = 0.0;
  real f0 = 0, V0 = 0, Vini = 0, fw = 0, L = 0, smooth_load_T = 0;
  int Fault_grid[4] = {0, 0, 0, 0};
  int Asp_grid[4] = {0, 0, 0, 0};
  real bi_vp1 = 6000, bi_vs1 = 3464, bi_rho1 = 2670;
  real bi_vp2 = 6000, bi_vs2 = 3464, bi_rho2 = 2670;
  real viscosity = 0, RupThres = 0;
  // runtime forms of the Makefile feature macros
  int usePML = 1, useCerjan = 0, FreeSurface = 1, NormalRefine = 1, Rupture = 1;
  int FaultSmooth = 0;  // Makefile:104 (off by default)

  // ---- derived
  int ni, nj, nk, nx, ny, nz, i0;
  long nxyz, nyz;
  int freenode;
  int isx1, isx2, isy1, isy2, isz1, isz2;

  // ---- inputs
  std::vector<real> gx, gy, gz;  // Fault_geometry x,y,z [nk][nj]
  bool have_geom = false;
  std::vector<real> coord3d;     // optional full coordinates [3][nxyz] (interior used)
  std::vector<real> topo;        // optional surface height [ni][nj] (extension)
  std::vector<real> m1d_depth, m1d_rho, m1d_vp, m1d_vs;  // Media1D layers
  std::vector<real> m3d_vp, m3d_vs, m3d_rho;            // Media3D [ni][nk][nj]
  std::vector<real> in_Tx, in_Ty, in_Tz, in_dTx, in_dTy, in_dTz;
  std::vector<real> in_mu_s, in_mu_d, in_Dc, in_C0, in_a, in_b, in_L, in_Vw, in_State, in_TP_hy;

  // ---- state
  std::vector<real> C, M;
  std::vector<real> W, mW, hW, tW;
  std::vector<real> matVx2Vz, matVy2Vz;
  std::vector<real> Ax, Bx, Dx, Ay, By, Dy, Az, Bz, Dz;
  std::vector<real> Pw[6], Pm[6], Ph[6], Pt[6];  // PML aux per face x1 x2 y1 y2 z1 z2
  std::vector<real> damp;

  // fault
  std::vector<real> FW, FmW, FhW, FtW;               // 9*2*nyz
  std::vector<real> FWs;                             // 9*2*nyz, smoothed T2*/T3* (smooth.cu)
  std::vector<real> T11, T12, T13, mT11, mT12, mT13, hT11, hT12, hT13;  // 7*nyz
  std::vector<real> rho_f, lam_f, mu_f;              // 2*nyz
  std::vector<real> D[2][3][3];                      // D(a+1)(b+1)_(m+1): [m][a][b] -> nyz*9
  std::vector<real> matT1toVx[2], matVytoVx[2], matVztoVx[2];  // nyz*9, m=0 '-', 1 '+'
  std::vector<real> matVx2Vz_f[2], matVy2Vz_f[2];              // ny*9
  std::vector<real> matP2Mf[3], matM2Pf[3];                    // ny*9
  std::vector<real> vec_n, vec_s1, vec_s2;                     // nyz*3
  // 2-D [nk][nj]
  std::vector<real> Tn, Ts1, Ts2, tTn, tTs1, tTs2, str_init_x, str_init_y, str_init_z;
  std::vector<real> T0x, T0y, T0z, dT0x, dT0y, dT0z, str_peak, Fmu_d, FDc, FC0;
  std::vector<real> Fa, Fb, FL, FVw, slip, hslip, mslip, tslip, Vs1, Vs2;
  std::vector<real> State, mState, hState, tState, init_t0, friction;
  std::vector<real> TP_T, TP_P, TP_hy;  // [TP_n][nk][nj] temperature and pore pressure rise; [nk][nj] hydraulic diffusivity
  std::vector<real> rsD_y, rsD_z, rup_sensor;
  std::vector<int> init_t0_flag, rup_index_y, rup_index_z, united, faultgrid, flag_rup;
  // Newton statistics (extension: iteration counts of the rate-state solve)
  long newton_calls = 0, newton_iters = 0;
  int newton_max_iter_seen = 0;
  bool setup_done = false;

  inline long P(int i, int j, int k) const { return j + (long)k * ny + (long)i * nyz; }
  real* Mf(int c) { return &M[(long)c * nxyz]; }
  real* Wf(std::vector<real>& v, int c) { return &v[(long)c * nxyz]; }

  // ---------------------------------------------------------------------------
  int set_param(const std::string& k, double v) {
#define SETI(name) if (k == #name) { name = (int)v; return 0; }
#define SETR(name) if (k == #name) { name = v; return 0; }
    SETI(NX) SETI(NY) SETI(NZ) SETR(TMAX) SETR(DT) SETR(DH)
    SETI(INPORT_GRID_TYPE) SETI(INPORT_STRESS_TYPE) SETI(Friction_type) SETI(TP_n)
    SETI(PML_N) SETR(PML_velocity) SETR(PML_bmax) SETR(PML_fc) SETI(DAMP_N)
    SETR(mu_s) SETR(mu_d) SETR(mu_0) SETR(Dc) SETR(C0)
    SETR(smooth_load_T) SETR(bi_vp1) SETR(bi_vs1) SETR(bi_rho1) SETR(bi_vp2) SETR(bi_vs2)
    SETR(bi_rho2) SETR(viscosity) SETR(RupThres)
    SETI(usePML) SETI(useCerjan) SETI(FreeSurface) SETI(NormalRefine) SETI(Rupture) SETI(FaultSmooth)
    // JSON key names that differ from the struct names, set_params.cu:222-247
    if (k == "RS_V0") { V0 = v; return 0; }
    if (k == "RS_Vini") { Vini = v; return 0; }
    if (k == "RS_f0") { f0 = v; return 0; }
    if (k == "RS_fw") { fw = v; return 0; }
    if (k == "RS_L") { L = v; return 0; }
    if (k == "rho1") { bi_rho1 = v; return 0; }  // set_params.cu:247-248 writes bi_rho1
    if (k == "vp1" || k == "vs1") return 0;      // stored but unused on the path
    if (k.rfind("Fault_grid", 0) == 0 && k.size() == 11) { Fault_grid[k[10] - '0'] = (int)v; return 0; }
    if (k.rfind("Asp_grid", 0) == 0 && k.size() == 9) { Asp_grid[k[8] - '0'] = (int)v; return 0; }
#undef SETI
#undef SETR
    return 1;
  }

  // ---------------------------------------------------------------------------
  void derive() {
    ni = NX; nj = NY; nk = NZ;  // single rank, set_params.cu:286-292
    nx = ni + 6; ny = nj + 6; nz = nk + 6;
    nxyz = (long)nx * ny * nz;
    nyz = (long)ny * nz;
    i0 = nx / 2;
    // absorbing faces of a single rank, main.cu:158-171
    freenode = FreeSurface ? 1 : 0;
    isx1 = isx2 = isy1 = isy2 = isz1 = 1;
    isz2 = FreeSurface ? 0 : 1;
  }

  // construct_coord, coord.cu:153-328 (single rank), then extend_crew_array,
  // metric.cu:248-313.
  void build_coord() {
    C.assign(3 * nxyz, 0.0);
    std::vector<real> xline(NX + 6);
    int ic = NX / 2 + 3;
    xline[ic] = 0;
    const int w1 = 10, w2 = 55;
    for (int i = ic + 1; i < NX + 6; i++) {
      int dist = std::abs(i - ic);
      real compr = dist < w1 ? 0.0 : (dist < w2 ? 1.0 - std::cos(PI_ * (i - (ic + w1)) / (real)(w2 - w1)) : 2.0);
      compr = 0.5 + 0.25 * compr;
      xline[i] = xline[i - 1] + DH * compr;
    }
    for (int i = ic - 1; i >= 0; i--) {
      int dist = std::abs(i - ic);
      real compr = dist < w1 ? 0.0 : (dist < w2 ? 1.0 - std::cos(PI_ * (i - (ic - w1)) / (real)(w2 - w1)) : 2.0);
      compr = 0.5 + 0.25 * compr;
      xline[i] = xline[i + 1] - DH * compr;
    }
    real* X = &C[0];
    real* Y = &C[nxyz];
    real* Z = &C[2 * nxyz];
    for (int i = 3; i < nx - 3; i++)
      for (int k = 3; k < nz - 3; k++)
        for (int j = 3; j < ny - 3; j++) {
          int gi = i - 3, gj = j - 3, gk = k - 3;
          real x0 = (-NX / 2) * DH, y0 = (-NY / 2) * DH, z0 = (1 - NZ) * DH;
          real x = gi * DH + x0, y = gj * DH + y0, z = gk * DH + z0;
          if (NormalRefine) x = xline[i];
          if (INPORT_GRID_TYPE == 1) {
            long f = (j - 3) + (long)(k - 3) * nj;
            x = gx[f] + gi * DH + x0;
            if (NormalRefine) x = gx[f] + xline[i];
            y = gy[f];
            z = gz[f];
          }
          if (!topo.empty()) z += topo[(long)gi * nj + gj] * (real)gk / (real)(NZ - 1);
          long p = P(i, j, k);
          X[p] = x; Y[p] = y; Z[p] = z;
          if (!coord3d.empty()) {  // full 3-D coordinate input (extension)
            X[p] = coord3d[p]; Y[p] = coord3d[nxyz + p]; Z[p] = coord3d[2 * nxyz + p];
          }
        }
    // extend_crew_array: linear extrapolation into the 3 ghost layers,
    // x then y then z, per component (metric.cu:263-311).
    int i1 = 3, i2 = nx - 4, j1 = 3, j2 = ny - 4, k1 = 3, k2 = nz - 4;
    for (int l = 0; l < 3; l++) {
      real* A = &C[l * nxyz];
      for (int n = 1; n <= 3; n++)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++) {
            A[P(i1 - n, j, k)] = 2.0 * A[P(i1, j, k)] - A[P(i1 + n, j, k)];
            A[P(i2 + n, j, k)] = 2.0 * A[P(i2, j, k)] - A[P(i2 - n, j, k)];
          }
      for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++)
          for (int n = 1; n <= 3; n++) {
            A[P(i, j1 - n, k)] = 2.0 * A[P(i, j1, k)] - A[P(i, j1 + n, k)];
            A[P(i, j2 + n, k)] = 2.0 * A[P(i, j2, k)] - A[P(i, j2 - n, k)];
          }
      for (int i = 0; i < nx; i++)
        for (int n = 1; n <= 3; n++)
          for (int j = 0; j < ny; j++) {
            A[P(i, j, k1 - n)] = 2.0 * A[P(i, j, k1)] - A[P(i, j, k1 + n)];
            A[P(i, j, k2 + n)] = 2.0 * A[P(i, j, k2)] - A[P(i, j, k2 - n)];
          }
    }
  }

  // cal_metric, metric.cu:15-171: 2nd-order central differences, cross
  // products, bimaterial media split at padded i < NX/2.
  void build_metric() {
    M.assign(13 * nxyz, 0.0);
    const real* X = &C[0];
    const real* Y = &C[nxyz];
    const real* Z = &C[2 * nxyz];
    real rDH = 1.0 / DH;
    real lam1 = bi_rho1 * (bi_vp1 * bi_vp1 - 2 * bi_vs1 * bi_vs1), miu1 = bi_rho1 * bi_vs1 * bi_vs1;
    real lam2 = bi_rho2 * (bi_vp2 * bi_vp2 - 2 * bi_vs2 * bi_vs2), miu2 = bi_rho2 * bi_vs2 * bi_vs2;
    for (int i = 3; i < nx - 3; i++)
      for (int k = 3; k < nz - 3; k++)
        for (int j = 3; j < ny - 3; j++) {
          long p = P(i, j, k);
          long sx = nyz, sy = 1, sz = ny;
          real v1[3] = {0.5 * rDH * (X[p + sx] - X[p - sx]), 0.5 * rDH * (Y[p + sx] - Y[p - sx]), 0.5 * rDH * (Z[p + sx] - Z[p - sx])};
          real v2[3] = {0.5 * rDH * (X[p + sy] - X[p - sy]), 0.5 * rDH * (Y[p + sy] - Y[p - sy]), 0.5 * rDH * (Z[p + sy] - Z[p - sy])};
          real v3[3] = {0.5 * rDH * (X[p + sz] - X[p - sz]), 0.5 * rDH * (Y[p + sz] - Y[p - sz]), 0.5 * rDH * (Z[p + sz] - Z[p - sz])};
          real g[3];
          cross3(v1, v2, g);
          real jac = dot3(g, v3);
          M[9 * nxyz + p] = jac;
          cross3(v2, v3, g);
          for (int a = 0; a < 3; a++) M[(0 + a) * nxyz + p] = g[a] / jac;
          cross3(v3, v1, g);
          for (int a = 0; a < 3; a++) M[(3 + a) * nxyz + p] = g[a] / jac;
          cross3(v1, v2, g);
          for (int a = 0; a < 3; a++) M[(6 + a) * nxyz + p] = g[a] / jac;
          bool side1 = i < NX / 2;  // metric.cu:140-150 (gi includes the +3 ghost offset)
          M[10 * nxyz + p] = side1 ? lam1 : lam2;
          M[11 * nxyz + p] = side1 ? miu1 : miu2;
          M[12 * nxyz + p] = side1 ? bi_rho1 : bi_rho2;
        }
    // init_media1d, media.cu:63-138
    if (!m1d_depth.empty()) {
      int nl = (int)m1d_depth.size();
      for (int i = 3; i < nx - 3; i++)
        for (int k = 3; k < nz - 3; k++)
          for (int j = 3; j < ny - 3; j++) {
            long p = P(i, j, k);
            real depth1 = -Z[p];
            real r = m1d_rho[nl - 1], a = m1d_vp[nl - 1], b = m1d_vs[nl - 1];
            for (int l = nl - 1; l >= 0; l--)
              if (depth1 > m1d_depth[l]) { r = m1d_rho[l]; a = m1d_vp[l]; b = m1d_vs[l]; break; }
            M[10 * nxyz + p] = r * (a * a - 2 * b * b);
            M[11 * nxyz + p] = r * b * b;
            M[12 * nxyz + p] = r;
          }
    }
    // init_media3d, media.cu:140-204 ([ni][nk][nj] without ghosts)
    if (!m3d_vp.empty()) {
      for (int i = 0; i < ni; i++)
        for (int k = 0; k < nk; k++)
          for (int j = 0; j < nj; j++) {
            long q = j + (long)k * nj + (long)i * nj * nk;
            long p = P(i + 3, j + 3, k + 3);
            real r = m3d_rho[q], a = m3d_vp[q], b = m3d_vs[q];
            M[10 * nxyz + p] = r * (a * a - 2 * b * b);
            M[11 * nxyz + p] = r * b * b;
            M[12 * nxyz + p] = r;
          }
    }
    // extend_Symm_array, metric.cu:173-246: mirror copy into ghosts.
    int i1 = 3, i2 = nx - 4, j1 = 3, j2 = ny - 4, k1 = 3, k2 = nz - 4;
    for (int l = 0; l < 13; l++) {
      real* A = &M[l * nxyz];
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int n = 1; n <= 3; n++) {
            A[P(i1 - n, j, k)] = A[P(i1 + n, j, k)];
            A[P(i2 + n, j, k)] = A[P(i2 - n, j, k)];
          }
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++)
          for (int n = 1; n <= 3; n++) {
            A[P(i, j1 - n, k)] = A[P(i, j1 + n, k)];
            A[P(i, j2 + n, k)] = A[P(i, j2 - n, k)];
          }
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          for (int n = 1; n <= 3; n++) {
            A[P(i, j, k1 - n)] = A[P(i, j, k1 + n)];
            A[P(i, j, k2 + n)] = A[P(i, j, k2 - n)];
          }
    }
  }

  // abs_init, init_pml.cu:18-155 (powers d:2, b:2, a:1).
  void build_pml() {
    Ax.assign(ni, 0); Bx.assign(ni, 1); Dx.assign(ni, 0);
    Ay.assign(nj, 0); By.assign(nj, 1); Dy.assign(nj, 0);
    Az.assign(nk, 0); Bz.assign(nk, 1); Dz.assign(nk, 0);
    if (!usePML) return;
    int N = PML_N;
    real Rpp = std::pow(10.0, -((std::log10((real)N) - 1.0) / std::log10(2.0) + 3.0));
    real amax = PI_ * PML_fc;
    real L0 = DH * (N - 1);
    real dmax = -PML_velocity / (2.0 * L0) * std::log(Rpp) * (2.0 + 1.0);
    auto fill = [&](std::vector<real>& A, std::vector<real>& B, std::vector<real>& Dd, int n, int lo, int hi) {
      std::vector<real> Lx(n, -1.0);
      if (lo) for (int i = 0; i < N; i++) Lx[i] = (N - 1 - i) * DH;
      if (hi) for (int i = 0; i < N; i++) Lx[n - N + i] = i * DH;
      for (int i = 0; i < n; i++) {
        real x = Lx[i];
        A[i] = x < 0 ? 0.0 : amax * (1.0 - std::pow(x / L0, 1.0));
        B[i] = x < 0 ? 1.0 : 1.0 + (PML_bmax - 1.0) * std::pow(x / L0, 2.0);
        Dd[i] = x < 0 ? 0.0 : dmax * std::pow(x / L0, 2.0);
        Dd[i] /= B[i];
      }
    };
    fill(Ax, Bx, Dx, ni, isx1, isx2);
    fill(Ay, By, Dy, nj, isy1, isy2);
    fill(Az, Bz, Dz, nk, isz1, isz2);
    long sz[6] = {(long)N * nj * nk, (long)N * nj * nk, (long)N * nk * ni, (long)N * nk * ni, (long)N * ni * nj, (long)N * ni * nj};
    int on[6] = {isx1, isx2, isy1, isy2, isz1, isz2};
    for (int f = 0; f < 6; f++)
      if (on[f]) { Pw[f].assign(9 * sz[f], 0); Pm[f].assign(9 * sz[f], 0); Ph[f].assign(9 * sz[f], 0); Pt[f].assign(9 * sz[f], 0); }
  }

  // init_cerjan_host, cerjan.cu:31-164.
  void build_cerjan() {
    damp.assign(nxyz, 1.0);
    if (!useCerjan) return;
    int ND = DAMP_N;
    real alpha = std::sqrt(-std::log(0.92)) / ND;
    int ni1 = 3, nj1 = 3, nk1 = 3, ni2 = nx - 3, nj2 = ny - 3, nk2 = nz - 3;
    auto val = [&](int d) { return std::exp(-std::pow(alpha * (ND - d), 2)); };
    for (int i = 0; i < ND; i++)
      for (int j = nj1; j < nj2; j++)
        for (int k = nk1; k < nk2; k++) damp[P(ni1 + i, j, k)] = val(i);
    for (int i = 0; i < ND; i++)
      for (int j = nj1; j < nj2; j++)
        for (int k = nk1; k < nk2; k++) damp[P(ni2 - i - 1, j, k)] = val(i);
    std::vector<real> tmp = damp;
    for (int j = 0; j < ND; j++)
      for (int i = ni1; i < ni2; i++)
        for (int k = nk1; k < nk2; k++) { long p = P(i, nj1 + j, k); damp[p] = std::min(val(j), tmp[p]); }
    for (int j = 0; j < ND; j++)
      for (int i = ni1; i < ni2; i++)
        for (int k = nk1; k < nk2; k++) { long p = P(i, nj2 - j - 1, k); damp[p] = std::min(val(j), tmp[p]); }
    tmp = damp;
    for (int k = 0; k < ND; k++)
      for (int i = ni1; i < ni2; i++)
        for (int j = nj1; j < nj2; j++) { long p = P(i, j, k + nk1); damp[p] = std::min(val(k), tmp[p]); }
    if (!freenode)
      for (int k = 0; k < ND; k++)
        for (int i = ni1; i < ni2; i++)
          for (int j = nj1; j < nj2; j++) { long p = P(i, j, nk2 - k - 1); damp[p] = std::min(val(k), tmp[p]); }
  }

  // init_wave_free_cu, fault_coef.cu:1320-1486: DzV = matVx2Vz DxV + matVy2Vz DyV
  // at the top interior layer.
  void build_free() {
    matVx2Vz.assign((long)nx * ny * 9, 0);
    matVy2Vz.assign((long)nx * ny * 9, 0);
    if (!freenode) return;
    int k = nz - 4;
    for (int i = 3; i < nx - 3; i++)
      for (int j = 3; j < ny - 3; j++) {
        real rb1 = 1.0 / Bx[i - 3], rb2 = 1.0 / By[j - 3], rb3 = 1.0 / Bz[k - 3];
        long p = P(i, j, k);
        real e[3][3];
        for (int a = 0; a < 9; a++) e[a / 3][a % 3] = M[a * nxyz + p];
        real lam = M[10 * nxyz + p], mu = M[11 * nxyz + p], l2m = lam + 2 * mu;
        real A[3][3], B[3][3], Cm[3][3];
        // traction on the zeta-surface: rows of A (e3-e3), B (e3-e1), C (e3-e2)
        auto blk = [&](int a, real s, real out[3][3]) {
          const real* ea = e[a];
          const real* e3 = e[2];
          out[0][0] = s * (l2m * e3[0] * ea[0] + mu * (e3[1] * ea[1] + e3[2] * ea[2]));
          out[0][1] = s * (lam * e3[0] * ea[1] + mu * e3[1] * ea[0]);
          out[0][2] = s * (lam * e3[0] * ea[2] + mu * e3[2] * ea[0]);
          out[1][0] = s * (lam * e3[1] * ea[0] + mu * e3[0] * ea[1]);
          out[1][1] = s * (l2m * e3[1] * ea[1] + mu * (e3[0] * ea[0] + e3[2] * ea[2]));
          out[1][2] = s * (lam * e3[1] * ea[2] + mu * e3[2] * ea[1]);
          out[2][0] = s * (lam * e3[2] * ea[0] + mu * e3[0] * ea[2]);
          out[2][1] = s * (lam * e3[2] * ea[1] + mu * e3[1] * ea[2]);
          out[2][2] = s * (l2m * e3[2] * ea[2] + mu * (e3[0] * ea[0] + e3[1] * ea[1]));
        };
        blk(2, rb3, A);
        blk(0, -rb1, B);
        blk(1, -rb2, Cm);
        inv3(A);
        real AB[3][3], AC[3][3];
        mm3(A, B, AB);
        mm3(A, Cm, AC);
        for (int a = 0; a < 9; a++) {
          matVx2Vz[(j + (long)i * ny) * 9 + a] = AB[a / 3][a % 3];
          matVy2Vz[(j + (long)i * ny) * 9 + a] = AC[a / 3][a % 3];
        }
      }
  }

  // ---------------------------------------------------------------------------
  // wave_deriv_cu, wave.cu:8-494: interior RHS with free-surface DzV and the
  // traction-image hV in the top three layers.
  void wave_deriv(int FX, int FY, int FZ) {
    real rDH = 1.0 / DH;
    const real* Vv[9];
    for (int c = 0; c < 9; c++) Vv[c] = Wf(W, c);
    real* XIX = Mf(0); real* XIY = Mf(1); real* XIZ = Mf(2);
    real* ETX = Mf(3); real* ETY = Mf(4); real* ETZ = Mf(5);
    real* ZTX = Mf(6); real* ZTY = Mf(7); real* ZTZ = Mf(8);
    real* JAC = Mf(9); real* LAM = Mf(10); real* MIU = Mf(11); real* RHO = Mf(12);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++)
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++) {
          long p = P(i + 3, j + 3, k + 3);
          real lam = LAM[p], mu = MIU[p], rrho = 1.0 / RHO[p], l2m = lam + 2 * mu;
          real dx[9], dy[9], dz[9];
          for (int c = 0; c < 9; c++) {
            dx[c] = FX * rDH * opL(Vv[c] + p, (long)FX * nyz, 1) ;
            dy[c] = FY * rDH * opL(Vv[c] + p, (long)FY, 1);
            dz[c] = FZ * rDH * opL(Vv[c] + p, (long)FZ * ny, 1);
          }
          // free-surface replacement of DzV, wave.cu:205-246
          if (freenode && k == nk - 1) {
            long ij = (j + 3 + (long)(i + 3) * ny) * 9;
            for (int a = 0; a < 3; a++) {
              const std::vector<real>& MX = MUT(11) ? matVy2Vz : matVx2Vz;  // MUT(11): maps swapped
              const std::vector<real>& MY = MUT(11) ? matVx2Vz : matVy2Vz;
              dz[a] = MX[ij + 3 * a + 0] * dx[0] + MX[ij + 3 * a + 1] * dx[1] + MX[ij + 3 * a + 2] * dx[2]
                    + MY[ij + 3 * a + 0] * dy[0] + MY[ij + 3 * a + 1] * dy[1] + MY[ij + 3 * a + 2] * dy[2];
            }
          } else if (freenode && k == nk - 2) {
            for (int a = 0; a < 3; a++) dz[a] = rDH * opL22(Vv[a] + p, ny, FZ);
          } else if (freenode && k == nk - 3) {
            for (int a = 0; a < 3; a++) dz[a] = rDH * opL24(Vv[a] + p, ny, FZ);
          }
          real h[9];
          real rb1 = 1.0 / Bx[i], rb2 = 1.0 / By[j], rb3 = 1.0 / Bz[k];
          real xi[3] = {XIX[p], XIY[p], XIZ[p]}, et[3] = {ETX[p], ETY[p], ETZ[p]}, zt[3] = {ZTX[p], ZTY[p], ZTZ[p]};
          auto rhs = [&](const real* d, const real* e, real* o) {
            // d: Vx Vy Vz Txx Tyy Tzz Txy Txz Tyz derivatives along one axis
            o[0] = (d[3] * e[0] + d[6] * e[1] + d[7] * e[2]) * rrho;
            o[1] = (d[6] * e[0] + d[4] * e[1] + d[8] * e[2]) * rrho;
            o[2] = (d[7] * e[0] + d[8] * e[1] + d[5] * e[2]) * rrho;
            o[3] = d[0] * e[0] * l2m + d[1] * e[1] * lam + d[2] * e[2] * lam;
            o[4] = d[0] * e[0] * lam + d[1] * e[1] * l2m + d[2] * e[2] * lam;
            o[5] = d[0] * e[0] * lam + d[1] * e[1] * lam + d[2] * e[2] * l2m;
            o[6] = (d[0] * e[1] + d[1] * e[0]) * mu;
            o[7] = (d[0] * e[2] + d[2] * e[0]) * mu;
            o[8] = (d[1] * e[2] + d[2] * e[1]) * mu;
          };
          real ox[9], oy[9], oz[9];
          rhs(dx, xi, ox); rhs(dy, et, oy); rhs(dz, zt, oz);
          for (int c = 0; c < 9; c++) h[c] = ox[c] * rb1 + oy[c] * rb2 + oz[c] * rb3;
          // traction image, wave.cu:398-487
          if (freenode && k >= nk - 3) {
            int ik = nk - 1 - k;
            real rrhojac = rrho / JAC[p];
            real v1[3][7], v2[3][7], v3[3][7];
            for (int l = 0; l < 7; l++) {
              int l3 = l - 3;
              long q = P(i + 3 + l3, j + 3, k + 3);
              const real* T = nullptr; (void)T;
              auto trac = [&](long q, const real* e0, const real* e1, const real* e2, real out[3]) {
                real a = e0[q], b = e1[q], c = e2[q], J = JAC[q];
                out[0] = (a * Vv[3][q] + b * Vv[6][q] + c * Vv[7][q]) * J;
                out[1] = (a * Vv[6][q] + b * Vv[4][q] + c * Vv[8][q]) * J;
                out[2] = (a * Vv[7][q] + b * Vv[8][q] + c * Vv[5][q]) * J;
              };
              real o[3];
              trac(q, XIX, XIY, XIZ, o);
              for (int a = 0; a < 3; a++) v1[a][l] = o[a];
              q = P(i + 3, j + 3 + l3, k + 3);
              trac(q, ETX, ETY, ETZ, o);
              for (int a = 0; a < 3; a++) v2[a][l] = o[a];
              q = P(i + 3, j + 3, k + 3 + l3);
              trac(q, ZTX, ZTY, ZTZ, o);
              for (int a = 0; a < 3; a++) v3[a][l] = o[a];
            }
            for (int a = 0; a < 3; a++) {
              v3[a][ik + 3] = 0.0;
              for (int l = 0; l < 3 - ik; l++) v3[a][6 - l] = (MUT(10) ? 1.0 : -1.0) * v3[a][l + 2 * ik];  // MUT(10): even image
            }
            for (int a = 0; a < 3; a++) {
              real d1 = FX * rDH * opL(&v1[a][3], FX, 1);
              real d2 = FY * rDH * opL(&v2[a][3], FY, 1);
              real d3 = FZ * rDH * opL(&v3[a][3], FZ, 1);
              h[a] = (d1 * rb1 + d2 * rb2 + d3 * rb3) * rrhojac;
            }
          }
          for (int c = 0; c < 9; c++) hW[c * nxyz + p] = h[c];
        }
  }

  // abs_deriv_{x,y,z}_cu, pmlx.cu:8-214, pmly.cu, pmlz.cu.
  void pml_deriv(int axis, int face, int Flag) {
    int N = PML_N;
    int idx0 = face == 0 ? 0 : (axis == 0 ? ni - N : axis == 1 ? nj - N : nk - N);
    int f = axis * 2 + face;
    if (Pw[f].empty()) return;
    int pni = axis == 0 ? N : ni, pnj = axis == 1 ? N : nj, pnk = axis == 2 ? N : nk;
    long ps = (long)pni * pnj * pnk;
    long s = axis == 0 ? nyz : axis == 1 ? 1 : ny;
    real rDH = 1.0 / DH;
    std::vector<real>& Bv = axis == 0 ? Bx : axis == 1 ? By : Bz;
    std::vector<real>& Dv = axis == 0 ? Dx : axis == 1 ? Dy : Dz;
    std::vector<real>& Av = axis == 0 ? Ax : axis == 1 ? Ay : Az;
    for (int i = 0; i < pni; i++)
      for (int k = 0; k < pnk; k++)
        for (int j = 0; j < pnj; j++) {
          int gi = i + (axis == 0 ? idx0 : 0), gj = j + (axis == 1 ? idx0 : 0), gk = k + (axis == 2 ? idx0 : 0);
          long p = P(gi + 3, gj + 3, gk + 3);
          real d[9];
          for (int c = 0; c < 9; c++) d[c] = opL(&W[c * nxyz + p], (long)Flag * s, 1) * Flag * rDH;
          int e0 = axis * 3;
          real xix = M[(e0 + 0) * nxyz + p], xiy = M[(e0 + 1) * nxyz + p], xiz = M[(e0 + 2) * nxyz + p];
          int li = axis == 0 ? gi : axis == 1 ? gj : gk;
          real b1 = Bv[li], d1 = Dv[li], ad1 = -(Av[li] + (MUT(15) ? -d1 : d1));  // MUT(15): sign of d in the decay
          real lam = M[10 * nxyz + p], mu = M[11 * nxyz + p], rrho = 1.0 / M[12 * nxyz + p], l2m = lam + 2 * mu;
          real rb1 = 1.0 / b1;
          real Dr[9];
          Dr[0] = (d[3] * xix + d[6] * xiy + d[7] * xiz) * rrho;
          Dr[1] = (d[6] * xix + d[4] * xiy + d[8] * xiz) * rrho;
          Dr[2] = (d[7] * xix + d[8] * xiy + d[5] * xiz) * rrho;
          Dr[3] = d[0] * xix * l2m + d[1] * xiy * lam + d[2] * xiz * lam;
          Dr[4] = d[0] * xix * lam + d[1] * xiy * l2m + d[2] * xiz * lam;
          Dr[5] = d[0] * xix * lam + d[1] * xiy * lam + d[2] * xiz * l2m;
          Dr[6] = (d[0] * xiy + d[1] * xix) * mu;
          Dr[7] = (d[0] * xiz + d[2] * xix) * mu;
          Dr[8] = (d[1] * xiz + d[2] * xiy) * mu;
          long pa = j + (long)k * pnj + (long)i * pnj * pnk;
          for (int c = 0; c < 9; c++) {
            hW[c * nxyz + p] += -Pw[f][c * ps + pa] * rb1;
            Ph[f][c * ps + pa] = ad1 * Pw[f][c * ps + pa] + d1 * Dr[c];
          }
          // free-surface correction of the aux RHS, pmlx.cu:185-210 (x and y only)
          if (axis < 2 && freenode && gk == nk - 1) {
            const std::vector<real>& mat = axis == 0 ? matVx2Vz : matVy2Vz;
            long ij = ((gj + 3) + (long)(gi + 3) * ny) * 9;
            real Dz1[3];
            for (int a = 0; a < 3; a++) Dz1[a] = mat[ij + 3 * a + 0] * d[0] + mat[ij + 3 * a + 1] * d[1] + mat[ij + 3 * a + 2] * d[2];
            real ztx = M[6 * nxyz + p], zty = M[7 * nxyz + p], ztz = M[8 * nxyz + p];
            real c0 = d1 * b1;
            Ph[f][3 * ps + pa] += c0 * (l2m * ztx * Dz1[0] + lam * zty * Dz1[1] + lam * ztz * Dz1[2]);
            Ph[f][4 * ps + pa] += c0 * (lam * ztx * Dz1[0] + l2m * zty * Dz1[1] + lam * ztz * Dz1[2]);
            Ph[f][5 * ps + pa] += c0 * (lam * ztx * Dz1[0] + lam * zty * Dz1[1] + l2m * ztz * Dz1[2]);
            Ph[f][6 * ps + pa] += c0 * mu * (MUT(16) ? ztx * Dz1[0] + zty * Dz1[1] : zty * Dz1[0] + ztx * Dz1[1]);  // MUT(16)
            Ph[f][7 * ps + pa] += c0 * mu * (ztz * Dz1[0] + ztx * Dz1[2]);
            Ph[f][8 * ps + pa] += c0 * mu * (ztz * Dz1[1] + zty * Dz1[2]);
          }
        }
  }

  // wave_rk_cu, rk4.cu:154-257: padded volume + PML aux.
  void wave_rk(int irk) {
    real rka, rkb;
    rk_coef(irk, DT, rka, rkb);
    long n = 9 * nxyz;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < n; q++) rk_syn(irk, rka, rkb, W[q], mW[q], tW[q], hW[q]);
    for (int f = 0; f < 6; f++)
      for (size_t q = 0; q < Pw[f].size(); q++) rk_syn(irk, rka, rkb, Pw[f][q], Pm[f][q], Pt[f][q], Ph[f][q]);
  }

  // apply_cerjan_cu over the whole interior, cerjan.cu:166-200, 398-410.
  void apply_cerjan() {
    for (int c = 0; c < 9; c++)
      for (int i = 3; i < nx - 3; i++)
        for (int k = 3; k < nz - 3; k++)
          for (int j = 3; j < ny - 3; j++) { long p = P(i, j, k); W[c * nxyz + p] *= damp[p]; }
  }

  // ===========================================================================
  // Fault
  // ===========================================================================
  long F2(int j, int k) const { return j + (long)k * nj; }
  long FS(int j1, int k1) const { return j1 + (long)k1 * ny; }  // with ghosts
  real& Fw(std::vector<real>& v, int c, int m, int j1, int k1) { return v[(long)c * 2 * nyz + (long)m * nyz + FS(j1, k1)]; }

  // init_fault_coef_cu, fault_coef.cu:26-1302 (parts used with TractionImg and
  // DxV_T1).  D_ab = J*(lam e_a (x) e_b + mu e_b (x) e_a + mu (e_a.e_b) I)
  // written element-wise as the reference does (fault_coef.cu:236-426).
  void build_fault_coef() {
    rho_f.assign(2 * nyz, 0); lam_f.assign(2 * nyz, 0); mu_f.assign(2 * nyz, 0);
    for (int m = 0; m < 2; m++) {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) D[m][a][b].assign(nyz * 9, 0);
      matT1toVx[m].assign(nyz * 9, 0); matVytoVx[m].assign(nyz * 9, 0); matVztoVx[m].assign(nyz * 9, 0);
      matVx2Vz_f[m].assign(ny * 9, 0); matVy2Vz_f[m].assign(ny * 9, 0);
    }
    for (int t = 0; t < 3; t++) { matP2Mf[t].assign(ny * 9, 0); matM2Pf[t].assign(ny * 9, 0); }
    vec_n.assign(nyz * 3, 0); vec_s1.assign(nyz * 3, 0); vec_s2.assign(nyz * 3, 0);
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        long p = P(i0, j, k);
        long pf = FS(j, k) * 9;
        real e[3][3];
        for (int a = 0; a < 9; a++) e[a / 3][a % 3] = M[a * nxyz + p];
        real jac = M[9 * nxyz + p];
        real mt[3][3];
        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) mt[a][b] = e[a][b];
        inv3(mt);
        real x_et[3] = {mt[0][1], mt[1][1], mt[2][1]};
        real s1[3] = {x_et[0], x_et[1], x_et[2]};
        real nr = norm3(s1);
        for (int a = 0; a < 3; a++) s1[a] /= nr;
        real n[3] = {e[0][0], e[0][1], e[0][2]};
        nr = norm3(n);
        for (int a = 0; a < 3; a++) n[a] /= nr;
        real s2[3];
        cross3(n, s1, s2);
        for (int a = 0; a < 3; a++) { vec_n[FS(j, k) * 3 + a] = n[a]; vec_s1[FS(j, k) * 3 + a] = s1[a]; vec_s2[FS(j, k) * 3 + a] = s2[a]; }
        rho_f[FS(j, k)] = M[12 * nxyz + p - nyz]; rho_f[nyz + FS(j, k)] = M[12 * nxyz + p + nyz];
        lam_f[FS(j, k)] = M[10 * nxyz + p - nyz]; lam_f[nyz + FS(j, k)] = M[10 * nxyz + p + nyz];
        mu_f[FS(j, k)] = M[11 * nxyz + p - nyz]; mu_f[nyz + FS(j, k)] = M[11 * nxyz + p + nyz];
        real Dl[2][3][3][3][3];
        for (int m = 0; m < 2; m++) {
          real lam = lam_f[m * nyz + FS(j, k)], mu = mu_f[m * nyz + FS(j, k)];
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
              real eab = dot3(e[a], e[b]);
              for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                  real v = lam * e[a][r] * e[b][c] + mu * e[a][c] * e[b][r] + (r == c ? mu * eab : 0.0);
                  Dl[m][a][b][r][c] = v * jac;
                  D[m][a][b][pf + 3 * r + c] = v * jac;
                }
            }
          // DxV_T1 maps, fault_coef.cu:674-712
          real D11[3][3], D12[3][3], D13[3][3], Dm[3][3], o[3][3];
          for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) { D11[r][c] = Dl[m][0][0][r][c]; D12[r][c] = Dl[m][0][1][r][c]; D13[r][c] = Dl[m][0][2][r][c]; }
          for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) Dm[r][c] = D11[r][c];
          inv3(Dm);
          for (int a = 0; a < 9; a++) matT1toVx[m][pf + a] = Dm[a / 3][a % 3];
          mm3(Dm, D12, o);
          for (int a = 0; a < 9; a++) matVytoVx[m][pf + a] = o[a / 3][a % 3];
          mm3(Dm, D13, o);
          for (int a = 0; a < 9; a++) matVztoVx[m][pf + a] = o[a / 3][a % 3];
        }
        // free-surface maps at the top interior row, fault_coef.cu:855-1010
        if (k == nz - 4) {
          real Vx2Vz[2][3][3], Vy2Vz[2][3][3];
          for (int m = 0; m < 2; m++) {
            real A[3][3], B[3][3], Cc[3][3];
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) {
              A[r][c] = Dl[m][2][2][r][c]; B[r][c] = -Dl[m][2][0][r][c]; Cc[r][c] = -Dl[m][2][1][r][c];
            }
            inv3(A);
            mm3(A, B, Vx2Vz[m]);
            mm3(A, Cc, Vy2Vz[m]);
            for (int a = 0; a < 9; a++) { matVx2Vz_f[m][j * 9 + a] = Vx2Vz[m][a / 3][a % 3]; matVy2Vz_f[m][j * 9 + a] = Vy2Vz[m][a / 3][a % 3]; }
          }
          real v1[3] = {e[0][0], e[0][1], e[0][2]};
          real v1n = norm3(v1);
          real v3[3];
          cross3(v1, x_et, v3);
          real h1 = dot3(v3, e[0]) / v1n, h2 = dot3(v3, e[1]) / v1n, h3 = dot3(v3, e[2]) / v1n;
          real miu = M[11 * nxyz + p];
          real Vx_free[2][3][3], Vy_free[2][3][3];
          for (int m = 0; m < 2; m++) {
            real D11f[3][3], D12f[3][3], D13f[3][3];
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) {
              real tv = v3[r] * v1[c];
              D11f[r][c] = Dl[m][0][0][r][c] + miu * h1 * tv;
              D12f[r][c] = Dl[m][0][1][r][c] + miu * h2 * tv;
              D13f[r][c] = Dl[m][0][2][r][c] + miu * h3 * tv;
            }
            real m1[3][3], m2[3][3];
            mm3(D13f, Vx2Vz[m], m1);
            mm3(D13f, Vy2Vz[m], m2);
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) { Vx_free[m][r][c] = D11f[r][c] + m1[r][c]; Vy_free[m][r][c] = D12f[r][c] + m2[r][c]; }
          }
          real Tv[3][3], o[3][3];
          for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) Tv[r][c] = Vx_free[0][r][c];
          inv3(Tv);
          mm3(Tv, Vx_free[1], o); for (int a = 0; a < 9; a++) matP2Mf[0][j * 9 + a] = o[a / 3][a % 3];
          mm3(Tv, Vy_free[1], o); for (int a = 0; a < 9; a++) matP2Mf[1][j * 9 + a] = o[a / 3][a % 3];
          mm3(Tv, Vy_free[0], o); for (int a = 0; a < 9; a++) matP2Mf[2][j * 9 + a] = o[a / 3][a % 3];
          for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) Tv[r][c] = Vx_free[1][r][c];
          inv3(Tv);
          if (MUT(18)) mm3(Vx_free[0], Tv, o);  // MUT(18): product in the wrong order
          else mm3(Tv, Vx_free[0], o);
          for (int a = 0; a < 9; a++) matM2Pf[0][j * 9 + a] = o[a / 3][a % 3];
          mm3(Tv, Vy_free[0], o); for (int a = 0; a < 9; a++) matM2Pf[1][j * 9 + a] = o[a / 3][a % 3];
          mm3(Tv, Vy_free[1], o); for (int a = 0; a < 9; a++) matM2Pf[2][j * 9 + a] = o[a / 3][a % 3];
        }
      }
  }

  // init_fault_cu, init.cu:160-849; nc_read_init_stress + init_stress_to_local,
  // init.cu:859-1022.
  void init_fault() {
    long n2 = (long)nj * nk;
    auto z = [&](std::vector<real>& v) { v.assign(n2, 0.0); };
    z(Tn); z(Ts1); z(Ts2); z(tTn); z(tTs1); z(tTs2); z(str_init_x); z(str_init_y); z(str_init_z);
    z(T0x); z(T0y); z(T0z); z(dT0x); z(dT0y); z(dT0z); z(str_peak); z(Fmu_d); z(FDc); z(FC0);
    z(Fa); z(Fb); z(FL); z(FVw); z(slip); z(hslip); z(mslip); z(tslip); z(Vs1); z(Vs2);
    z(State); z(mState); z(hState); z(tState); z(init_t0); z(friction); z(rsD_y); z(rsD_z); z(rup_sensor);
    init_t0_flag.assign(n2, 0); rup_index_y.assign(n2, 0); rup_index_z.assign(n2, 0);
    united.assign(n2, 0); faultgrid.assign(n2, 0); flag_rup.assign(n2, 0);
    // allocated and zeroed for every friction type (alloc.cu:111-119)
    TP_T.assign((long)std::max(TP_n, 1) * n2, 0.0); TP_P.assign((long)std::max(TP_n, 1) * n2, 0.0); TP_hy.assign(n2, 0.0);
    FW.assign(18 * nyz, 0); FmW.assign(18 * nyz, 0); FhW.assign(18 * nyz, 0); FtW.assign(18 * nyz, 0);
    for (auto* v : {&T11, &T12, &T13, &mT11, &mT12, &mT13, &hT11, &hT12, &hT13}) v->assign(7 * nyz, 0.0);
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        long pv = FS(j + 3, k + 3) * 3;
        int gj1 = j + 1, gk1 = k + 1;
        Tn[q] = -120.0e6;
        Ts1[q] = Tn[q] * mu_0;
        Ts2[q] = 0.0;
        if (gj1 >= Asp_grid[0] && gj1 <= Asp_grid[1] && gk1 >= Asp_grid[2] && gk1 <= Asp_grid[3])
          Ts1[q] = (mu_s + (mu_s - mu_d) * 0.001) * Tn[q];
        bool inF = gj1 >= Fault_grid[0] && gj1 <= Fault_grid[1] && gk1 >= Fault_grid[2] && gk1 <= Fault_grid[3];
        str_peak[q] = inF ? mu_s : 10000.;
        for (int a = 0; a < 3; a++) {
          real v = Tn[q] * vec_n[pv + a] + Ts1[q] * vec_s1[pv + a] + Ts2[q] * vec_s2[pv + a];
          (a == 0 ? str_init_x : a == 1 ? str_init_y : str_init_z)[q] = v;
        }
        rup_index_y[q] = (gj1 <= Fault_grid[0] + 3 || gj1 >= Fault_grid[1] - 3) ? 0 : 1;
        rup_index_z[q] = (gk1 <= Fault_grid[2] + 3 || gk1 >= Fault_grid[3] - 3) ? 0 : 1;
        // the second rule overrides the PML-based one, init.cu:795-816
        united[q] = (gj1 > 30 && gj1 < NY - 30 && gk1 > 30) ? 0 : 1;
        faultgrid[q] = (gj1 >= Fault_grid[0] + 3 && gj1 <= Fault_grid[1] - 3 && gk1 >= Fault_grid[2] + 3 && gk1 <= Fault_grid[3] - 3) ? 1 : 0;
        // rate-state: faultgrid from a-b, evaluated before a and b are read
        // (init.cu:835-847 runs before nc_read_init_stress), i.e. a = b = 0.
        if (Friction_type >= 1) faultgrid[q] = 0;
        init_t0[q] = -9999.9;
      }
    if (INPORT_STRESS_TYPE == 1) {
      T0x = in_Tx; T0y = in_Ty; T0z = in_Tz; dT0x = in_dTx; dT0y = in_dTy; dT0z = in_dTz;
      if (Friction_type == 0) { str_peak = in_mu_s; Fmu_d = in_mu_d; FDc = in_Dc; FC0 = in_C0; }
      else { Fa = in_a; Fb = in_b; FL = in_L; FVw = in_Vw; State = in_State; }
      if (Friction_type == 3) TP_hy = in_TP_hy;  // init.cu:999-1003
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++) {
          long q = F2(j, k), pv = FS(j + 3, k + 3) * 3;
          real T0[3] = {T0x[q], T0y[q], T0z[q]};
          Tn[q] = dot3(T0, &vec_n[pv]);
          Ts1[q] = dot3(T0, &vec_s1[pv]);
          Ts2[q] = dot3(T0, &vec_s2[pv]);
        }
    }
  }

  // cal_rup_sensor, rup_sensor.cu:7-97 (output field only; RupSensor is off).
  void rup_sensor_calc() {
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        int gj1 = j + 1, gk1 = k + 1;
        if (!(gj1 >= Fault_grid[0] + 3 && gj1 <= Fault_grid[1] - 3 && gk1 >= Fault_grid[2] + 3 && gk1 <= Fault_grid[3] - 3)) continue;
        long q = F2(j, k);
        if (united[q]) continue;
        rsD_y[q] = 0.25 * (-Ts1[q + 1] + 2.0 * Ts1[q] - Ts1[q - 1]);
        rsD_z[q] = 0.25 * (-Ts1[q + nj] + 2.0 * Ts1[q] - Ts1[q - nj]);
      }
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        int gj1 = j + 1, gk1 = k + 1;
        if (!(gj1 >= Fault_grid[0] + 3 && gj1 <= Fault_grid[1] - 3 && gk1 >= Fault_grid[2] + 3 && gk1 <= Fault_grid[3] - 3)) continue;
        long q = F2(j, k);
        real D1 = rsD_y[q] - rsD_y[q + 1], D2 = rsD_y[q] - rsD_y[q - 1];
        real mg = 0.5 * (D1 * D1 + D2 * D2);
        D1 = rsD_z[q] - (MUT(17) ? rsD_z : rsD_y)[q + nj];  // rup_sensor.cu:75-76 pairs Dz with Dy (MUT(17): Dz with Dz)
        D2 = rsD_z[q] - (MUT(17) ? rsD_z : rsD_y)[q - nj];
        mg += 0.5 * (D1 * D1 + D2 * D2);
        rup_sensor[q] = mg / (Ts1[q] * Ts1[q] + 1e-16) + 1e-16;
        if (united[q]) rup_sensor[q] = 0;
      }
  }

  // state_rk_cu, rk4.cu:791-938.
  void state_rk(int irk) {
    real rka, rkb;
    rk_coef(irk, DT, rka, rkb);
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        if (united[q]) continue;
        int j1 = j + 3, k1 = k + 3;
        long pv = FS(j1, k1) * 3;
        real dV[3];
        for (int a = 0; a < 3; a++) dV[a] = Fw(FW, a, 1, j1, k1) - Fw(FW, a, 0, j1, k1);
        real vs1 = dot3(dV, &vec_s1[pv]), vs2 = dot3(dV, &vec_s2[pv]);
        hslip[q] = std::sqrt(vs2 * vs2 + vs1 * vs1);
        real V = std::max(hslip[q], 1e-30);
        real st = State[q];
        real RS_a = Fa[q], RS_b = Fb[q], RS_Vw = FVw[q];
        real RS_f0 = f0, RS_V0 = V0, RS_L = FL[q];
        if (Friction_type == 1) {
          hState[q] = RS_b * RS_V0 / RS_L * (std::exp((RS_f0 - st) / RS_b) + (MUT(13) ? 1.0 : -1.0) * V / RS_V0);  // MUT(13)
        } else if (Friction_type == 2 || Friction_type == 3) {
          real flv = RS_f0 - (MUT(14) ? RS_a - RS_b : RS_b - RS_a) * std::log(V / RS_V0);  // MUT(14): a and b swapped
          real fss = fw + (flv - fw) / std::pow(1. + std::pow(V / RS_Vw, 8), 0.125);
          real psiss = RS_a * (std::log(std::sinh(fss / RS_a)) + std::log(2. * (RS_V0 / V)));
          hState[q] = -V / RS_L * (st - psiss);
        } else {
          hState[q] = 0.0;  // never written for slip weakening in the reference
        }
        if (irk == 0) {
          mState[q] = State[q]; tState[q] = mState[q] + rkb * hState[q];
          mslip[q] = slip[q]; tslip[q] = mslip[q] + rkb * hslip[q];
        } else if (irk == 1 || irk == 2) {
          tState[q] += rkb * hState[q];
          tslip[q] += rkb * hslip[q];
        }
        if (irk == 3) { State[q] = tState[q] + rkb * hState[q]; slip[q] = tslip[q] + rkb * hslip[q]; }
        else { State[q] = mState[q] + rka * hState[q]; slip[q] = mslip[q] + rka * hslip[q]; }
      }
  }

  // Common part of trial_sw_cu / trial_rs_cu: T1 on the planes i0-+1..3, the
  // in-plane divergence terms R and lumped masses Mrho
  // (trial_slipweakening.cu:130-312, trial_ratestate.cu:200-380).
  void trial_assemble(int j, int k, int FY, int FZ, real R[2][3], real Mrho[2], real& jac_last, real& rho_last) {
    int j1 = j + 3, k1 = k + 3;
    real rDH = 1.0 / DH, DH2 = DH * DH;
    long q = F2(j, k);
    int km = nk - k;
    long pf = FS(j1, k1);
    for (int m = 0; m < 2; m++) {
      real rho = rho_f[m * nyz + pf];
      real jac = 0;
      for (int l = 1; l <= 3; l++) {
        long p = P(i0 + (2 * m - 1) * l, j1, k1);
        real xix = M[p], xiy = M[nxyz + p], xiz = M[2 * nxyz + p];
        jac = M[9 * nxyz + p];
        real t1 = jac * (xix * W[3 * nxyz + p] + xiy * W[6 * nxyz + p] + xiz * W[7 * nxyz + p]);
        real t2 = jac * (xix * W[6 * nxyz + p] + xiy * W[4 * nxyz + p] + xiz * W[8 * nxyz + p]);
        real t3 = jac * (xix * W[7 * nxyz + p] + xiy * W[8 * nxyz + p] + xiz * W[5 * nxyz + p]);
        long pl = (3 + (2 * m - 1) * l) * nyz + pf;
        T11[pl] = t1; T12[pl] = t2; T13[pl] = t3;
      }
      real DyT[3], DzT[3];
      for (int a = 0; a < 3; a++) {
        const real* t2 = &FW[(3 + a) * 2 * nyz + m * nyz + pf];
        DyT[a] = (rup_index_y[q] % 7 ? opL22(t2, 1, FY) : opL(t2, 1, FY)) * rDH;
        real v[7];
        for (int l = -3; l <= 3; l++) v[l + 3] = FW[(6 + a) * 2 * nyz + m * nyz + FS(j1, k1 + l)];
        if (freenode && km <= 3) {
          v[km + 2] = 0.0;
          for (int l = km + 3; l < 7; l++) v[l] = -v[2 * (km + 2) - l];
        }
        DzT[a] = (rup_index_z[q] % 7 ? opL22(&v[3], 1, FZ) : opL(&v[3], 1, FZ)) * rDH;
      }
      std::vector<real>* TT[3] = {&T11, &T12, &T13};
      for (int a = 0; a < 3; a++) {
        std::vector<real>& T = *TT[a];
        // MUT(2): a_2 reads plane i0-1 instead of i0-2
        real r = m == 0 ? a_1 * T[2 * nyz + pf] + a_2 * T[(MUT(2) ? 2 : 1) * nyz + pf] + a_3 * T[0 * nyz + pf]
                        : a_1 * T[4 * nyz + pf] + a_2 * T[5 * nyz + pf] + a_3 * T[6 * nyz + pf];
        R[m][a] = 0.5 * ((2 * m - 1) * r + (DyT[a] + DzT[a]) * DH) * DH2;
      }
      Mrho[m] = 0.5 * jac * rho * DH * DH2;  // jac of i0-+3 (last l)
      jac_last = jac;
      rho_last = rho;
    }
  }

  // Shared tail of both trials: T1 writeback, hT1 at irk 3, viscosity,
  // tTs/tTn, rupture time (trial_slipweakening.cu:640-709).
  void trial_tail(int j, int k, int it, int irk, const real Tau[3], real Tau_n, bool ifchange, const real Trial[3],
                  real jacvec, real ratecheck) {
    int j1 = j + 3, k1 = k + 3;
    long q = F2(j, k), pf = FS(j1, k1), p1 = 3 * nyz + pf;
    long pv = pf * 3;
    if (ifchange) {
      T11[p1] = (Tau[0] - str_init_x[q]) * jacvec;
      T12[p1] = (Tau[1] - str_init_y[q]) * jacvec;
      T13[p1] = (Tau[2] - str_init_z[q]) * jacvec;
    } else {
      T11[p1] = Trial[0]; T12[p1] = Trial[MUT(19) ? 2 : 1]; T13[p1] = Trial[2];  // MUT(19): component index
    }
    real vis = viscosity * DT;
    real h1 = (T11[p1] - mT11[p1]) / DT, h2 = (T12[p1] - mT12[p1]) / DT, h3 = (T13[p1] - mT13[p1]) / DT;
    if (irk == 3) { hT11[pf] = h1; hT12[pf] = h2; hT13[pf] = h3; }
    T11[p1] += vis * h1; T12[p1] += vis * h2; T13[p1] += vis * h3;
    tTs1[q] = dot3(Tau, &vec_s1[pv]);
    tTs2[q] = dot3(Tau, &vec_s2[pv]);
    tTn[q] = Tau_n;
    if (!init_t0_flag[q] && ratecheck > 1e-3) {
      init_t0[q] = it * DT;
      init_t0_flag[q] = 1;
      flag_rup[q] = 1;
    }
  }

  real stage_time(int it, int irk) {
    return irk == 0 ? it * DT : (irk == 3 ? (it + 1) * DT : (it + 0.5) * DT);
  }

  // trial_sw_cu, trial_slipweakening.cu:23-732 (TractionImg on).
  void trial_sw(int it, int irk, int FX, int FY, int FZ) {
    (void)FX;
    real DH2 = DH * DH;
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        if (united[q]) continue;
        int j1 = j + 3, k1 = k + 3;
        long pf = FS(j1, k1), p1 = 3 * nyz + pf;
        if (irk == 0) { mT11[p1] = T11[p1]; mT12[p1] = T12[p1]; mT13[p1] = T13[p1]; }
        real R[2][3], Mr[2], jl, rl;
        trial_assemble(j, k, FY, FZ, R, Mr, jl, rl);
        real t = stage_time(it, irk);
        if (INPORT_STRESS_TYPE == 1) {
          real Gt = Gt_func(t, smooth_load_T);
          str_init_x[q] = T0x[q] + Gt * dT0x[q];
          str_init_y[q] = T0y[q] + Gt * dT0y[q];
          str_init_z[q] = T0z[q] + Gt * dT0z[q];
        }
        real dV[3];
        for (int a = 0; a < 3; a++) dV[a] = Fw(FmW, a, 1, j1, k1) - Fw(FmW, a, 0, j1, k1);
        real dt1 = DT;
        real Trial[3];
        for (int a = 0; a < 3; a++)
          Trial[a] = -(Mr[0] * Mr[1] * dV[a] / dt1 + Mr[0] * R[1][a] + (MUT(1) ? 1.0 : -1.0) * Mr[1] * R[0][a]) /
                     (DH2 * (Mr[0] * a_0p - a_0m * Mr[1])) * 2.0;  // MUT(1): sign of the minus-side term
        long pv = pf * 3;
        long pm = P(i0, j1, k1);
        real n[3] = {M[pm], M[nxyz + pm], M[2 * nxyz + pm]};
        real n0 = norm3(n);
        real jacvec = M[9 * nxyz + pm] * n0;
        for (int a = 0; a < 3; a++) n[a] /= n0;
        real TL[3] = {Trial[0] / jacvec + str_init_x[q], Trial[1] / jacvec + str_init_y[q], Trial[2] / jacvec + str_init_z[q]};
        real Tn_ = dot3(TL, n);
        real Ts[3] = {TL[0] - n[0] * Tn_, TL[1] - n[1] * Tn_, TL[2] - n[2] * Tn_};
        real Ts0 = norm3(Ts);
        real Tau_n, Tau_s;
        bool ifchange = false;
        if (Tn_ >= 0.0) { Tau_n = 0.0; ifchange = true; } else Tau_n = Tn_;
        real mu1 = str_peak[q], sl = slip[q], Dcq = FDc[q];
        real fr = sl <= Dcq ? mu1 - (mu1 - mu_d) * sl / Dcq : mu_d;  // scalar mu_d, trial_slipweakening.cu:519
        real Tau_c = -fr * Tau_n + FC0[q];
        if (Ts0 >= Tau_c) { Tau_s = Tau_c; flag_rup[q] = 1; ifchange = true; }
        else { Tau_s = Ts0; flag_rup[q] = 0; }
        real Tau[3];
        if (ifchange) {
          if (std::fabs(Ts0) < 1e-1) for (int a = 0; a < 3; a++) Tau[a] = Tau_n * n[a];
          else for (int a = 0; a < 3; a++) Tau[a] = Tau_n * n[a] + (Tau_s / Ts0) * Ts[a];
        } else {
          for (int a = 0; a < 3; a++) Tau[a] = TL[a];
        }
        (void)pv;
        trial_tail(j, k, it, irk, Tau, Tau_n, ifchange, Trial, jacvec, hslip[q]);
      }
  }

  // NRsearch, trial_ratestate.cu:20-57.
  int nr_search(real x, real psi, real a, real Vhat, real coef, real& xnew) {
    const int iter_max = 100;
    const real eps = 1e-12, RS_V0 = 1e-6;
    int iter = 0;
    xnew = x;
    while (iter < iter_max) {
      real f = 2.0 * std::exp(-psi / a) * RS_V0 * std::sinh(x) + coef * x - Vhat;
      real df = 2.0 * std::exp(-psi / a) * RS_V0 * std::cosh(x) + coef;
      real d = -f / (df + 1e-100);
      xnew = x + d;
      real err = std::fabs(d) / (std::fabs(x) + 1e-100);
      x = xnew;
      iter++;
      if (err < eps) break;
    }
    return iter;
  }

  // trial_rs_cu, trial_ratestate.cu:59-893 (SearchAbsolute, TractionImg).
  void trial_rs(int it, int irk, int FX, int FY, int FZ) {
    (void)FX;
    real DH2 = DH * DH;
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        if (united[q]) continue;
        int j1 = j + 3, k1 = k + 3;
        long pf = FS(j1, k1), p1 = 3 * nyz + pf, pv = pf * 3;
        if (irk == 0) { mT11[p1] = T11[p1]; mT12[p1] = T12[p1]; mT13[p1] = T13[p1]; }
        const real* s1 = &vec_s1[pv];
        const real* s2 = &vec_s2[pv];
        long pm = P(i0, j1, k1);
        real n[3] = {M[pm], M[nxyz + pm], M[2 * nxyz + pm]};
        real n0 = norm3(n);
        real jacvec = M[9 * nxyz + pm] * n0;
        for (int a = 0; a < 3; a++) n[a] /= n0;
        real R[2][3], Mr[2], jac, rho;
        trial_assemble(j, k, FY, FZ, R, Mr, jac, rho);
        real R1[2], R2[2], R3[2];
        for (int m = 0; m < 2; m++) { R1[m] = dot3(n, R[m]); R2[m] = dot3(s1, R[m]); R3[m] = dot3(s2, R[m]); }
        // hard-coded nucleation, trial_ratestate.cu:383-422
        real y1 = (real)(j - NY / 2);
        real z1 = (real)((NZ - 1 - k) - (int)(7.5e3 / DH));
        real r = std::sqrt(y1 * y1 + z1 * z1) * DH;
        real t = stage_time(it, irk);
        real Gt = Gt_func(t, smooth_load_T);
        real Fr = Fr_func(r, 3.0e3);
        str_init_x[q] = -120e6;
        str_init_y[q] = -25e6 * Fr * Gt + -75e6;
        str_init_z[q] = 0.0;
        if (INPORT_STRESS_TYPE == 1) {
          str_init_x[q] = T0x[q] + Gt * dT0x[q];
          str_init_y[q] = T0y[q] + Gt * dT0y[q];
          str_init_z[q] = T0z[q] + Gt * dT0z[q];
        }
        real T0[3] = {str_init_x[q], str_init_y[q], str_init_z[q]};
        real T0l[3] = {dot3(n, T0), dot3(s1, T0), dot3(s2, T0)};
        real dV[3];
        for (int a = 0; a < 3; a++) dV[a] = Fw(FmW, a, 1, j1, k1) - Fw(FmW, a, 0, j1, k1);
        real dt1 = DT;
        real Rl[3] = {dot3(n, dV), dot3(s1, dV), dot3(s2, dV)};
        Rl[1] -= Vini;  // trial_ratestate.cu:501
        real Tt[3];
        real den = DH2 * (Mr[0] * a_0p - a_0m * Mr[1]) * jacvec;
        Tt[0] = T0l[0] - 2.0 * (Mr[0] * Mr[1] * Rl[0] / dt1 + Mr[0] * R1[1] - Mr[1] * R1[0]) / den;
        Tt[1] = T0l[1] - 2.0 * (Mr[0] * Mr[1] * Rl[1] / dt1 + Mr[0] * R2[1] - Mr[1] * R2[0]) / den;
        Tt[2] = T0l[2] - 2.0 * (Mr[0] * Mr[1] * Rl[2] / dt1 + Mr[0] * R3[1] - Mr[1] * R3[0]) / den;
        real c = -(a_0p - a_0m) * dt1 * jacvec / (jac * rho * DH);  // plus-side jac(i0+3), rho
        real Vt[3] = {c * Tt[0], c * Tt[1], c * Tt[2]};
        real st = State[q], RS_a = Fa[q];
        real Tau_n = Tt[0] >= -1e-1 ? -1e-1 : Tt[0];
        // thermal pressurization: the pore pressure on the fault (level 0)
        // makes the effective normal stress, trial_ratestate.cu:636-644
        const real Pf = Friction_type == 3 ? TP_P[q] : 0.0;
        Tau_n += Pf;
        real coef = c * RS_a * std::fabs(Tau_n);
        real Vtn = std::sqrt(Vt[1] * Vt[1] + Vt[2] * Vt[2]);
        real Ttn = std::sqrt(Tt[1] * Tt[1] + Tt[2] * Tt[2]);
        real Rn = std::sqrt(Rl[1] * Rl[1] + Rl[2] * Rl[2]);
        const real RS_V0 = 1e-6;
        real x = std::asinh(std::exp(st / RS_a) * Rn / (2.0 * RS_V0));
        real xnew;
        int iters = nr_search(x, st, RS_a, Vtn, coef, xnew);
        newton_calls++; newton_iters += iters; newton_max_iter_seen = std::max(newton_max_iter_seen, iters);
        real V = 2.0 * RS_V0 * std::exp(-st / RS_a) * std::sinh(xnew);
        real T = RS_a * xnew * std::fabs(Tau_n);
        friction[q] = RS_a * xnew;
        Tau_n -= Pf;  // back to the total normal stress, trial_ratestate.cu:785-788
        Ttn = std::max(Ttn, 1e-8);
        real Tau_s1 = (MUT(9) ? Tt[2] : Tt[1]) / Ttn * T, Tau_s2 = Tt[2] / Ttn * T;  // MUT(9): s2 component for s1
        V = std::max(V, 1e-30);
        real Tau[3];
        for (int a = 0; a < 3; a++) Tau[a] = Tau_n * n[a] + Tau_s1 * s1[a] + Tau_s2 * s2[a];
        Tau[0] = str_init_x[q];  // trial_ratestate.cu:797
        real dummy[3] = {0, 0, 0};
        trial_tail(j, k, it, irk, Tau, dot3(Tau, n), true, dummy, jacvec, V);
      }
  }

  // fault2wave_cu, transform.cu:224-403.
  void fault2wave() {
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        if (united[F2(j, k)]) continue;
        int j1 = j + 3, k1 = k + 3;
        long p = P(i0, j1, k1), pf = FS(j1, k1);
        real mt[3][3] = {{M[p], M[3 * nxyz + p], M[6 * nxyz + p]},
                         {M[nxyz + p], M[4 * nxyz + p], M[7 * nxyz + p]},
                         {M[2 * nxyz + p], M[5 * nxyz + p], M[8 * nxyz + p]}};
        real rj = 1.0 / M[9 * nxyz + p];
        real tr[3][3];
        tr[0][0] = T11[3 * nyz + pf]; tr[1][0] = T12[3 * nyz + pf]; tr[2][0] = T13[3 * nyz + pf];
        for (int a = 0; a < 3; a++) {
          tr[a][1] = (FW[(3 + a) * 2 * nyz + pf] + FW[(3 + a) * 2 * nyz + nyz + pf]) * 0.5;
          tr[a][2] = (FW[(6 + a) * 2 * nyz + pf] + FW[(6 + a) * 2 * nyz + nyz + pf]) * 0.5;
        }
        inv3(mt);
        real sg[3][3];
        if (MUT(7)) mm3(mt, tr, sg);  // MUT(7): product in the wrong order
        else mm3(tr, mt, sg);
        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) sg[a][b] *= rj;
        for (int a = 0; a < 3; a++) W[a * nxyz + p] = (FW[a * 2 * nyz + pf] + FW[a * 2 * nyz + nyz + pf]) * 0.5;
        W[3 * nxyz + p] = sg[0][0];
        W[4 * nxyz + p] = sg[1][1];
        W[5 * nxyz + p] = sg[2][2];
        W[6 * nxyz + p] = (sg[0][1] + sg[1][0]) * 0.5;
        W[7 * nxyz + p] = (sg[0][2] + sg[2][0]) * 0.5;
        W[8 * nxyz + p] = (sg[1][2] + sg[2][1]) * 0.5;
      }
  }

  // fault_dvelo_cu, fault_dvelo.cu:17-389 (TractionImg on).
  void fault_dvelo(int FX, int FY, int FZ) {
    real rDH = 1.0 / DH;
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        if (united[q]) continue;
        int j1 = j + 3, k1 = k + 3;
        int km = nk - k;
        long pf = FS(j1, k1);
        for (int i = i0 - 3; i <= i0 + 3; i++) {
          int n = i0 - i;
          if (n == 0) continue;
          real v1[3][7], v2[3][7], v3[3][7];
          for (int l = -3; l <= 3; l++) {
            auto trac = [&](long p, int e0, real out[3]) {
              real a = M[e0 * nxyz + p], b = M[(e0 + 1) * nxyz + p], c = M[(e0 + 2) * nxyz + p], J = M[9 * nxyz + p];
              out[0] = J * (a * W[3 * nxyz + p] + b * W[6 * nxyz + p] + c * W[7 * nxyz + p]);
              out[1] = J * (a * W[6 * nxyz + p] + b * W[4 * nxyz + p] + c * W[8 * nxyz + p]);
              out[2] = J * (a * W[7 * nxyz + p] + b * W[8 * nxyz + p] + c * W[5 * nxyz + p]);
            };
            real o[3];
            trac(P(i + l, j1, k1), 0, o);
            for (int a = 0; a < 3; a++) v1[a][l + 3] = o[a];
            trac(P(i, j1 + l, k1), 3, o);
            for (int a = 0; a < 3; a++) v2[a][l + 3] = o[a];
            trac(P(i, j1, k1 + l), 6, o);
            for (int a = 0; a < 3; a++) v3[a][l + 3] = o[a];
          }
          v1[0][n + 3] = T11[3 * nyz + pf];
          v1[1][n + 3] = T12[3 * nyz + pf];
          v1[2][n + 3] = T13[3 * nyz + pf];
          for (int a = 0; a < 3; a++) {
            real* v = v1[a];
            if (n == 2) v[6] = 2.0 * v[5] - v[4];
            if (n == 1) { v[5] = MUT(3) ? v[4] : 2.0 * v[4] - v[3]; v[6] = 2.0 * v[4] - v[2]; }  // MUT(3): constant extrapolation
            if (n == -1) { v[0] = 2.0 * v[2] - v[4]; v[1] = 2.0 * v[2] - v[3]; }
            if (n == -2) v[0] = 2.0 * v[1] - v[2];
          }
          real DxT[3], DyT[3], DzT[3];
          for (int a = 0; a < 3; a++) DxT[a] = opL(&v1[a][3], 1, FX) * rDH;
          if (freenode && km <= 3)
            for (int a = 0; a < 3; a++) {
              v3[a][km + 2] = 0.0;
              for (int l = km + 3; l < 7; l++) v3[a][l] = -v3[a][2 * (km + 2) - l];
            }
          for (int a = 0; a < 3; a++) {
            DyT[a] = opL(&v2[a][3], 1, FY) * rDH;
            DzT[a] = opL(&v3[a][3], 1, FZ) * rDH;
          }
          long p = P(i, j1, k1);
          real rrj = 1.0 / (M[12 * nxyz + p] * M[9 * nxyz + p]);
          for (int a = 0; a < 3; a++) hW[a * nxyz + p] = (DxT[a] + DyT[a] + DzT[a]) * rrj;
        }
        // split nodes, fault_dvelo.cu:260-385
        for (int m = 0; m < 2; m++) {
          std::vector<real>* TT[3] = {&T11, &T12, &T13};
          real DxT[3], DyT[3], DzT[3];
          for (int a = 0; a < 3; a++) {
            std::vector<real>& T = *TT[a];
            // MUT(4): the minus side takes the plus side's fault coefficient
            DxT[a] = m == 0 ? rDH * ((MUT(4) ? a_0p : a_0m) * T[3 * nyz + pf] - a_1 * T[2 * nyz + pf] - a_2 * T[1 * nyz + pf] - a_3 * T[0 * nyz + pf])
                            : rDH * (a_0p * T[3 * nyz + pf] + a_1 * T[4 * nyz + pf] + a_2 * T[5 * nyz + pf] + a_3 * T[6 * nyz + pf]);
            real v2[7], v3[7];
            for (int l = -3; l <= 3; l++) {
              v2[l + 3] = FW[(3 + a) * 2 * nyz + m * nyz + FS(j1 + l, k1)];
              v3[l + 3] = FW[(6 + a) * 2 * nyz + m * nyz + FS(j1, k1 + l)];
            }
            if (freenode && km <= 3) {
              v3[km + 2] = 0.0;
              for (int l = km + 3; l < 7; l++) v3[l] = -v3[2 * (km + 2) - l];
              DzT[a] = opL(&v3[3], 1, FZ) * rDH;
            } else {
              DzT[a] = (rup_index_z[q] % 7 ? opL22(&v3[3], 1, FZ) : opL(&v3[3], 1, FZ)) * rDH;
            }
            DyT[a] = (rup_index_y[q] % 7 ? opL22(&v2[3], 1, FY) : opL(&v2[3], 1, FY)) * rDH;
          }
          real rrj = 1.0 / (rho_f[m * nyz + pf] * M[9 * nxyz + P(i0, j1, k1)]);
          for (int a = 0; a < 3; a++) FhW[a * 2 * nyz + m * nyz + pf] = (DxT[a] + DyT[a] + DzT[a]) * rrj;
        }
      }
  }

  // fault_dstrs_f_cu / fault_dstrs_b_cu, fault_dstrs_f.cu:17-1254 and
  // fault_dstrs_b.cu (DxV_T1 on).  FX = +1 selects _f, -1 selects _b.
  void fault_dstrs(int FX, int FY, int FZ) {
    real rDH = 1.0 / DH;
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        if (united[q]) continue;
        int j1 = j + 3, k1 = k + 3;
        int km = nk - k;
        long pf = FS(j1, k1);
        long pidx = pf * 9;
        auto wv = [&](int c, int i) -> const real* { return &W[c * nxyz + P(i, j1, k1)]; };
        auto fv = [&](int c, int m) -> const real* { return &FW[c * 2 * nyz + m * nyz + pf]; };
        real DxV[8][3], DyV[8][3], DzV[8][3];
        // one-sided x-derivatives off the fault
        for (int a = 0; a < 3; a++) {
          if (FX == 1) {
            DxV[7][a] = opLF(wv(a, i0 + 3), nyz) * rDH;
            DxV[6][a] = (MUT(5) ? -1.0 : 1.0) * opL24F(wv(a, i0 + 2), nyz) * rDH;  // MUT(5): sign
            DxV[5][a] = (wv(a, i0 + 1)[nyz] - wv(a, i0 + 1)[0]) * rDH;
            DxV[4][a] = (wv(a, i0 + 1)[0] - fv(a, 1)[0]) * rDH;
            // minus side uses the split-node value in place of i0
            real v5[5] = {wv(a, i0 - 4)[0], wv(a, i0 - 3)[0], wv(a, i0 - 2)[0], wv(a, i0 - 1)[0], fv(a, 0)[0]};
            DxV[0][a] = opLF(&v5[1], 1) * rDH;
            real v3[3] = {wv(a, i0 - 2)[0], wv(a, i0 - 1)[0], fv(a, 0)[0]};
            DxV[1][a] = opL24F(&v3[0], 1) * rDH;
            DxV[2][a] = (fv(a, 0)[0] - wv(a, i0 - 1)[0]) * rDH;
          } else {
            DxV[0][a] = opLB(wv(a, i0 - 3), nyz) * rDH;
            DxV[1][a] = opL24B(wv(a, i0 - 2), nyz) * rDH;
            DxV[2][a] = (wv(a, i0 - 1)[0] - wv(a, i0 - 1)[-nyz]) * rDH;
            DxV[3][a] = (fv(a, 0)[0] - wv(a, i0 - 1)[0]) * rDH;
            real v5[5] = {fv(a, 1)[0], wv(a, i0 + 1)[0], wv(a, i0 + 2)[0], wv(a, i0 + 3)[0], wv(a, i0 + 4)[0]};
            DxV[7][a] = opLB(&v5[3], 1) * rDH;
            real v3[3] = {fv(a, 1)[0], wv(a, i0 + 1)[0], wv(a, i0 + 2)[0]};
            DxV[6][a] = opL24B(&v3[2], 1) * rDH;
            DxV[5][a] = (wv(a, i0 + 1)[0] - fv(a, 1)[0]) * rDH;
          }
          // split-node y-derivatives
          for (int m = 0; m < 2; m++)
            DyV[3 + m][a] = (rup_index_y[q] % 7 ? opL22(fv(a, m), 1, FY) : opL(fv(a, m), 1, FY)) * rDH;
        }
        if (freenode && km == 1) {
          const real* VX1 = &matVx2Vz_f[0][j1 * 9]; const real* VY1 = &matVy2Vz_f[0][j1 * 9];
          const real* VX2 = &matVx2Vz_f[1][j1 * 9]; const real* VY2 = &matVy2Vz_f[1][j1 * 9];
          auto mv = [](const real* A, const real* x, real* y) { for (int a = 0; a < 3; a++) y[a] = A[3 * a] * x[0] + A[3 * a + 1] * x[1] + A[3 * a + 2] * x[2]; };
          real o1[3], o2[3], o3[3];
          if (FX == 1) {
            mv(&matP2Mf[0][j1 * 9], DxV[4], o1); mv(&matP2Mf[1][j1 * 9], DyV[4], o2); mv(&matP2Mf[2][j1 * 9], DyV[3], o3);
            for (int a = 0; a < 3; a++) DxV[3][a] = o1[a] + o2[a] + (MUT(12) ? 1.0 : -1.0) * o3[a];  // MUT(12): sign
          } else {
            mv(&matM2Pf[0][j1 * 9], DxV[3], o1); mv(&matM2Pf[1][j1 * 9], DyV[3], o2); mv(&matM2Pf[2][j1 * 9], DyV[4], o3);
            for (int a = 0; a < 3; a++) DxV[4][a] = o1[a] + o2[a] - o3[a];
          }
          mv(VX1, DxV[3], o1); mv(VY1, DyV[3], o2);
          for (int a = 0; a < 3; a++) DzV[3][a] = o1[a] + o2[a];
          mv(VX2, DxV[4], o1); mv(VY2, DyV[4], o2);
          for (int a = 0; a < 3; a++) DzV[4][a] = o1[a] + o2[a];
        } else {
          for (int a = 0; a < 3; a++)
            for (int m = 0; m < 2; m++) {
              const real* f = fv(a, m);
              if (freenode && km == 2) DzV[3 + m][a] = opL22(f, ny, FZ) * rDH;
              else if (freenode && km == 3) DzV[3 + m][a] = opL24(f, ny, FZ) * rDH;
              else DzV[3 + m][a] = (rup_index_z[q] % 7 ? opL22(f, ny, FZ) : opL(f, ny, FZ)) * rDH;
            }
          // DxV_T1: DxV = D11^-1 (dT1/dt - D12 DyV - D13 DzV), fault_dstrs_f.cu:236-268
          real dtT1[3] = {0, 0, 0};
          if (faultgrid[q]) { dtT1[0] = hT11[pf]; dtT1[1] = hT12[pf]; dtT1[2] = hT13[pf]; }
          for (int m = 0; m < 2; m++) {
            const real* A = &matT1toVx[m][pidx];
            const real* B = &matVytoVx[m][pidx];
            const real* Cc = &matVztoVx[m][pidx];
            for (int a = 0; a < 3; a++)
              DxV[3 + m][a] = (A[3 * a] * dtT1[0] + A[3 * a + 1] * dtT1[1] + A[3 * a + 2] * dtT1[2])
                            - (B[3 * a] * DyV[3 + m][0] + B[3 * a + 1] * DyV[3 + m][1] + B[3 * a + 2] * DyV[3 + m][2])
                            - (Cc[3 * a] * DzV[3 + m][0] + Cc[3 * a + 1] * DzV[3 + m][1] + Cc[3 * a + 2] * DzV[3 + m][2]);
          }
        }
        // split-node stress rates hT2*, hT3*, fault_dstrs_f.cu:874-1007
        for (int m = 0; m < 2; m++) {
          for (int row = 0; row < 2; row++) {
            const real* A = &D[m][row + 1][0][pidx];
            const real* B = &D[m][row + 1][MUT(6) ? 2 : 1][pidx];  // MUT(6): DyV takes D_a3
            const real* Cc = &D[m][row + 1][2][pidx];
            for (int a = 0; a < 3; a++) {
              real v = A[3 * a] * DxV[3 + m][0] + A[3 * a + 1] * DxV[3 + m][1] + A[3 * a + 2] * DxV[3 + m][2]
                     + B[3 * a] * DyV[3 + m][0] + B[3 * a + 1] * DyV[3 + m][1] + B[3 * a + 2] * DyV[3 + m][2]
                     + Cc[3 * a] * DzV[3 + m][0] + Cc[3 * a + 1] * DzV[3 + m][1] + Cc[3 * a + 2] * DzV[3 + m][2];
              FhW[(3 + 3 * row + a) * 2 * nyz + m * nyz + pf] = v;
            }
          }
        }
        // off-fault stress rates at i0-3..-1, i0+1..+3, fault_dstrs_f.cu:1009-1250
        for (int nn = 0; nn < 8; nn++) {
          int l = nn < 4 ? nn - 3 : nn - 4;
          if (l == 0) continue;
          int i = i0 + l;
          long p = P(i, j1, k1);
          for (int a = 0; a < 3; a++) DyV[nn][a] = opL(&W[a * nxyz + p], 1, FY) * rDH;
          long ij = (j1 + (long)i * ny) * 9;
          for (int a = 0; a < 3; a++) {
            if (freenode && km == 1)
              DzV[nn][a] = matVx2Vz[ij + 3 * a] * DxV[nn][0] + matVx2Vz[ij + 3 * a + 1] * DxV[nn][1] + matVx2Vz[ij + 3 * a + 2] * DxV[nn][2]
                         + matVy2Vz[ij + 3 * a] * DyV[nn][0] + matVy2Vz[ij + 3 * a + 1] * DyV[nn][1] + matVy2Vz[ij + 3 * a + 2] * DyV[nn][2];
            else if (freenode && km == 2) DzV[nn][a] = opL22(&W[a * nxyz + p], ny, FZ) * rDH;
            else if (freenode && km == 3) DzV[nn][a] = opL24(&W[a * nxyz + p], ny, FZ) * rDH;
            else DzV[nn][a] = opL(&W[a * nxyz + p], ny, FZ) * rDH;
          }
          real lam = M[10 * nxyz + p], mu = M[11 * nxyz + p], l2m = lam + 2 * mu;
          real xi[3] = {M[p], M[nxyz + p], M[2 * nxyz + p]};
          real et[3] = {M[3 * nxyz + p], M[4 * nxyz + p], M[5 * nxyz + p]};
          real zt[3] = {M[6 * nxyz + p], M[7 * nxyz + p], M[8 * nxyz + p]};
          const real* dx = DxV[nn]; const real* dy = DyV[nn]; const real* dz = DzV[nn];
          hW[3 * nxyz + p] = l2m * xi[0] * dx[0] + lam * xi[1] * dx[1] + lam * xi[2] * dx[2]
                           + l2m * et[0] * dy[0] + lam * et[1] * dy[1] + lam * et[2] * dy[2]
                           + l2m * zt[0] * dz[0] + lam * zt[1] * dz[1] + lam * zt[2] * dz[2];
          hW[4 * nxyz + p] = lam * xi[0] * dx[0] + l2m * xi[1] * dx[1] + lam * xi[2] * dx[2]
                           + lam * et[0] * dy[0] + l2m * et[1] * dy[1] + lam * et[2] * dy[2]
                           + lam * zt[0] * dz[0] + l2m * zt[1] * dz[1] + lam * zt[2] * dz[2];
          hW[5 * nxyz + p] = lam * xi[0] * dx[0] + lam * xi[1] * dx[1] + l2m * xi[2] * dx[2]
                           + lam * et[0] * dy[0] + lam * et[1] * dy[1] + l2m * et[2] * dy[2]
                           + lam * zt[0] * dz[0] + lam * zt[1] * dz[1] + l2m * zt[2] * dz[2];
          hW[6 * nxyz + p] = (xi[1] * dx[0] + xi[0] * dx[1] + et[1] * dy[0] + et[0] * dy[1] + zt[1] * dz[0] + zt[0] * dz[1]) * mu;
          hW[7 * nxyz + p] = (xi[2] * dx[0] + xi[0] * dx[2] + et[2] * dy[0] + et[0] * dy[2] + zt[2] * dz[0] + zt[0] * dz[2]) * mu;
          hW[8 * nxyz + p] = (xi[2] * dx[1] + xi[1] * dx[2] + et[2] * dy[1] + et[1] * dy[2] + zt[2] * dz[1] + zt[1] * dz[2]) * mu;
        }
      }
  }

  // fault_rk_cu + fault_rk_stage2_cu, rk4.cu:466-789.
  void fault_rk(int irk) {
    real rka, rkb;
    rk_coef(irk, DT, rka, rkb);
    real vis = viscosity * DT;
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        int j1 = j + 3, k1 = k + 3;
        long pf = FS(j1, k1);
        if (!united[q]) {
          for (int m = 0; m < 2; m++)
            for (int c = 0; c < 9; c++) {
              long x = c * 2 * nyz + m * nyz + pf;
              rk_syn(irk, rka, rkb, FW[x], FmW[x], FtW[x], FhW[x]);
              if (c >= 3) FW[x] += vis * FhW[x];
            }
          long pv = pf * 3;
          real dV[3];
          for (int a = 0; a < 3; a++) dV[a] = FW[a * 2 * nyz + nyz + pf] - FW[a * 2 * nyz + pf];
          real vs1 = dot3(dV, &vec_s1[pv]), vs2 = dot3(dV, &vec_s2[pv]);
          hslip[q] = std::sqrt(vs2 * vs2 + vs1 * vs1);
        }
      }
    // stage2 runs on every node, united or not (rk4.cu:519-567)
    real rkb1 = (MUT(8) ? rka : rkb) / DT;  // MUT(8): stage weight of the wavefield update
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = F2(j, k);
        long pf = FS(j + 3, k + 3), pv = pf * 3;
        real dV[3];
        for (int a = 0; a < 3; a++) dV[a] = FW[a * 2 * nyz + nyz + pf] - FW[a * 2 * nyz + pf];
        real vs1 = dot3(dV, &vec_s1[pv]), vs2 = dot3(dV, &vec_s2[pv]);
        hslip[q] = std::sqrt(vs2 * vs2 + vs1 * vs1);
        Vs1[q] = vs1; Vs2[q] = vs2;
        if (irk == 0) { Ts1[q] = rkb1 * tTs1[q]; Ts2[q] = rkb1 * tTs2[q]; Tn[q] = rkb1 * tTn[q]; }
        else { Ts1[q] += rkb1 * tTs1[q]; Ts2[q] += rkb1 * tTs2[q]; Tn[q] += rkb1 * tTn[q]; }
      }
  }

  // ---------------------------------------------------------------------------
  // One RK stage in the order of main.cu:448-516.
  // The fault's part of a stage before the volume derivative
  // (main.cu:455-468): state_rk, the trial, fault2wave.
  void fault_trial(int it, int irk, int FX, int FY, int FZ) {
    state_rk(irk);
    if (Friction_type == 0) trial_sw(it, irk, FX, FY, FZ);
    else trial_rs(it, irk, FX, FY, FZ);
    fault2wave();
  }

  void stage(int it, int irk, int FX, int FY, int FZ) {
    if (Rupture) fault_trial(it, irk, FX, FY, FZ);
    wave_deriv(FX, FY, FZ);
    if (usePML) {
      if (isx1) pml_deriv(0, 0, FX);
      if (isx2) pml_deriv(0, 1, FX);
      if (isy1) pml_deriv(1, 0, FY);
      if (isy2) pml_deriv(1, 1, FY);
      if (isz1) pml_deriv(2, 0, FZ);
      if (isz2) pml_deriv(2, 1, FZ);
    }
    if (Rupture) {
      fault_dvelo(FX, FY, FZ);
      fault_dstrs(FX, FY, FZ);
    }
    wave_rk(irk);
    if (Rupture) {
      fault_rk(irk);
      fault2wave();
    }
    if (useCerjan) apply_cerjan();
  }

  // One time step, main.cu:420-535 (FD flag table main.cu:86-94).
  // smooth_gauss_volume (smooth.cu:13-367), run once per step after the
  // four stages when FaultSmooth is set (main.cu:517-523).
  void smooth_gauss() {
    const int HW = 3;  // SMOOTH_HALF_WIDTH (smooth.cu:7)
    FWs.resize(18 * nyz, 0.0);
    // smooth_gauss_volume0 (smooth.cu:15-62): mirror the split nodes, the
    // fault-plane T1 and the wavefield planes i0-3..i0+3 above the surface.
    if (freenode) {
      const int k1 = nk - 1 + 3;
      for (int j = 0; j < nj; j++) {
        const int j1 = j + 3;
        for (int l = 0; l < 9; l++)
          for (int m = 0; m < 2; m++)
            for (int n = 1; n <= HW; n++) {
              real* f = &FW[(long)l * 2 * nyz + m * nyz];
              f[(k1 + n) * ny + j1] = 2.0 * f[k1 * ny + j1] - f[(k1 - n) * ny + j1];
            }
        for (int n = 1; n <= HW; n++)
          for (std::vector<real>* T : {&T11, &T12, &T13}) {
            real* f = &(*T)[3 * nyz];
            f[(k1 + n) * ny + j1] = 2.0 * f[k1 * ny + j1] - f[(k1 - n) * ny + j1];
          }
        for (int l = 0; l < 9; l++)
          for (int i1 = i0 - HW; i1 <= i0 + HW; i1++)
            for (int n = 1; n <= HW; n++)
              W[l * nxyz + P(i1, j1, k1 + n)] = 2.0 * W[l * nxyz + P(i1, j1, k1)] - W[l * nxyz + P(i1, j1, k1 - n)];
      }
    }
    // smooth_gauss_volume1 (smooth.cu:64-296): half-volume Gaussian average
    // of the T2/T3 tractions; width 0.7 and side weight 0.4 are the
    // reference's hard-coded values (smooth.cu:72-73).
    const real width2 = 0.7 * 0.7, pmax1 = 0.4;
    real wp1[HW + 1][2 * HW + 1][2 * HW + 1], wp2[HW + 1][2 * HW + 1][2 * HW + 1];
    real sum1 = 0, sum2 = 0;
    for (int ii = 0; ii <= HW; ii++)
      for (int a = 0; a <= 2 * HW; a++)
        for (int b = 0; b <= 2 * HW; b++) {
          const real r2 = (real)(ii * ii + (a - HW) * (a - HW) + (b - HW) * (b - HW)) / width2;
          const real w = std::exp(-r2) * (ii == 0 ? 1.0 : pmax1);
          wp1[ii][a][b] = w; wp2[ii][a][b] = w;
          sum1 += w; sum2 += w;
        }
    for (int ii = 0; ii <= HW; ii++)
      for (int a = 0; a <= 2 * HW; a++)
        for (int b = 0; b <= 2 * HW; b++) { wp1[ii][a][b] /= sum1; wp2[ii][a][b] /= sum2; }
#pragma omp parallel for
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        const int j1 = j + 3, k1 = k + 3;
        for (int m = 0; m < 2; m++) {
          real Ts[6] = {0, 0, 0, 0, 0, 0};
          auto& wp = m == 0 ? wp1 : wp2;
          for (int a = 0; a <= 2 * HW; a++)
            for (int b = 0; b <= 2 * HW; b++) {
              const long q = (long)(k1 + a - HW) * ny + (j1 + b - HW);
              for (int c = 0; c < 6; c++) Ts[c] += FW[(3 + c) * 2 * nyz + m * nyz + q] * wp[0][a][b];
            }
          for (int l = 1; l <= HW; l++)
            for (int a = 0; a <= 2 * HW; a++)
              for (int b = 0; b <= 2 * HW; b++) {
                const long p = P(m == 0 ? i0 - l : i0 + l, j1 + b - HW, k1 + a - HW);
                real E[3][3], S[3][3], T[3][3];
                for (int r = 0; r < 3; r++)
                  for (int cc = 0; cc < 3; cc++) E[r][cc] = M[(3 * cc + r) * nxyz + p];  // columns grad xi, eta, zeta
                const real sxx = W[3 * nxyz + p], syy = W[4 * nxyz + p], szz = W[5 * nxyz + p];
                const real sxy = W[6 * nxyz + p], sxz = W[7 * nxyz + p], syz = W[8 * nxyz + p];
                const real Sm[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}};
                for (int r = 0; r < 3; r++)
                  for (int cc = 0; cc < 3; cc++) { S[r][cc] = Sm[r][cc]; T[r][cc] = 0; }
                for (int r = 0; r < 3; r++)
                  for (int cc = 0; cc < 3; cc++)
                    for (int q = 0; q < 3; q++) T[r][cc] += S[r][q] * E[q][cc];
                const real w = M[9 * nxyz + p] * wp[l][a][b];
                for (int r = 0; r < 3; r++) { Ts[r] += T[r][1] * w; Ts[3 + r] += T[r][2] * w; }
              }
          for (int c = 0; c < 6; c++) FWs[(3 + c) * 2 * nyz + m * nyz + (long)k1 * ny + j1] = Ts[c];
        }
      }
    // smooth_gauss_volume2 (smooth.cu:298-345): blend by min(slip/Dc, 1)
    // inside Fault_grid, the smoothed value outside.
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        const int gj = j + 1, gk = k + 1;
        real beta = 1.0;
        if (gj >= Fault_grid[0] && gj <= Fault_grid[1] && gk >= Fault_grid[2] && gk <= Fault_grid[3])
          beta = std::min(slip[k * nj + j] / Dc, (real)1.0);
        const long pf = (long)(k + 3) * ny + (j + 3);
        for (int l = 3; l < 9; l++)
          for (int m = 0; m < 2; m++) {
            const long x = (long)l * 2 * nyz + m * nyz + pf;
            FW[x] = beta * FWs[x] + (1.0 - beta) * FW[x];
          }
      }
  }

  // thermpress_cu, thermpress.cu:22-233: per fault node, the temperature
  // and pore-pressure rises on a 1-D grid of TP_n points normal to the fault
  // (x = i dh, mirror symmetric at both ends), driven by the frictional
  // heating tau V spread as a Gaussian of half-width h, advanced with
  // explicit Euler substeps over DT.  tau uses the effective normal stress
  // of the last step's output traction and friction coefficient.
  void thermpress() {
    const real Lamb = 0.1e6, rho_c = 2.7e6, h = 20e-3, alpha_th = 1.0e-6;
    const real h2 = h * h, SQRT2PI = 2.506628274631000;
    const long n2 = (long)nj * nk;
    std::vector<real> dT(TP_n), dP(TP_n);
    for (long q = 0; q < n2; q++) {
      if (united[q]) continue;
      const real rate = std::sqrt(Vs1[q] * Vs1[q] + Vs2[q] * Vs2[q]);
      const real alpha_hy = TP_hy[q];
      const real alpha_max = std::max(alpha_th, alpha_hy);
      const real w = std::sqrt(4.0 * alpha_max * 25.0);
      // TP_n = 1: dh is infinite and the diffusion terms vanish (the
      // reference then reads one level past the array, times 1/inf)
      const real dh = TP_n > 1 ? w / (real)(TP_n - 1) : INFINITY;
      const real dh2 = dh * dh;
      const int nsub = (int)(DT / (0.1 * dh2 / alpha_max)) + 1;
      const real dt = DT / (real)nsub;
      for (int s = 0; s < nsub; s++) {
        const real tau = -friction[q] * (Tn[q] + TP_P[q]);
        const real tauV = tau * rate;
        for (int i = 0; i < TP_n; i++) {
          const real x = TP_n > 1 ? i * dh : 0.0;
          const real omega = std::exp(-0.5 * x * x / h2) * (tauV / (rho_c * h * SQRT2PI));
          const int im = i == 0 ? 1 : i - 1, ip = i == TP_n - 1 ? i - 1 : i + 1;
          real lapT = 0.0, lapP = 0.0;
          if (TP_n > 1) {
            lapT = (TP_T[ip * n2 + q] - 2.0 * TP_T[i * n2 + q] + TP_T[im * n2 + q]) / dh2;
            lapP = (TP_P[ip * n2 + q] - 2.0 * TP_P[i * n2 + q] + TP_P[im * n2 + q]) / dh2;
          }
          dT[i] = alpha_th * lapT + omega;
          dP[i] = alpha_hy * lapP + Lamb * dT[i];
        }
        for (int i = 0; i < TP_n; i++) {
          TP_T[i * n2 + q] += dt * dT[i];
          TP_P[i * n2 + q] += dt * dP[i];
        }
      }
    }
  }

  void step(int it) {
    static const int FD[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};
    int FX = FD[it % 8][0], FY = FD[it % 8][1], FZ = FD[it % 8][2];
    if (Rupture) rup_sensor_calc();
    if (Rupture && Friction_type == 3) thermpress();  // main.cu:445-447
    for (int irk = 0; irk < 4; irk++) {
      stage(it, irk, FX, FY, FZ);
      FX = -FX; FY = -FY; FZ = -FZ;
    }
    if (Rupture && FaultSmooth) smooth_gauss();
  }

  int setup() {
    derive();
    if (INPORT_GRID_TYPE == 1 && !have_geom) return 3;
    build_coord();
    build_metric();
    build_pml();
    build_cerjan();
    W.assign(9 * nxyz, 0); mW.assign(9 * nxyz, 0); hW.assign(9 * nxyz, 0); tW.assign(9 * nxyz, 0);
    if (Rupture) {
      build_fault_coef();
      if (INPORT_STRESS_TYPE == 1 && in_Tx.empty()) return 4;
      init_fault();
    }
    build_free();
    setup_done = true;
    return 0;
  }

  std::vector<real>* field(const std::string& n) {
    if (n == "W") return &W; if (n == "hW") return &hW; if (n == "mW") return &mW; if (n == "tW") return &tW;
    if (n == "M") return &M; if (n == "C") return &C; if (n == "damp") return &damp;
    if (n == "matVx2Vz") return &matVx2Vz; if (n == "matVy2Vz") return &matVy2Vz;
    if (n == "Bx") return &Bx; if (n == "By") return &By; if (n == "Bz") return &Bz;
    if (n == "Ax") return &Ax; if (n == "Ay") return &Ay; if (n == "Az") return &Az;
    if (n == "Dx") return &Dx; if (n == "Dy") return &Dy; if (n == "Dz") return &Dz;
    if (n == "FW") return &FW; if (n == "FhW") return &FhW; if (n == "FmW") return &FmW; if (n == "FWs") return &FWs;
    if (n == "T11") return &T11; if (n == "T12") return &T12; if (n == "T13") return &T13;
    if (n == "hT11") return &hT11; if (n == "hT12") return &hT12; if (n == "hT13") return &hT13;
    if (n == "Tn") return &Tn; if (n == "Ts1") return &Ts1; if (n == "Ts2") return &Ts2;
    if (n == "tTn") return &tTn; if (n == "tTs1") return &tTs1; if (n == "tTs2") return &tTs2;
    if (n == "Vs1") return &Vs1; if (n == "Vs2") return &Vs2; if (n == "slip") return &slip;
    if (n == "hslip") return &hslip; if (n == "State") return &State; if (n == "hState") return &hState; if (n == "init_t0") return &init_t0;
    if (n == "friction") return &friction; if (n == "rup_sensor") return &rup_sensor;
    if (n == "TP_T") return &TP_T; if (n == "TP_P") return &TP_P; if (n == "TP_hy") return &TP_hy;
    if (n == "str_init_x") return &str_init_x; if (n == "str_init_y") return &str_init_y; if (n == "str_init_z") return &str_init_z;
    if (n == "str_peak") return &str_peak; if (n == "vec_n") return &vec_n; if (n == "vec_s1") return &vec_s1; if (n == "vec_s2") return &vec_s2;
    if (n == "rho_f") return &rho_f;
    if (n == "matT1toVxm") return &matT1toVx[0]; if (n == "matT1toVxp") return &matT1toVx[1];
    if (n == "matVytoVxm") return &matVytoVx[0]; if (n == "matVytoVxp") return &matVytoVx[1];
    if (n == "matVztoVxm") return &matVztoVx[0]; if (n == "matVztoVxp") return &matVztoVx[1];
    if (n == "matVx2Vz1") return &matVx2Vz_f[0]; if (n == "matVx2Vz2") return &matVx2Vz_f[1];
    if (n == "matVy2Vz1") return &matVy2Vz_f[0]; if (n == "matVy2Vz2") return &matVy2Vz_f[1];
    if (n == "matPlus2Min1f") return &matP2Mf[0]; if (n == "matPlus2Min2f") return &matP2Mf[1]; if (n == "matPlus2Min3f") return &matP2Mf[2];
    if (n == "matMin2Plus1f") return &matM2Pf[0]; if (n == "matMin2Plus2f") return &matM2Pf[1]; if (n == "matMin2Plus3f") return &matM2Pf[2];
    const char* Dn[2][3][3] = {{{"D11_1", "D12_1", "D13_1"}, {"D21_1", "D22_1", "D23_1"}, {"D31_1", "D32_1", "D33_1"}},
                               {{"D11_2", "D12_2", "D13_2"}, {"D21_2", "D22_2", "D23_2"}, {"D31_2", "D32_2", "D33_2"}}};
    for (int m = 0; m < 2; m++) for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) if (n == Dn[m][a][b]) return &D[m][a][b];
    const char* pn[6] = {"PMLx1", "PMLx2", "PMLy1", "PMLy2", "PMLz1", "PMLz2"};
    for (int f = 0; f < 6; f++) if (n == pn[f]) return &Pw[f];
    return nullptr;
  }
};

// =============================================================================
// C ABI for the test harness.
// =============================================================================
extern "C" {

void* orc_create() { return new Oracle(); }
void orc_destroy(void* h) { delete (Oracle*)h; }
int orc_set_param(void* h, const char* key, double v) { return ((Oracle*)h)->set_param(key, v); }

// Fault_geometry x,y,z [nk][nj] (coord.cu:71-151)
int orc_set_geometry(void* h, const double* x, const double* y, const double* z, long n) {
  Oracle* o = (Oracle*)h;
  o->gx.assign(x, x + n); o->gy.assign(y, y + n); o->gz.assign(z, z + n);
  o->have_geom = true;
  return 0;
}
int orc_set_coord3d(void* h, const double* c, long n) {
  ((Oracle*)h)->coord3d.assign(c, c + n);
  return 0;
}
int orc_set_topography(void* h, const double* t, long n) {
  ((Oracle*)h)->topo.assign(t, t + n);
  return 0;
}
int orc_set_media1d(void* h, int n, const double* depth, const double* rho, const double* vp, const double* vs) {
  Oracle* o = (Oracle*)h;
  o->m1d_depth.assign(depth, depth + n); o->m1d_rho.assign(rho, rho + n);
  o->m1d_vp.assign(vp, vp + n); o->m1d_vs.assign(vs, vs + n);
  return 0;
}
int orc_set_media3d(void* h, const double* vp, const double* vs, const double* rho, long n) {
  Oracle* o = (Oracle*)h;
  o->m3d_vp.assign(vp, vp + n); o->m3d_vs.assign(vs, vs + n); o->m3d_rho.assign(rho, rho + n);
  return 0;
}
// Fault_init_stress variables [nk][nj] (init.cu:904-1022)
int orc_set_stress(void* h, const char* name, const double* v, long n) {
  Oracle* o = (Oracle*)h;
  std::string s(name);
  std::vector<double>* t = nullptr;
  if (s == "Tx") t = &o->in_Tx; else if (s == "Ty") t = &o->in_Ty; else if (s == "Tz") t = &o->in_Tz;
  else if (s == "dTx") t = &o->in_dTx; else if (s == "dTy") t = &o->in_dTy; else if (s == "dTz") t = &o->in_dTz;
  else if (s == "mu_s") t = &o->in_mu_s; else if (s == "mu_d") t = &o->in_mu_d; else if (s == "Dc") t = &o->in_Dc;
  else if (s == "C0") t = &o->in_C0; else if (s == "a") t = &o->in_a; else if (s == "b") t = &o->in_b;
  else if (s == "L") t = &o->in_L; else if (s == "Vw") t = &o->in_Vw; else if (s == "State") t = &o->in_State;
  else if (s == "TP_hy") t = &o->in_TP_hy;
  if (!t) return 1;
  t->assign(v, v + n);
  return 0;
}
int orc_setup(void* h) { return ((Oracle*)h)->setup(); }
int orc_step(void* h, int it) { ((Oracle*)h)->step(it); return 0; }
int orc_stage(void* h, int it, int irk, int fx, int fy, int fz) { ((Oracle*)h)->stage(it, irk, fx, fy, fz); return 0; }
int orc_fault_trial(void* h, int it, int irk, int fx, int fy, int fz) {
  ((Oracle*)h)->fault_trial(it, irk, fx, fy, fz);
  return 0;
}
// fault_dvelo + fault_dstrs_{f,b} (main.cu:486-491)
int orc_fault_deriv(void* h, int fx, int fy, int fz) {
  ((Oracle*)h)->fault_dvelo(fx, fy, fz);
  ((Oracle*)h)->fault_dstrs(fx, fy, fz);
  return 0;
}
int orc_rup_sensor(void* h) { ((Oracle*)h)->rup_sensor_calc(); return 0; }
int orc_state_rk(void* h, int irk) { ((Oracle*)h)->state_rk(irk); return 0; }
int orc_set_mutation(int m) { g_mutation = m; return 0; }
int orc_thermpress(void* h) { ((Oracle*)h)->thermpress(); return 0; }
int orc_smooth(void* h) { ((Oracle*)h)->smooth_gauss(); return 0; }
// kernel-level entry points for unit parity tests
int orc_wave_deriv(void* h, int fx, int fy, int fz) { ((Oracle*)h)->wave_deriv(fx, fy, fz); return 0; }
int orc_pml_deriv(void* h, int axis, int face, int flag) { ((Oracle*)h)->pml_deriv(axis, face, flag); return 0; }
int orc_wave_rk(void* h, int irk) { ((Oracle*)h)->wave_rk(irk); return 0; }
int orc_apply_cerjan(void* h) { ((Oracle*)h)->apply_cerjan(); return 0; }
int orc_nr_search(void* h, double x, double psi, double a, double vhat, double coef, double* xnew) {
  return ((Oracle*)h)->nr_search(x, psi, a, vhat, coef, *xnew);
}
long orc_newton_stats(void* h, long* iters, int* maxit) {
  Oracle* o = (Oracle*)h;
  *iters = o->newton_iters; *maxit = o->newton_max_iter_seen;
  return o->newton_calls;
}
long orc_size(void* h, const char* name) {
  Oracle* o = (Oracle*)h;
  if (std::string(name) == "dims") return 7;
  std::vector<double>* v = o->field(name);
  if (v) return (long)v->size();
  std::string s(name);
  if (s == "united" || s == "faultgrid" || s == "rup_index_y" || s == "rup_index_z" || s == "init_t0_flag" || s == "flag_rup")
    return (long)o->united.size();
  return -1;
}
int orc_get(void* h, const char* name, double* out, long n) {
  Oracle* o = (Oracle*)h;
  std::string s(name);
  if (s == "dims") {
    double d[7] = {(double)o->ni, (double)o->nj, (double)o->nk, (double)o->nx, (double)o->ny, (double)o->nz, (double)o->i0};
    std::memcpy(out, d, sizeof(double) * std::min<long>(n, 7));
    return 0;
  }
  std::vector<int>* iv = nullptr;
  if (s == "united") iv = &o->united; else if (s == "faultgrid") iv = &o->faultgrid;
  else if (s == "rup_index_y") iv = &o->rup_index_y; else if (s == "rup_index_z") iv = &o->rup_index_z;
  else if (s == "init_t0_flag") iv = &o->init_t0_flag; else if (s == "flag_rup") iv = &o->flag_rup;
  if (iv) {
    if ((long)iv->size() != n) return 2;
    for (long q = 0; q < n; q++) out[q] = (*iv)[q];
    return 0;
  }
  std::vector<double>* v = o->field(s);
  if (!v) return 1;
  if ((long)v->size() != n) return 2;
  std::memcpy(out, v->data(), sizeof(double) * n);
  return 0;
}
int orc_put(void* h, const char* name, const double* in, long n) {
  Oracle* o = (Oracle*)h;
  std::vector<double>* v = o->field(name);
  if (!v) return 1;
  if ((long)v->size() != n) return 2;
  std::memcpy(v->data(), in, sizeof(double) * n);
  return 0;
}

}  // extern "C"

// Setup in double precision: grid on the host, metric and media on the GPU,
// absorbing-layer profiles, fault coefficients and fault initial state on
// the host, then upload to HBM.
// Replaces construct_coord (coord.cu:153-328), extend_crew_array /
// cal_metric / extend_Symm_array (metric.cu), init_media1d/3d (media.cu),
// abs_init (init_pml.cu), init_cerjan_host (cerjan.cu:31-164),
// init_fault_coef_cu (fault_coef.cu:26-1302), init_fault_cu (init.cu:160-849),
// nc_read_init_stress (init.cu:904-1022) and init_wave_free_cu
// (fault_coef.cu:1320-1486).  Setup runs once and is not on the timed path.
#include "cgfdm_internal.h"
#include "mlp.h"
#include "netcdf_classic.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cgfdm {

namespace {

const double PI_ = 3.141592653589793238463;

__host__ __device__ inline void inv3(double m[3][3]) {
  double v[3][3];
  v[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  v[0][1] = m[2][1] * m[0][2] - m[0][1] * m[2][2];
  v[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  v[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  v[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
  v[1][2] = m[1][0] * m[0][2] - m[0][0] * m[1][2];
  v[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  v[2][1] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
  v[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  double det = v[0][0] * m[0][0] + v[0][1] * m[1][0] + v[0][2] * m[2][0];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) m[a][b] = v[a][b] / det;
}
__host__ __device__ inline void mul3(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) C[a][b] = A[a][0] * B[0][b] + A[a][1] * B[1][b] + A[a][2] * B[2][b];
}
__host__ __device__ inline void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__host__ __device__ inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__host__ __device__ inline double nrm(const double* a) { return sqrt(dot(a, a)); }

// Distance from x0 to the plane through x1 x2 x3 (mathFuncs.cu:73-98).
__host__ __device__ inline double dist_point_plane(const double* x0, const double* x1, const double* x2, const double* x3) {
  double u[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  double v[3] = {x3[0] - x1[0], x3[1] - x1[1], x3[2] - x1[2]};
  double p[3];
  cross(u, v, p);
  return fabs(dot(p, x0) - dot(p, x1)) / nrm(p);
}

// Metric and media of one interior point (cal_metric, metric.cu:74-109;
// bimaterial split metric.cu:123-150; init_media1d/3d, media.cu:63-204) in
// FP64 from the FP64 coordinates: the same function runs on the GPU for
// the whole grid and on the host for the planes its fault and free-surface
// coefficients read.  out: xi(3) eta(3) zeta(3) jac lam mu rho.
struct MetricIn {
  const double* X;
  const double* Y;
  const double* Z;
  long nyz;
  int ny, nz, nx, nj, nk, NX;
  double rDH;
  double lam1, miu1, rho1, lam2, miu2, rho2;
  int nl;
  const float* d1;
  const float* r1;
  const float* p1;
  const float* s1;
  const float* vp3;
  const float* vs3;
  const float* rho3;
};
__host__ __device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__host__ __device__ inline double dot3d(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// index of the interior point whose values a ghost point mirrors
// (extend_Symm_array, metric.cu:173-246, applied along x, then y, then z)
__host__ __device__ inline int mirror_idx(int a, int n) { return a < 3 ? 6 - a : (a > n - 4 ? 2 * (n - 4) - a : a); }
// o: xi(3) eta(3) zeta(3) jac lam mu rho; *entry (if given): the point's
// media palette entry, the layer of a 1-D model or the half of a two-half
// one (0 with 3-D media, which have no palette)
__host__ __device__ inline void metric_point(const MetricIn& m, int i, int j, int k, double* o, int* entry = nullptr) {
  const long q = j + (long)k * m.ny + (long)i * m.nyz;
  const long s[3] = {m.nyz, 1, (long)m.ny};
  double t[3][3];  // t[axis][component] = d(x,y,z)/d(xi,eta,zeta)
  for (int a = 0; a < 3; a++) {
    t[a][0] = 0.5 * m.rDH * (m.X[q + s[a]] - m.X[q - s[a]]);
    t[a][1] = 0.5 * m.rDH * (m.Y[q + s[a]] - m.Y[q - s[a]]);
    t[a][2] = 0.5 * m.rDH * (m.Z[q + s[a]] - m.Z[q - s[a]]);
  }
  double g[3];
  cross3(t[0], t[1], g);
  const double jac = dot3d(g, t[2]);
  o[9] = jac;
  cross3(t[1], t[2], g);
  for (int a = 0; a < 3; a++) o[a] = g[a] / jac;
  cross3(t[2], t[0], g);
  for (int a = 0; a < 3; a++) o[3 + a] = g[a] / jac;
  cross3(t[0], t[1], g);
  for (int a = 0; a < 3; a++) o[6 + a] = g[a] / jac;
  double lam = i < m.NX / 2 ? m.lam1 : m.lam2, miu = i < m.NX / 2 ? m.miu1 : m.miu2, rho = i < m.NX / 2 ? m.rho1 : m.rho2;
  int e = i < m.NX / 2 ? 0 : 1;
  if (m.nl > 0) {
    const double depth1 = -m.Z[q];
    double r = m.r1[m.nl - 1], va = m.p1[m.nl - 1], vb = m.s1[m.nl - 1];
    e = m.nl - 1;
    for (int l = m.nl - 1; l >= 0; l--)
      if (depth1 > m.d1[l]) { r = m.r1[l]; va = m.p1[l]; vb = m.s1[l]; e = l; break; }
    lam = r * (va * va - 2 * vb * vb); miu = r * vb * vb; rho = r;
  }
  if (m.vp3) {
    const long q3 = (j - 3) + (long)(k - 3) * m.nj + (long)(i - 3) * m.nj * m.nk;
    const double r = m.rho3[q3], va = m.vp3[q3], vb = m.vs3[q3];
    lam = r * (va * va - 2 * vb * vb); miu = r * vb * vb; rho = r;
    e = 0;
  }
  o[10] = lam; o[11] = miu; o[12] = rho;
  if (entry) *entry = e;
}

__device__ inline float metric_store(double v, float*) { return (float)v; }
__device__ inline uint16_t metric_store(double v, uint16_t*) {  // bf16, round to nearest even
  const uint32_t u = __float_as_uint((float)v);
  return (uint16_t)((u + 0x7fff + ((u >> 16) & 1)) >> 16);
}

// The 13 metric and media words of every stored point (ghosts mirrored), in
// the device layout, and the largest P-wave speed of the interior as the
// bits of a positive double.  One thread per point, j fastest.  NF = 18 is
// the bfloat16-deviation form: the bfloat16 of each metric word less its
// field's reference value ref.r[c], two to a 32-bit word (slots c and c+1,
// field c in the low half) for the nine coefficients, the media palette
// entry in slot 9, lambda, mu, rho as three FP32 planes in slots 10-15, and
// the Jacobian's deviation in slot 16 (slot 17 zero).
struct MetRef {
  double r[10];
};
// The planes ib .. ib+np-1 are written, plane ib first.
template <class OT, int NF = 13>
__global__ void __launch_bounds__(256) metric_media_kernel(MetricIn m, OT* out, long pyz, int py, int jpad,
                                                           unsigned long long* vpmax_bits, MetRef ref, int ib, int np) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)np * m.nz * m.ny;
  double vp = 0.0;
  if (idx < total) {
    const int j = (int)(idx % m.ny);
    const long r = idx / m.ny;
    const int k = (int)(r % m.nz), i = ib + (int)(r / m.nz), io = i - ib;
    double v[13];
    int entry = 0;
    metric_point(m, mirror_idx(i, m.nx), mirror_idx(j, m.ny), mirror_idx(k, m.nz), v, &entry);
    const long base = (long)k * py + jpad + j;
    if constexpr (NF == 18) {
      auto word = [&](int s) { return reinterpret_cast<uint32_t*>(out + ((long)io * 18 + s) * pyz) + base; };
      for (int c = 0; c < 8; c += 2)  // two deviations per 32-bit word, field c in the low half
        *word(c) = metric_store(v[c] - ref.r[c], out) | ((uint32_t)metric_store(v[c + 1] - ref.r[c + 1], out) << 16);
      *word(8) = metric_store(v[8] - ref.r[8], out) | ((uint32_t)entry << 16);
      *word(16) = metric_store(v[9] - ref.r[9], out);
      float* media = reinterpret_cast<float*>(out + ((long)io * 18 + 10) * pyz);
      for (int c = 10; c < 13; c++) media[(long)(c - 10) * pyz + base] = (float)v[c];
    } else {
      for (int c = 0; c < 13; c++) out[((long)io * 13 + c) * pyz + base] = metric_store(v[c], out);
    }
    if (i >= 3 && i < m.nx - 3 && j >= 3 && j < m.ny - 3 && k >= 3 && k < m.nz - 3)
      vp = sqrt((v[10] + 2 * v[11]) / v[12]);
  }
  for (int o = 32; o > 0; o >>= 1) vp = fmax(vp, __shfl_xor(vp, o));
  if ((threadIdx.x & 63) == 0 && vp > 0.0) atomicMax(vpmax_bits, (unsigned long long)__double_as_longlong(vp));
}

// The media palette of the bfloat16-deviation form: with check 0 every
// stored point writes its lambda, mu, rho to its palette entry (the points
// of one entry hold the same words, so the racing writes agree); with check
// 1 every point compares its entry with its own words, and *bad counts the
// points that differ or name no entry.
__global__ void __launch_bounds__(256) media_palette_kernel(const uint16_t* met, long pyz, int py, int jpad, int nx, int ny,
                                                            int nz, float4* pal, int npal, int check, unsigned* bad) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)nx * nz * ny) return;
  const int j = (int)(idx % ny);
  const long r = idx / ny;
  const int k = (int)(r % nz), i = (int)(r / nz);
  const long base = (long)k * py + jpad + j;
  const uint32_t e = reinterpret_cast<const uint32_t*>(met + ((long)i * 18 + 8) * pyz)[base] >> 16;
  const float* media = reinterpret_cast<const float*>(met + ((long)i * 18 + 10) * pyz);
  const float4 v = make_float4(media[base], media[pyz + base], media[2 * pyz + base], 0.0f);
  if (e >= (uint32_t)npal) {
    atomicAdd(bad, 1u);
    return;
  }
  if (!check) {
    pal[e] = v;
    return;
  }
  const float4 w = pal[e];
  if (__float_as_uint(w.x) != __float_as_uint(v.x) || __float_as_uint(w.y) != __float_as_uint(v.y) ||
      __float_as_uint(w.z) != __float_as_uint(v.z))
    atomicAdd(bad, 1u);
}

// OR of each of the five 32-bit words (coefficient pairs, the fifth with
// the palette entry) of the bfloat16-deviation form over every stored
// point, sign bits left out: a word that is +-0 everywhere (a coefficient
// equal to its reference value on the whole grid, as the Cartesian y of a
// grid bent only in x and z gives for four of the nine) decodes to the
// reference values, as a word that is not loaded (0) does.
// A grid-stride loop over the points, then one atomic per block and word
// (an atomic per wavefront on the same five words would serialize).
__global__ void __launch_bounds__(256) metric_pair_or_kernel(const uint16_t* met, long pyz, int py, int jpad, int nx, int ny,
                                                             int nz, unsigned* any) {
  __shared__ unsigned part[4][5];
  unsigned v[5] = {0u, 0u, 0u, 0u, 0u};
  const long total = (long)nx * nz * ny;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const int j = (int)(idx % ny);
    const long r = idx / ny;
    const int k = (int)(r % nz), i = (int)(r / nz);
    const long base = (long)k * py + jpad + j;
    for (int p = 0; p < 5; p++) v[p] |= reinterpret_cast<const uint32_t*>(met + ((long)i * 18 + 2 * p) * pyz)[base] & 0x7fff7fffu;
  }
  for (int p = 0; p < 5; p++) {
    unsigned x = v[p];
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][p] = x;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const unsigned x = part[0][threadIdx.x] | part[1][threadIdx.x] | part[2][threadIdx.x] | part[3][threadIdx.x];
    if (x) atomicOr(any + threadIdx.x, x);
  }
}

// ---- setup passes on the device, FP64 (one thread per point or node)

// Curvilinear coordinates of the interior points (construct_coord,
// coord.cu:153-328): the x line (grid refinement normal to the fault, built
// on the host), the fault-surface geometry of INPORT_GRID_TYPE 1 and the
// topography stretched over the depth.  Ghost points follow in
// coord_ghost_kernel.
struct CoordIn {
  const double* xline;  // NX + 6 padded x positions (NormalRefine)
  const float* gx;      // [nk][nj] fault geometry (INPORT_GRID_TYPE 1)
  const float* gy;
  const float* gz;
  const float* topo;    // [ni][nj] surface height, or null
  int nx, ny, nz, nj, NX, NY, NZ;
  long nyz;
  double DH;
  int normal_refine, grid_type;
};
__global__ void __launch_bounds__(256) coord_kernel(CoordIn c, double* X, double* Y, double* Z) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const int mj = c.ny - 6, mk = c.nz - 6;
  if (idx >= (long)(c.nx - 6) * mk * mj) return;
  const int j = 3 + (int)(idx % mj);
  const long r = idx / mj;
  const int k = 3 + (int)(r % mk), i = 3 + (int)(r / mk);
  const int gi = i - 3, gj = j - 3, gk = k - 3;
  const double x0 = (-c.NX / 2) * c.DH, y0 = (-c.NY / 2) * c.DH, z0 = (1 - c.NZ) * c.DH;
  double x = gi * c.DH + x0, y = gj * c.DH + y0, z = gk * c.DH + z0;
  if (c.normal_refine) x = c.xline[i];
  if (c.grid_type == 1) {
    const long f = (j - 3) + (long)(k - 3) * c.nj;
    x = (c.normal_refine ? c.xline[i] : gi * c.DH + x0) + c.gx[f];
    y = c.gy[f];
    z = c.gz[f];
  }
  if (c.topo) z += c.topo[(long)gi * c.nj + gj] * (double)gk / (double)(c.NZ - 1);
  const long q = j + (long)k * c.ny + (long)i * c.nyz;
  X[q] = x; Y[q] = y; Z[q] = z;
}

// Linear extrapolation into the three ghost layers along one axis of every
// coordinate array (metric.cu:248-313): x for all (j, k), then y for all
// (i, k), then z for all (i, j), each pass on the previous one's output.
// blockIdx.y is the coordinate component.
__global__ void __launch_bounds__(256) coord_ghost_kernel(double* C, long nxyz, int nx, int ny, int nz, long nyz, int axis) {
  const long line = (long)blockIdx.x * 256 + threadIdx.x;
  double* A = C + (long)blockIdx.y * nxyz;
  const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
  long base, stride;
  if (axis == 0) {  // line over (j, k)
    if (line >= (long)ny * nz) return;
    base = line; stride = nyz;
  } else if (axis == 1) {  // line over (i, k)
    if (line >= (long)nx * nz) return;
    base = (line / nz) * nyz + (line % nz) * ny; stride = 1;
  } else {  // line over (i, j)
    if (line >= (long)nx * ny) return;
    base = (line / ny) * nyz + line % ny; stride = ny;
  }
  for (int m = 1; m <= 3; m++) {
    A[base + (3 - m) * stride] = 2.0 * A[base + 3 * stride] - A[base + (3 + m) * stride];
    A[base + (n - 4 + m) * stride] = 2.0 * A[base + (n - 4) * stride] - A[base + (n - 4 - m) * stride];
  }
}

// Keys that order doubles of either sign like unsigned integers, for
// atomic min / max.
__device__ inline unsigned long long order_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double from_order_key(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double v;
  std::memcpy(&v, &b, sizeof v);
  return v;
}

// Smallest distance from an interior point to the planes through its axis
// neighbours (cal_range_steph, coord.cu:330-391), for the CFL check.
__global__ void __launch_bounds__(256) hmin_kernel(const double* X, const double* Y, const double* Z, int nx, int ny, int nz,
                                                   long nyz, unsigned long long* out) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const int mj = ny - 6, mk = nz - 6;
  double h = 1e38;
  if (idx < (long)(nx - 6) * mk * mj) {
    const int j = 3 + (int)(idx % mj);
    const long r = idx / mj;
    const int k = 3 + (int)(r % mk), i = 3 + (int)(r / mk);
    const long q = j + (long)k * ny + (long)i * nyz;
    const double x0[3] = {X[q], Y[q], Z[q]};
    for (int ii = -1; ii <= 1; ii += 2)
      for (int jj = -1; jj <= 1; jj += 2)
        for (int kk = -1; kk <= 1; kk += 2) {
          const long q1 = q + ii * nyz, q2 = q + jj, q3 = q + (long)kk * ny;
          const double x1[3] = {X[q1], Y[q1], Z[q1]}, x2[3] = {X[q2], Y[q2], Z[q2]}, x3[3] = {X[q3], Y[q3], Z[q3]};
          h = fmin(h, dist_point_plane(x0, x1, x2, x3));
        }
  }
  for (int o = 32; o > 0; o >>= 1) h = fmin(h, __shfl_xor(h, o));
  if ((threadIdx.x & 63) == 0) atomicMin(out, order_key(h));
}

// Gathers na x nb doubles src[base + a * sa + b * sb] of each of the three
// coordinate arrays into dst[comp][a][b] (the output files' coordinate slices).
__global__ void __launch_bounds__(256) gather_kernel(const double* C, long nxyz, long base, long sa, long sb, int na, int nb,
                                                     float* dst) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)na * nb) return;
  const long a = idx / nb, b = idx % nb;
  for (int comp = 0; comp < 3; comp++) dst[comp * (long)na * nb + idx] = (float)C[comp * nxyz + base + a * sa + b * sb];
}

// Per-field range of the metric words over a sample of the stored points
// (every 4th index along each axis and the last, ghosts mirrored): the
// reference value of each field is its midrange (MetricBF16 = 2).
// out[2 q] min key, out[2 q + 1] max key of field q.
__global__ void __launch_bounds__(256) metric_range_kernel(MetricIn m, const int* ai, const int* aj, const int* ak, int ni_,
                                                           int nj_, int nk_, unsigned long long* out) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const bool ok = idx < (long)ni_ * nj_ * nk_;
  double v[13];
  for (int q = 0; q < 13; q++) v[q] = 0.0;
  if (ok) {
    const int j = aj[idx % nj_];
    const long r = idx / nj_;
    metric_point(m, ai[r / nk_], j, ak[r % nk_], v);
  }
  for (int q = 0; q < 10; q++) {
    unsigned long long lo = ok ? order_key(v[q]) : ~0ull, hi = ok ? order_key(v[q]) : 0ull;
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&out[2 * q], lo);
      atomicMax(&out[2 * q + 1], hi);
    }
  }
}

// Split-node coefficients of one node (j, k) of the padded fault plane
// (fault_coef.cu:26-1302): the unit normal and strike/dip vectors, the
// scaled Jacobian, the media of both sides, the traction matrices D_ab of
// both sides and the maps T1 -> Vx, Vy -> Vx, Vz -> Vx, and on the row
// under the free surface the split-node velocity maps Vx, Vy -> Vz and the
// side-to-side maps P2M / M2P (fault_coef.cu:855-1010).
struct FaultCoefOut {
  float* rho_f;   // [2][nyz]
  float* Dm;      // [2 * 9][nyz][9]
  float* T1x;     // [2][nyz][9]
  float* Vyx;
  float* Vzx;
  float* Vx2Vzf;  // [2][ny][9]
  float* Vy2Vzf;
  float* P2M;     // [3][ny][9]
  float* M2P;
  float* vn;      // [nyz][3]
  float* vs1;
  float* vs2;
  float* jacvec;  // [nyz]
  float* jac0;
  double* nd;     // [nyz][9]: n, s1, s2 in FP64 (the initial stress projection)
};
__global__ void __launch_bounds__(256) fault_coef_kernel(MetricIn md, int i0, FaultCoefOut o) {
  const long pf = (long)blockIdx.x * 256 + threadIdx.x;
  const int ny = md.ny, nz = md.nz, nx = md.nx;
  const long nyz = md.nyz;
  if (pf >= nyz) return;
  const int j = (int)(pf % ny), k = (int)(pf / ny);
  double v0[13], vm[13], vp[13];
  metric_point(md, mirror_idx(i0, nx), mirror_idx(j, ny), mirror_idx(k, nz), v0);
  metric_point(md, mirror_idx(i0 - 1, nx), mirror_idx(j, ny), mirror_idx(k, nz), vm);
  metric_point(md, mirror_idx(i0 + 1, nx), mirror_idx(j, ny), mirror_idx(k, nz), vp);
  double e[3][3];
  for (int a = 0; a < 9; a++) e[a / 3][a % 3] = v0[a];
  const double jac = v0[9];
  double mt[3][3];
  for (int a = 0; a < 9; a++) mt[a / 3][a % 3] = e[a / 3][a % 3];
  inv3(mt);
  const double x_et[3] = {mt[0][1], mt[1][1], mt[2][1]};
  double s1[3] = {x_et[0], x_et[1], x_et[2]}, n[3] = {e[0][0], e[0][1], e[0][2]}, s2[3];
  const double ns1 = nrm(s1), nn = nrm(n);
  for (int a = 0; a < 3; a++) { s1[a] /= ns1; n[a] /= nn; }
  cross(n, s1, s2);
  for (int a = 0; a < 3; a++) {
    o.vn[pf * 3 + a] = (float)n[a]; o.vs1[pf * 3 + a] = (float)s1[a]; o.vs2[pf * 3 + a] = (float)s2[a];
    o.nd[pf * 9 + a] = n[a]; o.nd[pf * 9 + 3 + a] = s1[a]; o.nd[pf * 9 + 6 + a] = s2[a];
  }
  o.jacvec[pf] = (float)(jac * nn);
  o.jac0[pf] = (float)jac;
  const double lamf[2] = {vm[10], vp[10]}, muf[2] = {vm[11], vp[11]};
  o.rho_f[pf] = (float)vm[12];
  o.rho_f[nyz + pf] = (float)vp[12];
  const bool surf = k == nz - 4;
  double V2Z[2][3][3], VY2Z[2][3][3], Dfb[2][3][3][3];  // surface row: the D_1b of both sides
  for (int m = 0; m < 2; m++) {
    double D1[3][3][3];  // D_1b, b = 0..2
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        const double eab = dot(e[a], e[b]);
        double Dab[3][3];
        for (int r = 0; r < 3; r++)
          for (int cc = 0; cc < 3; cc++) {
            const double v = jac * (lamf[m] * e[a][r] * e[b][cc] + muf[m] * e[a][cc] * e[b][r] + (r == cc ? muf[m] * eab : 0.0));
            Dab[r][cc] = v;
            o.Dm[(((long)(m * 9 + a * 3 + b)) * nyz + pf) * 9 + r * 3 + cc] = (float)v;
          }
        if (a == 0)
          for (int r = 0; r < 9; r++) D1[b][r / 3][r % 3] = Dab[r / 3][r % 3];
        if (surf && a == 2 && b == 2) {
          double A[3][3];
          for (int r = 0; r < 9; r++) A[r / 3][r % 3] = Dab[r / 3][r % 3];
          inv3(A);
          for (int r = 0; r < 9; r++) V2Z[m][r / 3][r % 3] = A[r / 3][r % 3];  // (D_33)^-1 for now
        }
        if (surf && a == 2 && b == 0)
          for (int r = 0; r < 9; r++) VY2Z[m][r / 3][r % 3] = -Dab[r / 3][r % 3];  // -D_31 for now
        if (surf && a == 2 && b == 1)
          for (int r = 0; r < 9; r++) Dfb[m][0][r / 3][r % 3] = -Dab[r / 3][r % 3];  // -D_32 for now
      }
    double D11[3][3], oo[3][3];
    for (int r = 0; r < 9; r++) D11[r / 3][r % 3] = D1[0][r / 3][r % 3];
    inv3(D11);
    for (int a = 0; a < 9; a++) o.T1x[(m * nyz + pf) * 9 + a] = (float)D11[a / 3][a % 3];
    mul3(D11, D1[1], oo);
    for (int a = 0; a < 9; a++) o.Vyx[(m * nyz + pf) * 9 + a] = (float)oo[a / 3][a % 3];
    mul3(D11, D1[2], oo);
    for (int a = 0; a < 9; a++) o.Vzx[(m * nyz + pf) * 9 + a] = (float)oo[a / 3][a % 3];
    if (surf) {
      // Vx -> Vz = D_33^-1 (-D_31), Vy -> Vz = D_33^-1 (-D_32)
      double Ainv[3][3], B[3][3], Cc[3][3];
      for (int r = 0; r < 9; r++) {
        Ainv[r / 3][r % 3] = V2Z[m][r / 3][r % 3];
        B[r / 3][r % 3] = VY2Z[m][r / 3][r % 3];
        Cc[r / 3][r % 3] = Dfb[m][0][r / 3][r % 3];
      }
      mul3(Ainv, B, V2Z[m]);
      mul3(Ainv, Cc, VY2Z[m]);
      for (int a = 0; a < 9; a++) {
        o.Vx2Vzf[((long)m * ny + j) * 9 + a] = (float)V2Z[m][a / 3][a % 3];
        o.Vy2Vzf[((long)m * ny + j) * 9 + a] = (float)VY2Z[m][a / 3][a % 3];
      }
      for (int b = 0; b < 3; b++)
        for (int r = 0; r < 9; r++) Dfb[m][b][r / 3][r % 3] = D1[b][r / 3][r % 3];
    }
  }
  if (!surf) return;
  double v1[3] = {e[0][0], e[0][1], e[0][2]}, v3[3];
  const double v1n = nrm(v1);
  cross(v1, x_et, v3);
  const double h[3] = {dot(v3, e[0]) / v1n, dot(v3, e[1]) / v1n, dot(v3, e[2]) / v1n};
  const double miu = v0[11];
  double Vxf[2][3][3], Vyf[2][3][3];
  for (int m = 0; m < 2; m++) {
    double Df[3][3][3];
    for (int b = 0; b < 3; b++)
      for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) Df[b][r][cc] = Dfb[m][b][r][cc] + miu * h[b] * v3[r] * v1[cc];
    double m1[3][3], m2[3][3];
    mul3(Df[2], V2Z[m], m1);
    mul3(Df[2], VY2Z[m], m2);
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++) { Vxf[m][r][cc] = Df[0][r][cc] + m1[r][cc]; Vyf[m][r][cc] = Df[1][r][cc] + m2[r][cc]; }
  }
  double Ti[3][3], oo[3][3];
  for (int r = 0; r < 9; r++) Ti[r / 3][r % 3] = Vxf[0][r / 3][r % 3];
  inv3(Ti);
  mul3(Ti, Vxf[1], oo); for (int a = 0; a < 9; a++) o.P2M[(0L * ny + j) * 9 + a] = (float)oo[a / 3][a % 3];
  mul3(Ti, Vyf[1], oo); for (int a = 0; a < 9; a++) o.P2M[(1L * ny + j) * 9 + a] = (float)oo[a / 3][a % 3];
  mul3(Ti, Vyf[0], oo); for (int a = 0; a < 9; a++) o.P2M[(2L * ny + j) * 9 + a] = (float)oo[a / 3][a % 3];
  for (int r = 0; r < 9; r++) Ti[r / 3][r % 3] = Vxf[1][r / 3][r % 3];
  inv3(Ti);
  mul3(Ti, Vxf[0], oo); for (int a = 0; a < 9; a++) o.M2P[(0L * ny + j) * 9 + a] = (float)oo[a / 3][a % 3];
  mul3(Ti, Vyf[0], oo); for (int a = 0; a < 9; a++) o.M2P[(1L * ny + j) * 9 + a] = (float)oo[a / 3][a % 3];
  mul3(Ti, Vyf[1], oo); for (int a = 0; a < 9; a++) o.M2P[(2L * ny + j) * 9 + a] = (float)oo[a / 3][a % 3];
}

// Free-surface velocity maps of the top row (init_wave_free_cu,
// fault_coef.cu:1320-1486): with K_b the traction on the zeta surface from
// the derivatives along axis b, scaled by the PML 1/beta of that axis,
// DzV = K_3^-1 (-K_1) DxV + K_3^-1 (-K_2) DyV.  One thread per (i, j).
__global__ void __launch_bounds__(256) wave_free_kernel(MetricIn md, const double* B0, const double* B1, const double* B2,
                                                        float* mVx, float* mVy) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const int nx = md.nx, ny = md.ny, nz = md.nz;
  if (idx >= (long)(nx - 6) * (ny - 6)) return;
  const int j = 3 + (int)(idx % (ny - 6)), i = 3 + (int)(idx / (ny - 6)), k = nz - 4;
  const double rb[3] = {1.0 / B0[i - 3], 1.0 / B1[j - 3], 1.0 / B2[k - 3]};
  double v[13];
  metric_point(md, mirror_idx(i, nx), mirror_idx(j, ny), mirror_idx(k, nz), v);
  double e[3][3];
  for (int a = 0; a < 9; a++) e[a / 3][a % 3] = v[a];
  const double lam = v[10], mu = v[11];
  double Ainv[3][3], Bm[3][3], Cm[3][3];
  for (int b = 0; b < 3; b++) {
    const double eb3 = dot(e[2], e[b]);
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++) {
        const double K = rb[b] * (lam * e[2][r] * e[b][cc] + mu * e[2][cc] * e[b][r] + (r == cc ? mu * eb3 : 0.0));
        if (b == 2) Ainv[r][cc] = K;
        else if (b == 0) Bm[r][cc] = -K;
        else Cm[r][cc] = -K;
      }
  }
  inv3(Ainv);
  double AB[3][3], AC[3][3];
  mul3(Ainv, Bm, AB);
  mul3(Ainv, Cm, AC);
  for (int a = 0; a < 9; a++) {
    mVx[(j + (long)i * ny) * 9 + a] = (float)AB[a / 3][a % 3];
    mVy[(j + (long)i * ny) * 9 + a] = (float)AC[a / 3][a % 3];
  }
}

template <class T>
T* dmalloc(Ctx& c, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
  if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) {
    (void)hipFree(p);
    return nullptr;
  }
  c.allocations.push_back(p);
  return (T*)p;
}
template <class T>
bool upload(T* dst, const std::vector<T>& src) {
  return hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

bool read_media1d(const std::string& path, Ctx& c, std::string& err) {
  std::ifstream in(path);
  if (!in) { err = "can not open input media dat " + path; return false; }
  std::string line;
  std::getline(in, line);  // comment
  int n = 0;
  in >> n;
  std::getline(in, line);
  std::getline(in, line);  // header
  for (int l = 0; l < n; l++) {
    float d, r, a, b;
    if (!(in >> d >> r >> a >> b)) { err = "bad Media1D file " + path; return false; }
    c.m1d_depth.push_back(d); c.m1d_rho.push_back(r); c.m1d_vp.push_back(a); c.m1d_vs.push_back(b);
  }
  return true;
}

}  // namespace

int host_setup(Ctx& c, std::string& err) {
  Params& p = c.p;
  p.derive();
  if (p.PX * p.PY * p.PZ != 1) { err = "only one rank (PX=PY=PZ=1) is supported: the whole domain lives on one GPU"; return 1; }
  Dims& d = c.d;
  d.ni = p.ni; d.nj = p.nj; d.nk = p.nk; d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
  d.nxyz = (long)d.nx * d.ny * d.nz;
  d.nyz = (long)d.ny * d.nz;
  d.i0 = d.nx / 2;
  d.jpad = (32 - (interior_lo_j(p) + 3) % 32) % 32;
  d.py = (d.jpad + d.ny + 31) / 32 * 32;
  // a power-of-two row pitch (4 KB rows for 1000 columns) puts the rows a
  // block reads at once, and the same words of the wavefield buffers, on
  // fewer memory channels: one more 128-byte line per row (1600 x 1000 x 500:
  // 127.6 against 130.5 ms per step, interleaved runs after a warm-up run)
  if ((d.py & (d.py - 1)) == 0) d.py += 32;
  d.pyz = (long)d.py * d.nz;
  if (p.NX < 7 || p.NY < 7 || p.NZ < 7) { err = "NX, NY and NZ must be at least 7"; return 1; }
  if (p.MetricBF16 < 0 || p.MetricBF16 > 2) { err = "MetricBF16 must be 0, 1 or 2"; return 1; }
  const int nx = d.nx, ny = d.ny, nz = d.nz, ni = d.ni, nj = d.nj, nk = d.nk;
  const long nxyz = d.nxyz, nyz = d.nyz;
  const int freenode = p.FreeSurface ? 1 : 0;
  int on[6] = {1, 1, 1, 1, 1, freenode ? 0 : 1};  // single rank: every face absorbs except a free top

  // ---- inputs from files when not given as arrays
  if (p.INPORT_GRID_TYPE == 1 && c.gx.empty()) {
    NcFile nf;
    if (!nf.open_read(p.Fault_geometry, err)) return 4;
    if (!nf.read_float("x", c.gx, err) || !nf.read_float("y", c.gy, err) || !nf.read_float("z", c.gz, err)) return 4;
    if ((long)c.gx.size() < (long)nj * nk) { err = "Fault_geometry smaller than NY*NZ"; return 4; }
    // hyperslab start {0,0}, count {nk,nj} of arrays shaped (NZ,NY)
    c.gx.resize((size_t)nj * nk); c.gy.resize((size_t)nj * nk); c.gz.resize((size_t)nj * nk);
  }
  if (!p.Media1D.empty() && c.m1d_depth.empty())
    if (!read_media1d(p.Media1D, c, err)) return 2;
  if (!p.Media3D.empty() && c.m3d_vp.empty()) {  // media.cu:157-167: vp, vs, rho (NX, NZ, NY)
    NcFile nf;
    if (!nf.open_read(p.Media3D, err)) return 4;
    if (!nf.read_float("vp", c.m3d_vp, err) || !nf.read_float("vs", c.m3d_vs, err) || !nf.read_float("rho", c.m3d_rho, err))
      return 4;
  }
  if (!c.m3d_vp.empty() && (c.m3d_vp.size() < (size_t)ni * nj * nk || c.m3d_vs.size() < (size_t)ni * nj * nk ||
                            c.m3d_rho.size() < (size_t)ni * nj * nk)) {
    err = "Media3D smaller than NX*NY*NZ";
    return 4;
  }
  if (p.Rupture && p.INPORT_STRESS_TYPE == 1 && c.stress_in.find("Tx") == c.stress_in.end()) {
    NcFile nf;
    if (!nf.open_read(p.Fault_init_stress, err)) return 4;
    std::vector<std::string> names = {"Tx", "Ty", "Tz", "dTx", "dTy", "dTz"};
    if (p.Friction_type == 0) for (auto s : {"mu_s", "mu_d", "Dc", "C0"}) names.push_back(s);
    else for (auto s : {"a", "b", "L", "Vw", "State"}) names.push_back(s);
    if (p.Friction_type == 3) names.push_back("TP_hy");  // init.cu:999-1003
    for (auto& s : names) {
      std::vector<float> v;
      if (!nf.read_float(s, v, err)) return 4;
      v.resize((size_t)nj * nk);
      c.stress_in[s] = v;
    }
  }

  // Device copies of the setup inputs and the FP64 coordinates, freed at the
  // end of the setup (the run needs neither).
  struct DevTemps {
    std::vector<void*> v;
    ~DevTemps() { for (void* q : v) (void)hipFree(q); }
  } tmp;
  auto dput = [&](const void* src, size_t bytes) -> void* {
    void* q = nullptr;
    if (bytes == 0) return nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return nullptr;
    tmp.v.push_back(q);
    if (src && hipMemcpy(q, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return q;
  };
  auto blocks_for = [](long n) { return (unsigned)((n + 255) / 256); };

  // ---- coordinates (coord.cu:153-328) and ghost extrapolation (metric.cu:248-313), on the device
  if (p.INPORT_GRID_TYPE == 1 && c.coord3d.empty() &&
      (c.gx.size() < (size_t)nj * nk || c.gy.size() < (size_t)nj * nk || c.gz.size() < (size_t)nj * nk)) {
    err = "fault geometry smaller than NY*NZ";
    return 4;
  }
  if (!c.topo.empty() && c.topo.size() < (size_t)ni * nj) { err = "topography smaller than NX*NY"; return 4; }
  if (!c.coord3d.empty() && c.coord3d.size() < 3 * (size_t)nxyz) { err = "3-D coordinates smaller than 3 x the padded grid"; return 4; }
  double* dC = (double*)dput(nullptr, 3 * nxyz * sizeof(double));
  if (!dC) { err = "hipMalloc failed for the coordinates"; return 6; }
  {
    std::vector<double> xline(p.NX + 6);
    int ic = p.NX / 2 + 3;
    xline[ic] = 0;
    auto spacing = [&](int i, int ref) {
      int dist = std::abs(i - ic);
      double cp = dist < 10 ? 0.0 : (dist < 55 ? 1.0 - std::cos(PI_ * (i - ref) / 45.0) : 2.0);
      return p.DH * (0.5 + 0.25 * cp);
    };
    for (int i = ic + 1; i < p.NX + 6; i++) xline[i] = xline[i - 1] + spacing(i, ic + 10);
    for (int i = ic - 1; i >= 0; i--) xline[i] = xline[i + 1] - spacing(i, ic - 10);
    bool ok = true;
    if (!c.coord3d.empty()) {  // interior points as given; the ghosts follow below
      ok = hipMemcpy(dC, c.coord3d.data(), 3 * nxyz * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    } else {
      CoordIn ci{};
      ci.xline = (const double*)dput(xline.data(), xline.size() * sizeof(double));
      if (p.INPORT_GRID_TYPE == 1) {
        ci.gx = (const float*)dput(c.gx.data(), c.gx.size() * sizeof(float));
        ci.gy = (const float*)dput(c.gy.data(), c.gy.size() * sizeof(float));
        ci.gz = (const float*)dput(c.gz.data(), c.gz.size() * sizeof(float));
        ok = ci.gx && ci.gy && ci.gz;
      }
      if (!c.topo.empty()) { ci.topo = (const float*)dput(c.topo.data(), c.topo.size() * sizeof(float)); ok = ok && ci.topo; }
      ci.nx = nx; ci.ny = ny; ci.nz = nz; ci.nj = nj; ci.NX = p.NX; ci.NY = p.NY; ci.NZ = p.NZ;
      ci.nyz = nyz; ci.DH = p.DH; ci.normal_refine = p.NormalRefine; ci.grid_type = p.INPORT_GRID_TYPE;
      ok = ok && ci.xline;
      if (ok) hipLaunchKernelGGL(coord_kernel, dim3(blocks_for((long)(nx - 6) * (ny - 6) * (nz - 6))), dim3(256), 0, 0, ci, dC,
                                 dC + nxyz, dC + 2 * nxyz);
    }
    if (ok) {
      const long lines[3] = {(long)ny * nz, (long)nx * nz, (long)nx * ny};
      for (int ax = 0; ax < 3; ax++)
        hipLaunchKernelGGL(coord_ghost_kernel, dim3(blocks_for(lines[ax]), 3), dim3(256), 0, 0, dC, nxyz, nx, ny, nz, nyz, ax);
    }
    if (!ok || hipGetLastError() != hipSuccess) { err = "coordinates on the device failed"; return 6; }
  }
  const double* X = dC;
  const double* Y = dC + nxyz;
  const double* Z = dC + 2 * nxyz;

  // ---- grid-step range for the CFL check (cal_range_steph, coord.cu:330-391)
  {
    unsigned long long* dh = (unsigned long long*)dput(nullptr, sizeof(unsigned long long));
    const unsigned long long init = ~0ull;
    unsigned long long bits = 0;
    bool ok = dh && hipMemcpy(dh, &init, sizeof init, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) hipLaunchKernelGGL(hmin_kernel, dim3(blocks_for((long)(nx - 6) * (ny - 6) * (nz - 6))), dim3(256), 0, 0, X, Y, Z, nx, ny,
                               nz, nyz, dh);
    ok = ok && hipGetLastError() == hipSuccess && hipMemcpy(&bits, dh, sizeof bits, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) { err = "grid-step range on the device failed"; return 6; }
    c.hmin = from_order_key(bits);
  }
  const double hmin = c.hmin;

  // coordinates of the output files: the fault plane and the three wave
  // slices (io_fault.cu:37-67, io_wave_*.cu)
  {
    const int kz = std::min(std::max(p.EXPORT_WAVE_SLICE_Z, 0), nk - 1);
    const int jy = std::min(std::max(p.EXPORT_WAVE_SLICE_Y, 0), nj - 1);
    const int ix = std::min(std::max(p.EXPORT_WAVE_SLICE_X, 0), ni - 1);
    auto P3 = [&](int i, int j, int k) { return (long)j + (long)k * ny + (long)i * nyz; };
    // (base, stride a, stride b, na, nb) of the fault plane (nk, nj), xy (ni, nj), xz (ni, nk), yz (nk, nj)
    const long base[4] = {P3(d.i0, 3, 3), P3(3, 3, kz + 3), P3(3, jy + 3, 3), P3(ix + 3, 3, 3)};
    const long sa[4] = {(long)ny, nyz, nyz, (long)ny}, sb[4] = {1, 1, (long)ny, 1};
    const int na[4] = {nk, ni, ni, nk}, nb[4] = {nj, nj, nk, nj};
    for (int s = 0; s < 4; s++) {
      const long n = (long)na[s] * nb[s];
      float* dst = (float*)dput(nullptr, 3 * n * sizeof(float));
      c.out_coord[s].resize(3 * n);
      bool ok = dst != nullptr;
      if (ok) hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, X, nxyz, base[s], sa[s], sb[s], na[s], nb[s], dst);
      ok = ok && hipGetLastError() == hipSuccess &&
           hipMemcpy(c.out_coord[s].data(), dst, 3 * n * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
      if (!ok) { err = "coordinate slices failed"; return 6; }
    }
  }

  // ---- metric + media (metric.cu:15-171, extend_Symm_array metric.cu:173-246,
  // media.cu:63-204): on the GPU, FP64 from the FP64 coordinates, rounded
  // once into the device layout (FP32 or BF16).  The fault coefficients and
  // the free-surface maps below evaluate the same point function on the
  // device for the planes they read.
  const long pyz = d.pyz;
  MetricIn md{};
  md.X = X; md.Y = Y; md.Z = Z;
  md.nyz = nyz; md.ny = ny; md.nz = nz; md.nx = nx; md.nj = nj; md.nk = nk; md.NX = p.NX;
  md.rDH = 1.0 / p.DH;
  md.lam1 = p.bi_rho1 * (p.bi_vp1 * p.bi_vp1 - 2 * p.bi_vs1 * p.bi_vs1); md.miu1 = p.bi_rho1 * p.bi_vs1 * p.bi_vs1;
  md.lam2 = p.bi_rho2 * (p.bi_vp2 * p.bi_vp2 - 2 * p.bi_vs2 * p.bi_vs2); md.miu2 = p.bi_rho2 * p.bi_vs2 * p.bi_vs2;
  md.rho1 = p.bi_rho1; md.rho2 = p.bi_rho2;
  md.nl = (int)c.m1d_depth.size();
  {
    const size_t n1 = c.m1d_depth.size() * sizeof(float);
    md.d1 = (const float*)dput(c.m1d_depth.data(), n1); md.r1 = (const float*)dput(c.m1d_rho.data(), n1);
    md.p1 = (const float*)dput(c.m1d_vp.data(), n1); md.s1 = (const float*)dput(c.m1d_vs.data(), n1);
    if (!c.m3d_vp.empty()) {
      const size_t n3 = c.m3d_vp.size() * sizeof(float);
      md.vp3 = (const float*)dput(c.m3d_vp.data(), n3); md.vs3 = (const float*)dput(c.m3d_vs.data(), n3);
      md.rho3 = (const float*)dput(c.m3d_rho.data(), n3);
    }
    unsigned long long* dvp = (unsigned long long*)dput(nullptr, sizeof(unsigned long long));
    bool ok = dvp && (md.nl == 0 || (md.d1 && md.r1 && md.p1 && md.s1)) && (c.m3d_vp.empty() || (md.vp3 && md.vs3 && md.rho3));
    const long nmet = (long)nx * (p.MetricBF16 == 2 ? 18 : 13) * pyz;
    void* met = nullptr;
    if (ok) met = p.MetricBF16 ? (void*)dmalloc<uint16_t>(c, nmet) : (void*)dmalloc<float>(c, nmet);
    ok = ok && met && hipMemset(dvp, 0, sizeof(unsigned long long)) == hipSuccess;
    MetRef ref{};
    if (ok && p.MetricBF16 == 2) {
      // reference value of each metric field: the midrange over every 4th
      // stored point along each axis (and the last), ghosts mirrored as the
      // device stores them, rounded to FP32 (the kernels add it in FP32)
      auto axis = [](int n) {
        std::vector<int> v;
        for (int q = 0; q < n; q += 4) v.push_back(mirror_idx(q, n));
        v.push_back(mirror_idx(n - 1, n));
        return v;
      };
      const std::vector<int> ai = axis(nx), aj = axis(ny), ak = axis(nz);
      std::vector<unsigned long long> keys(20);
      for (int q = 0; q < 10; q++) { keys[2 * q] = ~0ull; keys[2 * q + 1] = 0ull; }
      const int* dai = (const int*)dput(ai.data(), ai.size() * sizeof(int));
      const int* daj = (const int*)dput(aj.data(), aj.size() * sizeof(int));
      const int* dak = (const int*)dput(ak.data(), ak.size() * sizeof(int));
      unsigned long long* dk = (unsigned long long*)dput(keys.data(), keys.size() * sizeof(unsigned long long));
      ok = dai && daj && dak && dk;
      const long ns = (long)ai.size() * aj.size() * ak.size();
      if (ok) hipLaunchKernelGGL(metric_range_kernel, dim3(blocks_for(ns)), dim3(256), 0, 0, md, dai, daj, dak, (int)ai.size(),
                                 (int)aj.size(), (int)ak.size(), dk);
      ok = ok && hipGetLastError() == hipSuccess &&
           hipMemcpy(keys.data(), dk, keys.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) == hipSuccess;
      for (int q = 0; ok && q < 10; q++) {
        c.met_ref[q] = (float)(0.5 * (from_order_key(keys[2 * q]) + from_order_key(keys[2 * q + 1])));
        ref.r[q] = c.met_ref[q];
      }
    }
    if (ok) {
      const unsigned blocks = blocks_for((long)nx * nz * ny);
      if (p.MetricBF16 == 2)
        hipLaunchKernelGGL((metric_media_kernel<uint16_t, 18>), dim3(blocks), dim3(256), 0, 0, md, (uint16_t*)met, pyz, d.py, d.jpad, dvp, ref, 0, nx);
      else if (p.MetricBF16)
        hipLaunchKernelGGL(metric_media_kernel<uint16_t>, dim3(blocks), dim3(256), 0, 0, md, (uint16_t*)met, pyz, d.py, d.jpad, dvp, ref, 0, nx);
      else
        hipLaunchKernelGGL(metric_media_kernel<float>, dim3(blocks), dim3(256), 0, 0, md, (float*)met, pyz, d.py, d.jpad, dvp, ref, 0, nx);
      // the media palette: a 1-D layered model (one entry per layer) or the
      // two halves of a bimaterial one, used when every point's FP32 media
      // words equal its entry's
      c.met_pal = nullptr;
      c.met_npal = 0;
      const int npal = md.nl > 0 ? md.nl : 2;
      if (ok && p.MetricBF16 == 2 && p.MetricPalette && c.m3d_vp.empty() && npal <= kMaxMediaPalette) {
        float4* pal = dmalloc<float4>(c, npal);
        unsigned* bad = (unsigned*)dput(nullptr, sizeof(unsigned));
        unsigned nbad = 1;
        if (pal && bad && hipMemset(bad, 0, sizeof(unsigned)) == hipSuccess) {
          for (int check = 0; check < 2; check++)
            hipLaunchKernelGGL(media_palette_kernel, dim3(blocks), dim3(256), 0, 0, (const uint16_t*)met, pyz, d.py, d.jpad, nx,
                               ny, nz, pal, npal, check, bad);
          if (hipGetLastError() != hipSuccess || hipMemcpy(&nbad, bad, sizeof nbad, hipMemcpyDeviceToHost) != hipSuccess)
            ok = false;
        }
        if (nbad == 0) { c.met_pal = (float*)pal; c.met_npal = npal; }
      }
      // the coefficient words that are 0 on the whole grid (MetricSkipConstant)
      c.met_skip = 0;
      if (ok && p.MetricBF16 == 2 && p.MetricSkipConstant) {
        unsigned* any = (unsigned*)dput(nullptr, 5 * sizeof(unsigned));
        unsigned h[5] = {1u, 1u, 1u, 1u, 1u};
        if (any && hipMemset(any, 0, 5 * sizeof(unsigned)) == hipSuccess) {
          hipLaunchKernelGGL(metric_pair_or_kernel, dim3(std::min(blocks, 4096u)), dim3(256), 0, 0, (const uint16_t*)met, pyz,
                             d.py, d.jpad, nx, ny, nz, any);
          if (hipGetLastError() != hipSuccess || hipMemcpy(h, any, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
        }
        for (int q = 0; q < 5; q++)
          if (h[q] == 0u) c.met_skip |= 1u << q;
      }
      if (p.MetricBF16 && p.Rupture) {
        // FP32 metric and media of the planes the fault kernels read (i0-6..i0+6):
        // the fault normal and traction stay FP32 under bfloat16 storage
        const int fb = std::max(d.i0 - 6, 0), fe = std::min(d.i0 + 6, nx - 1), fnp = fe - fb + 1;
        c.met_fault = dmalloc<float>(c, (size_t)fnp * 13 * pyz);
        if (c.met_fault) {
          c.met_fault_base = c.met_fault - (long)fb * 13 * pyz;
          hipLaunchKernelGGL(metric_media_kernel<float>, dim3(blocks_for((long)fnp * nz * ny)), dim3(256), 0, 0, md, c.met_fault,
                             pyz, d.py, d.jpad, dvp, ref, fb, fnp);
        }
        ok = c.met_fault != nullptr;
      }
      unsigned long long bits = 0;
      ok = ok && hipGetLastError() == hipSuccess && hipMemcpy(&bits, dvp, sizeof bits, hipMemcpyDeviceToHost) == hipSuccess;
      double vpm;
      std::memcpy(&vpm, &bits, sizeof vpm);
      c.vpmax = vpm;
    }
    if (!ok) { err = "metric computation on the device failed"; return 6; }
    c.met = met;
  }
  // CFL (main.cu:302-345)
  const double vpmax = c.vpmax;
  float dtmax = (float)(1.3 * hmin / vpmax);
  if (!((float)p.DT < dtmax)) {
    char b[256];
    std::snprintf(b, sizeof b, "Serious Error: DT = %10.2e > dtmax = %10.2e (sec) do not satisfy stability condition", p.DT, (double)dtmax);
    err = b;
    return 5;
  }

  // ---- PML profiles (init_pml.cu:18-155)
  std::vector<double> Ap[3], Bp[3], Dp[3];
  {
    int n3[3] = {ni, nj, nk};
    int N = p.PML_N;
    double Rpp = std::pow(10.0, -((std::log10((double)N) - 1.0) / std::log10(2.0) + 3.0));
    double amax = PI_ * p.PML_fc, L0 = p.DH * (N - 1);
    double dmax = -p.PML_velocity / (2.0 * L0) * std::log(Rpp) * 3.0;
    for (int a = 0; a < 3; a++) {
      int n = n3[a];
      Ap[a].assign(n, 0.0); Bp[a].assign(n, 1.0); Dp[a].assign(n, 0.0);
      if (!p.usePML) continue;
      std::vector<double> Lx(n, -1.0);
      if (on[2 * a]) for (int i = 0; i < N && i < n; i++) Lx[i] = (N - 1 - i) * p.DH;
      if (on[2 * a + 1]) for (int i = 0; i < N && n - N + i >= 0; i++) Lx[n - N + i] = i * p.DH;
      for (int i = 0; i < n; i++) {
        double x = Lx[i];
        if (x < 0) continue;
        double r = x / L0;
        Ap[a][i] = amax * (1.0 - r);
        Bp[a][i] = 1.0 + (p.PML_bmax - 1.0) * r * r;
        Dp[a][i] = dmax * r * r / Bp[a][i];
      }
    }
  }

  // ---- free-surface velocity maps (init_wave_free_cu, fault_coef.cu:1320-1486), on the device
  c.matVx2Vz = dmalloc<float>(c, nx * (long)ny * 9);
  c.matVy2Vz = dmalloc<float>(c, nx * (long)ny * 9);
  if (!c.matVx2Vz || !c.matVy2Vz || hipMemset(c.matVx2Vz, 0, nx * (long)ny * 9 * sizeof(float)) != hipSuccess ||
      hipMemset(c.matVy2Vz, 0, nx * (long)ny * 9 * sizeof(float)) != hipSuccess) {
    err = "hipMalloc failed for the free-surface maps";
    return 6;
  }
  if (freenode) {
    const double* dB[3];
    for (int a = 0; a < 3; a++) dB[a] = (const double*)dput(Bp[a].data(), Bp[a].size() * sizeof(double));
    bool ok = dB[0] && dB[1] && dB[2];
    if (ok) hipLaunchKernelGGL(wave_free_kernel, dim3(blocks_for((long)(nx - 6) * (ny - 6))), dim3(256), 0, 0, md, dB[0], dB[1], dB[2],
                               c.matVx2Vz, c.matVy2Vz);
    if (!ok || hipGetLastError() != hipSuccess) { err = "free-surface maps on the device failed"; return 6; }
  }

  // ---- device allocations: 4 wavefield buffers (step start S, stage A/B, accumulator T)
  const long nvol = (long)nx * 9 * pyz;  // device wavefield words (padded rows)
  // buffer q starts q * WaveBufferSkew words into its allocation, so that
  // the same word of the four buffers, which a stage reads and writes
  // together, sits at different offsets from the allocations' 2 MB-aligned
  // bases (1600 x 1000 x 500, interleaved runs: 127.4, 127.2, 126.4 and 125.8
  // ms per step for skews of 0, 64, 1024 and 16384 words)
  const long skew = p.WaveBufferSkew;
  if (skew < 0 || skew % 32 != 0 || skew > (1L << 20)) {
    err = "WaveBufferSkew must be a multiple of 32 words from 0 to 2^20";
    return 1;
  }
  float** bufs[4] = {&c.S, &c.A, &c.B, &c.Tacc};
  for (int q = 0; q < 4; q++) {
    float* b = dmalloc<float>(c, nvol + q * skew);
    *bufs[q] = b ? b + q * skew : nullptr;
  }
  if (!c.S || !c.A || !c.B || !c.Tacc) { err = "hipMalloc failed for the wavefield"; return 6; }
  c.cur = c.S;

  // PML
  c.pml.N = p.PML_N;
  for (int a = 0; a < 3; a++) {
    std::vector<float> fa(Ap[a].begin(), Ap[a].end()), fb(Bp[a].begin(), Bp[a].end()), fd(Dp[a].begin(), Dp[a].end());
    c.pml.A[a] = dmalloc<float>(c, fa.size()); upload(c.pml.A[a], fa);
    c.pml.B[a] = dmalloc<float>(c, fb.size()); upload(c.pml.B[a], fb);
    c.pml.D[a] = dmalloc<float>(c, fd.size()); upload(c.pml.D[a], fd);
  }
  for (int f = 0; f < 6; f++) {
    c.pml.on[f] = p.usePML ? on[f] : 0;
    if (!c.pml.on[f]) continue;
    int a = f / 2;
    long n = 9L * p.PML_N * (a == 0 ? (long)nj * nk : a == 1 ? (long)nk * ni : (long)ni * nj);
    c.pml.Y[f] = dmalloc<float>(c, n);
    c.pml.S[f] = dmalloc<float>(c, n);
    c.pml.T[f] = dmalloc<float>(c, n);
    if (!c.pml.Y[f] || !c.pml.S[f] || !c.pml.T[f]) { err = "hipMalloc failed for PML"; return 6; }
  }

  // Cerjan: damp(i,j,k) = min of three 1-D profiles over the interior
  // (init_cerjan_host builds the same cube face by face with min).
  {
    int n3[3] = {nx, ny, nz};
    int ND = p.DAMP_N;
    float alpha = std::sqrt(-std::log(0.92f)) / ND;
    for (int a = 0; a < 3; a++) {
      std::vector<float> prof(n3[a], 1.0f);
      if (p.useCerjan) {
        for (int i = 0; i < ND && 3 + i < n3[a] - 3; i++) prof[3 + i] = std::exp(-std::pow(alpha * (ND - i), 2.0f));
        bool hi = a < 2 || !freenode;
        if (hi) for (int i = 0; i < ND && n3[a] - 4 - i >= 3; i++) prof[n3[a] - 4 - i] = std::exp(-std::pow(alpha * (ND - i), 2.0f));
      }
      c.cj[a] = dmalloc<float>(c, prof.size());
      upload(c.cj[a], prof);
    }
  }

  // ---- fault
  if (p.Rupture) {
    FaultDev& F = c.f;
    const int i0 = d.i0;
    long n2 = (long)nj * nk;
    // split-node coefficients (fault_coef.cu:26-1302) on the device, straight
    // into the fault's device arrays; the FP64 unit vectors come back for the
    // initial stress projection below
    FaultCoefOut fo{};
    fo.rho_f = F.rho_f = dmalloc<float>(c, 2 * nyz);
    fo.Dm = F.Dmat = dmalloc<float>(c, 2L * 9 * nyz * 9);
    fo.T1x = F.matT1toVx = dmalloc<float>(c, 2 * nyz * 9);
    fo.Vyx = F.matVytoVx = dmalloc<float>(c, 2 * nyz * 9);
    fo.Vzx = F.matVztoVx = dmalloc<float>(c, 2 * nyz * 9);
    fo.Vx2Vzf = F.matVx2Vz_f = dmalloc<float>(c, 2L * ny * 9);
    fo.Vy2Vzf = F.matVy2Vz_f = dmalloc<float>(c, 2L * ny * 9);
    fo.P2M = F.matP2Mf = dmalloc<float>(c, 3L * ny * 9);
    fo.M2P = F.matM2Pf = dmalloc<float>(c, 3L * ny * 9);
    fo.vn = F.vec_n = dmalloc<float>(c, nyz * 3);
    fo.vs1 = F.vec_s1 = dmalloc<float>(c, nyz * 3);
    fo.vs2 = F.vec_s2 = dmalloc<float>(c, nyz * 3);
    fo.jacvec = F.jacvec = dmalloc<float>(c, nyz);
    fo.jac0 = F.jac_i0 = dmalloc<float>(c, nyz);
    fo.nd = (double*)dput(nullptr, nyz * 9 * sizeof(double));
    std::vector<double> nd(nyz * 9);
    {
      bool ok = fo.rho_f && fo.Dm && fo.T1x && fo.Vyx && fo.Vzx && fo.Vx2Vzf && fo.Vy2Vzf && fo.P2M && fo.M2P && fo.vn && fo.vs1 &&
                fo.vs2 && fo.jacvec && fo.jac0 && fo.nd;
      // the surface maps exist for the row under the free surface only
      for (float* z : {fo.Vx2Vzf, fo.Vy2Vzf}) ok = ok && hipMemset(z, 0, 2L * ny * 9 * sizeof(float)) == hipSuccess;
      for (float* z : {fo.P2M, fo.M2P}) ok = ok && hipMemset(z, 0, 3L * ny * 9 * sizeof(float)) == hipSuccess;
      if (ok) hipLaunchKernelGGL(fault_coef_kernel, dim3(blocks_for(nyz)), dim3(256), 0, 0, md, i0, fo);
      ok = ok && hipGetLastError() == hipSuccess &&
           hipMemcpy(nd.data(), fo.nd, nd.size() * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
      if (!ok) { err = "fault coefficients on the device failed"; return 6; }
    }
    std::vector<double> vnd(nyz * 3), vs1d(nyz * 3), vs2d(nyz * 3);
    for (long q = 0; q < nyz; q++)
      for (int a = 0; a < 3; a++) { vnd[q * 3 + a] = nd[q * 9 + a]; vs1d[q * 3 + a] = nd[q * 9 + 3 + a]; vs2d[q * 3 + a] = nd[q * 9 + 6 + a]; }
    // initial fault state (init_fault_cu, init.cu:160-849)
    std::vector<float> Tn(n2), Ts1(n2), Ts2(n2), six(n2), siy(n2), siz(n2), peak(n2), initt0(n2, -9999.9f);
    std::vector<int> riy(n2), riz(n2), uni(n2), fg(n2);
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++) {
        long q = j + (long)k * nj, pv = (j + 3 + (long)(k + 3) * ny) * 3;
        int gj1 = j + 1, gk1 = k + 1;
        double tn = -120.0e6, ts1 = tn * p.mu_0, ts2 = 0.0;
        if (gj1 >= p.Asp_grid[0] && gj1 <= p.Asp_grid[1] && gk1 >= p.Asp_grid[2] && gk1 <= p.Asp_grid[3])
          ts1 = (p.mu_s + (p.mu_s - p.mu_d) * 0.001) * tn;
        bool inF = gj1 >= p.Fault_grid[0] && gj1 <= p.Fault_grid[1] && gk1 >= p.Fault_grid[2] && gk1 <= p.Fault_grid[3];
        peak[q] = inF ? (float)p.mu_s : 10000.f;
        double si[3];
        for (int a = 0; a < 3; a++) si[a] = tn * vnd[pv + a] + ts1 * vs1d[pv + a] + ts2 * vs2d[pv + a];
        six[q] = (float)si[0]; siy[q] = (float)si[1]; siz[q] = (float)si[2];
        Tn[q] = (float)tn; Ts1[q] = (float)ts1; Ts2[q] = (float)ts2;
        riy[q] = (gj1 <= p.Fault_grid[0] + 3 || gj1 >= p.Fault_grid[1] - 3) ? 0 : 1;
        riz[q] = (gk1 <= p.Fault_grid[2] + 3 || gk1 >= p.Fault_grid[3] - 3) ? 0 : 1;
        uni[q] = (gj1 > 30 && gj1 < p.NY - 30 && gk1 > 30) ? 0 : 1;  // init.cu:806-816 overrides the PML rule
        fg[q] = (gj1 >= p.Fault_grid[0] + 3 && gj1 <= p.Fault_grid[1] - 3 && gk1 >= p.Fault_grid[2] + 3 && gk1 <= p.Fault_grid[3] - 3) ? 1 : 0;
        // rate-state: a-b is tested before a and b are read (init.cu:835-847
        // precedes nc_read_init_stress), so it sees zeros.
        if (p.Friction_type >= 1) fg[q] = 0;
      }
    auto get_in = [&](const char* nm) -> std::vector<float> {
      auto it = c.stress_in.find(nm);
      if (it == c.stress_in.end()) return std::vector<float>(n2, 0.f);
      std::vector<float> v = it->second;
      v.resize(n2);
      return v;
    };
    std::vector<float> T0x = get_in("Tx"), T0y = get_in("Ty"), T0z = get_in("Tz");
    std::vector<float> dT0x = get_in("dTx"), dT0y = get_in("dTy"), dT0z = get_in("dTz");
    std::vector<float> Dcv = get_in("Dc"), C0v = get_in("C0"), av = get_in("a"), bv = get_in("b"), Lv = get_in("L"), Vwv = get_in("Vw");
    std::vector<double> State(n2, 0.0);
    if (p.INPORT_STRESS_TYPE == 1) {
      if (p.Friction_type == 0) peak = get_in("mu_s");
      else { std::vector<float> s = get_in("State"); for (long q = 0; q < n2; q++) State[q] = s[q]; }
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++) {
          long q = j + (long)k * nj, pv = (j + 3 + (long)(k + 3) * ny) * 3;
          double T0[3] = {T0x[q], T0y[q], T0z[q]};
          Tn[q] = (float)dot(T0, &vnd[pv]); Ts1[q] = (float)dot(T0, &vs1d[pv]); Ts2[q] = (float)dot(T0, &vs2d[pv]);
        }
    }
#define UPF(dst, vec) do { F.dst = dmalloc<float>(c, (vec).size()); upload(F.dst, vec); } while (0)
#define UPI(dst, vec) do { F.dst = dmalloc<int>(c, (vec).size()); upload(F.dst, vec); } while (0)
    F.W = dmalloc<float>(c, 18 * nyz); F.mW = dmalloc<float>(c, 18 * nyz);
    F.hW = dmalloc<float>(c, 18 * nyz); F.tW = dmalloc<float>(c, 18 * nyz);
    F.T11 = dmalloc<float>(c, 7 * nyz); F.T12 = dmalloc<float>(c, 7 * nyz); F.T13 = dmalloc<float>(c, 7 * nyz);
    F.mT11 = dmalloc<float>(c, nyz); F.mT12 = dmalloc<float>(c, nyz); F.mT13 = dmalloc<float>(c, nyz);
    F.hT11 = dmalloc<float>(c, nyz); F.hT12 = dmalloc<float>(c, nyz); F.hT13 = dmalloc<float>(c, nyz);
    UPF(Tn, Tn); UPF(Ts1, Ts1); UPF(Ts2, Ts2);
    F.tTn = dmalloc<float>(c, n2); F.tTs1 = dmalloc<float>(c, n2); F.tTs2 = dmalloc<float>(c, n2);
    UPF(str_init_x, six); UPF(str_init_y, siy); UPF(str_init_z, siz);
    UPF(T0x, T0x); UPF(T0y, T0y); UPF(T0z, T0z); UPF(dT0x, dT0x); UPF(dT0y, dT0y); UPF(dT0z, dT0z);
    UPF(str_peak, peak); UPF(Dc, Dcv); UPF(C0, C0v); UPF(a, av); UPF(b, bv); UPF(L, Lv); UPF(Vw, Vwv);
    F.slip = dmalloc<double>(c, n2); F.mslip = dmalloc<double>(c, n2); F.tslip = dmalloc<double>(c, n2);
    F.State = dmalloc<double>(c, n2); upload(F.State, State);
    F.mState = dmalloc<double>(c, n2); F.tState = dmalloc<double>(c, n2);
    F.hslip = dmalloc<float>(c, n2); F.Vs1 = dmalloc<float>(c, n2); F.Vs2 = dmalloc<float>(c, n2);
    UPF(init_t0, initt0);
    F.friction = dmalloc<float>(c, n2); F.rsD_y = dmalloc<float>(c, n2); F.rsD_z = dmalloc<float>(c, n2); F.rup_sensor = dmalloc<float>(c, n2);
    F.init_t0_flag = dmalloc<int>(c, n2); F.flag_rup = dmalloc<int>(c, n2);
    UPI(rup_index_y, riy); UPI(rup_index_z, riz); UPI(united, uni); UPI(faultgrid, fg);
    F.newton_stats = dmalloc<unsigned int>(c, 5 * n2 + kNewtonHistBins);
    if (p.Friction_type >= 1) {
      F.rs_carry = dmalloc<double>(c, 10 * n2);
      F.rs_carryf = dmalloc<float>(c, 9 * n2);
      if (!F.rs_carry || !F.rs_carryf) { err = "hipMalloc failed for the rate-state hand-over"; return 6; }
    }
    if (p.Friction_type == 3) {  // thermal pressurization, zero initial rises (alloc.cu:111-119)
      F.TP_n = std::max(p.TP_n, 1);
      F.TP_T = dmalloc<double>(c, (size_t)F.TP_n * n2);
      F.TP_P = dmalloc<double>(c, (size_t)F.TP_n * n2);
      std::vector<float> hy = get_in("TP_hy");
      std::vector<double> hyd(hy.begin(), hy.end());
      F.TP_hy = dmalloc<double>(c, n2);
      if (!F.TP_T || !F.TP_P || !F.TP_hy) { err = "hipMalloc failed for thermal pressurization"; return 6; }
      upload(F.TP_hy, hyd);
    }
    if (p.NewtonMode == 2) {
      if (p.Friction_type < 1) { err = "NewtonMode 2 needs rate-and-state friction"; return 1; }
      if (c.mlp_blob.empty() && !p.NewtonMLP.empty()) {
        FILE* fp = std::fopen(p.NewtonMLP.c_str(), "rb");
        if (!fp) { err = "cannot open NewtonMLP " + p.NewtonMLP; return 2; }
        c.mlp_blob.resize(kMlpParams + 1);
        const size_t got = std::fread(c.mlp_blob.data(), sizeof(float), kMlpParams + 1, fp);
        std::fclose(fp);
        if (got != (size_t)kMlpParams) { err = "NewtonMLP " + p.NewtonMLP + " does not hold 4555 float32 values"; return 1; }
        c.mlp_blob.resize(kMlpParams);
      }
      if (c.mlp_blob.empty()) { err = "NewtonMode 2 needs network weights (NewtonMLP or cgfdm_set_newton_mlp)"; return 1; }
      if (mlp_upload(c, c.mlp_blob.data(), (long)c.mlp_blob.size())) { err = "hipMalloc failed for the Newton network"; return 6; }
    }
    if (p.FaultSmooth) {
      // half-volume Gaussian weights of smooth_gauss_volume1 (smooth.cu:70-120):
      // exp(-r^2/0.7^2), off-plane layers x 0.4, each side normalised to 1
      std::vector<float> w(2 * 4 * 49);
      std::vector<double> wd(4 * 49);
      double sum = 0;
      for (int ii = 0; ii < 4; ii++)
        for (int a = 0; a < 7; a++)
          for (int b = 0; b < 7; b++) {
            const double r2 = (double)(ii * ii + (a - 3) * (a - 3) + (b - 3) * (b - 3)) / (0.7 * 0.7);
            wd[(ii * 7 + a) * 7 + b] = std::exp(-r2) * (ii == 0 ? 1.0 : 0.4);
            sum += wd[(ii * 7 + a) * 7 + b];
          }
      for (int s = 0; s < 2; s++)
        for (int q = 0; q < 4 * 49; q++) w[s * 4 * 49 + q] = (float)(wd[q] / sum);
      F.smooth_w = dmalloc<float>(c, w.size());
      F.Ws = dmalloc<float>(c, 18 * nyz);
      if (!F.smooth_w || !F.Ws || !upload(F.smooth_w, w)) { err = "hipMalloc failed for the fault smoothing"; return 6; }
    }
#undef UPF
#undef UPI
    c.slabH = dmalloc<float>(c, 7 * 9 * pyz);
    if (!c.slabH || !F.newton_stats) { err = "hipMalloc failed for the fault"; return 6; }
  }
  c.hW = nullptr;
  if (upload_wave_cold(c)) { err = "hipMalloc/hipMemcpy failed for the stage-kernel tables"; return 6; }
  if (hipDeviceSynchronize() != hipSuccess) { err = "HIP error during setup upload"; return 6; }
  return 0;
}

}  // namespace cgfdm

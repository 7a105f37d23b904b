// Device-side check of an FP32 field: the number of non-finite words, the
// number of nonzero words in the three ghost layers of a volume array (the
// reference keeps them zero on one rank), and the largest magnitude of the
// finite interior words.  One read of the array, no host copy: the property
// checks of the 800 M-point workload and the run's guard against a solution
// that has turned to NaN use it.
#include "cgfdm_internal.h"

#include <cstring>

namespace cgfdm {

namespace {

constexpr int kCheckBlock = 256;

// counts[0] non-finite words, counts[1] nonzero ghost words, low 32 bits of
// counts[2] the bit pattern of the largest finite magnitude (non-negative
// floats order like their bit patterns)
__global__ void __launch_bounds__(kCheckBlock) field_check_kernel(const float* __restrict__ a, int nf, int f0, int fn,
                                                                   int nx, int ny, int nz, int py, int jpad, long pyz,
                                                                   long flat_n, unsigned long long* counts) {
  float v = 0.0f;
  bool valid, ghost = false;
  if (flat_n > 0) {
    const long idx = (long)blockIdx.x * kCheckBlock + threadIdx.x;
    valid = idx < flat_n;
    if (valid) v = a[idx];
  } else {
    const int j = blockIdx.x * kCheckBlock + threadIdx.x, k = blockIdx.y;
    const int i = blockIdx.z / fn, c = f0 + blockIdx.z % fn;
    valid = j < ny;
    if (valid) v = a[((long)i * nf + c) * pyz + (long)k * py + jpad + j];
    ghost = i < 3 || i >= nx - 3 || k < 3 || k >= nz - 3 || j < 3 || j >= ny - 3;
  }
  const bool bad = valid && !isfinite(v);
  const bool gnz = valid && ghost && v != 0.0f;
  float m = (valid && !bad && !ghost) ? fabsf(v) : 0.0f;
  const unsigned long long nbad = __popcll(__ballot(bad));
  const unsigned long long ngh = __popcll(__ballot(gnz));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (__lane_id() == 0) {
    if (nbad) atomicAdd(&counts[0], nbad);
    if (ngh) atomicAdd(&counts[1], ngh);
    if (m > 0.0f) atomicMax((unsigned int*)&counts[2], __float_as_uint(m));
  }
}

__global__ void clear_counts_kernel(unsigned long long* counts) {
  if (threadIdx.x < 3) counts[threadIdx.x] = 0ull;
}

}  // namespace

int field_check_host(Ctx& c, const float* ptr, int nf, int f0, int fn, long flat_n, long long* nonfinite,
                     long long* ghost_nonzero, double* maxabs, std::string& err) {
  if (!c.check_counts) {
    void* p = nullptr;
    if (hipMalloc(&p, 3 * sizeof(unsigned long long)) != hipSuccess) { err = "hipMalloc of the check counters"; return 1; }
    c.allocations.push_back(p);
    c.check_counts = (unsigned long long*)p;
  }
  hipLaunchKernelGGL(clear_counts_kernel, dim3(1), dim3(64), 0, c.stream, c.check_counts);
  if (flat_n > 0) {
    const long blocks = (flat_n + kCheckBlock - 1) / kCheckBlock;
    hipLaunchKernelGGL(field_check_kernel, dim3((unsigned)blocks), dim3(kCheckBlock), 0, c.stream, ptr, 0, 0, 1, 0, 0,
                       0, 0, 0, 0L, flat_n, c.check_counts);
  } else {
    const long zdim = (long)c.d.nx * fn;
    if (zdim > 65535 || c.d.nz > 65535) { err = "field too large for the check grid"; return 1; }
    const dim3 grid((unsigned)((c.d.ny + kCheckBlock - 1) / kCheckBlock), (unsigned)c.d.nz, (unsigned)zdim);
    hipLaunchKernelGGL(field_check_kernel, grid, dim3(kCheckBlock), 0, c.stream, ptr, nf, f0, fn, c.d.nx, c.d.ny, c.d.nz,
                       c.d.py, c.d.jpad, c.d.pyz, 0L, c.check_counts);
  }
  if (hipError_t e = hipGetLastError()) { err = std::string("field check launch: ") + hipGetErrorString(e); return 1; }
  unsigned long long h[3];
  if (hipMemcpyAsync(h, c.check_counts, sizeof(h), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
      hipStreamSynchronize(c.stream) != hipSuccess) {
    err = "field check copy failed";
    return 1;
  }
  unsigned int bits = (unsigned int)(h[2] & 0xffffffffull);
  float mx;
  std::memcpy(&mx, &bits, sizeof(mx));
  if (nonfinite) *nonfinite = (long long)h[0];
  if (ghost_nonzero) *ghost_nonzero = (long long)h[1];
  if (maxabs) *maxabs = (double)mx;
  return 0;
}

}  // namespace cgfdm

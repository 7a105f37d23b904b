// Streaming-bandwidth probe: the practical HBM ceiling of a kernel that reads
// NR arrays and writes NW arrays once each, with the access form of the stage
// kernels (non-temporal 16-byte loads and stores, fully coalesced).  The
// stage kernels read about two words per word they write (stage 0 21:18,
// stages 1-2 39:18, stage 3 30:9), so the 2:1 form is their ceiling; 1:1 is a
// copy and 1:0 a pure read.  Measurement infrastructure only: nothing on the
// solver path calls it.
#include "cgfdm_internal.h"

namespace cgfdm {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kProbeUnroll = 4;
constexpr int kProbeBlock = 256;

template <int NR, int NW>
__global__ void __launch_bounds__(kProbeBlock) stream_probe_kernel(const v4f* __restrict__ r0,
                                                                   const v4f* __restrict__ r1,
                                                                   const v4f* __restrict__ r2,
                                                                   v4f* __restrict__ w0, v4f* __restrict__ w1,
                                                                   v4f* __restrict__ sink, long n) {
  const long base = (long)blockIdx.x * kProbeBlock * kProbeUnroll + threadIdx.x;
  v4f v[kProbeUnroll];
#pragma unroll
  for (int u = 0; u < kProbeUnroll; u++) {
    const long i = base + (long)u * kProbeBlock;
    v4f s = {0.f, 0.f, 0.f, 0.f};
    if (i < n) {
      if (NR > 0) s = __builtin_nontemporal_load(&r0[i]);
      if (NR > 1) {
        const v4f b = __builtin_nontemporal_load(&r1[i]);
        s += b;
      }
      if (NR > 2) {
        const v4f b = __builtin_nontemporal_load(&r2[i]);
        s += b;
      }
    }
    v[u] = s;
  }
#pragma unroll
  for (int u = 0; u < kProbeUnroll; u++) {
    const long i = base + (long)u * kProbeBlock;
    if (i >= n) continue;
    if (NW > 0) __builtin_nontemporal_store(v[u], &w0[i]);
    if (NW > 1) __builtin_nontemporal_store(v[u], &w1[i]);
    // read-only form: keep the loads live with a store that never happens
    // for finite data (the arrays hold zeros)
    if (NW == 0 && v[u].x != v[u].x) sink[0] = v[u];
  }
}

template <int NR, int NW>
void launch_probe(v4f* const* bufs, long n, hipStream_t st) {
  const long per_block = (long)kProbeBlock * kProbeUnroll;
  const unsigned grid = (unsigned)((n + per_block - 1) / per_block);
  hipLaunchKernelGGL((stream_probe_kernel<NR, NW>), dim3(grid), dim3(kProbeBlock), 0, st, bufs[0], bufs[1], bufs[2],
                     bufs[3], bufs[4], bufs[5], n);
}

}  // namespace

int stream_probe_host(int nread, int nwrite, size_t bytes_per_array, int reps, double* gbps, std::string& err) {
  using Fn = void (*)(v4f* const*, long, hipStream_t);
  Fn fn = nullptr;
  if (nread == 1 && nwrite == 0) fn = launch_probe<1, 0>;
  else if (nread == 1 && nwrite == 1) fn = launch_probe<1, 1>;
  else if (nread == 2 && nwrite == 1) fn = launch_probe<2, 1>;
  else if (nread == 3 && nwrite == 1) fn = launch_probe<3, 1>;
  else if (nread == 2 && nwrite == 2) fn = launch_probe<2, 2>;
  if (!fn) { err = "stream probe forms: 1:0, 1:1, 2:1, 3:1, 2:2 (arrays read : arrays written)"; return 1; }
  const long n = (long)(bytes_per_array / sizeof(v4f));
  if (n <= 0 || n > (1L << 31) || reps <= 0) { err = "stream probe: bad size or repetitions"; return 1; }
  v4f* bufs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  const int narr = nread + nwrite;
  for (int b = 0; b < narr && !rc; b++) {
    const int slot = b < nread ? b : 3 + (b - nread);
    if (hipMalloc((void**)&bufs[slot], (size_t)n * sizeof(v4f)) != hipSuccess ||
        hipMemset(bufs[slot], 0, (size_t)n * sizeof(v4f)) != hipSuccess) {
      err = "stream probe: device allocation failed";
      rc = 1;
    }
  }
  if (!rc && hipMalloc((void**)&bufs[5], sizeof(v4f)) != hipSuccess) { err = "stream probe: allocation failed"; rc = 1; }
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
    err = "stream probe: event creation failed";
    rc = 1;
  }
  if (!rc) {
    fn(bufs, n, nullptr);  // warm-up
    fn(bufs, n, nullptr);
    const hipError_t r0 = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; r++) fn(bufs, n, nullptr);
    const hipError_t r1 = hipEventRecord(e1, nullptr);
    float ms = 0.f;
    if (r0 != hipSuccess || r1 != hipSuccess || hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.f)) {
      err = "stream probe: timing failed";
      rc = 1;
    } else {
      *gbps = (double)narr * (double)n * sizeof(v4f) * reps / (ms * 1e-3) / 1e9;
    }
  }
  if (e0) (void)hipEventDestroy(e0);  // best effort: teardown of the probe
  if (e1) (void)hipEventDestroy(e1);
  for (v4f* b : bufs)
    if (b) (void)hipFree(b);
  return rc;
}

}  // namespace cgfdm

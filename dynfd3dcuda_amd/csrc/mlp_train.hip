// Learned Newton guess: device copy of the network, evaluation, gradient and
// training kernels (extension; no counterpart in the reference).
//
// The network and its parameter blob are described in mlp.h.  Training keeps
// FP32 master parameters P (the blob layout) and, before every batch, packs
// them into the device copy the rate-state kernel uses, so the forward pass
// of training has exactly the roundings of inference: layer-2 weights and
// first hidden layer in bf16, FP32 elsewhere and in all sums.  The backward
// pass runs in FP32 and treats the bf16 rounding as identity (gradients with
// respect to P).  Loss: mean over samples of (o - t')^2 with
// o the network output before de-standardisation and t' = (t - ymean)/yscale.
#include "cgfdm_internal.h"
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>
#include <vector>

namespace cgfdm {

// Scratch of cgfdm_mlp_forward_noalloc: the packed network (an MlpDev) in
// static device memory, one copy per GPU, held until the process ends, so
// that call allocates nothing.  Every call on a GPU overwrites that GPU's
// copy: calls that overlap on one GPU race.
constexpr size_t kMlpScratchBytes = 32768;
static_assert(sizeof(MlpDev) <= kMlpScratchBytes, "the scratch holds one packed network");
static __device__ __align__(256) unsigned char scratch[kMlpScratchBytes];

// offsets of the parameter blob
constexpr int kOffW1 = 10, kOffB1 = kOffW1 + kMlpH * kMlpIn, kOffW2 = kOffB1 + kMlpH,
              kOffB2 = kOffW2 + kMlpH * kMlpH, kOffW3 = kOffB2 + kMlpH, kOffB3 = kOffW3 + kMlpH;
static_assert(kOffB3 + 1 == kMlpParams, "blob layout");

// P (FP32 blob) -> MlpDev, one thread per packed element (4096 threads)
__global__ void __launch_bounds__(256) mlp_pack_kernel(const float* __restrict__ P, MlpDev* __restrict__ m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kMlpH * kMlpH) m->W2[t >> 6][t & 63] = f2bf(P[kOffW2 + t]);
  if (t < kMlpH * kMlpIn) m->W1[t / kMlpIn][t % kMlpIn] = P[kOffW1 + t];
  if (t < kMlpH) {
    m->b1[t] = P[kOffB1 + t];
    m->b2[t] = P[kOffB2 + t];
    m->w3[t] = P[kOffW3 + t];
  }
  if (t < 4) {
    m->xmean[t] = P[t];
    m->xscale[t] = P[4 + t];
  }
  if (t == 0) {
    m->ymean = P[8];
    m->yscale = P[9];
    m->b3 = P[kOffB3];
  }
}

// Network output for n samples (feat [n][4]): mode 0 with the wavefront
// matrix-instruction form used in the rate-state kernel, mode 1 scalar.
__global__ void __launch_bounds__(64) mlp_forward_kernel(const MlpDev* __restrict__ m, const float* __restrict__ feat,
                                                         long n, float* __restrict__ out, int mode) {
  const long s = (long)blockIdx.x * 64 + threadIdx.x;
  const bool ok = s < n;
  float f[4];
  for (int i = 0; i < 4; i++) f[i] = ok ? feat[4 * s + i] : 0.0f;
  const float y = mode == 0 ? mlp_guess_wave(m, f) : mlp_forward_scalar(m, f);
  if (ok) out[s] = y;
}

struct GradArgs {
  const MlpDev* m;
  const float* feat;    // [n][4]
  const float* target;  // [n] log10 slip rate
  const int* idx;       // sample order, or null for 0..count-1
  long first, count;    // samples idx[first .. first+count)
  float gscale;         // 2 / count
  float lscale;         // loss weight of one squared error
  float* grad;          // [kMlpParams], accumulated
  float* loss;          // accumulated
};

// Gradient of the loss over a batch.  A block is one wavefront; lane l owns
// hidden unit l of both layers for the accumulation (W1[l][:], b1[l],
// W2[l][:], b2[l], w3[l]) and lane 0 also b3.  Per chunk of 64 samples every
// lane runs forward and backward for its own sample in registers, and the
// per-sample outer products are summed through LDS.
__global__ void __launch_bounds__(64) mlp_grad_kernel(GradArgs a) {
  __shared__ float bufA[64][65];  // per-sample rows: h2, then dz2 ... (padded: conflict-free row writes)
  __shared__ float bufB[64][65];  // h1, then dz1
  __shared__ float xs[64][4];
  __shared__ float dos[64];
  const MlpDev* __restrict__ m = a.m;
  const int l = threadIdx.x;
  float gW2[kMlpH], gW1[4] = {0, 0, 0, 0}, gb1 = 0, gb2 = 0, gw3 = 0, gb3 = 0, lsum = 0;
  for (int i = 0; i < kMlpH; i++) gW2[i] = 0.0f;
  const long nchunk = (a.count + 63) / 64;
  for (long ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const long r = ch * 64 + l;
    const bool ok = r < a.count;
    long s = 0;
    if (ok) s = a.idx ? (long)a.idx[a.first + r] : a.first + r;
    float x[4];
    for (int i = 0; i < 4; i++) x[i] = ok ? (a.feat[4 * s + i] - m->xmean[i]) * m->xscale[i] : 0.0f;
    float h1[kMlpH];
#pragma unroll
    for (int h = 0; h < kMlpH; h++) h1[h] = mlp_h1(m, h, x);
    float o = m->b3;
    float h2[kMlpH];
#pragma unroll
    for (int h = 0; h < kMlpH; h++) {
      float z = 0.0f;
#pragma unroll
      for (int i = 0; i < kMlpH; i++) z += bf2f(m->W2[h][i]) * h1[i];
      h2[h] = fmaxf(z + m->b2[h], 0.0f);
      o += m->w3[h] * h2[h];
    }
    const float tn = ok ? (a.target[s] - m->ymean) / m->yscale : 0.0f;
    const float e = ok ? o - tn : 0.0f;
    lsum += e * e;
    const float dout = a.gscale * e;
    for (int i = 0; i < 4; i++) xs[l][i] = x[i];
    dos[l] = dout;
#pragma unroll
    for (int h = 0; h < kMlpH; h++) {
      bufA[l][h] = h2[h];
      bufB[l][h] = h1[h];
    }
    __syncthreads();
    for (int q = 0; q < 64; q++) {  // w3, b3
      gw3 += dos[q] * bufA[q][l];
      gb3 += dos[q];
    }
    __syncthreads();
    // dz2 of this lane's sample
    float dz2[kMlpH];
#pragma unroll
    for (int h = 0; h < kMlpH; h++) dz2[h] = h2[h] > 0.0f ? dout * m->w3[h] : 0.0f;
#pragma unroll
    for (int h = 0; h < kMlpH; h++) bufA[l][h] = dz2[h];
    __syncthreads();
    for (int q = 0; q < 64; q++) {  // W2 row l, b2[l]
      const float d = bufA[q][l];
      gb2 += d;
#pragma unroll
      for (int i = 0; i < kMlpH; i++) gW2[i] += d * bufB[q][i];
    }
    __syncthreads();
    // dz1 of this lane's sample: (W2^T dz2) masked by the first layer's ReLU
    float d1[kMlpH];
#pragma unroll
    for (int i = 0; i < kMlpH; i++) d1[i] = 0.0f;
#pragma unroll
    for (int h = 0; h < kMlpH; h++)
#pragma unroll
      for (int i = 0; i < kMlpH; i++) d1[i] += bf2f(m->W2[h][i]) * dz2[h];
#pragma unroll
    for (int i = 0; i < kMlpH; i++) bufB[l][i] = h1[i] > 0.0f ? d1[i] : 0.0f;
    __syncthreads();
    for (int q = 0; q < 64; q++) {  // W1 row l, b1[l]
      const float d = bufB[q][l];
      gb1 += d;
      for (int i = 0; i < 4; i++) gW1[i] += d * xs[q][i];
    }
    __syncthreads();
  }
  float* g = a.grad;
  for (int i = 0; i < 4; i++) atomicAdd(&g[kOffW1 + l * kMlpIn + i], gW1[i]);
  atomicAdd(&g[kOffB1 + l], gb1);
#pragma unroll
  for (int i = 0; i < kMlpH; i++) atomicAdd(&g[kOffW2 + l * kMlpH + i], gW2[i]);
  atomicAdd(&g[kOffB2 + l], gb2);
  atomicAdd(&g[kOffW3 + l], gw3);
  if (l == 0) atomicAdd(&g[kOffB3], gb3);
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if (l == 0) atomicAdd(a.loss, lsum * a.lscale);
}

__global__ void __launch_bounds__(256) mlp_fill_kernel(float* __restrict__ p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0f;
}

// Adam (Kingma & Ba 2015) on the trainable entries of P (all but the
// standardisation constants), bias-corrected, step t >= 1.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ g,
                                                       float* __restrict__ m1, float* __restrict__ m2, float lr,
                                                       float c1, float c2) {
  const int t = kOffW1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kMlpParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[t];
  const float u = b1 * m1[t] + (1.0f - b1) * gi;
  const float v = b2 * m2[t] + (1.0f - b2) * gi * gi;
  m1[t] = u;
  m2[t] = v;
  P[t] -= lr * (u / c1) / (sqrtf(v / c2) + eps);
}

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* alloc(size_t n) {
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) p = nullptr;
    return (T*)p;
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() {  // s stays null on failure, which the callers check
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

int hip_ok(hipError_t e, const char* what, std::string& err) {
  if (e == hipSuccess) return 0;
  err = std::string(what) + ": " + hipGetErrorString(e);
  return 1;
}

void pack_on(hipStream_t s, const float* P, MlpDev* m) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(16), dim3(256), 0, s, P, m);
}

void grad_on(hipStream_t s, const MlpDev* m, const float* feat, const float* target, const int* idx, long first,
             long count, float lscale, float* grad, float* loss) {
  GradArgs a;
  a.m = m; a.feat = feat; a.target = target; a.idx = idx; a.first = first; a.count = count;
  a.gscale = 2.0f / (float)count;
  a.lscale = lscale;
  a.grad = grad; a.loss = loss;
  const long nchunk = (count + 63) / 64;
  const int grid = (int)std::min<long>(nchunk, 2048);
  hipLaunchKernelGGL(mlp_fill_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, grad, kMlpParams);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(grid), dim3(64), 0, s, a);
}

}  // namespace

int mlp_upload(Ctx& c, const float* blob, long n) {
  if (n != kMlpParams) return 1;
  MlpDev h;
  mlp_pack(blob, &h);
  if (!c.mlp_dev && hipMalloc(&c.mlp_dev, sizeof(MlpDev)) != hipSuccess) {
    c.mlp_dev = nullptr;
    return 1;
  }
  return hipMemcpy(c.mlp_dev, &h, sizeof(MlpDev), hipMemcpyHostToDevice) != hipSuccess;
}

namespace {
constexpr int kScratchDevices = 64;
void* g_scratch_dev[kScratchDevices] = {};  // device address of `scratch` on each GPU
std::mutex g_scratch_mu;
}  // namespace

// The address of this GPU's `scratch`, looked up once per device with that
// device current (the host symbol is only a stand-in for the device array).
int mlp_scratch_address(void** ptr, std::string& err) {
  int dev = 0;
  if (hip_ok(hipGetDevice(&dev), "hipGetDevice", err)) return 1;
  if (dev < 0 || dev >= kScratchDevices) { err = "device index out of range for the scratch table"; return 1; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)", err)) return 1;
    g_scratch_dev[dev] = p;
  }
  *ptr = g_scratch_dev[dev];
  return 0;
}

// cgfdm_mlp_forward with the caller's device arrays and the static scratch:
// the same pack and forward kernels, queued on `s`, no allocation.
int mlp_forward_noalloc_host(const float* P, const float* feat, long n, float* out, int mode, hipStream_t s,
                             std::string& err) {
  void* sp = nullptr;
  if (mlp_scratch_address(&sp, err)) return 1;
  MlpDev* m = (MlpDev*)sp;
  pack_on(s, P, m);
  if (n > 0) hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, m, feat, n, out, mode);
  return hip_ok(hipGetLastError(), "mlp forward launch", err);
}

int mlp_forward_host(const float* blob, const float* feat, long n, float* out, int mode, std::string& err) {
  Stream st;
  DevBuf bP, bm, bf, bo;
  float* P = bP.alloc<float>(kMlpParams);
  MlpDev* m = bm.alloc<MlpDev>(1);
  float* f = bf.alloc<float>(4 * n);
  float* o = bo.alloc<float>(n);
  if (!st.s || !P || !m || !f || !o) { err = "hipMalloc failed"; return 1; }
  if (hip_ok(hipMemcpyAsync(P, blob, kMlpParams * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err)) return 1;
  if (n > 0 && hip_ok(hipMemcpyAsync(f, feat, 4 * n * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err)) return 1;
  pack_on(st.s, P, m);
  if (n > 0) hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st.s, m, f, n, o, mode);
  if (n > 0 && hip_ok(hipMemcpyAsync(out, o, n * sizeof(float), hipMemcpyDeviceToHost, st.s), "copy", err)) return 1;
  return hip_ok(hipStreamSynchronize(st.s), "mlp forward", err);
}

int mlp_grad_host(const float* blob, const float* feat, const float* target, long n, float* grad, float* loss,
                  std::string& err) {
  Stream st;
  DevBuf bP, bm, bf, bt, bg, bl;
  float* P = bP.alloc<float>(kMlpParams);
  MlpDev* m = bm.alloc<MlpDev>(1);
  float* f = bf.alloc<float>(4 * n);
  float* t = bt.alloc<float>(n);
  float* g = bg.alloc<float>(kMlpParams);
  float* l = bl.alloc<float>(1);
  if (!st.s || !P || !m || !f || !t || !g || !l) { err = "hipMalloc failed"; return 1; }
  if (hip_ok(hipMemcpyAsync(P, blob, kMlpParams * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err) ||
      hip_ok(hipMemcpyAsync(f, feat, 4 * n * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err) ||
      hip_ok(hipMemcpyAsync(t, target, n * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err))
    return 1;
  pack_on(st.s, P, m);
  hipLaunchKernelGGL(mlp_fill_kernel, dim3(1), dim3(64), 0, st.s, l, 1);
  grad_on(st.s, m, f, t, nullptr, 0, n, 1.0f / (float)n, g, l);
  if (hip_ok(hipMemcpyAsync(grad, g, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, st.s), "copy", err) ||
      hip_ok(hipMemcpyAsync(loss, l, sizeof(float), hipMemcpyDeviceToHost, st.s), "copy", err))
    return 1;
  return hip_ok(hipStreamSynchronize(st.s), "mlp grad", err);
}

int mlp_train_host(float* blob, const float* feat, const float* target, long n, int epochs, int batch, float lr,
                   unsigned seed, float* loss_hist, std::string& err) {
  Stream st;
  DevBuf bP, bm, bf, bt, bg, bl, bm1, bm2, bi;
  float* P = bP.alloc<float>(kMlpParams);
  MlpDev* m = bm.alloc<MlpDev>(1);
  float* f = bf.alloc<float>(4 * n);
  float* t = bt.alloc<float>(n);
  float* g = bg.alloc<float>(kMlpParams);
  float* l = bl.alloc<float>(epochs);
  float* m1 = bm1.alloc<float>(kMlpParams);
  float* m2 = bm2.alloc<float>(kMlpParams);
  int* idx = bi.alloc<int>(n);
  if (!st.s || !P || !m || !f || !t || !g || !l || !m1 || !m2 || !idx) { err = "hipMalloc failed"; return 1; }
  if (hip_ok(hipMemcpyAsync(P, blob, kMlpParams * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err) ||
      hip_ok(hipMemcpyAsync(f, feat, 4 * n * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err) ||
      hip_ok(hipMemcpyAsync(t, target, n * sizeof(float), hipMemcpyHostToDevice, st.s), "copy", err))
    return 1;
  const int fb = (kMlpParams + 255) / 256;
  hipLaunchKernelGGL(mlp_fill_kernel, dim3(fb), dim3(256), 0, st.s, m1, kMlpParams);
  hipLaunchKernelGGL(mlp_fill_kernel, dim3(fb), dim3(256), 0, st.s, m2, kMlpParams);
  hipLaunchKernelGGL(mlp_fill_kernel, dim3((epochs + 255) / 256), dim3(256), 0, st.s, l, epochs);
  std::mt19937 rng(seed);
  std::vector<int> perm(n);
  for (long i = 0; i < n; i++) perm[i] = (int)i;
  long step = 0;
  const double beta1 = 0.9, beta2 = 0.999;  // as in mlp_adam_kernel
  for (int e = 0; e < epochs; e++) {
    // cosine decay of the step size from lr to lr/20 over the epochs
    const double ph = epochs > 1 ? (double)e / (epochs - 1) : 0.0;
    const float lr_e = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * ph))));
    std::shuffle(perm.begin(), perm.end(), rng);
    // the order buffer is rewritten only after the previous epoch's kernels have read it
    if (hip_ok(hipStreamSynchronize(st.s), "mlp train", err) ||
        hip_ok(hipMemcpyAsync(idx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, st.s), "copy", err))
      return 1;
    for (long first = 0; first < n; first += batch) {
      const long cnt = std::min<long>(batch, n - first);
      pack_on(st.s, P, m);
      grad_on(st.s, m, f, t, idx, first, cnt, 1.0f / (float)n, g, l + e);
      step++;
      const float c1 = (float)(1.0 - std::pow(beta1, (double)step)), c2 = (float)(1.0 - std::pow(beta2, (double)step));
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams - kOffW1 + 255) / 256), dim3(256), 0, st.s, P, g, m1, m2,
                         lr_e, c1, c2);
    }
  }
  if (hip_ok(hipGetLastError(), "mlp train launch", err) ||
      hip_ok(hipMemcpyAsync(blob, P, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, st.s), "copy", err) ||
      hip_ok(hipMemcpyAsync(loss_hist, l, epochs * sizeof(float), hipMemcpyDeviceToHost, st.s), "copy", err))
    return 1;
  return hip_ok(hipStreamSynchronize(st.s), "mlp train", err);
}

}  // namespace cgfdm

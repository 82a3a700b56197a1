// The learned part of the barostat: a 16 -> 64 -> 64 -> 1 tanh MLP that
// predicts the part of a volume move's energy change that the first-order
// (scaling-derivative) estimate misses. Inference and training run on
// MI355X matrix cores: every layer product is v_mfma_f32_16x16x32_bf16 (the
// 16-deep input layer v_mfma_f32_16x16x16_bf16) over BF16 operands staged in
// LDS, accumulating in FP32; biases, activations,
// the loss and Adam run in FP32 on FP32 master weights.
//
// No counterpart in the reference (its barostat evaluates every move
// exactly, webgpu/simulation.js:995-1070). The network only screens moves:
// the exact energy still makes the final accept (delayed acceptance, in
// context.hip).
//
// One workgroup (4 waves) works on a tile of 64 samples. Parameter layout,
// FP32, SURROGATE_PARAMS floats:
//   W1[16][64] (input-major), b1[64], W2[64][64] (input-major), b2[64], W3[64], b3.
#include "surrogate.h"

namespace rmd {

constexpr int TILE = 64;      // samples per workgroup pass
constexpr int THREADS = 256;  // 4 waves

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short toBf16(float f) {
  // round to nearest even (finite inputs)
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// One 16x16 output tile of A[M][K] * B[K][N] with A row-major (lda) and B
// given transposed, Bt[N][K] (ldb), both BF16 in LDS; K a multiple of 16.
// K in steps of 32 takes gfx950's v_mfma_f32_16x16x32_bf16 (lane l holds
// A[m0 + l%16][k + 8*(l/16) .. +7] and B[k + 8*(l/16) .. +7][n0 + l%16]); a
// remaining 16 takes v_mfma_f32_16x16x16_bf16 (4 elements per lane at
// k + 4*(l/16)). Both leave D[m0 + 4*(l/16) + i][n0 + l%16] in element i of
// the accumulator.
__device__ __forceinline__ f32x4 mfmaTile(const unsigned short *A, int lda, const unsigned short *Bt, int ldb, int K,
                                          int m0, int n0) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  int k = 0;
  for (; k + 32 <= K; k += 32) {
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(A + (m0 + r) * lda + k + 8 * q);
    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Bt + (n0 + r) * ldb + k + 8 * q);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  for (; k < K; k += 16) {
    const bf16x4 a = *reinterpret_cast<const bf16x4 *>(A + (m0 + r) * lda + k + 4 * q);
    const bf16x4 b = *reinterpret_cast<const bf16x4 *>(Bt + (n0 + r) * ldb + k + 4 * q);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// (Every BF16 operand array starts on a 16-byte boundary: the matrix
// products read it 8 or 16 bytes per lane.)
struct Lds {
  // weights (BF16 operand copies) and FP32 biases / last layer
  alignas(16) unsigned short w1t[64 * 16];    // [out][in]
  alignas(16) unsigned short w2t[64 * 64];    // [out][in]
  alignas(16) unsigned short w2n[64 * 64];    // [in][out]
  float b1[64], b2[64], w3[64], b3;
  // activations for one tile
  alignas(16) unsigned short x[TILE * 16];    // [sample][in]
  alignas(16) unsigned short xt[16 * TILE];   // [in][sample]
  alignas(16) unsigned short a1[TILE * 64];   // [sample][unit]
  alignas(16) unsigned short a1t[64 * TILE];  // [unit][sample]
  float a1f[TILE * 64];
  float a2f[TILE * 64];
  float y[TILE];
  float dy[TILE];
  alignas(16) unsigned short dz2[TILE * 64];  // [sample][unit]
  alignas(16) unsigned short dz2t[64 * TILE]; // [unit][sample]
  alignas(16) unsigned short dz1t[64 * TILE]; // [unit][sample]
  float grad[SURROGATE_PARAMS];
  float lossPart[THREADS / 64];
};

__device__ void loadWeights(Lds &s, const float *p) {
  const float *W1 = p, *B1 = p + 1024, *W2 = p + 1088, *B2 = p + 5184, *W3 = p + 5248;
  for (int k = threadIdx.x; k < 1024; k += blockDim.x) {
    const int i = k / 64, o = k % 64;
    s.w1t[o * 16 + i] = toBf16(W1[k]);
  }
  for (int k = threadIdx.x; k < 4096; k += blockDim.x) {
    const int i = k / 64, o = k % 64;
    const unsigned short v = toBf16(W2[k]);
    s.w2t[o * 64 + i] = v;
    s.w2n[k] = v;
  }
  for (int k = threadIdx.x; k < 64; k += blockDim.x) {
    s.b1[k] = B1[k];
    s.b2[k] = B2[k];
    s.w3[k] = W3[k];
  }
  if (threadIdx.x == 0) s.b3 = p[5312];
}

// Loads samples [first, first + count) of x (16 floats each) into the tile;
// rows past `count` are zero.
__device__ void loadInputs(Lds &s, const float *x, int first, int count) {
  for (int k = threadIdx.x; k < TILE * 16; k += blockDim.x) {
    const int b = k / 16, i = k % 16;
    const unsigned short v = b < count ? toBf16(x[(size_t)(first + b) * 16 + i]) : (unsigned short)0;
    s.x[k] = v;
    s.xt[i * TILE + b] = v;
  }
}

// Forward pass of one tile: a1, a2 (FP32 and BF16 copies) and y.
__device__ void forward(Lds &s) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, row4 = 4 * (lane >> 4);
  // layer 1: [64 x 16] x [16 x 64]; 16 tiles, 4 per wave
  for (int t = wave; t < 16; t += 4) {
    const int m0 = 16 * (t / 4), n0 = 16 * (t % 4);
    const f32x4 acc = mfmaTile(s.x, 16, s.w1t, 16, 16, m0, n0);
    for (int i = 0; i < 4; ++i) {
      const int b = m0 + row4 + i, o = n0 + col;
      const float a = tanhf(acc[i] + s.b1[o]);
      s.a1f[b * 64 + o] = a;
      const unsigned short h = toBf16(a);
      s.a1[b * 64 + o] = h;
      s.a1t[o * TILE + b] = h;
    }
  }
  __syncthreads();
  // layer 2: [64 x 64] x [64 x 64]
  for (int t = wave; t < 16; t += 4) {
    const int m0 = 16 * (t / 4), n0 = 16 * (t % 4);
    const f32x4 acc = mfmaTile(s.a1, 64, s.w2t, 64, 64, m0, n0);
    for (int i = 0; i < 4; ++i) {
      const int b = m0 + row4 + i, o = n0 + col;
      s.a2f[b * 64 + o] = tanhf(acc[i] + s.b2[o]);
    }
  }
  __syncthreads();
  // layer 3: one output per sample, FP32 (a 64-long dot product)
  for (int b = threadIdx.x; b < TILE; b += blockDim.x) {
    float y = s.b3;
    for (int o = 0; o < 64; ++o) y += s.a2f[b * 64 + o] * s.w3[o];
    s.y[b] = y;
  }
  __syncthreads();
}

// Backward pass of one tile for the loss sum_b (y_b - t_b)^2 / n over the
// first `count` rows, accumulating parameter gradients into s.grad (already
// zeroed) and returning the tile's loss contribution (thread 0).
__device__ float backward(Lds &s, const float *target, int first, int count, int n) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, row4 = 4 * (lane >> 4);
  float loss = 0.f;
  for (int b = threadIdx.x; b < TILE; b += blockDim.x) {
    float d = 0.f;
    if (b < count) {
      const float e = s.y[b] - target[first + b];
      loss += e * e / n;
      d = 2.f * e / n;
    }
    s.dy[b] = d;
  }
  loss = waveSum(loss);
  if (lane == 0) s.lossPart[wave] = loss;
  __syncthreads();
  float *gW1 = s.grad, *gB1 = s.grad + 1024, *gW2 = s.grad + 1088, *gB2 = s.grad + 5184, *gW3 = s.grad + 5248;
  // dZ2 = dy * w3 * (1 - a2^2); gW3 = a2^T dy; gb3 = sum dy
  for (int k = threadIdx.x; k < TILE * 64; k += blockDim.x) {
    const int b = k / 64, o = k % 64;
    const float a = s.a2f[k];
    const unsigned short h = toBf16(s.dy[b] * s.w3[o] * (1.f - a * a));
    s.dz2[k] = h;
    s.dz2t[o * TILE + b] = h;
  }
  for (int o = threadIdx.x; o < 64; o += blockDim.x) {
    float g = 0.f;
    for (int b = 0; b < TILE; ++b) g += s.a2f[b * 64 + o] * s.dy[b];
    gW3[o] += g;
  }
  if (threadIdx.x == 0) {
    float g = 0.f;
    for (int b = 0; b < TILE; ++b) g += s.dy[b];
    s.grad[5312] += g;
  }
  __syncthreads();
  // gW2[i][o] = sum_b a1[b][i] dz2[b][o]: A = a1t [i][b], Bt = dz2t [o][b]
  for (int t = wave; t < 16; t += 4) {
    const int m0 = 16 * (t / 4), n0 = 16 * (t % 4);
    const f32x4 acc = mfmaTile(s.a1t, TILE, s.dz2t, TILE, TILE, m0, n0);
    for (int i = 0; i < 4; ++i) gW2[(m0 + row4 + i) * 64 + n0 + col] += acc[i];
  }
  // gb2 = sum_b dz2 (from the BF16 values the products used)
  for (int o = threadIdx.x; o < 64; o += blockDim.x) {
    float g = 0.f;
    for (int b = 0; b < TILE; ++b) g += __uint_as_float((unsigned int)s.dz2[b * 64 + o] << 16);
    gB2[o] += g;
  }
  // dA1[b][i] = sum_o dz2[b][o] w2[i][o]: A = dz2 [b][o], Bt = w2n [i][o];
  // dZ1 = dA1 * (1 - a1^2) stored [unit][sample]
  for (int t = wave; t < 16; t += 4) {
    const int m0 = 16 * (t / 4), n0 = 16 * (t % 4);
    const f32x4 acc = mfmaTile(s.dz2, 64, s.w2n, 64, 64, m0, n0);
    for (int i = 0; i < 4; ++i) {
      const int b = m0 + row4 + i, u = n0 + col;
      const float a = s.a1f[b * 64 + u];
      s.dz1t[u * TILE + b] = toBf16(acc[i] * (1.f - a * a));
    }
  }
  __syncthreads();
  // gW1[j][i] = sum_b x[b][j] dz1[b][i]: A = xt [j][b], Bt = dz1t [i][b];
  // one 16-wide column tile per wave
  {
    const int n0 = 16 * wave;
    const f32x4 acc = mfmaTile(s.xt, TILE, s.dz1t, TILE, TILE, 0, n0);
    for (int i = 0; i < 4; ++i) gW1[(row4 + i) * 64 + n0 + col] += acc[i];
  }
  for (int u = threadIdx.x; u < 64; u += blockDim.x) {
    float g = 0.f;
    for (int b = 0; b < TILE; ++b) g += __uint_as_float((unsigned int)s.dz1t[u * TILE + b] << 16);
    gB1[u] += g;
  }
  __syncthreads();
  float total = 0.f;
  for (int w = 0; w < THREADS / 64; ++w) total += s.lossPart[w];
  return total;
}

// y = f(x) for n samples, one workgroup per tile of 64.
__global__ void __launch_bounds__(THREADS) surrogateForwardKernel(const float *params, const float *x, int n, float *y) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  Lds &s = *reinterpret_cast<Lds *>(smem);
  const int first = blockIdx.x * TILE;
  const int count = min(TILE, n - first);
  loadWeights(s, params);
  loadInputs(s, x, first, count);
  __syncthreads();
  forward(s);
  for (int b = threadIdx.x; b < count; b += blockDim.x) y[first + b] = s.y[b];
}

// Loss and gradient over all n samples (one workgroup walks the tiles).
__global__ void __launch_bounds__(THREADS) surrogateGradientKernel(const float *params, const float *x, const float *t,
                                                                   int n, float *loss, float *grad) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  Lds &s = *reinterpret_cast<Lds *>(smem);
  loadWeights(s, params);
  for (int k = threadIdx.x; k < SURROGATE_PARAMS; k += blockDim.x) s.grad[k] = 0.f;
  float total = 0.f;
  for (int first = 0; first < n; first += TILE) {
    const int count = min(TILE, n - first);
    __syncthreads();
    loadInputs(s, x, first, count);
    __syncthreads();
    forward(s);
    total += backward(s, t, first, count, n);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < SURROGATE_PARAMS; k += blockDim.x) grad[k] = s.grad[k];
  if (threadIdx.x == 0) *loss = total;
}

__device__ __forceinline__ unsigned int hashIndex(unsigned int v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}

// `steps` Adam steps on mini-batches of 64 samples drawn from the n samples
// with a counter hash (seed, step); params, m and v (FP32, in global memory)
// are updated in place, and the mean loss of the last 16 steps is returned.
__global__ void __launch_bounds__(THREADS) surrogateTrainKernel(float *params, float *m, float *v, const float *x,
                                                                const float *t, int n, int steps, float lr,
                                                                unsigned int seed, int stepOffset, float *loss) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  Lds &s = *reinterpret_cast<Lds *>(smem);
  __shared__ int pick[TILE];
  __shared__ float batchX[TILE * 16];
  __shared__ float batchT[TILE];
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  float recent = 0.f;
  for (int step = 0; step < steps; ++step) {
    __syncthreads();
    loadWeights(s, params);
    for (int k = threadIdx.x; k < SURROGATE_PARAMS; k += blockDim.x) s.grad[k] = 0.f;
    for (int b = threadIdx.x; b < TILE; b += blockDim.x)
      pick[b] = (int)(hashIndex(seed ^ hashIndex((unsigned)(stepOffset + step) * 0x9e3779b1u + b)) % (unsigned)n);
    __syncthreads();
    for (int k = threadIdx.x; k < TILE * 16; k += blockDim.x) batchX[k] = x[(size_t)pick[k / 16] * 16 + k % 16];
    for (int b = threadIdx.x; b < TILE; b += blockDim.x) batchT[b] = t[pick[b]];
    __syncthreads();
    loadInputs(s, batchX, 0, TILE);
    __syncthreads();
    forward(s);
    const float l = backward(s, batchT, 0, TILE, TILE);
    if (step >= steps - 16) recent += l;
    // Adam on the FP32 masters (bias-corrected)
    const int it = stepOffset + step + 1;
    const float c1 = 1.f - powf(beta1, (float)it), c2 = 1.f - powf(beta2, (float)it);
    for (int k = threadIdx.x; k < SURROGATE_PARAMS; k += blockDim.x) {
      const float g = s.grad[k];
      const float mk = beta1 * m[k] + (1.f - beta1) * g;
      const float vk = beta2 * v[k] + (1.f - beta2) * g * g;
      m[k] = mk;
      v[k] = vk;
      params[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    }
    __threadfence_block();
  }
  if (threadIdx.x == 0) *loss = recent / (float)(steps < 16 ? steps : 16);
}

static size_t ldsBytes() { return sizeof(Lds); }

void surrogateConfigure() {
  (void)hipFuncSetAttribute((const void *)surrogateForwardKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)ldsBytes());
  (void)hipFuncSetAttribute((const void *)surrogateGradientKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)ldsBytes());
  (void)hipFuncSetAttribute((const void *)surrogateTrainKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)ldsBytes());
}

void launchSurrogateForward(const float *params, const float *x, int n, float *y, hipStream_t s) {
  if (n <= 0) return;
  surrogateForwardKernel<<<(n + TILE - 1) / TILE, THREADS, ldsBytes(), s>>>(params, x, n, y);
}

void launchSurrogateGradient(const float *params, const float *x, const float *t, int n, float *loss, float *grad,
                             hipStream_t s) {
  surrogateGradientKernel<<<1, THREADS, ldsBytes(), s>>>(params, x, t, n, loss, grad);
}

void launchSurrogateTrain(float *params, float *m, float *v, const float *x, const float *t, int n, int steps, float lr,
                          unsigned int seed, int stepOffset, float *loss, hipStream_t s) {
  if (n <= 0 || steps <= 0) return;
  surrogateTrainKernel<<<1, THREADS, ldsBytes(), s>>>(params, m, v, x, t, n, steps, lr, seed, stepOffset, loss);
}

}  // namespace rmd

// ============================================================================
// The standalone network behind the C ABI (rmd_surrogate_*): FP32 master
// parameters and Adam moments on one device, with the kernels above.
// ============================================================================

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "host.h"

#define SURROGATE_HIP(call)                                                                        \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw rmd::Error(RMD_E_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(e_));   \
  } while (0)

// The staging of rmd_surrogate_forward_noalloc: the inputs (16 floats per
// sample) followed by the outputs of up to RMD_SURROGATE_NOALLOC_SAMPLES
// samples, 1.1 MiB that each device the library runs on holds for the life of
// the process. Host code never passes this name to a kernel or a copy: it is
// a host-side stand-in there. It uses the device address hipGetSymbolAddress
// returns for the current device (noallocScratchFor below).
static __device__ __align__(256) float scratch[RMD_SURROGATE_NOALLOC_SAMPLES * (RMD_SURROGATE_INPUTS + 1)];

struct rmd_surrogate {
  int device = 0;
  hipStream_t stream = nullptr;
  float *params = nullptr, *m = nullptr, *v = nullptr;
  float *scalar = nullptr;  // loss
  float *grad = nullptr;
  float *smallX = nullptr, *smallY = nullptr;  // inference buffers for up to 64 samples
  int adamSteps = 0;
  uint32_t seed = 0;
};

namespace rmd {

// Uniform initial weights scaled by sqrt(6 / (fan_in + fan_out)), from the
// counter hash; biases zero.
std::vector<float> surrogateInitialParams(uint32_t seed) {
  std::vector<float> p(SURROGATE_PARAMS, 0.f);
  auto fill = [&](int offset, int fanIn, int fanOut, uint32_t stream) {
    const double scale = std::sqrt(6.0 / (fanIn + fanOut));
    for (int k = 0; k < fanIn * fanOut; ++k)
      p[offset + k] = (float)(scale * (2.0 * uniformAt(seed, (uint32_t)k, stream, 0x1ea7u) - 1.0));
  };
  fill(0, 16, 64, 1);
  fill(1088, 64, 64, 2);
  fill(5248, 64, 1, 3);
  return p;
}

namespace {

template <typename T>
T *deviceCopy(const T *src, size_t n, hipStream_t s) {
  T *d = nullptr;
  SURROGATE_HIP(hipMalloc(&d, std::max<size_t>(1, n) * sizeof(T)));
  if (n) SURROGATE_HIP(hipMemcpyAsync(d, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

// The device address of `scratch` on `device` (which must be current),
// looked up once per device and kept.
float *noallocScratchFor(int device) {
  constexpr int MAX_DEVICES = 64;
  static std::mutex lock;
  static float *address[MAX_DEVICES] = {nullptr};
  if (device < 0 || device >= MAX_DEVICES) throw Error(RMD_E_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> hold(lock);
  if (!address[device]) {
    void *p = nullptr;
    SURROGATE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(RMD_E_DEVICE, "hipGetSymbolAddress returned no address for the staging array");
    address[device] = static_cast<float *>(p);
  }
  return address[device];
}

void surrogateConfigureOnce() {
  static bool done = false;
  if (!done) {
    surrogateConfigure();
    done = true;
  }
}

}  // namespace
}  // namespace rmd

#define NET_GUARD(net, ...)                                    \
  do {                                                         \
    if (!(net)) {                                              \
      rmd::setGlobalError("null network");                     \
      return RMD_E_INVALID;                                    \
    }                                                          \
    try {                                                      \
      SURROGATE_HIP(hipSetDevice((net)->device));              \
      __VA_ARGS__;                                             \
      return RMD_OK;                                           \
    } catch (const rmd::Error &e) {                            \
      rmd::setGlobalError(e.what());                           \
      return e.code;                                           \
    } catch (const std::exception &e) {                        \
      rmd::setGlobalError(e.what());                           \
      return RMD_E_INVALID;                                    \
    }                                                          \
  } while (0)

extern "C" {

int rmd_surrogate_create(int32_t device, uint32_t seed, rmd_surrogate **out) {
  if (!out) return RMD_E_INVALID;
  rmd_surrogate *net = new rmd_surrogate();
  net->device = device;
  net->seed = seed;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device) {
    delete net;
    rmd::setGlobalError("no HIP device is available; this library runs on an MI355X (gfx950)");
    return RMD_E_DEVICE;
  }
  const int status = [&]() -> int {
    NET_GUARD(net, {
      rmd::surrogateConfigureOnce();
      SURROGATE_HIP(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
      const std::vector<float> p = rmd::surrogateInitialParams(seed);
      const std::vector<float> zero(rmd::SURROGATE_PARAMS, 0.f);
      net->params = rmd::deviceCopy(p.data(), p.size(), net->stream);
      net->m = rmd::deviceCopy(zero.data(), zero.size(), net->stream);
      net->v = rmd::deviceCopy(zero.data(), zero.size(), net->stream);
      net->grad = rmd::deviceCopy(zero.data(), zero.size(), net->stream);
      net->scalar = rmd::deviceCopy(zero.data(), 4, net->stream);
      net->smallX = rmd::deviceCopy(zero.data(), 64 * 16, net->stream);
      net->smallY = rmd::deviceCopy(zero.data(), 64, net->stream);
      SURROGATE_HIP(hipStreamSynchronize(net->stream));
      rmd::noallocScratchFor(device);
    });
  }();
  if (status != RMD_OK) {
    rmd_surrogate_free(net);
    return status;
  }
  *out = net;
  return RMD_OK;
}

void rmd_surrogate_free(rmd_surrogate *net) {
  if (!net) return;
  (void)hipSetDevice(net->device);
  for (float *p : {net->params, net->m, net->v, net->grad, net->scalar, net->smallX, net->smallY})
    if (p) (void)hipFree(p);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

int rmd_surrogate_get_params(rmd_surrogate *net, float *params) {
  NET_GUARD(net, {
    SURROGATE_HIP(hipMemcpyAsync(params, net->params, rmd::SURROGATE_PARAMS * sizeof(float), hipMemcpyDeviceToHost,
                                 net->stream));
    SURROGATE_HIP(hipStreamSynchronize(net->stream));
  });
}

int rmd_surrogate_set_params(rmd_surrogate *net, const float *params) {
  NET_GUARD(net, {
    SURROGATE_HIP(hipMemcpyAsync(net->params, params, rmd::SURROGATE_PARAMS * sizeof(float), hipMemcpyHostToDevice,
                                 net->stream));
    SURROGATE_HIP(hipMemsetAsync(net->m, 0, rmd::SURROGATE_PARAMS * sizeof(float), net->stream));
    SURROGATE_HIP(hipMemsetAsync(net->v, 0, rmd::SURROGATE_PARAMS * sizeof(float), net->stream));
    net->adamSteps = 0;
    SURROGATE_HIP(hipStreamSynchronize(net->stream));
  });
}

int rmd_surrogate_forward(rmd_surrogate *net, const float *x, int32_t n, float *y) {
  NET_GUARD(net, {
    if (n < 0 || (n > 0 && (!x || !y))) throw rmd::Error(RMD_E_INVALID, "bad arguments");
    if (n == 0) return RMD_OK;
    const bool small = n <= 64;
    float *dx = small ? net->smallX : nullptr, *dy = small ? net->smallY : nullptr;
    if (small) {
      SURROGATE_HIP(hipMemcpyAsync(dx, x, (size_t)n * 16 * sizeof(float), hipMemcpyHostToDevice, net->stream));
    } else {
      dx = rmd::deviceCopy(x, (size_t)n * 16, net->stream);
      SURROGATE_HIP(hipMalloc(&dy, (size_t)n * sizeof(float)));
    }
    rmd::launchSurrogateForward(net->params, dx, n, dy, net->stream);
    SURROGATE_HIP(hipGetLastError());
    SURROGATE_HIP(hipMemcpyAsync(y, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    SURROGATE_HIP(hipStreamSynchronize(net->stream));
    if (!small) {
      (void)hipFree(dx);
      (void)hipFree(dy);
    }
  });
}

// rmd_surrogate_forward with the device's static staging array in place of
// allocated buffers: the same kernel on the same inputs, so the same outputs.
int rmd_surrogate_forward_noalloc(rmd_surrogate *net, const float *x, int32_t n, float *y) {
  NET_GUARD(net, {
    if (n < 0 || (n > 0 && (!x || !y))) throw rmd::Error(RMD_E_INVALID, "bad arguments");
    if (n > RMD_SURROGATE_NOALLOC_SAMPLES)
      throw rmd::Error(RMD_E_INVALID, "rmd_surrogate_forward_noalloc takes at most " +
                                          std::to_string(RMD_SURROGATE_NOALLOC_SAMPLES) + " samples per call (" +
                                          std::to_string(n) + " given); split the batch or use rmd_surrogate_forward");
    if (n == 0) return RMD_OK;
    float *dx = rmd::noallocScratchFor(net->device);
    float *dy = dx + (size_t)RMD_SURROGATE_NOALLOC_SAMPLES * RMD_SURROGATE_INPUTS;
    SURROGATE_HIP(hipMemcpyAsync(dx, x, (size_t)n * RMD_SURROGATE_INPUTS * sizeof(float), hipMemcpyHostToDevice,
                                 net->stream));
    rmd::launchSurrogateForward(net->params, dx, n, dy, net->stream);
    SURROGATE_HIP(hipGetLastError());
    SURROGATE_HIP(hipMemcpyAsync(y, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    SURROGATE_HIP(hipStreamSynchronize(net->stream));
  });
}

int rmd_surrogate_gradient(rmd_surrogate *net, const float *x, const float *t, int32_t n, float *loss, float *grad) {
  NET_GUARD(net, {
    if (n <= 0 || !x || !t || !loss || !grad) throw rmd::Error(RMD_E_INVALID, "bad arguments");
    float *dx = rmd::deviceCopy(x, (size_t)n * 16, net->stream);
    float *dt = rmd::deviceCopy(t, (size_t)n, net->stream);
    rmd::launchSurrogateGradient(net->params, dx, dt, n, net->scalar, net->grad, net->stream);
    SURROGATE_HIP(hipGetLastError());
    SURROGATE_HIP(hipMemcpyAsync(grad, net->grad, rmd::SURROGATE_PARAMS * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    SURROGATE_HIP(hipMemcpyAsync(loss, net->scalar, sizeof(float), hipMemcpyDeviceToHost, net->stream));
    SURROGATE_HIP(hipStreamSynchronize(net->stream));
    (void)hipFree(dx);
    (void)hipFree(dt);
  });
}

int rmd_surrogate_train(rmd_surrogate *net, const float *x, const float *t, int32_t n, int32_t steps,
                        float learning_rate, float *loss) {
  NET_GUARD(net, {
    if (n <= 0 || steps <= 0 || !x || !t) throw rmd::Error(RMD_E_INVALID, "bad arguments");
    float *dx = rmd::deviceCopy(x, (size_t)n * 16, net->stream);
    float *dt = rmd::deviceCopy(t, (size_t)n, net->stream);
    rmd::launchSurrogateTrain(net->params, net->m, net->v, dx, dt, n, steps, learning_rate, net->seed,
                              net->adamSteps, net->scalar, net->stream);
    SURROGATE_HIP(hipGetLastError());
    net->adamSteps += steps;
    float l = 0.f;
    SURROGATE_HIP(hipMemcpyAsync(&l, net->scalar, sizeof(float), hipMemcpyDeviceToHost, net->stream));
    SURROGATE_HIP(hipStreamSynchronize(net->stream));
    if (loss) *loss = l;
    (void)hipFree(dx);
    (void)hipFree(dt);
  });
}

}  // extern "C"

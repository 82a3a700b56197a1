// Device-side helpers shared by every kernel file: fixed-point force
// accumulation, the FP32 minimum-image distance test (in a fixed operation
// order, without contraction, so a CPU emulation reproduces it bit for bit),
// an FP32 erfc/erf, and wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rmd {

// Forces are accumulated as int64 at 2^32 units per kJ/mol/nm: order
// independent, so every run gives the same bits.
constexpr double FORCE_SCALE = 4294967296.0;
constexpr double FORCE_UNSCALE = 1.0 / 4294967296.0;
// PME charge grid: int64 at 2^32 units per e.
constexpr double GRID_SCALE = 4294967296.0;
constexpr double GRID_UNSCALE = 1.0 / 4294967296.0;

constexpr float ONE_4PI_EPS0_F = 138.935456f;
constexpr float TWO_OVER_SQRT_PI_F = 1.1283791670955126f;

// Device geometry for one box, kept in device memory so a captured block of
// steps reads the current box: f32 rows, per-axis inverse widths (the image
// test multiplies by them), the f32 reciprocal rows for fractional
// coordinates, the squared cutoffs, the FP64 box and reciprocal rows, and the
// cell-grid shape.
struct BoxF {
  float a[3], b[3], c[3];
  float inv[3];
  float recip[9];
  float rc2, rl2;
  float cutoff;
  int periodic;
  double box64[9];
  double recip64[9];
  int cells[3];
  int cellTotal;
  int reach;  // stencil cells either side (cells are at least rl / reach wide)
  int cellSearch;  // the list search runs one wave per cell (else one per atom)
};

__device__ __forceinline__ long long toFixed(double v) { return static_cast<long long>(__double2ll_rn(v * FORCE_SCALE)); }

__device__ __forceinline__ void atomicAddFixed(long long *address, long long value) {
  atomicAdd(reinterpret_cast<unsigned long long *>(address), static_cast<unsigned long long>(value));
}

// Minimum image and squared distance. The order of operations is fixed and
// contraction is off, so oracle_r2_f32 reproduces every bit: image along c,
// then b, then a (the reference's kernel order,
// webgpu/kernels/wgsl-common.js:81-88), with s the nearest integer to
// d * inv (one rounding instruction; the reference's floor(d / L + 0.5)
// differs only where d / L is within rounding of a half-integer, where both
// images lie half a box away, beyond the list cutoff).
__device__ __forceinline__ float minimumImageR2(float &dx, float &dy, float &dz, const BoxF &b, float &s0, float &s1,
                                                float &s2) {
#pragma clang fp contract(off)
  s2 = rintf(dz * b.inv[2]);
  dx = dx - s2 * b.c[0];
  dy = dy - s2 * b.c[1];
  dz = dz - s2 * b.c[2];
  s1 = rintf(dy * b.inv[1]);
  dx = dx - s1 * b.b[0];
  dy = dy - s1 * b.b[1];
  s0 = rintf(dx * b.inv[0]);
  dx = dx - s0 * b.a[0];
  float xx = dx * dx;
  float yy = dy * dy;
  float zz = dz * dz;
  float sum = xx + yy;
  return sum + zz;
}

// The same minimum image for a rectangular box: the terms of the general form
// that multiply an off-diagonal row entry are products with exact zeros and
// leave every bit as it is, so they are left out.
__device__ __forceinline__ float minimumImageR2Rect(float &dx, float &dy, float &dz, const BoxF &b, float &s0,
                                                    float &s1, float &s2) {
#pragma clang fp contract(off)
  s2 = rintf(dz * b.inv[2]);
  dz = dz - s2 * b.c[2];
  s1 = rintf(dy * b.inv[1]);
  dy = dy - s1 * b.b[1];
  s0 = rintf(dx * b.inv[0]);
  dx = dx - s0 * b.a[0];
  float xx = dx * dx;
  float yy = dy * dy;
  float zz = dz * dz;
  float sum = xx + yy;
  return sum + zz;
}

template <bool RECT>
__device__ __forceinline__ float imageR2(float &dx, float &dy, float &dz, const BoxF &b, float &s0, float &s1,
                                         float &s2) {
  return RECT ? minimumImageR2Rect(dx, dy, dz, b, s0, s1, s2) : minimumImageR2(dx, dy, dz, b, s0, s1, s2);
}

__device__ __forceinline__ float minimumImageR2(float &dx, float &dy, float &dz, const BoxF &b) {
  float s0, s1, s2;
  return minimumImageR2(dx, dy, dz, b, s0, s1, s2);
}

// The squared distance of the same image in FP64: the FP32 box rows differ
// from the FP64 box by up to half an ulp, a bias every boundary-crossing pair
// shares, so the energies take the separation from the positions and the
// FP64 rows with the images the FP32 test chose.
template <bool RECT>
__device__ __forceinline__ double imageR2Double(float4 pi, float4 pj, const BoxF &b, float s0, float s1, float s2) {
  const double *r = b.box64;
  if (RECT) {  // the same sums without the off-diagonal rows' zero terms
    const double ex = ((double)pi.x - (double)pj.x) - s0 * r[0];
    const double ey = ((double)pi.y - (double)pj.y) - s1 * r[4];
    const double ez = ((double)pi.z - (double)pj.z) - s2 * r[8];
    return ex * ex + ey * ey + ez * ez;
  }
  const double ex = ((double)pi.x - (double)pj.x) - s2 * r[6] - s1 * r[3] - s0 * r[0];
  const double ey = ((double)pi.y - (double)pj.y) - s2 * r[7] - s1 * r[4] - s0 * r[1];
  const double ez = ((double)pi.z - (double)pj.z) - s2 * r[8] - s1 * r[5] - s0 * r[2];
  return ex * ex + ey * ey + ez * ez;
}

// A pointer into global memory typed as such: loads through it are
// global_load (in-order completion counter) rather than flat loads, whose
// completion the compiler can only wait for as a whole. For pointers the
// compiler cannot place itself, such as those read from a view in memory.
// (The host pass of the same source, which never runs device code, sees a
// plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using GlobalPtr = const __attribute__((address_space(1))) T *;
template <typename T>
using GlobalOut = __attribute__((address_space(1))) T *;
#else
template <typename T>
using GlobalPtr = const T *;
template <typename T>
using GlobalOut = T *;
#endif
template <typename T>
__device__ __forceinline__ GlobalPtr<T> globalPtr(const T *p) {
  return (GlobalPtr<T>)p;
}
template <typename T>
__device__ __forceinline__ GlobalOut<T> globalOut(T *p) {
  return (GlobalOut<T>)p;
}

// A raw buffer over `bytes` bytes at p, for buffer loads with 32-bit byte
// offsets: the descriptor lives in scalar registers (the pointer must be the
// same in every lane), and a load past `bytes` (or at a negative offset, taken
// as unsigned) returns zero instead of touching memory.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bufferOf(const T *p, unsigned int bytes) {
  const unsigned long long a = (unsigned long long)p;
  const unsigned int lo = __builtin_amdgcn_readfirstlane((unsigned int)a);
  const unsigned int hi = __builtin_amdgcn_readfirstlane((unsigned int)(a >> 32));
  void *base = (void *)(((unsigned long long)hi << 32) | lo);
  constexpr int RAW_BUFFER_FLAGS = 0x00020000;  // the gfx9 raw-buffer format word
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes),
                                           RAW_BUFFER_FLAGS);
}

// A value the same in every lane of the wave (read from lane 0 or the first
// active lane), held in a scalar register.
__device__ __forceinline__ float uniformFloat(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniformDouble(double x) {
  const long long bits = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(bits & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(bits >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The fields of a box the pair tests read (rows, inverse lengths, rc^2,
// rl^2, and with `rows64` the FP64 rows), in scalar registers: a box is
// uniform over a workgroup. Call with every lane active.
__device__ __forceinline__ BoxF uniformPairBox(const BoxF &g, bool rows64) {
  BoxF b;
  for (int k = 0; k < 3; ++k) {
    b.a[k] = uniformFloat(g.a[k]);
    b.b[k] = uniformFloat(g.b[k]);
    b.c[k] = uniformFloat(g.c[k]);
    b.inv[k] = uniformFloat(g.inv[k]);
  }
  b.rc2 = uniformFloat(g.rc2);
  b.rl2 = uniformFloat(g.rl2);
  if (rows64)
    for (int k = 0; k < 9; ++k) b.box64[k] = uniformDouble(g.box64[k]);
  return b;
}

// erfc(x) for x >= 0 as exp(-x^2) * t * P(t), t = 1/(1 + x/2). P is a
// degree-8 relative-error minimax fit of erfcx(x)/t over x in [0, 4] made for
// this file by erfc_fit.py (max relative fit error 1.2e-8; the f32
// evaluation's mean error -3e-9). x^2 is split into hi + lo with an FMA so
// the exponential's argument is exact.
// Also returns exp(-x^2), which the force needs.
__device__ __forceinline__ float erfcWithGaussian(float x, float &gaussian) {
  const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, x, 1.0f));  // 1 ulp
  float p = -4.261279106e-02f;
  p = fmaf(p, t, 1.989041418e-01f);
  p = fmaf(p, t, -2.959853113e-01f);
  p = fmaf(p, t, 3.935875744e-02f);
  p = fmaf(p, t, 1.554778367e-01f);
  p = fmaf(p, t, 1.158371866e-01f);
  p = fmaf(p, t, 2.683745921e-01f);
  p = fmaf(p, t, 2.782751322e-01f);
  p = fmaf(p, t, 2.823704481e-01f);
  const float hi = x * x;
  const float lo = fmaf(x, x, -hi);
  const float e = expf(-hi) * (1.0f - lo);
  gaussian = e;
  return (e * t) * p;
}

// erf(x) for x >= 0: an odd polynomial x*Q(x^2) below 0.8 (a degree-5
// relative-error fit in x^2 made for this file by erfc_fit.py, max fit
// error 3e-9), and 1 - erfc above.
__device__ __forceinline__ float erfWithGaussian(float x, float &gaussian) {
  if (x < 0.8f) {
    const float x2 = x * x;
    float s = -6.535268622e-04f;
    s = fmaf(s, x2, 5.085136276e-03f);
    s = fmaf(s, x2, -2.682081051e-02f);
    s = fmaf(s, x2, 1.128311306e-01f);
    s = fmaf(s, x2, -3.761260211e-01f);
    s = fmaf(s, x2, 1.128379107e+00f);
    const float lo = fmaf(x, x, -x2);
    gaussian = expf(-x2) * (1.0f - lo);
    return x * s;
  }
  float g;
  const float c = erfcWithGaussian(x, g);
  gaussian = g;
  return 1.0f - c;
}

// Wave64 reductions.
template <typename T>
__device__ __forceinline__ T waveSum(T v) {
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}
__device__ __forceinline__ float waveMax(float v) {
  for (int offset = 32; offset > 0; offset >>= 1) v = fmaxf(v, __shfl_xor(v, offset, 64));
  return v;
}

}  // namespace rmd

// Smooth PME reciprocal space: order-5 B-spline charge spread, a real-to-
// complex 3D FFT done as line transforms in LDS, the convolution with the
// Ewald influence function (energy read off it in FP64), the inverse
// transform and the force gather.
//
// Replaces encodePme (webgpu/kernels/pme.js:137-463) and encodeFft3d
// (webgpu/kernels/pme-fft.js:42-197). Conventions are the reference's
// (cpu/forces/pme.js:11-15): the forward transform carries exp(+2 pi i m k/K),
// the B-spline moduli divide the convolution, grid index for weight k is
// (floor(u) + k - order + 1) mod K, and the inverse is unnormalised.
//
// MI355X shape: the charge grid is int64 fixed point (exact, order
// independent atomics in L2); the z transform is real-to-complex with two
// real lines packed in one complex transform; the convolution is fused into
// the x transform so the spectrum is read and written once per pass; grid
// sizes may carry a factor 7 (2^a 3^b 5^c 7^d).
#include "kernels.h"

namespace rmd {

constexpr int ORDER = 5;
constexpr int LINES_PER_BLOCK = 4;  // one wave64 per line

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__device__ __forceinline__ float2 twiddle(const float2 *tw, int index, int sign) {
  float2 w = tw[index];
  if (sign < 0) w.y = -w.y;
  return w;
}

// A Stockham autosort FFT of one line of length K held in LDS, performed by
// the `nt` threads of a group; `rad` lists the radices (0-terminated). Every
// group of the block must run the same K (the stage barrier is block wide).
// Returns the buffer holding the result.
__device__ float2 *fftLine(float2 *a, float2 *b, int K, const int *rad, const float2 *tw, int sign, int t, int nt) {
  float2 *src = a, *dst = b;
  int Ns = 1;
  for (int s = 0; s < 8 && rad[s] > 0; ++s) {
    const int R = rad[s];
    const int L = K / R;
    const int twStride = K / (Ns * R);
    for (int j = t; j < L; j += nt) {
      const int k = j % Ns;
      float2 v[7];
      for (int q = 0; q < R; ++q) v[q] = src[j + q * L];
      for (int q = 1; q < R; ++q) v[q] = cmul(v[q], twiddle(tw, (k * q * twStride) % K, sign));
      const int base = (j / Ns) * Ns * R + k;
      if (R == 2) {
        dst[base] = make_float2(v[0].x + v[1].x, v[0].y + v[1].y);
        dst[base + Ns] = make_float2(v[0].x - v[1].x, v[0].y - v[1].y);
      } else if (R == 4) {
        // radix-4 with the +/- i rotation chosen by the sign
        const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        // i*sign*d13
        const float2 r13 = sign > 0 ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
        dst[base] = make_float2(s02.x + s13.x, s02.y + s13.y);
        dst[base + Ns] = make_float2(d02.x + r13.x, d02.y + r13.y);
        dst[base + 2 * Ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
        dst[base + 3 * Ns] = make_float2(d02.x - r13.x, d02.y - r13.y);
      } else {
        const int step = K / R;
        for (int p = 0; p < R; ++p) {
          float2 acc = v[0];
          for (int q = 1; q < R; ++q) {
            const float2 w = twiddle(tw, (p * q * step) % K, sign);
            const float2 m = cmul(v[q], w);
            acc.x += m.x;
            acc.y += m.y;
          }
          dst[base + p * Ns] = acc;
        }
      }
    }
    __syncthreads();
    float2 *swap = src;
    src = dst;
    dst = swap;
    Ns *= R;
  }
  return src;
}

// Order-5 cardinal B-spline weights and derivatives at an offset in [0, 1)
// (shared/pme/parameters.js:121-149), in FP32.
__device__ __forceinline__ void bspline5(float w, float *m, float *dm) {
  m[0] = 1.f - w;
  m[1] = w;
  m[2] = 0.f;
  m[3] = 0.f;
  m[4] = 0.f;
  // k = 3
  m[2] = 0.5f * w * m[1];
  m[1] = 0.5f * ((w + 1.f) * m[0] + (2.f - w) * m[1]);
  m[0] = 0.5f * (1.f - w) * m[0];
  // k = 4
  const float third = 1.0f / 3.0f;
  m[3] = third * w * m[2];
  m[2] = third * ((w + 1.f) * m[1] + (3.f - w) * m[2]);
  m[1] = third * ((w + 2.f) * m[0] + (2.f - w) * m[1]);
  m[0] = third * (1.f - w) * m[0];
  dm[0] = -m[0];
  dm[1] = m[0] - m[1];
  dm[2] = m[1] - m[2];
  dm[3] = m[2] - m[3];
  dm[4] = m[3];
  // k = 5
  m[4] = 0.25f * w * m[3];
  m[3] = 0.25f * ((w + 1.f) * m[2] + (4.f - w) * m[3]);
  m[2] = 0.25f * ((w + 2.f) * m[1] + (3.f - w) * m[2]);
  m[1] = 0.25f * ((w + 3.f) * m[0] + (2.f - w) * m[1]);
  m[0] = 0.25f * (1.f - w) * m[0];
}

// Fractional coordinate along one axis in FP64 (the charge's position on the
// grid must not lose bits to a large coordinate), split into the base index
// and the FP32 offset.
__device__ __forceinline__ void gridCoordinate(const PmeDev &p, const double *recip, int d, float4 pos, int &base,
                                               float &offset) {
  const double s = (double)pos.x * recip[d] + (double)pos.y * recip[3 + d] + (double)pos.z * recip[6 + d];
  const double wrapped = s - floor(s);
  const double u = wrapped * p.K[d];
  int b = (int)floor(u);
  double off = u - b;
  if (b >= p.K[d]) b -= p.K[d];
  // A non-finite coordinate must not turn into an out-of-range index.
  if (!(b >= 0 && b < p.K[d])) {
    b = 0;
    off = 0.0;
  }
  base = b;
  offset = (float)off;
}

__global__ void __launch_bounds__(256) pmeSpreadKernel(Dev d, PmeDev p) {
  const int atom = blockIdx.x * blockDim.x + threadIdx.x;
  if (atom >= d.n) return;
  const float4 pos = d.pos[atom];
  const float q = pos.w;
  if (q == 0.f) return;
  int base[3];
  float w[3][ORDER], dw[ORDER];
  for (int a = 0; a < 3; ++a) {
    float off;
    gridCoordinate(p, d.box->recip64, a, pos, base[a], off);
    bspline5(off, w[a], dw);
  }
  int ix[ORDER], iy[ORDER], iz[ORDER];
  for (int k = 0; k < ORDER; ++k) {
    ix[k] = (base[0] + k - ORDER + 1 + p.K[0]) % p.K[0];
    iy[k] = (base[1] + k - ORDER + 1 + p.K[1]) % p.K[1];
    iz[k] = (base[2] + k - ORDER + 1 + p.K[2]) % p.K[2];
  }
  for (int a = 0; a < ORDER; ++a) {
    const float qa = q * w[0][a];
    for (int b = 0; b < ORDER; ++b) {
      const float qab = qa * w[1][b];
      long long *row = p.qgrid + ((size_t)ix[a] * p.K[1] + iy[b]) * p.K[2];
      for (int c = 0; c < ORDER; ++c) {
        const long long v = __double2ll_rn((double)(qab * w[2][c]) * GRID_SCALE);
        atomicAddFixed(row + iz[c], v);
      }
    }
  }
}

// z pass, forward: two real lines per complex transform; the fixed-point
// charges are read, converted and cleared for the next step.
__global__ void __launch_bounds__(256) pmeZForwardKernel(PmeDev p) {
  extern __shared__ float2 lds[];
  const int K = p.K[2];
  const int group = threadIdx.x / 64, t = threadIdx.x % 64;
  float2 *a = lds + group * 2 * K, *b = a + K;
  const int lines = p.K[0] * p.K[1];
  const int pair = blockIdx.x * LINES_PER_BLOCK + group;
  const int l0 = 2 * pair, l1 = 2 * pair + 1;
  const bool active = l0 < lines;
  if (active) {
    for (int k = t; k < K; k += 64) {
      const size_t i0 = (size_t)l0 * K + k;
      const float re = (float)((double)p.qgrid[i0] * GRID_UNSCALE);
      p.qgrid[i0] = 0;
      float im = 0.f;
      if (l1 < lines) {
        const size_t i1 = (size_t)l1 * K + k;
        im = (float)((double)p.qgrid[i1] * GRID_UNSCALE);
        p.qgrid[i1] = 0;
      }
      a[k] = make_float2(re, im);
    }
  }
  __syncthreads();
  float2 *r = fftLine(a, b, K, p.radices[2], p.twiddle[2], +1, t, 64);
  if (!active) return;
  for (int m = t; m < p.K2h; m += 64) {
    const float2 z = r[m];
    const float2 zc = r[(K - m) % K];
    // A = (Z[m] + conj Z[-m]) / 2,  B = (Z[m] - conj Z[-m]) / (2i)
    p.cgrid[(size_t)l0 * p.K2h + m] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
    if (l1 < lines) p.cgrid[(size_t)l1 * p.K2h + m] = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
  }
}

// y pass, either direction, in place on the half spectrum.
__global__ void __launch_bounds__(256) pmeYKernel(PmeDev p, int sign) {
  extern __shared__ float2 lds[];
  const int K = p.K[1];
  const int group = threadIdx.x / 64, t = threadIdx.x % 64;
  float2 *a = lds + group * 2 * K, *b = a + K;
  const int lines = p.K[0] * p.K2h;
  const int line = blockIdx.x * LINES_PER_BLOCK + group;
  const bool active = line < lines;
  const int x = line / p.K2h, m = line % p.K2h;
  const size_t base = (size_t)x * p.K[1] * p.K2h + m;
  if (active)
    for (int k = t; k < K; k += 64) a[k] = p.cgrid[base + (size_t)k * p.K2h];
  __syncthreads();
  float2 *r = fftLine(a, b, K, p.radices[1], p.twiddle[1], sign, t, 64);
  if (!active) return;
  for (int k = t; k < K; k += 64) p.cgrid[base + (size_t)k * p.K2h] = r[k];
}

// x pass: forward transform, convolution with the influence function (the
// reciprocal energy is summed here in FP64, each half-spectrum column
// weighted by its multiplicity in the full spectrum), inverse transform.
__global__ void __launch_bounds__(256) pmeXConvolveKernel(Dev d, PmeDev p, int energies) {
  extern __shared__ float2 lds[];
  const int K = p.K[0];
  const int group = threadIdx.x / 64, t = threadIdx.x % 64;
  float2 *a = lds + group * 2 * K, *b = a + K;
  const int lines = p.K[1] * p.K2h;
  const int line = blockIdx.x * LINES_PER_BLOCK + group;
  const bool active = line < lines;
  const int y = line / p.K2h, m = line % p.K2h;
  const size_t base = (size_t)y * p.K2h + m;
  const size_t stride = (size_t)p.K[1] * p.K2h;
  if (active)
    for (int k = t; k < K; k += 64) a[k] = p.cgrid[base + k * stride];
  __syncthreads();
  float2 *r = fftLine(a, b, K, p.radices[0], p.twiddle[0], +1, t, 64);
  double energy = 0.0;
  if (active) {
    const double weight = (m == 0 || (p.K[2] % 2 == 0 && m == p.K[2] / 2)) ? 1.0 : 2.0;
    for (int k = t; k < K; k += 64) {
      const float g = p.influence[base + k * stride];
      const float2 f = r[k];
      energy += weight * (double)g * ((double)f.x * f.x + (double)f.y * f.y);
      r[k] = make_float2(2.f * g * f.x, 2.f * g * f.y);
    }
  }
  __syncthreads();
  float2 *other = (r == a) ? b : a;
  float2 *s = fftLine(r, other, K, p.radices[0], p.twiddle[0], -1, t, 64);
  if (active)
    for (int k = t; k < K; k += 64) p.cgrid[base + k * stride] = s[k];
  if (energies) {
    __shared__ double partial[4];
    energy = waveSum(energy);
    if (t == 0) partial[group] = energy;
    __syncthreads();
    if (threadIdx.x == 0) d.slots[E_RECIPROCAL * MAX_SLOTS + blockIdx.x] = partial[0] + partial[1] + partial[2] + partial[3];
  }
}

// z pass, inverse: Hermitian extension of two half spectra into one complex
// line, transform, real and imaginary parts are the two real lines.
__global__ void __launch_bounds__(256) pmeZInverseKernel(PmeDev p) {
  extern __shared__ float2 lds[];
  const int K = p.K[2];
  const int group = threadIdx.x / 64, t = threadIdx.x % 64;
  float2 *a = lds + group * 2 * K, *b = a + K;
  const int lines = p.K[0] * p.K[1];
  const int pair = blockIdx.x * LINES_PER_BLOCK + group;
  const int l0 = 2 * pair, l1 = 2 * pair + 1;
  const bool active = l0 < lines;
  if (active) {
    for (int k = t; k < K; k += 64) {
      float2 A, B = make_float2(0.f, 0.f);
      if (k < p.K2h) {
        A = p.cgrid[(size_t)l0 * p.K2h + k];
        if (l1 < lines) B = p.cgrid[(size_t)l1 * p.K2h + k];
      } else {
        A = p.cgrid[(size_t)l0 * p.K2h + (K - k)];
        A.y = -A.y;
        if (l1 < lines) {
          B = p.cgrid[(size_t)l1 * p.K2h + (K - k)];
          B.y = -B.y;
        }
      }
      a[k] = make_float2(A.x - B.y, A.y + B.x);
    }
  }
  __syncthreads();
  float2 *r = fftLine(a, b, K, p.radices[2], p.twiddle[2], -1, t, 64);
  if (!active) return;
  for (int k = t; k < K; k += 64) {
    p.rgrid[(size_t)l0 * K + k] = r[k].x;
    if (l1 < lines) p.rgrid[(size_t)l1 * K + k] = r[k].y;
  }
}

// Force gather (webgpu/kernels/pme.js:281-336; cpu/forces/pme.js:222-257):
// dE/dr = q * sum (dW/dr) * convolved grid, chained through the reciprocal
// rows.
__global__ void __launch_bounds__(256) pmeGatherKernel(Dev d, PmeDev p) {
  const int atom = blockIdx.x * blockDim.x + threadIdx.x;
  if (atom >= d.n) return;
  const float4 pos = d.pos[atom];
  if (pos.w == 0.f) return;
  int base[3];
  float w[3][ORDER], dw[3][ORDER];
  for (int a = 0; a < 3; ++a) {
    float off;
    gridCoordinate(p, d.box->recip64, a, pos, base[a], off);
    bspline5(off, w[a], dw[a]);
  }
  int ix[ORDER], iy[ORDER], iz[ORDER];
  for (int k = 0; k < ORDER; ++k) {
    ix[k] = (base[0] + k - ORDER + 1 + p.K[0]) % p.K[0];
    iy[k] = (base[1] + k - ORDER + 1 + p.K[1]) % p.K[1];
    iz[k] = (base[2] + k - ORDER + 1 + p.K[2]) % p.K[2];
  }
  float gu = 0.f, gv = 0.f, gw = 0.f;
  for (int a = 0; a < ORDER; ++a) {
    for (int b = 0; b < ORDER; ++b) {
      const float *row = p.rgrid + ((size_t)ix[a] * p.K[1] + iy[b]) * p.K[2];
      const float wxy = w[0][a] * w[1][b], dxy = dw[0][a] * w[1][b], xdy = w[0][a] * dw[1][b];
      for (int c = 0; c < ORDER; ++c) {
        const float v = row[iz[c]];
        gu += dxy * w[2][c] * v;
        gv += xdy * w[2][c] * v;
        gw += wxy * dw[2][c] * v;
      }
    }
  }
  const double q = pos.w;
  const double u = (double)gu * p.K[0] * q, v = (double)gv * p.K[1] * q, ww = (double)gw * p.K[2] * q;
  const double *r = d.box->recip64;
  const double fx = -(u * r[0] + v * r[1] + ww * r[2]);
  const double fy = -(u * r[3] + v * r[4] + ww * r[5]);
  const double fz = -(u * r[6] + v * r[7] + ww * r[8]);
  atomicAddFixed(&d.force[atom], __double2ll_rn(fx * FORCE_SCALE));
  atomicAddFixed(&d.force[d.nPad + atom], __double2ll_rn(fy * FORCE_SCALE));
  atomicAddFixed(&d.force[2 * d.nPad + atom], __double2ll_rn(fz * FORCE_SCALE));
}

// The influence function over the half spectrum, in FP64, stored as f32:
// pref * exp(-pi^2 m^2 / alpha^2) / (m^2 * B(m)) with the reference's row
// convention for the wave vector (cpu/forces/pme.js:176-209).
__global__ void pmeInfluenceKernel(PmeDev p, const BoxF *box, const double *mod0, const double *mod1,
                                   const double *mod2, double alpha) {
  const double *recip = box->recip64;
  const double *bx = box->box64;
  const double volume = bx[0] * bx[4] * bx[8];
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)p.K[0] * p.K[1] * p.K2h;
  if (idx >= total) return;
  const int m = idx % p.K2h;
  const int y = (idx / p.K2h) % p.K[1];
  const int x = idx / ((size_t)p.K2h * p.K[1]);
  const int kx = x <= p.K[0] / 2 ? x : x - p.K[0];
  const int ky = y <= p.K[1] / 2 ? y : y - p.K[1];
  const int kz = m;
  if (kx == 0 && ky == 0 && kz == 0) {
    p.influence[idx] = 0.f;
    return;
  }
  const double mx = kx * recip[0] + ky * recip[1] + kz * recip[2];
  const double my = kx * recip[3] + ky * recip[4] + kz * recip[5];
  const double mz = kx * recip[6] + ky * recip[7] + kz * recip[8];
  const double m2 = mx * mx + my * my + mz * mz;
  const double factor = M_PI * M_PI / (alpha * alpha);
  const double pref = 138.935456 / (2.0 * M_PI * volume);
  const double modulus = mod0[x] * mod1[y] * mod2[m];
  p.influence[idx] = (float)(pref * exp(-factor * m2) / (m2 * modulus));
}

void pmeComputeInfluence(const PmeDev &p, const BoxF *box, const double *moduli[3], double alpha, hipStream_t s) {
  const size_t total = (size_t)p.K[0] * p.K[1] * p.K2h;
  pmeInfluenceKernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(p, box, moduli[0], moduli[1], moduli[2], alpha);
}

static size_t ldsBytes(int K) { return (size_t)LINES_PER_BLOCK * 2 * K * sizeof(float2); }

int launchPme(const Dev &d, const PmeDev &p, int energies, hipStream_t s, int *slotsUsed) {
  const int atomBlocks = (d.n + 255) / 256;
  pmeSpreadKernel<<<atomBlocks, 256, 0, s>>>(d, p);
  const int zPairs = (p.K[0] * p.K[1] + 1) / 2;
  const int zBlocks = (zPairs + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeZForwardKernel<<<zBlocks, 256, ldsBytes(p.K[2]), s>>>(p);
  const int yBlocks = (p.K[0] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeYKernel<<<yBlocks, 256, ldsBytes(p.K[1]), s>>>(p, +1);
  const int xBlocks = (p.K[1] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeXConvolveKernel<<<xBlocks, 256, ldsBytes(p.K[0]), s>>>(d, p, energies);
  pmeYKernel<<<yBlocks, 256, ldsBytes(p.K[1]), s>>>(p, -1);
  pmeZInverseKernel<<<zBlocks, 256, ldsBytes(p.K[2]), s>>>(p);
  pmeGatherKernel<<<atomBlocks, 256, 0, s>>>(d, p);
  if (slotsUsed) *slotsUsed = xBlocks;
  return xBlocks;
}

// Debug/verification entry: the forward R2C 3D transform of a real grid.
__global__ void realToFixedKernel(PmeDev p, const float *real) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)p.K[0] * p.K[1] * p.K[2]) return;
  p.qgrid[i] = __double2ll_rn((double)real[i] * GRID_SCALE);
}

__global__ void copyComplexKernel(PmeDev p, float2 *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)p.K[0] * p.K[1] * p.K2h) return;
  out[i] = p.cgrid[i];
}

// Forward x transform alone (the verification path has no convolution).
__global__ void __launch_bounds__(256) pmeXForwardKernel(PmeDev p) {
  extern __shared__ float2 lds[];
  const int K = p.K[0];
  const int group = threadIdx.x / 64, t = threadIdx.x % 64;
  float2 *a = lds + group * 2 * K, *b = a + K;
  const int lines = p.K[1] * p.K2h;
  const int line = blockIdx.x * LINES_PER_BLOCK + group;
  const bool active = line < lines;
  const size_t base = (size_t)(line / p.K2h) * p.K2h + line % p.K2h;
  const size_t stride = (size_t)p.K[1] * p.K2h;
  if (active)
    for (int k = t; k < K; k += 64) a[k] = p.cgrid[base + k * stride];
  __syncthreads();
  float2 *r = fftLine(a, b, K, p.radices[0], p.twiddle[0], +1, t, 64);
  if (!active) return;
  for (int k = t; k < K; k += 64) p.cgrid[base + k * stride] = r[k];
}

void launchFftForward(const PmeDev &p, const float *real, float2 *out, hipStream_t s) {
  const size_t n = (size_t)p.K[0] * p.K[1] * p.K[2];
  realToFixedKernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, real);
  const int zPairs = (p.K[0] * p.K[1] + 1) / 2;
  pmeZForwardKernel<<<(zPairs + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK, 256, ldsBytes(p.K[2]), s>>>(p);
  const int yBlocks = (p.K[0] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeYKernel<<<yBlocks, 256, ldsBytes(p.K[1]), s>>>(p, +1);
  const int xBlocks = (p.K[1] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeXForwardKernel<<<xBlocks, 256, ldsBytes(p.K[0]), s>>>(p);
  const size_t nc = (size_t)p.K[0] * p.K[1] * p.K2h;
  copyComplexKernel<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(p, out);
}

}  // namespace rmd

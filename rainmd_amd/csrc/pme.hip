// Smooth PME reciprocal space: order-5 B-spline charge spread, a real-to-
// complex 3D FFT done in LDS, the convolution with the Ewald influence
// function (energy read off it in FP64), the inverse transform and the force
// gather.
//
// Replaces encodePme (webgpu/kernels/pme.js:137-463) and encodeFft3d
// (webgpu/kernels/pme-fft.js:42-197). Conventions are the reference's
// (cpu/forces/pme.js:11-15): the forward transform carries exp(+2 pi i m k/K),
// the B-spline moduli divide the convolution, grid index for weight k is
// (floor(u) + k - order + 1) mod K, and the inverse is unnormalised.
//
// MI355X shape. The default "plane" pipeline is three launches:
//   1. one workgroup per x plane gathers the charges that land on its plane
//      into an int64 fixed-point plane in LDS (exact and order independent,
//      no global atomics), then runs the z (real-to-complex, two real lines
//      per complex transform) and y transforms of that plane in LDS;
//   2. x lines: forward transform, convolution with the influence function
//      (and the energy), inverse transform;
//   3. one workgroup per x plane runs the inverse y and z transforms into a
//      real potential plane in LDS and adds each atom's share of the force
//      from that plane.
// Grids whose plane does not fit in LDS use the "line" pipeline: a global
// int64 charge grid filled with atomics and one launch per axis pass.
// Grid sizes may carry a factor 7 (2^a 3^b 5^c 7^d).
#include "pme_plane.h"

namespace rmd {

// The atoms by the base index of their x support, rebuilt each evaluation by
// one workgroup per system (at most BIN_MAX_ATOMS atoms): each charged atom
// takes its rank in its bin by an LDS atomic count, the counts become
// offsets, and each atom is written at its offset plus rank. The order inside
// a bin varies from run to run; nothing depends on it (the spread
// accumulates in fixed point, the gather is per atom), so the binned planes
// give the scanned planes' results bit for bit.
constexpr int BIN_THREADS = 1024;
constexpr int BIN_MAX_K = 1024;  // grid sizes are at most 1024

__device__ __forceinline__ void pmeBinBuild(const Dev &d, const PmeDev &p) {
  __shared__ int start[BIN_MAX_K];
  __shared__ int waveTotal[BIN_THREADS / 64];
  const int K0 = p.K[0], t = threadIdx.x, nt = blockDim.x;
  for (int b = t; b < K0; b += nt) start[b] = 0;
  __syncthreads();
  const double *recip = d.box->recip64;
  for (int a = t; a < d.n; a += nt) {
    const float4 pos = d.pos[a];
    int2 e = make_int2(-1, 0);
    if (pos.w != 0.f) {
      int bx;
      float ox;
      gridCoordinate(p, recip, 0, pos, bx, ox);
      e = make_int2(bx, atomicAdd(&start[bx], 1));
    }
    p.atomBin[a] = e;
  }
  __syncthreads();
  // exclusive prefix over the counts, in place: a few bins per thread, then
  // the waves in order
  const int per = (K0 + nt - 1) / nt;
  const int lo = t * per, hi = min(K0, lo + per);
  int sum = 0;
  for (int b = lo; b < hi; ++b) sum += start[b];
  const int lane = t & 63, wave = t >> 6;
  int inclusive = sum;
  for (int offset = 1; offset < 64; offset <<= 1) {
    const int v = __shfl_up(inclusive, offset, 64);
    if (lane >= offset) inclusive += v;
  }
  if (lane == 63) waveTotal[wave] = inclusive;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += waveTotal[w];
  int offset = before + inclusive - sum;
  for (int b = lo; b < hi; ++b) {
    const int c = start[b];
    start[b] = offset;
    p.binStart[b] = offset;
    offset += c;
  }
  if (t == nt - 1) p.binStart[K0] = before + inclusive;
  __syncthreads();
  for (int a = t; a < d.n; a += nt) {
    const int2 e = p.atomBin[a];
    if (e.x >= 0) p.binAtoms[start[e.x] + e.y] = a;
  }
}

__global__ void __launch_bounds__(BIN_THREADS) pmeBinKernel(Dev d, PmeDev p) { pmeBinBuild(d, p); }
__global__ void __launch_bounds__(BIN_THREADS) pmeBinBatchKernel(const BatchView *__restrict__ v,
                                                                 const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0 || !v[t.x].p.binned) return;
  pmeBinBuild(v[t.x].d, v[t.x].p);
}

bool pmeBinsFit(const PmeDev &p, int atoms) {
  return p.planes && p.K[0] >= ORDER && p.K[0] <= BIN_MAX_K && atoms <= BIN_MAX_ATOMS;
}

bool pmeBinsFitWide(const PmeDev &p) { return p.planes && p.K[0] >= ORDER && p.K[0] <= BIN_MAX_K; }

// The same bins for a large single system, in three launches. 1: each group
// of BIN_GROUP_ATOMS atoms counts its charged atoms per bin in LDS (an atom's
// rank in its group's bin comes from the LDS atomic) and writes the counts,
// bin-major (entry b * binGroups + g).
constexpr int BIN_WIDE_THREADS = 256;

__global__ void __launch_bounds__(BIN_WIDE_THREADS) pmeBinCountKernel(Dev d, PmeDev p) {
  __shared__ int count[BIN_MAX_K];
  const int K0 = p.K[0], G = p.binGroups, t = threadIdx.x;
  for (int b = t; b < K0; b += BIN_WIDE_THREADS) count[b] = 0;
  __syncthreads();
  const double *recip = d.box->recip64;
  const int first = blockIdx.x * BIN_GROUP_ATOMS, last = min(d.n, first + BIN_GROUP_ATOMS);
  for (int a = first + t; a < last; a += BIN_WIDE_THREADS) {
    const float4 pos = d.pos[a];
    int2 e = make_int2(-1, 0);
    if (pos.w != 0.f) {
      int bx;
      float ox;
      gridCoordinate(p, recip, 0, pos, bx, ox);
      e = make_int2(bx, atomicAdd(&count[bx], 1));
    }
    p.atomBin[a] = e;
  }
  __syncthreads();
  for (int b = t; b < K0; b += BIN_WIDE_THREADS) p.binTable[(size_t)b * G + blockIdx.x] = count[b];
}

// 2: one workgroup turns the counts into offsets within each bin (the groups
// in order; a wave scans a bin's counts 64 at a time) and the bins' starts.
__global__ void __launch_bounds__(BIN_THREADS) pmeBinOffsetsKernel(PmeDev p) {
  __shared__ int total[BIN_MAX_K];
  __shared__ int waveTotal[BIN_THREADS / 64];
  const int K0 = p.K[0], G = p.binGroups, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, waves = BIN_THREADS / 64;
  for (int b = wave; b < K0; b += waves) {
    int *row = p.binTable + (size_t)b * G;
    int carry = 0;
    for (int g0 = 0; g0 < G; g0 += 64) {
      const int g = g0 + lane;
      const int v = g < G ? row[g] : 0;
      int inclusive = v;
      for (int offset = 1; offset < 64; offset <<= 1) {
        const int u = __shfl_up(inclusive, offset, 64);
        if (lane >= offset) inclusive += u;
      }
      if (g < G) row[g] = carry + inclusive - v;
      carry += __shfl(inclusive, 63, 64);
    }
    if (lane == 0) total[b] = carry;
  }
  __syncthreads();
  // the bins' starts: one bin per thread (K0 <= BIN_MAX_K = BIN_THREADS)
  const int sum = t < K0 ? total[t] : 0;
  int inclusive = sum;
  for (int offset = 1; offset < 64; offset <<= 1) {
    const int u = __shfl_up(inclusive, offset, 64);
    if (lane >= offset) inclusive += u;
  }
  if (lane == 63) waveTotal[wave] = inclusive;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += waveTotal[w];
  const int start = before + inclusive - sum;
  if (t < K0) p.binStart[t] = start;
  if (t == K0 - 1) p.binStart[K0] = start + sum;
}

// 3: each atom goes to its bin's start plus its group's offset in the bin
// plus its rank.
__global__ void __launch_bounds__(BIN_WIDE_THREADS) pmeBinScatterKernel(Dev d, PmeDev p) {
  const int G = p.binGroups;
  const int first = blockIdx.x * BIN_GROUP_ATOMS, last = min(d.n, first + BIN_GROUP_ATOMS);
  for (int a = first + threadIdx.x; a < last; a += BIN_WIDE_THREADS) {
    const int2 e = p.atomBin[a];
    if (e.x >= 0) p.binAtoms[p.binStart[e.x] + p.binTable[(size_t)e.x * G + blockIdx.x] + e.y] = a;
  }
}

// The bins of a binned system, by whichever build it was given.
static void launchPmeBins(const Dev &d, const PmeDev &p, hipStream_t s) {
  if (!p.binned) return;
  if (p.binGroups == 0) {
    pmeBinKernel<<<1, BIN_THREADS, 0, s>>>(d, p);
    return;
  }
  pmeBinCountKernel<<<p.binGroups, BIN_WIDE_THREADS, 0, s>>>(d, p);
  pmeBinOffsetsKernel<<<1, BIN_THREADS, 0, s>>>(p);
  pmeBinScatterKernel<<<p.binGroups, BIN_WIDE_THREADS, 0, s>>>(d, p);
}

__global__ void __launch_bounds__(PLANE_THREADS) pmePlaneForwardKernel(Dev d, PmeDev p, const float *real) {
  extern __shared__ float2 lds[];
  pmePlaneForward(d, p, real, blockIdx.x, lds);
}

__global__ void __launch_bounds__(PLANE_THREADS) pmePlaneForwardBatchKernel(const BatchView *__restrict__ v,
                                                                          const int2 *__restrict__ task) {
  extern __shared__ float2 lds[];
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  pmePlaneForward<SCAN_BATCHED>(v[t.x].d, v[t.x].p, nullptr, t.y, lds);
}

__global__ void __launch_bounds__(XLINE_THREADS) pmeXLinesKernel(Dev d, PmeDev p, int convolve, int energies) {
  extern __shared__ float2 lds[];
  pmeXLines(d, p, convolve, energies, blockIdx.x, lds);
}

__global__ void __launch_bounds__(XLINE_THREADS) pmeXLinesBatchKernel(const BatchView *__restrict__ v,
                                                                    const int2 *__restrict__ task, int energies) {
  extern __shared__ float2 lds[];
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  pmeXLines(v[t.x].d, v[t.x].p, 1, energies, t.y, lds);
}

// A large binned system's planes in workgroups of WIDE_PLANE_THREADS: its
// plane's LDS (153 KB at 112^3) leaves room for one workgroup per CU, and K0
// planes occupy fewer CUs than the chip has, so each plane's spread,
// transforms and gather get the CU's four SIMDs at four waves each. (SCAN 3
// keeps a round of the scan list within its LDS at this width.) Every single
// system gathers in such workgroups - K0 planes leave most CUs idle whatever
// the system's size, and a plane's atoms are gathered twice as wide (2854
// atoms, 30^3: 18.3 -> 17.1 us per chain). The results are the 512-thread
// kernels' bit for bit: the plane is an integer sum, each transform output
// and each atom's share of a plane has one formula, and forces are integer
// sums.
constexpr int WIDE_PLANE_THREADS = 1024;

__global__ void __launch_bounds__(WIDE_PLANE_THREADS) pmePlaneForwardWideKernel(Dev d, PmeDev p) {
  extern __shared__ float2 lds[];
  pmePlaneForward<PLANE_LIST / WIDE_PLANE_THREADS>(d, p, nullptr, blockIdx.x, lds);
}

__global__ void __launch_bounds__(WIDE_PLANE_THREADS) pmePlaneGatherKernel(Dev d, PmeDev p, int folded) {
  extern __shared__ float2 lds[];
  pmePlaneGather<PLANE_LIST / WIDE_PLANE_THREADS>(d, p, blockIdx.x, lds, folded);
}

static void launchPlaneForward(const Dev &d, const PmeDev &p, size_t lds, hipStream_t s) {
  if (p.binned)
    pmePlaneForwardWideKernel<<<p.K[0], WIDE_PLANE_THREADS, lds, s>>>(d, p);
  else
    pmePlaneForwardKernel<<<p.K[0], PLANE_THREADS, lds, s>>>(d, p, nullptr);
}

static void launchPlaneGather(const Dev &d, const PmeDev &p, size_t lds, hipStream_t s, int folded = 0) {
  pmePlaneGatherKernel<<<p.K[0], WIDE_PLANE_THREADS, lds, s>>>(d, p, folded);
}

// (Six waves per SIMD: three workgroups per CU, the register count the body
// needs without spills.)
__global__ void __launch_bounds__(PLANE_THREADS) __attribute__((amdgpu_waves_per_eu(6)))
pmePlaneGatherBatchKernel(const BatchView *__restrict__ v, const int2 *__restrict__ task) {
  extern __shared__ float2 lds[];
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  pmePlaneGather<SCAN_BATCHED>(v[t.x].d, v[t.x].p, t.y, lds);
}

// The gather of a binned system, one thread per charged atom in bin order (a
// wave's atoms share an x slab of the grid; the cell order is not used: the
// list rebuild may be rewriting it meanwhile on another stream): each of the
// atom's ORDER planes' shares is computed from the potential grid with the
// plane gather's own arithmetic and rounded to fixed point as it would be
// there, and the shares are summed in integers - one set of force atomics per
// atom instead of one per plane, and the same force bit for bit (the grid has
// at least ORDER planes, so each weight lands on a plane of its own).
__device__ __forceinline__ void pmeAtomGather(const Dev &d, const PmeDev &p, int k) {
  if (k >= p.binStart[p.K[0]]) return;
  const int a = p.binAtoms[k];
  const float4 pos = d.pos[a];
  const int K0 = p.K[0], K1 = p.K[1], K2 = p.K[2];
  const double *recip = d.box->recip64;
  int bx, by, bz;
  float ox, oy, oz;
  gridCoordinate(p, recip, 0, pos, bx, ox);
  float w[ORDER], dw[ORDER], wy[ORDER], dwy[ORDER], wz[ORDER], dwz[ORDER];
  bspline5(ox, w, dw);
  gridCoordinate(p, recip, 1, pos, by, oy);
  gridCoordinate(p, recip, 2, pos, bz, oz);
  bspline5(oy, wy, dwy);
  bspline5(oz, wz, dwz);
  int iz[ORDER], iy[ORDER];
  for (int c = 0; c < ORDER; ++c) {
    iz[c] = gridIndex(bz, c, K2);
    iy[c] = gridIndex(by, c, K1);
  }
  const GlobalPtr<float> grid = globalPtr(p.rgrid);
  long long total[3] = {0, 0, 0};
  for (int r = 0; r < ORDER; ++r) {
    const size_t plane = (size_t)gridIndex(bx, r, K0) * K1;
    const float wx = 0.f + w[r], dwx = 0.f + dw[r];
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int b = 0; b < ORDER; ++b) {
      const GlobalPtr<float> row = grid + (plane + iy[b]) * K2;
      float r0 = 0.f, r2 = 0.f;
      for (int c = 0; c < ORDER; ++c) {
        const float v = row[iz[c]];
        r0 += wz[c] * v;
        r2 += dwz[c] * v;
      }
      s0 += wy[b] * r0;
      s1 += dwy[b] * r0;
      s2 += wy[b] * r2;
    }
    long long f[3];
    gridForceFixed(d, p, pos.w, dwx * s0, wx * s1, wx * s2, f);
    for (int k = 0; k < 3; ++k) total[k] += f[k];
  }
  atomicAddFixed(&d.force[a], total[0]);
  atomicAddFixed(&d.force[d.nPad + a], total[1]);
  atomicAddFixed(&d.force[2 * d.nPad + a], total[2]);
}

__global__ void __launch_bounds__(256) pmeAtomGatherKernel(Dev d, PmeDev p) {
  pmeAtomGather(d, p, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) pmeAtomGatherBatchKernel(const BatchView *__restrict__ v,
                                                                const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0 || !v[t.x].p.binned) return;
  pmeAtomGather(v[t.x].d, v[t.x].p, t.y * blockDim.x + threadIdx.x);
}

// ============================================================================
// Line pipeline (grids whose plane does not fit in LDS)
// ============================================================================

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
    ix[k] = gridIndex(base[0], k, p.K[0]);
    iy[k] = gridIndex(base[1], k, p.K[1]);
    iz[k] = gridIndex(base[2], k, p.K[2]);
  }
  for (int a = 0; a < ORDER; ++a)
    for (int b = 0; b < ORDER; ++b) {
      long long *row = p.qgrid + ((size_t)ix[a] * p.K[1] + iy[b]) * p.K[2];
      for (int c = 0; c < ORDER; ++c) atomicAddFixed(row + iz[c], chargeToFixed(q, w[0][a], w[1][b], w[2][c]));
    }
}

// z pass, forward: two real lines per complex transform; the fixed-point
// charges are read, converted and cleared for the next step.
__global__ void __launch_bounds__(256) pmeZForwardKernel(PmeDev p) {
  extern __shared__ float2 lds[];
  const int K = p.K[2];
  float2 *tw = lds, *A = tw + K, *B = A + LINES_PER_BLOCK * K;
  loadTwiddles(tw, p.twiddle[2], K);
  const int lines = p.K[0] * p.K[1];
  const int pairs = (lines + 1) / 2;
  const int first = blockIdx.x * LINES_PER_BLOCK;
  const int count = min(LINES_PER_BLOCK, pairs - first);
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    const int l0 = 2 * (first + g), l1 = l0 + 1;
    const size_t i0 = (size_t)l0 * K + k;
    const float re = (float)((double)p.qgrid[i0] * GRID_UNSCALE);
    p.qgrid[i0] = 0;
    float im = 0.f;
    if (l1 < lines) {
      const size_t i1 = (size_t)l1 * K + k;
      im = (float)((double)p.qgrid[i1] * GRID_UNSCALE);
      p.qgrid[i1] = 0;
    }
    A[idx] = make_float2(re, im);
  }
  __syncthreads();
  float2 *r = fftBatch(A, B, count, K, p.radices[2], tw, +1);
  for (int idx = threadIdx.x; idx < count * p.K2h; idx += blockDim.x) {
    const int g = idx / p.K2h, m = idx - g * p.K2h;
    const int l0 = 2 * (first + g), l1 = l0 + 1;
    const float2 z = r[g * K + m], zc = r[g * K + (m == 0 ? 0 : K - m)];
    p.cgrid[(size_t)l0 * p.K2h + m] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
    if (l1 < lines) p.cgrid[(size_t)l1 * p.K2h + m] = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
  }
}

// y pass, either direction, in place on the half spectrum.
__global__ void __launch_bounds__(256) pmeYKernel(PmeDev p, int sign) {
  extern __shared__ float2 lds[];
  const int K = p.K[1];
  float2 *tw = lds, *A = tw + K, *B = A + LINES_PER_BLOCK * K;
  loadTwiddles(tw, p.twiddle[1], K);
  const int lines = p.K[0] * p.K2h;
  const int first = blockIdx.x * LINES_PER_BLOCK;
  const int count = min(LINES_PER_BLOCK, lines - first);
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    const int line = first + g, x = line / p.K2h, m = line - x * p.K2h;
    A[idx] = p.cgrid[(size_t)x * p.K[1] * p.K2h + m + (size_t)k * p.K2h];
  }
  __syncthreads();
  float2 *r = fftBatch(A, B, count, K, p.radices[1], tw, sign);
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    const int line = first + g, x = line / p.K2h, m = line - x * p.K2h;
    p.cgrid[(size_t)x * p.K[1] * p.K2h + m + (size_t)k * p.K2h] = r[idx];
  }
}

// x pass: forward transform, convolution with the influence function (energy
// in FP64), inverse transform.
__global__ void __launch_bounds__(256) pmeXConvolveKernel(Dev d, PmeDev p, int convolve, int energies) {
  extern __shared__ float2 lds[];
  const int K = p.K[0];
  float2 *tw = lds, *A = tw + K, *B = A + LINES_PER_BLOCK * K;
  loadTwiddles(tw, p.twiddle[0], K);
  const int lines = p.K[1] * p.K2h;
  const int first = blockIdx.x * LINES_PER_BLOCK;
  const int count = min(LINES_PER_BLOCK, lines - first);
  const size_t stride = (size_t)p.K[1] * p.K2h;
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    A[idx] = p.cgrid[(size_t)(first + g) + k * stride];
  }
  __syncthreads();
  float2 *r = fftBatch(A, B, count, K, p.radices[0], tw, +1);
  double energy = 0.0, dscale = 0.0;
  if (convolve) {
    const double *recip = d.box->recip64;
    for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
      const int g = idx / K, k = idx - g * K;
      const int line = first + g, y = line / p.K2h, m = line - y * p.K2h;
      const double weight = (m == 0 || (p.K[2] % 2 == 0 && m == p.K[2] / 2)) ? 1.0 : 2.0;
      const float gi = p.influence[(size_t)line + k * stride];
      const float2 f = r[idx];
      const double e = weight * (double)gi * ((double)f.x * f.x + (double)f.y * f.y);
      energy += e;
      if (energies) dscale += e * (2.0 * p.piOverAlpha2 * waveVector2(p, recip, k, y, m) - 1.0);
      r[idx] = make_float2(2.f * gi * f.x, 2.f * gi * f.y);
    }
    __syncthreads();
    r = fftBatch(r, r == A ? B : A, count, K, p.radices[0], tw, -1);
  }
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    p.cgrid[(size_t)(first + g) + k * stride] = r[idx];
  }
  if (convolve && energies) storeReciprocal(d, energy, dscale, blockIdx.x);
}

// z pass, inverse: Hermitian extension of two half spectra into one complex
// line, transform, real and imaginary parts are the two real lines.
__global__ void __launch_bounds__(256) pmeZInverseKernel(PmeDev p) {
  extern __shared__ float2 lds[];
  const int K = p.K[2];
  float2 *tw = lds, *A = tw + K, *B = A + LINES_PER_BLOCK * K;
  loadTwiddles(tw, p.twiddle[2], K);
  const int lines = p.K[0] * p.K[1];
  const int pairs = (lines + 1) / 2;
  const int first = blockIdx.x * LINES_PER_BLOCK;
  const int count = min(LINES_PER_BLOCK, pairs - first);
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    const int l0 = 2 * (first + g), l1 = l0 + 1;
    float2 a, b = make_float2(0.f, 0.f);
    if (k < p.K2h) {
      a = p.cgrid[(size_t)l0 * p.K2h + k];
      if (l1 < lines) b = p.cgrid[(size_t)l1 * p.K2h + k];
    } else {
      a = p.cgrid[(size_t)l0 * p.K2h + (K - k)];
      a.y = -a.y;
      if (l1 < lines) {
        b = p.cgrid[(size_t)l1 * p.K2h + (K - k)];
        b.y = -b.y;
      }
    }
    A[idx] = make_float2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  float2 *r = fftBatch(A, B, count, K, p.radices[2], tw, -1);
  for (int idx = threadIdx.x; idx < count * K; idx += blockDim.x) {
    const int g = idx / K, k = idx - g * K;
    const int l0 = 2 * (first + g), l1 = l0 + 1;
    p.rgrid[(size_t)l0 * K + k] = r[idx].x;
    if (l1 < lines) p.rgrid[(size_t)l1 * K + k] = r[idx].y;
  }
}

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
    ix[k] = gridIndex(base[0], k, p.K[0]);
    iy[k] = gridIndex(base[1], k, p.K[1]);
    iz[k] = gridIndex(base[2], k, p.K[2]);
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
  addGridForce(d, p, atom, pos.w, gu, gv, gw);
}

// The influence function over the half spectrum, in FP64, stored as f32:
// pref * exp(-pi^2 m^2 / alpha^2) / (m^2 * B(m)) with the reference's row
// convention for the wave vector (cpu/forces/pme.js:176-209).
__device__ __forceinline__ void pmeInfluence(const PmeDev &p, const BoxF *box, const double *mod0, const double *mod1,
                                             const double *mod2, double alpha, size_t idx) {
  const double *recip = box->recip64;
  const double *bx = box->box64;
  const double volume = bx[0] * bx[4] * bx[8];
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

// Batched: the replicas whose box changed.
__global__ void __launch_bounds__(256) pmeInfluenceBatchKernel(const BatchView *__restrict__ v,
                                                               const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0 || !v[t.x].boxChanged) return;
  const BatchView &b = v[t.x];
  pmeInfluence(b.p, b.d.box, b.moduli[0], b.moduli[1], b.moduli[2], b.alpha,
               (size_t)t.y * blockDim.x + threadIdx.x);
}

// The folded gathers' kernel X(l, c) = sum_k 2 g(k, c) exp(-2 pi i k l / K0)
// (PmeDev::xFolded), one thread per (l, c), summed in FP64 over ascending k,
// the roots from the x axis' twiddle table (exp(+2 pi i r / K0) at the
// numerator r = k l reduced modulo K0); for a rectangular box g is even in k
// and X real (the sine terms cancel), stored as its real part.
__global__ void pmeXKernelKernel(PmeDev p) {
  const int K0 = p.K[0], PK = p.K[1] * p.K2h;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)K0 * PK) return;
  const int l = (int)(idx / PK), c = (int)(idx - (size_t)l * PK);
  // the column's influence values loaded together (K0 <= X_FOLD_MAX_K), not
  // one dependent round trip per term
  float g[X_FOLD_MAX_K];
#pragma unroll
  for (int k = 0; k < X_FOLD_MAX_K; ++k) g[k] = k < K0 ? p.influence[(size_t)k * PK + c] : 0.f;
  double re = 0.0, im = 0.0;
  int r = 0;  // k l mod K0
#pragma unroll
  for (int k = 0; k < X_FOLD_MAX_K; ++k) {
    if (k >= K0) break;
    const float2 w = p.twiddle[0][r];
    re += 2.0 * (double)g[k] * (double)w.x;
    im -= 2.0 * (double)g[k] * (double)w.y;
    r += l;
    r -= r >= K0 ? K0 : 0;
  }
  if (p.xReal)
    p.xKernel[idx] = (float)re;
  else
    reinterpret_cast<float2 *>(p.xKernel)[idx] = make_float2((float)re, (float)im);
}

// A single system's box change in one launch: the box by value, stored to
// the device box by the first thread, and the influence function from it.
__global__ void pmeSetBoxAndInfluenceKernel(PmeDev p, BoxF box, BoxF *dst, const double *mod0, const double *mod1,
                                            const double *mod2, double alpha) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) *dst = box;
  pmeInfluence(p, &box, mod0, mod1, mod2, alpha, idx);
}

__global__ void pmeSetBoxAndInfluenceFromKernel(PmeDev p, const BoxF *box, BoxF *dst, const double *mod0,
                                                const double *mod1, const double *mod2, double alpha) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) *dst = *box;
  pmeInfluence(p, box, mod0, mod1, mod2, alpha, idx);
}

void pmeSetBoxAndInfluenceFrom(const PmeDev &p, const BoxF *box, BoxF *dst, const double *moduli[3], double alpha,
                               hipStream_t s) {
  const size_t total = (size_t)p.K[0] * p.K[1] * p.K2h;
  pmeSetBoxAndInfluenceFromKernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(p, box, dst, moduli[0], moduli[1],
                                                                                  moduli[2], alpha);
  if (p.xKernel) pmeXKernelKernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(p);
}

void pmeSetBoxAndInfluence(const PmeDev &p, const BoxF &box, BoxF *dst, const double *moduli[3], double alpha,
                           hipStream_t s) {
  const size_t total = (size_t)p.K[0] * p.K[1] * p.K2h;
  pmeSetBoxAndInfluenceKernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(p, box, dst, moduli[0], moduli[1],
                                                                              moduli[2], alpha);
  if (p.xKernel) pmeXKernelKernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(p);
}

// ---- configuration and launchers --------------------------------------------

// LDS: twiddles, two complex work buffers (the second also holds the int64
// charge plane), then the compacted atom list and its count.
static size_t listBytes() { return (size_t)PLANE_LIST * sizeof(PlaneEntry) + 4 * sizeof(int); }
size_t planeForwardLds(const PmeDev &p) {
  return (size_t)(p.K[1] + p.K[2] + p.planeSpan + p.planeRegion) * sizeof(float2) + listBytes();
}
size_t planeGatherLds(const PmeDev &p) {
  return (size_t)(p.K[1] + p.K[2]) * sizeof(float2) + 2 * (size_t)p.planeSpan * sizeof(float2) + listBytes();
}
size_t xLinesLds(const PmeDev &p) { return (size_t)p.K[0] * sizeof(float2) * (1 + 2 * XLINES); }
static size_t lineLds(int K) { return (size_t)K * sizeof(float2) * (1 + 2 * LINES_PER_BLOCK); }

bool pmeConfigure(PmeDev &p, int mode, size_t maxLds) {
  const int P = (p.K[1] + 1) / 2;
  const int a = P * p.K[2], b = p.K2h * p.K[1];
  p.planeSpan = a > b ? a : b;
  const int planeWords = (p.K[1] * p.K[2] * (int)sizeof(long long) + (int)sizeof(float2) - 1) / (int)sizeof(float2);
  p.planeRegion = p.planeSpan > planeWords ? p.planeSpan : planeWords;
  const size_t need = std::max(planeForwardLds(p), planeGatherLds(p));
  const bool fits = need <= maxLds && xLinesLds(p) <= maxLds;
  p.planes = (mode == PME_PLANES || mode == PME_AUTO) && fits ? 1 : 0;
  if (p.planes) {
    (void)hipFuncSetAttribute((const void *)pmePlaneForwardKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)planeForwardLds(p));
    (void)hipFuncSetAttribute((const void *)pmePlaneGatherKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)planeGatherLds(p));
    (void)hipFuncSetAttribute((const void *)pmePlaneForwardWideKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)planeForwardLds(p));
  }
  return mode != PME_PLANES || fits;
}

static int xLineBlocks(const PmeDev &p) { return p.K[1] * ((p.K2h + XLINES - 1) / XLINES); }
int pmeXLineBlocks(const PmeDev &p) { return xLineBlocks(p); }
// (A binned system's planes never use the scan list: its LDS is left out.)
size_t pmePlaneForwardLds(const PmeDev &p) { return planeForwardLds(p) - (p.binned ? listBytes() : 0); }
size_t pmeXLinesLds(const PmeDev &p) { return xLinesLds(p); }
size_t pmePlaneGatherLds(const PmeDev &p) { return planeGatherLds(p) - (p.binned ? listBytes() : 0); }

// Batched plane pipeline; the LDS sizes are the largest over the replicas and
// the kernels' LDS limits were raised to them at batch creation.
void launchPmeBatch(const BatchView *v, const TaskTable &planes, const TaskTable &xlines, const TaskTable &replicas,
                    const TaskTable &atoms256, int energies, size_t forwardLds, size_t xlinesLds, size_t gatherLds,
                    hipStream_t s) {
  pmeBinBatchKernel<<<replicas.blocks, BIN_THREADS, 0, s>>>(v, replicas.task);
  pmePlaneForwardBatchKernel<<<planes.blocks, PLANE_THREADS, forwardLds, s>>>(v, planes.task);
  pmeXLinesBatchKernel<<<xlines.blocks, XLINE_THREADS, xlinesLds, s>>>(v, xlines.task, energies);
  pmePlaneGatherBatchKernel<<<planes.blocks, PLANE_THREADS, gatherLds, s>>>(v, planes.task);
  pmeAtomGatherBatchKernel<<<atoms256.blocks, 256, 0, s>>>(v, atoms256.task);
}

void launchInfluenceBatch(const BatchView *v, const TaskTable &t, hipStream_t s) {
  pmeInfluenceBatchKernel<<<t.blocks, 256, 0, s>>>(v, t.task);
}

void pmeBatchLdsLimits(size_t forwardLds, size_t xlinesLds, size_t gatherLds) {
  (void)hipFuncSetAttribute((const void *)pmePlaneForwardBatchKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)forwardLds);
  (void)hipFuncSetAttribute((const void *)pmeXLinesBatchKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)xlinesLds);
  (void)hipFuncSetAttribute((const void *)pmePlaneGatherBatchKernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)gatherLds);
}

// The plane pipeline after its forward planes: x lines, then the plane
// gathers, or, for a force-only evaluation of a folding system, the gathers
// alone; returns the x lines' energy slots.
static int launchAfterForward(const Dev &d, const PmeDev &p, int energies, hipStream_t s) {
  const int xBlocks = xLineBlocks(p);
  const int folded = p.xFolded && !energies && !p.binned;
  if (!folded) pmeXLinesKernel<<<xBlocks, XLINE_THREADS, xLinesLds(p), s>>>(d, p, 1, energies);
  launchPlaneGather(d, p, planeGatherLds(p), s, folded);
  return xBlocks;
}

int launchPmeAfterForward(const Dev &d, const PmeDev &p, int energies, hipStream_t s) {
  return launchAfterForward(d, p, energies, s);
}

int launchPme(const Dev &d, const PmeDev &p, int energies, hipStream_t s, int *slotsUsed) {
  if (p.planes) {
    launchPmeBins(d, p, s);
    launchPlaneForward(d, p, planeForwardLds(p), s);
    const int xBlocks = launchAfterForward(d, p, energies, s);
    if (p.binned) pmeAtomGatherKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d, p);
    if (slotsUsed) *slotsUsed = xBlocks;
    return xBlocks;
  }
  const int atomBlocks = (d.n + 255) / 256;
  pmeSpreadKernel<<<atomBlocks, 256, 0, s>>>(d, p);
  const int zPairs = (p.K[0] * p.K[1] + 1) / 2;
  const int zBlocks = (zPairs + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeZForwardKernel<<<zBlocks, 256, lineLds(p.K[2]), s>>>(p);
  const int yBlocks = (p.K[0] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeYKernel<<<yBlocks, 256, lineLds(p.K[1]), s>>>(p, +1);
  const int xBlocks = (p.K[1] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
  pmeXConvolveKernel<<<xBlocks, 256, lineLds(p.K[0]), s>>>(d, p, 1, energies);
  pmeYKernel<<<yBlocks, 256, lineLds(p.K[1]), s>>>(p, -1);
  pmeZInverseKernel<<<zBlocks, 256, lineLds(p.K[2]), s>>>(p);
  pmeGatherKernel<<<atomBlocks, 256, 0, s>>>(d, p);
  if (slotsUsed) *slotsUsed = xBlocks;
  return xBlocks;
}

// One stage of launchPme's chain (no energies), for per-stage timing: 0 the
// spread with the forward transforms it carries, 1 the remaining transforms
// with the convolution (nothing where a force-only chain folds the x lines
// into the gathers), 2 the inverse transforms with the gather.
void launchPmeStage(const Dev &d, const PmeDev &p, hipStream_t s, int stage) {
  if (p.planes) {
    const int folded = p.xFolded && !p.binned;
    if (stage == 0) {
      launchPmeBins(d, p, s);
      launchPlaneForward(d, p, planeForwardLds(p), s);
    } else if (stage == 1) {
      if (!folded) pmeXLinesKernel<<<xLineBlocks(p), XLINE_THREADS, xLinesLds(p), s>>>(d, p, 1, 0);
    } else {
      launchPlaneGather(d, p, planeGatherLds(p), s, folded);
      if (p.binned) pmeAtomGatherKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d, p);
    }
    return;
  }
  const int atomBlocks = (d.n + 255) / 256;
  if (stage == 0) {
    pmeSpreadKernel<<<atomBlocks, 256, 0, s>>>(d, p);
  } else if (stage == 1) {
    const int zPairs = (p.K[0] * p.K[1] + 1) / 2;
    const int zBlocks = (zPairs + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
    pmeZForwardKernel<<<zBlocks, 256, lineLds(p.K[2]), s>>>(p);
    const int yBlocks = (p.K[0] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
    pmeYKernel<<<yBlocks, 256, lineLds(p.K[1]), s>>>(p, +1);
    const int xBlocks = (p.K[1] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
    pmeXConvolveKernel<<<xBlocks, 256, lineLds(p.K[0]), s>>>(d, p, 1, 0);
    pmeYKernel<<<yBlocks, 256, lineLds(p.K[1]), s>>>(p, -1);
    pmeZInverseKernel<<<zBlocks, 256, lineLds(p.K[2]), s>>>(p);
  } else {
    pmeGatherKernel<<<atomBlocks, 256, 0, s>>>(d, p);
  }
}

bool pmeTransformsFolded(const PmeDev &p) { return p.planes && p.xFolded && !p.binned; }

// Verification entry: the forward R2C 3D transform of a real grid, through
// the same kernels as the active pipeline.
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

void launchFftForward(const Dev &d, const PmeDev &p, const float *real, float2 *out, hipStream_t s) {
  if (p.planes) {
    pmePlaneForwardKernel<<<p.K[0], PLANE_THREADS, planeForwardLds(p), s>>>(d, p, real);
    pmeXLinesKernel<<<xLineBlocks(p), XLINE_THREADS, xLinesLds(p), s>>>(d, p, 0, 0);
  } else {
    const size_t n = (size_t)p.K[0] * p.K[1] * p.K[2];
    realToFixedKernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, real);
    const int zPairs = (p.K[0] * p.K[1] + 1) / 2;
    pmeZForwardKernel<<<(zPairs + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK, 256, lineLds(p.K[2]), s>>>(p);
    const int yBlocks = (p.K[0] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
    pmeYKernel<<<yBlocks, 256, lineLds(p.K[1]), s>>>(p, +1);
    const int xBlocks = (p.K[1] * p.K2h + LINES_PER_BLOCK - 1) / LINES_PER_BLOCK;
    pmeXConvolveKernel<<<xBlocks, 256, lineLds(p.K[0]), s>>>(d, p, 0, 0);
  }
  const size_t nc = (size_t)p.K[0] * p.K[1] * p.K2h;
  copyComplexKernel<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(p, out);
}

}  // namespace rmd

// Device bodies of the PME plane pipeline's first stage and the helpers the
// pipeline shares (complex arithmetic, the LDS FFT, B-splines, grid
// coordinates, fixed-point charges and forces): included by the PME kernels
// and by the single-system launch that runs the direct space, the bonded terms
// and the PME forward planes together.
#pragma once

#include "kernels.h"

// Multiply-adds are fused only within one expression (decided in the front
// end), never across statements by the backend: each kernel body then rounds
// the same way wherever it is inlined, so a replica batch reproduces the
// single-system kernels bit for bit.
#pragma clang fp contract(on)

namespace rmd {

constexpr int ORDER = 5;
constexpr int PLANE_THREADS = 512;
constexpr int XLINE_THREADS = PLANE_THREADS;  // the x lines run in the plane kernels' workgroup shape
constexpr int XLINES = 16;                     // x lines per workgroup in the plane pipeline
constexpr int LINES_PER_BLOCK = 4;  // lines per workgroup in the line pipeline

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a + sign * i * b
__device__ __forceinline__ float2 caddi(float2 a, float2 b, int sign) {
  return sign > 0 ? make_float2(a.x - b.y, a.y + b.x) : make_float2(a.x + b.y, a.y - b.x);
}
__device__ __forceinline__ float2 csubi(float2 a, float2 b, int sign) { return caddi(a, b, -sign); }

// In-register DFT of length R (2, 3, 4, 5 or 7): y_p = sum_q v_q exp(sign 2 pi i p q / R).
__device__ __forceinline__ void smallDft(float2 *v, int R, int sign) {
  if (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = caddi(d02, d13, sign);
    v[3] = csubi(d02, d13, sign);
  } else if (R == 3) {
    const float c1 = -0.5f, s1 = 0.866025403784438647f;
    const float2 t1 = cadd(v[1], v[2]);
    const float2 t2 = make_float2(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
    const float2 t3 = cscale(csub(v[1], v[2]), s1);
    v[0] = cadd(v[0], t1);
    v[1] = caddi(t2, t3, sign);
    v[2] = csubi(t2, t3, sign);
  } else if (R == 5) {
    const float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;
    const float s1 = 0.951056516295153572f, s2 = 0.587785252292473129f;
    const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const float2 x0 = v[0];
    const float2 r1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 r2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    const float2 i1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const float2 i2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = caddi(r1, i1, sign);
    v[4] = csubi(r1, i1, sign);
    v[2] = caddi(r2, i2, sign);
    v[3] = csubi(r2, i2, sign);
  } else {  // R == 7
    const float C1 = 0.623489801858733531f, C2 = -0.222520933956314404f, C3 = -0.900968867902419126f;
    const float S1 = 0.781831482468029809f, S2 = 0.974927912181823607f, S3 = 0.433883739117558120f;
    const float2 a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    const float2 a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    const float2 a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    const float2 x0 = v[0];
    auto comb = [](float2 x, float p, float2 u, float q, float2 w, float r, float2 z) {
      return make_float2(x.x + p * u.x + q * w.x + r * z.x, x.y + p * u.y + q * w.y + r * z.y);
    };
    const float2 zero = make_float2(0.f, 0.f);
    const float2 r1 = comb(x0, C1, a1, C2, a2, C3, a3), i1 = comb(zero, S1, b1, S2, b2, S3, b3);
    const float2 r2 = comb(x0, C2, a1, C3, a2, C1, a3), i2 = comb(zero, S2, b1, -S3, b2, -S1, b3);
    const float2 r3 = comb(x0, C3, a1, C1, a2, C2, a3), i3 = comb(zero, S3, b1, -S1, b2, S2, b3);
    v[0] = cadd(x0, cadd(a1, cadd(a2, a3)));
    v[1] = caddi(r1, i1, sign);
    v[6] = csubi(r1, i1, sign);
    v[2] = caddi(r2, i2, sign);
    v[5] = csubi(r2, i2, sign);
    v[3] = caddi(r3, i3, sign);
    v[4] = csubi(r3, i3, sign);
  }
}

// One radix-R pass of a Stockham autosort FFT over `lines` lines of length K
// (line l at src + l * K, in LDS); `Ns` is the product of the radices done.
// R is a compile-time constant, so the butterfly stays in registers.
template <int R>
__device__ __forceinline__ void fftStage(const float2 *src, float2 *dst, int lines, int K, int Ns, const float2 *tw,
                                         int sign) {
  const int L = K / R;
  const int twStride = K / (Ns * R);
  const int total = lines * L;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int line = idx / L, j = idx - line * L;
    const float2 *in = src + line * K;
    float2 *out = dst + line * K;
    const int k = j % Ns;
    float2 v[8];
    v[0] = in[j];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      float2 w = tw[k * q * twStride];
      if (sign < 0) w.y = -w.y;
      v[q] = cmul(in[j + q * L], w);
    }
    smallDft(v, R, sign);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) out[base + q * Ns] = v[q];
  }
}

// Stockham autosort FFT of `lines` lines of length K stored line-major in LDS
// (line l at src + l * K), by all threads of the block. `rad` lists the
// radices (0-terminated), `tw` holds exp(+2 pi i t / K) for t < K in LDS.
// Every stage ends with a block barrier; returns the buffer with the result.
__device__ __forceinline__ float2 *fftBatch(float2 *src, float2 *dst, int lines, int K, const int *rad,
                                            const float2 *tw, int sign) {
  int Ns = 1;
  for (int s = 0; s < 8 && rad[s] > 0; ++s) {
    const int R = rad[s];
    switch (R) {
      case 2: fftStage<2>(src, dst, lines, K, Ns, tw, sign); break;
      case 3: fftStage<3>(src, dst, lines, K, Ns, tw, sign); break;
      case 4: fftStage<4>(src, dst, lines, K, Ns, tw, sign); break;
      case 5: fftStage<5>(src, dst, lines, K, Ns, tw, sign); break;
      default: fftStage<7>(src, dst, lines, K, Ns, tw, sign); break;
    }
    __syncthreads();
    float2 *swap = src;
    src = dst;
    dst = swap;
    Ns *= R;
  }
  return src;
}

__device__ __forceinline__ void loadTwiddles(float2 *dst, const float2 *src, int K) {
  for (int k = threadIdx.x; k < K; k += blockDim.x) dst[k] = src[k];
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

// Grid index of weight k for base b: (b + k - order + 1) mod K.
__device__ __forceinline__ int gridIndex(int b, int k, int K) {
  int g = b + k - (ORDER - 1);
  while (g < 0) g += K;
  while (g >= K) g -= K;
  return g;
}

// The first weight index k with gridIndex(b, k, K) == plane, or ORDER if the
// atom's support misses the plane (later matches are k + K, k + 2K, ...).
__device__ __forceinline__ int planeWeight(int b, int plane, int K) {
  int r = (plane - b + ORDER - 1) % K;
  if (r < 0) r += K;
  return r;
}

// round-to-nearest-even of v to an integer, exact for |v| < 2^51: adding
// 1.5 * 2^52 puts v's integer part in the low mantissa bits.
__device__ __forceinline__ long long roundToInt64(double v) {
  const double magic = 6755399441055744.0;
  return (long long)(__double_as_longlong(v + magic) - __double_as_longlong(magic));
}

// |q w w w| <= |q| < 2^19 here, far inside the exact range.
__device__ __forceinline__ long long chargeToFixed(float q, float wx, float wy, float wz) {
  const float qab = (q * wx) * wy;
  return roundToInt64((double)(qab * wz) * GRID_SCALE);
}

// Force on one atom from the derivatives of its fractional-coordinate
// gradient (gu, gv, gw per grid unit), chained through the reciprocal rows.
// The same force in fixed point, without adding it.
__device__ __forceinline__ void gridForceFixed(const Dev &d, const PmeDev &p, double q, float gu, float gv, float gw,
                                               long long f[3]) {
  const double u = (double)gu * p.K[0] * q, v = (double)gv * p.K[1] * q, ww = (double)gw * p.K[2] * q;
  const double *r = d.box->recip64;
  const double fx = -(u * r[0] + v * r[1] + ww * r[2]);
  const double fy = -(u * r[3] + v * r[4] + ww * r[5]);
  const double fz = -(u * r[6] + v * r[7] + ww * r[8]);
  f[0] = __double2ll_rn(fx * FORCE_SCALE);
  f[1] = __double2ll_rn(fy * FORCE_SCALE);
  f[2] = __double2ll_rn(fz * FORCE_SCALE);
}

__device__ __forceinline__ void addGridForce(const Dev &d, const PmeDev &p, int atom, double q, float gu, float gv,
                                             float gw) {
  long long f[3];
  gridForceFixed(d, p, q, gu, gv, gw, f);
  atomicAddFixed(&d.force[atom], f[0]);
  atomicAddFixed(&d.force[d.nPad + atom], f[1]);
  atomicAddFixed(&d.force[2 * d.nPad + atom], f[2]);
}

// |m|^2 of the wave vector of half-spectrum point (x, y, m), reciprocal rows
// (the influence function's convention).
__device__ __forceinline__ double waveVector2(const PmeDev &p, const double *recip, int x, int y, int m) {
  const int kx = x <= p.K[0] / 2 ? x : x - p.K[0];
  const int ky = y <= p.K[1] / 2 ? y : y - p.K[1];
  const double mx = kx * recip[0] + ky * recip[1] + m * recip[2];
  const double my = kx * recip[3] + ky * recip[4] + m * recip[5];
  const double mz = kx * recip[6] + ky * recip[7] + m * recip[8];
  return mx * mx + my * my + mz * mz;
}

// The block's two energy sums into their slot rows.
__device__ void storeReciprocal(const Dev &d, double energy, double dscale, int block) {
  __shared__ double partial[2][16];
  energy = waveSum(energy);
  dscale = waveSum(dscale);
  const int waves = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    partial[0][threadIdx.x >> 6] = energy;
    partial[1][threadIdx.x >> 6] = dscale;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double e = 0.0, s = 0.0;
    for (int w = 0; w < waves; ++w) {
      e += partial[0][w];
      s += partial[1][w];
    }
    d.slots[E_RECIPROCAL * MAX_SLOTS + block] = e;
    d.slots[E_DSCALE_RECIPROCAL * MAX_SLOTS + block] = s;
  }
}

// ============================================================================
// Plane pipeline
// ============================================================================

// The binned atoms that reach plane i: bins i .. i + ORDER - 1 (mod K0), one
// contiguous range of binAtoms or two where the bins wrap.
__device__ __forceinline__ void planeRanges(const PmeDev &p, int i, int2 range[2]) {
  const int K0 = p.K[0], last = i + ORDER - 1;
  if (last < K0) {
    range[0] = make_int2(p.binStart[i], p.binStart[last + 1]);
    range[1] = make_int2(0, 0);
  } else {
    range[0] = make_int2(p.binStart[i], p.binStart[K0]);
    range[1] = make_int2(p.binStart[0], p.binStart[last - K0 + 1]);
  }
}

// One round of a plane's atom scan: of atoms [a0, a0 + PLANE_THREADS * SCAN),
// those whose x support meets `plane` are appended to `list` as 16 bits,
// (atom - a0) << 3 | weight index, wave-aggregated into the shared `count`.
// The expensive per-atom work then runs over the compacted list with every
// lane busy (about 5 in K0 atoms land on a given plane). Single systems scan
// six atoms per thread: a 3,072-atom system in one round (each round is a
// global-load latency and three barriers on the kernel's serial path), in
// less LDS than 2,048 int entries. The batched kernels scan four: their
// members mostly take the binned path, and the wider scan's registers would
// cost them a workgroup per CU.
constexpr int SCAN_PER_THREAD = 6;  // the widest; sizes the list in LDS
constexpr int SCAN_BATCHED = 4;
constexpr int PLANE_LIST = PLANE_THREADS * SCAN_PER_THREAD;
static_assert(PLANE_LIST <= (1 << 13), "a list entry holds the atom's offset in its round in 13 bits");
typedef unsigned short PlaneEntry;

// The round's positions, loaded apart from their use so that a kernel can
// issue its first round's loads at its very start, beside its other loads.
template <int SCAN>
__device__ __forceinline__ void planeScanLoad(const Dev &d, int a0, float4 batch[SCAN]) {
#pragma unroll
  for (int u = 0; u < SCAN; ++u) {
    const int a = a0 + u * blockDim.x + threadIdx.x;
    batch[u] = a < d.n ? d.pos[a] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

template <int SCAN>
__device__ __forceinline__ void planeScanTake(const Dev &d, const PmeDev &p, int plane, int a0, const float4 batch[SCAN],
                                              PlaneEntry *list, int *count) {
  const double *recip = d.box->recip64;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int u = 0; u < SCAN; ++u) {
    const int a = a0 + u * blockDim.x + threadIdx.x;
    int r = ORDER;
    if (a < d.n && batch[u].w != 0.f) {
      int bx;
      float ox;
      gridCoordinate(p, recip, 0, batch[u], bx, ox);
      r = planeWeight(bx, plane, p.K[0]);
    }
    const bool take = r < ORDER;
    const unsigned long long mask = __ballot(take);
    int base = 0;
    if (lane == (int)(__ffsll((long long)mask) - 1)) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, __ffsll((long long)mask) - 1, 64);
    if (take) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (PlaneEntry)(((a - a0) << 3) | r);
  }
}

template <int SCAN>
__device__ void planeScan(const Dev &d, const PmeDev &p, int plane, int a0, PlaneEntry *list, int *count) {
  float4 batch[SCAN];
  planeScanLoad<SCAN>(d, a0, batch);
  planeScanTake<SCAN>(d, p, plane, a0, batch, list, count);
}



// 1. Spread onto plane i and forward z, y transforms of the plane. With
// `real` set, the plane is read from a real K0*K1*K2 grid instead (the FFT
// verification entry).
template <int SCAN = SCAN_PER_THREAD>
__device__ __forceinline__ void pmePlaneForward(const Dev &d, const PmeDev &p, const float *real, int i, float2 *lds) {
  const int K0 = p.K[0], K1 = p.K[1], K2 = p.K[2], K2h = p.K2h;
  const int S = p.planeSpan, P = (K1 + 1) / 2;
  const int t = threadIdx.x, nt = blockDim.x;
  float2 *tw1 = lds, *tw2 = tw1 + K1;
  float2 *A = tw2 + K2, *B = A + S;
  long long *plane = reinterpret_cast<long long *>(B);
  PlaneEntry *list = reinterpret_cast<PlaneEntry *>(B + p.planeRegion);
  int *listCount = reinterpret_cast<int *>(list + PLANE_LIST);
  // the first scan round's positions, in flight with the twiddles' loads
  const bool scanning = !real && !p.binned;
  float4 first[SCAN];
  if (scanning) planeScanLoad<SCAN>(d, 0, first);
  loadTwiddles(tw1, p.twiddle[1], K1);
  loadTwiddles(tw2, p.twiddle[2], K2);
  if (real) {
    for (int idx = t; idx < P * K2; idx += nt) {
      const int pr = idx / K2, k = idx - pr * K2;
      const float *row = real + ((size_t)i * K1 + 2 * pr) * K2;
      A[idx] = make_float2(row[k], 2 * pr + 1 < K1 ? row[K2 + k] : 0.f);
    }
  } else {
    for (int idx = t; idx < K1 * K2; idx += nt) plane[idx] = 0;
    const double *recip = d.box->recip64;
    // atom a's charge onto this plane; r: the first weight index that lands
    // on the plane, or -1 to derive it from a binned atom's base index (in
    // i .. i + ORDER - 1 mod K0)
    auto spread = [&](int a, int r) {
      const float4 pos = d.pos[a];
      int bx, by, bz;
      float ox, oy, oz;
      gridCoordinate(p, recip, 0, pos, bx, ox);
      if (r < 0) {
        r = i - bx + ORDER - 1;
        r += r < 0 ? K0 : 0;
        r -= r >= K0 ? K0 : 0;
      }
      float w[ORDER], wy[ORDER], wz[ORDER], dw[ORDER];
      bspline5(ox, w, dw);
      float wx = 0.f;
      for (int q = r; q < ORDER; q += K0) wx += w[q];
      gridCoordinate(p, recip, 1, pos, by, oy);
      gridCoordinate(p, recip, 2, pos, bz, oz);
      bspline5(oy, wy, dw);
      bspline5(oz, wz, dw);
      int iz[ORDER];
      for (int c = 0; c < ORDER; ++c) iz[c] = gridIndex(bz, c, K2);
      for (int b = 0; b < ORDER; ++b) {
        unsigned long long *row = reinterpret_cast<unsigned long long *>(plane + gridIndex(by, b, K1) * K2);
        for (int c = 0; c < ORDER; ++c)
          atomicAdd(row + iz[c], (unsigned long long)chargeToFixed(pos.w, wx, wy[b], wz[c]));
      }
    };
    __syncthreads();
    if (p.binned) {
      int2 range[2];
      planeRanges(p, i, range);
      for (int h = 0; h < 2; ++h)
        for (int k = range[h].x + t; k < range[h].y; k += nt) spread(p.binAtoms[k], -1);
    } else {
      for (int a0 = 0; a0 < d.n; a0 += nt * SCAN) {
        if (t == 0) *listCount = 0;
        __syncthreads();
        if (a0 == 0)
          planeScanTake<SCAN>(d, p, i, 0, first, list, listCount);
        else
          planeScan<SCAN>(d, p, i, a0, list, listCount);
        __syncthreads();
        const int count = *listCount;
        for (int k = t; k < count; k += nt) spread(a0 + (list[k] >> 3), list[k] & 7);
        __syncthreads();
      }
    }
    __syncthreads();
    // Two real z lines per complex line.
    for (int idx = t; idx < P * K2; idx += nt) {
      const int pr = idx / K2, k = idx - pr * K2;
      const long long *row = plane + 2 * pr * K2;
      const float re = (float)((double)row[k] * GRID_UNSCALE);
      const float im = 2 * pr + 1 < K1 ? (float)((double)row[K2 + k] * GRID_UNSCALE) : 0.f;
      A[idx] = make_float2(re, im);
    }
  }
  __syncthreads();
  float2 *Z = fftBatch(A, B, P, K2, p.radices[2], tw2, +1);
  float2 *Y = (Z == A) ? B : A;
  // Split the packed spectra: line 2p = (Z[m] + conj Z[-m]) / 2, line 2p+1 =
  // (Z[m] - conj Z[-m]) / (2i); y lines laid out m-major.
  for (int idx = t; idx < K2h * K1; idx += nt) {
    const int m = idx / K1, j = idx - m * K1;
    const float2 *zl = Z + (j >> 1) * K2;
    const float2 z = zl[m], zc = zl[m == 0 ? 0 : K2 - m];
    Y[idx] = (j & 1) == 0 ? make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y))
                          : make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
  }
  __syncthreads();
  float2 *R = fftBatch(Y, Z, K2h, K1, p.radices[1], tw1, +1);
  float2 *out = p.cgrid + (size_t)i * K1 * K2h;
  for (int idx = t; idx < K1 * K2h; idx += nt) {
    const int j = idx / K2h, m = idx - j * K2h;
    out[idx] = R[m * K1 + j];
  }
}

// 2. x lines (j, m0 .. m0 + XLINES - 1): forward transform, then, with
// `convolve`, the influence function (energy in FP64, each half-spectrum
// column weighted by its multiplicity in the full spectrum) and the inverse
// transform.
__device__ __forceinline__ void pmeXLines(const Dev &d, const PmeDev &p, int convolve, int energies, int block,
                                          float2 *lds) {
  const int K0 = p.K[0], K1 = p.K[1], K2h = p.K2h;
  const int chunks = (K2h + XLINES - 1) / XLINES;
  const int j = block / chunks, m0 = (block - j * chunks) * XLINES;
  const int lines = min(XLINES, K2h - m0);
  const int t = threadIdx.x, nt = blockDim.x;
  float2 *tw = lds, *A = tw + K0, *B = A + XLINES * K0;
  loadTwiddles(tw, p.twiddle[0], K0);
  const size_t stride = (size_t)K1 * K2h;
  float2 *col = p.cgrid + (size_t)j * K2h + m0;
  for (int idx = t; idx < lines * K0; idx += nt) {
    const int x = idx / lines, l = idx - x * lines;
    A[l * K0 + x] = col[x * stride + l];
  }
  __syncthreads();
  float2 *R = fftBatch(A, B, lines, K0, p.radices[0], tw, +1);
  double energy = 0.0, dscale = 0.0;
  if (convolve) {
    const float *g0 = p.influence + (size_t)j * K2h + m0;
    const double *recip = d.box->recip64;
    for (int idx = t; idx < lines * K0; idx += nt) {
      const int x = idx / lines, l = idx - x * lines;
      const int m = m0 + l;
      const double weight = (m == 0 || (p.K[2] % 2 == 0 && m == p.K[2] / 2)) ? 1.0 : 2.0;
      const float g = g0[x * stride + l];
      const float2 f = R[l * K0 + x];
      const double e = weight * (double)g * ((double)f.x * f.x + (double)f.y * f.y);
      energy += e;
      if (energies) dscale += e * (2.0 * p.piOverAlpha2 * waveVector2(p, recip, x, j, m) - 1.0);
      R[l * K0 + x] = make_float2(2.f * g * f.x, 2.f * g * f.y);
    }
    __syncthreads();
    R = fftBatch(R, R == A ? B : A, lines, K0, p.radices[0], tw, -1);
  }
  for (int idx = t; idx < lines * K0; idx += nt) {
    const int x = idx / lines, l = idx - x * lines;
    col[x * stride + l] = R[l * K0 + x];
  }
  if (convolve && energies) storeReciprocal(d, energy, dscale, block);
}

// 2'. The x lines folded into plane i's gather (PmeDev::xFolded): column c
// of the plane is sum_x X(i - x, c) F(x, c) over the forward planes F, in
// ascending x, each thread one column at a time. The loads of RUN planes are
// issued together (the planes were written by other workgroups of the
// previous launch; the kernel rows are read-only). REAL: X is real.
template <bool REAL>
__device__ __forceinline__ void planeXFold(const PmeDev &p, int i, float2 *A) {
  const int K0 = p.K[0], K1 = p.K[1], K2h = p.K2h, PK = K1 * K2h;
  constexpr int RUN = 16;
  const float *Xr = p.xKernel;
  const float2 *Xc = reinterpret_cast<const float2 *>(p.xKernel);
  for (int c = threadIdx.x; c < PK; c += blockDim.x) {
    const int j = c / K2h, m = c - j * K2h;
    float re = 0.f, im = 0.f;
    for (int x0 = 0; x0 < K0; x0 += RUN) {
      float2 f[RUN], k[RUN];
#pragma unroll
      for (int u = 0; u < RUN; ++u) {
        // past the last plane: the last plane's loads, weighted zero
        const bool in = x0 + u < K0;
        const int x = in ? x0 + u : K0 - 1;
        int l = i - x;
        l += l < 0 ? K0 : 0;
        const float2 fx = p.cgrid[(size_t)x * PK + c];
        const float2 kx = REAL ? make_float2(Xr[(size_t)l * PK + c], 0.f) : Xc[(size_t)l * PK + c];
        f[u] = in ? fx : make_float2(0.f, 0.f);
        k[u] = in ? kx : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < RUN; ++u) {
        re = fmaf(k[u].x, f[u].x, re);
        im = fmaf(k[u].x, f[u].y, im);
        if (!REAL) {
          re = fmaf(-k[u].y, f[u].y, re);
          im = fmaf(k[u].y, f[u].x, im);
        }
      }
    }
    A[m * K1 + j] = make_float2(re, im);
  }
}

// 3. Inverse y and z transforms of plane i into the real potential, then each
// atom whose support meets the plane takes that plane's share of its force
// (webgpu/kernels/pme.js:281-336; cpu/forces/pme.js:222-257). With `folded`
// the plane's spectrum is formed from the forward planes (planeXFold) instead
// of read from the x lines' output.
template <int SCAN = SCAN_PER_THREAD>
__device__ __forceinline__ void pmePlaneGather(const Dev &d, const PmeDev &p, int i, float2 *lds, int folded = 0) {
  const int K0 = p.K[0], K1 = p.K[1], K2 = p.K[2], K2h = p.K2h;
  const int S = p.planeSpan, P = (K1 + 1) / 2;
  const int t = threadIdx.x, nt = blockDim.x;
  float2 *tw1 = lds, *tw2 = tw1 + K1;
  float2 *A = tw2 + K2, *B = A + S;
  PlaneEntry *list = reinterpret_cast<PlaneEntry *>(B + S);
  int *listCount = reinterpret_cast<int *>(list + PLANE_LIST);
  // The first scan round's positions are loaded with the plane's spectrum and
  // its list taken before the transforms (the list has LDS of its own), so
  // the positions' load latency overlaps the spectrum's.
  float4 first[SCAN];
  if (!p.binned) planeScanLoad<SCAN>(d, 0, first);
  loadTwiddles(tw1, p.twiddle[1], K1);
  loadTwiddles(tw2, p.twiddle[2], K2);
  if (folded) {
    if (p.xReal)
      planeXFold<true>(p, i, A);
    else
      planeXFold<false>(p, i, A);
  } else {
    const float2 *in = p.cgrid + (size_t)i * K1 * K2h;
    for (int idx = t; idx < K1 * K2h; idx += nt) {
      const int j = idx / K2h, m = idx - j * K2h;
      A[m * K1 + j] = in[idx];
    }
  }
  if (!p.binned && t == 0) *listCount = 0;
  __syncthreads();
  if (!p.binned) planeScanTake<SCAN>(d, p, i, 0, first, list, listCount);
  float2 *Y = fftBatch(A, B, K2h, K1, p.radices[1], tw1, -1);
  float2 *O = (Y == A) ? B : A;
  // Hermitian extension of lines 2p and 2p+1 packed into one complex line.
  for (int idx = t; idx < P * K2; idx += nt) {
    const int pr = idx / K2, k = idx - pr * K2;
    const int j0 = 2 * pr, j1 = 2 * pr + 1;
    float2 a, b = make_float2(0.f, 0.f);
    if (k < K2h) {
      a = Y[k * K1 + j0];
      if (j1 < K1) b = Y[k * K1 + j1];
    } else {
      a = Y[(K2 - k) * K1 + j0];
      a.y = -a.y;
      if (j1 < K1) {
        b = Y[(K2 - k) * K1 + j1];
        b.y = -b.y;
      }
    }
    O[idx] = make_float2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  float2 *Z = fftBatch(O, Y, P, K2, p.radices[2], tw2, -1);
  float *phi = reinterpret_cast<float *>(Z == A ? B : A);
  for (int idx = t; idx < P * K2; idx += nt) {
    const int pr = idx / K2, k = idx - pr * K2;
    phi[2 * pr * K2 + k] = Z[idx].x;
    if (2 * pr + 1 < K1) phi[(2 * pr + 1) * K2 + k] = Z[idx].y;
  }
  __syncthreads();
  if (p.binned) {
    // a binned system gathers per atom from the potential grid
    // (pmeAtomGather)
    float *out = p.rgrid + (size_t)i * K1 * K2;
    for (int idx = t; idx < K1 * K2; idx += nt) out[idx] = phi[idx];
    return;
  }
  const double *recip = d.box->recip64;
  for (int a0 = 0; a0 < d.n; a0 += nt * SCAN) {
    if (a0 > 0) {  // round 0's list was taken before the transforms
      if (t == 0) *listCount = 0;
      __syncthreads();
      planeScan<SCAN>(d, p, i, a0, list, listCount);
    }
    __syncthreads();
    const int count = *listCount;
    for (int k = t; k < count; k += nt) {
      const int a = a0 + (list[k] >> 3), r = list[k] & 7;
      const float4 pos = d.pos[a];
      int bx, by, bz;
      float ox, oy, oz;
      gridCoordinate(p, recip, 0, pos, bx, ox);
      float w[ORDER], dw[ORDER], wy[ORDER], dwy[ORDER], wz[ORDER], dwz[ORDER];
      bspline5(ox, w, dw);
      float wx = 0.f, dwx = 0.f;
      for (int q = r; q < ORDER; q += K0) {
        wx += w[q];
        dwx += dw[q];
      }
      gridCoordinate(p, recip, 1, pos, by, oy);
      gridCoordinate(p, recip, 2, pos, bz, oz);
      bspline5(oy, wy, dwy);
      bspline5(oz, wz, dwz);
      int iz[ORDER];
      for (int c = 0; c < ORDER; ++c) iz[c] = gridIndex(bz, c, K2);
      float s0 = 0.f, s1 = 0.f, s2 = 0.f;
      for (int b = 0; b < ORDER; ++b) {
        const float *row = phi + gridIndex(by, b, K1) * K2;
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
      addGridForce(d, p, a, pos.w, dwx * s0, wx * s1, wx * s2);
    }
    __syncthreads();
  }
}

}  // namespace rmd

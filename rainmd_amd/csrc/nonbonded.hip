// LJ + erfc-screened real-space Ewald over the Verlet list.
//
// Replaces directSpaceShader (webgpu/kernels/nonbonded.js:30-109). The pair
// arithmetic is the reference's (A/r^12 - B/r^6, qq*erfc(ar)/r, and the
// Gaussian term of the screened-Coulomb derivative, nonbonded.js:68-98) in
// FP32. The machinery is MI355X's: a group of `LANES` lanes of a wave64
// shares one atom's list, each lane's partial force is converted to int64
// fixed point and the group reduces it exactly, so the sum does not depend
// on the lane count or the order. Energies (a separate instantiation, used on
// the steps that report them) are evaluated and accumulated per lane in FP64
// from the FP32 distance - a small box's LJ total is a large cancellation
// that FP32 terms and sums would not keep to 1e-6 - halved (each pair appears
// in both atoms' lists) and reduced in FP64 into per-block slots.
#include "kernels.h"

// Multiply-adds fuse only within one expression, so each body rounds the same
// way wherever it is inlined (the single-system and the batched kernels).
#pragma clang fp contract(on)

namespace rmd {

constexpr int LDS_TYPES = 64;  // type-pair tables up to 64^2 entries live in LDS

template <int LANES, bool ENERGIES>
__device__ __forceinline__ void nonbonded(const Dev &d, float alpha, int block) {
  // The LJ type-pair table in LDS: one lookup per pair, off the global
  // latency chain (list index -> position and type -> table entry).
  __shared__ float2 ldsTable[LDS_TYPES * LDS_TYPES];
  const int tableSize = d.types * d.types;
  const bool tableInLds = d.types <= LDS_TYPES;
  if (tableInLds)
    for (int k = threadIdx.x; k < tableSize; k += blockDim.x) ldsTable[k] = d.lj[k];
  __syncthreads();
  const float2 *table = tableInLds ? ldsTable : d.lj;
  const BoxF &box = *d.box;
  const int tid = block * blockDim.x + threadIdx.x;
  const int row = tid / LANES;  // a slot: rows and list entries follow the cell order
  const int sub = tid % LANES;
  const bool valid = row < d.n;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  double elj = 0.0, ecoul = 0.0, dscale = 0.0;
  if (valid) {
    const float4 pi = d.slotPos[row];
    const int ti = d.slotType[row] * d.types;
    const float qi = ONE_4PI_EPS0_F * pi.w;
    const float gaussScale = TWO_OVER_SQRT_PI_F * alpha;
    const int count = d.nbrCount[row];
    const int *list = d.nbrList + (size_t)row * d.cap;
    // Software pipeline: the neighbour slot two iterations ahead and the
    // position and type one iteration ahead are in flight while a pair is
    // computed. Past the end the prefetch reads the row's own slot (in bounds).
    int k = sub;
    int jAhead = k + LANES < count ? list[k + LANES] : row;
    int jNext = k < count ? list[k] : row;
    float4 pNext = d.slotPos[jNext];
    int tNext = d.slotType[jNext];
    while (k < count) {
      const float4 pj = pNext;
      const int tj = tNext;
      k += LANES;
      jNext = jAhead;
      jAhead = k + LANES < count ? list[k + LANES] : row;
      pNext = d.slotPos[jNext];
      tNext = d.slotType[jNext];
      float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      float s0, s1, s2;
      const float r2 = minimumImageR2(dx, dy, dz, box, s0, s1, s2);
      if (r2 > box.rc2) continue;
      const float2 ab = table[ti + tj];
      const float invR2 = __builtin_amdgcn_rcpf(r2);
      const float invR = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * invR;
      const float invR6 = invR2 * invR2 * invR2;
      const float lj12 = ab.x * invR6;
      float scalar = (12.0f * lj12 - 6.0f * ab.y) * invR6 * invR2;
      const float qq = qi * pj.w;
      float gaussian;
      const float screened = erfcWithGaussian(alpha * r, gaussian);
      scalar += qq * (screened * invR + gaussScale * gaussian) * invR2;
      if (ENERGIES) {
        // FP64 at the same image: the FP32 box rows, the hardware
        // reciprocals' one-ulp bias and the FP32 table would otherwise be
        // systematic over the box, and a water box's LJ total is a small
        // difference of large sums
        const double2 ab64 = d.lj64[ti + tj];
        const double ir2 = 1.0 / imageR2Double(pi, pj, box, s0, s1, s2);
        const double ir6 = ir2 * ir2 * ir2;
        elj += (ab64.x * ir6 - ab64.y) * ir6;
        ecoul += (double)(qq * screened) * sqrt(ir2);
        dscale -= (double)(scalar * r2);  // r dU/dr
      }
      fx += scalar * dx;
      fy += scalar * dy;
      fz += scalar * dz;
    }
  }
  // Exact group reduction in fixed point.
  long long ix = __double2ll_rn((double)fx * FORCE_SCALE);
  long long iy = __double2ll_rn((double)fy * FORCE_SCALE);
  long long iz = __double2ll_rn((double)fz * FORCE_SCALE);
#pragma unroll
  for (int offset = LANES / 2; offset > 0; offset >>= 1) {
    ix += __shfl_xor(ix, offset, LANES);
    iy += __shfl_xor(iy, offset, LANES);
    iz += __shfl_xor(iz, offset, LANES);
  }
  if (valid && sub == 0) {
    const int atom = d.cellAtoms[row];
    atomicAddFixed(&d.force[atom], ix);
    atomicAddFixed(&d.force[d.nPad + atom], iy);
    atomicAddFixed(&d.force[2 * d.nPad + atom], iz);
  }
  if (ENERGIES) {
    __shared__ double partial[3][4];
    double a = waveSum(0.5 * elj);
    double b = waveSum(0.5 * ecoul);
    double c = waveSum(0.5 * dscale);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      partial[0][wave] = a;
      partial[1][wave] = b;
      partial[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double sa = 0, sb = 0, sc = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        sa += partial[0][w];
        sb += partial[1][w];
        sc += partial[2][w];
      }
      d.slots[E_LJ_DIRECT * MAX_SLOTS + block] = sa;
      d.slots[E_COUL_DIRECT * MAX_SLOTS + block] = sb;
      d.slots[E_DSCALE_DIRECT * MAX_SLOTS + block] = sc;
    }
  }
}

template <int LANES, bool ENERGIES>
__global__ void __launch_bounds__(256) nonbondedKernel(Dev d, float alpha) {
  nonbonded<LANES, ENERGIES>(d, alpha, blockIdx.x);
}

// Batched.
template <int LANES, bool ENERGIES>
__global__ void __launch_bounds__(256) nonbondedBatchKernel(const BatchView *__restrict__ v,
                                                            const int2 *__restrict__ task, float alpha) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  nonbonded<LANES, ENERGIES>(v[t.x].d, alpha, t.y);
}

template <bool ENERGIES>
static void launchLanes(const Dev &d, float alpha, int lanes, int blocks, hipStream_t s) {
  switch (lanes) {
    case 1: nonbondedKernel<1, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
    case 2: nonbondedKernel<2, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
    case 4: nonbondedKernel<4, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
    case 8: nonbondedKernel<8, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
    case 16: nonbondedKernel<16, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
    case 32: nonbondedKernel<32, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
    default: nonbondedKernel<64, ENERGIES><<<blocks, 256, 0, s>>>(d, alpha); break;
  }
}

int launchNonbonded(const Dev &d, float alpha, int lanes, int energies, hipStream_t s) {
  const long long threads = (long long)d.n * lanes;
  const int blocks = (int)((threads + 255) / 256);
  if (energies)
    launchLanes<true>(d, alpha, lanes, blocks, s);
  else
    launchLanes<false>(d, alpha, lanes, blocks, s);
  return blocks;
}

template <bool ENERGIES>
static void launchBatchLanes(const BatchView *v, const TaskTable &t, int lanes, float alpha, hipStream_t s) {
  switch (lanes) {
    case 8: nonbondedBatchKernel<8, ENERGIES><<<t.blocks, 256, 0, s>>>(v, t.task, alpha); break;
    case 16: nonbondedBatchKernel<16, ENERGIES><<<t.blocks, 256, 0, s>>>(v, t.task, alpha); break;
    case 32: nonbondedBatchKernel<32, ENERGIES><<<t.blocks, 256, 0, s>>>(v, t.task, alpha); break;
    default: nonbondedBatchKernel<64, ENERGIES><<<t.blocks, 256, 0, s>>>(v, t.task, alpha); break;
  }
}

void launchNonbondedBatch(const BatchView *v, const TaskTable &t, int lanes, float alpha, int energies, hipStream_t s) {
  if (energies)
    launchBatchLanes<true>(v, t, lanes, alpha, s);
  else
    launchBatchLanes<false>(v, t, lanes, alpha, s);
}

}  // namespace rmd

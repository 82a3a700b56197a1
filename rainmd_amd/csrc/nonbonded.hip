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
#include "bonded_terms.h"
#include "kernels.h"
#include "pme_plane.h"
#include "search.h"

// Multiply-adds fuse only within one expression, so each body rounds the same
// way wherever it is inlined (the single-system and the batched kernels).
#pragma clang fp contract(on)

namespace rmd {

// Type-pair tables up to 90^2 entries (64.8 KB) are copied to LDS, sized to
// the system; larger ones are read from global memory.
constexpr int LDS_TYPES = 90;

inline bool tableInLds(int types) { return types <= LDS_TYPES; }
inline size_t tableLdsBytes(int types) { return tableInLds(types) ? (size_t)types * types * sizeof(float2) : 0; }

// G(w) = (erf(z) - 2 z / sqrt(pi) exp(-w)) / z^3 with w = z^2 = (alpha r)^2,
// as P(w) / Q(w) of degree 5 (a relative-error fit on [0, 12] made for this
// file by ewald_force_fit.py: 6.0e-8 in FP64, 4e-7 evaluated in FP32; Q >= 1
// there). The screened Coulomb force scalar of a pair is
// qq (1/r^3 - alpha^3 G(w)): no exponential, no erfc.
constexpr float EWALD_G_WMAX = 12.0f;
__device__ __forceinline__ float ewaldG(float w) {
  float p = -3.737052279e-07f;
  p = fmaf(p, w, 3.548286411e-05f);
  p = fmaf(p, w, 1.812391234e-04f);
  p = fmaf(p, w, 1.684099673e-02f);
  p = fmaf(p, w, -1.957190408e-02f);
  p = fmaf(p, w, 7.522528230e-01f);
  float q = 1.580766103e-04f;
  q = fmaf(q, w, 2.452907735e-03f);
  q = fmaf(q, w, 2.431532518e-02f);
  q = fmaf(q, w, 1.524804702e-01f);
  q = fmaf(q, w, 5.739843349e-01f);
  q = fmaf(q, w, 1.0f);
  return p * __builtin_amdgcn_rcpf(q);
}

bool ewaldRationalFits(double alpha, double cutoff) { return alpha * cutoff * alpha * cutoff <= EWALD_G_WMAX; }

// One pair: the force scalar F/r (added to the lane's force with the
// separation), and with ENERGIES the pair's LJ, screened Coulomb and r dU/dr
// in FP64. A pair that is not `live` (an entry past the row's count) or
// beyond the cutoff adds nothing. The force-only variant computes every pair
// and selects zero for those (a branch around them would be taken by whole
// waves almost never, and lets the compiler sink the pair's loads behind
// it); the energy variant branches around its FP64 work.
// With RATIONAL (alpha rc within the fit) the Coulomb force takes ewaldG;
// else erfc and the Gaussian.
template <bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
__device__ __forceinline__ void pairTerm(const float4 &pi, float qi, int ti, const float4 &pj, int tj, bool live,
                                         const BoxF &box, float alpha, float alpha2, float alpha3,
                                         const float2 *ldsTable,
                                         GlobalPtr<float2> table, const double2 *lj64, float &fx, float &fy, float &fz,
                                         double &elj, double &ecoul, double &dscale) {
  float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  float s0, s1, s2;
  const float r2 = imageR2<RECT>(dx, dy, dz, box, s0, s1, s2);
  const bool inside = live && r2 <= box.rc2;
  if (ENERGIES && !inside) return;
  const float2 ab = LDS_TABLE ? ldsTable[ti + tj] : table[ti + tj];
  const float invR = __builtin_amdgcn_rsqf(r2);
  const float invR2 = invR * invR;
  const float invR6 = invR2 * invR2 * invR2;
  const float lj12 = ab.x * invR6;
  float scalar = (12.0f * lj12 - 6.0f * ab.y) * invR6 * invR2;
  const float qq = qi * pj.w;
  float screened = 0.f;  // erfc(alpha r), for the energy
  float coulomb;
  if (RATIONAL) {
    coulomb = invR * invR2 - alpha3 * ewaldG(alpha2 * r2);
    if (ENERGIES) {
      float gaussian;
      screened = erfcWithGaussian(alpha * (r2 * invR), gaussian);
    }
  } else {
    float gaussian;
    screened = erfcWithGaussian(alpha * (r2 * invR), gaussian);
    coulomb = (screened * invR + TWO_OVER_SQRT_PI_F * alpha * gaussian) * invR2;
  }
  scalar += qq * coulomb;
  if (ENERGIES) {
    // FP64 at the same image: the FP32 box rows, the hardware reciprocals'
    // one-ulp bias and the FP32 table would otherwise be systematic over the
    // box, and a water box's LJ total is a small difference of large sums
    // 1/r from the hardware FP64 reciprocal square root and one Newton step
    // (relative error about the square of the estimate's), in place of a
    // correctly rounded division and square root
    const double2 ab64 = lj64[ti + tj];
    const double r2d = imageR2Double<RECT>(pi, pj, box, s0, s1, s2);
    const double y = __builtin_amdgcn_rsq(r2d);
    const double ir = fma(0.5 * y, fma(-r2d * y, y, 1.0), y);
    const double ir2 = ir * ir;
    const double ir6 = ir2 * ir2 * ir2;
    elj += (ab64.x * ir6 - ab64.y) * ir6;
    ecoul += (double)(qq * screened) * ir;
    dscale -= (double)(scalar * r2);  // r dU/dr
  }
  // a select, not a product: a pair that is not computed may hold r2 = 0
  const float f = inside ? scalar : 0.0f;
  fx += f * dx;
  fy += f * dy;
  fz += f * dz;
}

template <int LANES, bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
__device__ __forceinline__ void nonbonded(const Dev &d, float alpha, int block, float2 *ldsTable) {
  const float alpha2 = alpha * alpha, alpha3 = alpha2 * alpha;
  // The LJ type-pair table in LDS: one lookup per pair, off the global
  // latency chain (list index -> position and type -> table entry).
  const GlobalPtr<float2> table = globalPtr(d.lj);
  if (LDS_TABLE) {
    const int tableSize = d.types * d.types;
    for (int k = threadIdx.x; k < tableSize; k += blockDim.x) ldsTable[k] = table[k];
    __syncthreads();
  }
  const BoxF box = uniformPairBox(*d.box, ENERGIES);
  const GlobalPtr<float4> slotPos = globalPtr(d.slotPos);
  const GlobalPtr<int> slotType = globalPtr(d.slotType);
  const __amdgpu_buffer_rsrc_t posBuffer = bufferOf(d.slotPos, (unsigned int)d.n * 16u);
  const __amdgpu_buffer_rsrc_t typeBuffer = bufferOf(d.slotType, (unsigned int)d.n * 4u);
  const int tid = block * blockDim.x + threadIdx.x;
  const int row = tid / LANES;  // a slot: rows and list entries follow the cell order
  const int sub = tid % LANES;
  const bool valid = row < d.n;
  // the list rows of this wave (64 / LANES consecutive rows from firstRow)
  const int firstRow = __builtin_amdgcn_readfirstlane(row);
  const int waveRows = min(64 / LANES, max(0, d.n - firstRow));
  const __amdgpu_buffer_rsrc_t listBuffer =
      bufferOf(d.nbrList + (size_t)firstRow * d.cap, (unsigned int)waveRows * (unsigned int)d.cap * 4u);
  float fx = 0.f, fy = 0.f, fz = 0.f;
  double elj = 0.0, ecoul = 0.0, dscale = 0.0;
  if (valid) {
    const float4 pi = slotPos[row];
    const int ti = slotType[row] * d.types;
    const float qi = ONE_4PI_EPS0_F * pi.w;
    const int count = globalPtr(d.nbrCount)[row];
    // UNROLL entries per iteration, k + u * LANES: their indices are loaded
    // together, then their positions and types together, then the pairs are
    // computed, so a lane waits two memory round trips per UNROLL pairs and
    // the other waves of the SIMD fill the wait. Every load is issued
    // unconditionally, through buffer descriptors that return zero for any
    // offset outside them: the list's over this wave's rows (an entry past a
    // row's count reads unwritten memory or the next row's entries) and the
    // slots' (whatever slot such an entry holds). Such a pair is loaded but
    // not computed.
    constexpr int UNROLL = 4;
    const unsigned int rowOffset = (unsigned int)(row - firstRow) * (unsigned int)d.cap * 4u;
    for (int k = sub; k < count; k += UNROLL * LANES) {
      int j[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
        j[u] = (int)__builtin_amdgcn_raw_buffer_load_b32(listBuffer, rowOffset + (unsigned int)(k + u * LANES) * 4u, 0,
                                                         0);
      float4 p[UNROLL];
      int t[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        p[u] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(posBuffer, j[u] * 16, 0, 0));
        t[u] = (int)__builtin_amdgcn_raw_buffer_load_b32(typeBuffer, j[u] * 4, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
        pairTerm<ENERGIES, LDS_TABLE, RECT, RATIONAL>(pi, qi, ti, p[u], t[u], k + u * LANES < count, box, alpha, alpha2,
                                                      alpha3, ldsTable, table, d.lj64, fx, fy, fz, elj, ecoul, dscale);
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
    __shared__ double partial[3][8];  // up to 512 threads
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
    // One slot per 256 threads, numbered as a 256-thread launch numbers its
    // workgroups, whatever this launch's workgroup size: the slot sums, and
    // so the energies, are bit for bit those of the 256-thread kernels.
    if ((threadIdx.x & 255) == 0) {
      const int group = threadIdx.x >> 8, slot = block * (int)(blockDim.x >> 8) + group;
      const int waves = min(4, (int)(blockDim.x >> 6) - 4 * group);
      double sa = 0, sb = 0, sc = 0;
      for (int w = 4 * group; w < 4 * group + waves; ++w) {
        sa += partial[0][w];
        sb += partial[1][w];
        sc += partial[2][w];
      }
      if (slot < (int)(((long long)d.n * LANES + 255) / 256)) {
        d.slots[E_LJ_DIRECT * MAX_SLOTS + slot] = sa;
        d.slots[E_COUL_DIRECT * MAX_SLOTS + slot] = sb;
        d.slots[E_DSCALE_DIRECT * MAX_SLOTS + slot] = sc;
      }
    }
  }
}

template <int LANES, bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
__global__ void __launch_bounds__(256) nonbondedKernel(Dev d, float alpha) {
  extern __shared__ float2 ldsTable[];
  nonbonded<LANES, ENERGIES, LDS_TABLE, RECT, RATIONAL>(d, alpha, blockIdx.x, ldsTable);
}

// One system's direct space, bonded terms and PME forward planes in one launch
// of PLANE_THREADS-thread workgroups: [0, pairBlocks) run the direct space,
// [pairBlocks, planesFrom) the bonded terms, and each workgroup after that one
// x plane of the PME forward stage, each with the same body as its own
// kernel. A captured step runs its force branches one after another, so
// separate launches cost the sum of their durations, where one launch
// overlaps them. With `search`, a rebuild's one-wave-per-atom list search
// runs here as well, each direct-space workgroup searching the rows it then
// computes (no workgroup reads another's rows). (The cell build and the x
// lines stay out: joining them behind in-launch waits between workgroups -
// bounded waits, device-scope release and acquire across the XCDs' caches -
// made the 2854-atom step 5 to 25 us slower, and its rebuild twice as long,
// for dispatch gaps of about 1.6 us per dependent launch.)
template <int LANES, bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
__global__ void __launch_bounds__(PLANE_THREADS) forcesJoinedKernel(Dev d, PmeDev p, float alpha, double alphaD,
                                                                   int terms, int pairBlocks, int planesFrom,
                                                                   int search) {
  extern __shared__ float2 lds[];
  const int b = blockIdx.x;
  if (b < pairBlocks) {
    if (search && d.sc->rebuild) {
      // rows [b R, (b + 1) R), R = PLANE_THREADS / LANES, eight per pass:
      // searchBlock's block k holds rows 8 k .. 8 k + 7, one per wave
      constexpr int WAVES = PLANE_THREADS / 64, ROWS = PLANE_THREADS / LANES;
      static_assert(ROWS % WAVES == 0, "whole passes of rows");
      for (int pass = 0; pass < ROWS / WAVES; ++pass) searchBlock<RECT, WAVES>(d, b * (ROWS / WAVES) + pass);
      __syncthreads();  // the workgroup's lists before its direct space
    }
    nonbonded<LANES, ENERGIES, LDS_TABLE, RECT, RATIONAL>(d, alpha, b, lds);
  } else if (b < planesFrom)
    bonded(d, alphaD, terms, ENERGIES ? 1 : 0, b - pairBlocks);
  else
    pmePlaneForward(d, p, nullptr, b - planesFrom, lds);
}

// Batched.
template <int LANES, bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
__global__ void __launch_bounds__(256) nonbondedBatchKernel(const BatchView *__restrict__ v,
                                                            const int2 *__restrict__ task, float alpha) {
  extern __shared__ float2 ldsTable[];
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  nonbonded<LANES, ENERGIES, LDS_TABLE, RECT, RATIONAL>(v[t.x].d, alpha, t.y, ldsTable);
}

// The bonded terms and PME forward planes joined to a single system's
// direct-space launch (p null: none joined).
struct JoinedPart {
  const PmeDev *p;
  double alpha;
  int terms;
  int search;
};

template <bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
static void launchLanes(const Dev &d, float alpha, int lanes, int blocks, size_t lds, hipStream_t s,
                        const JoinedPart &j) {
  if constexpr (LDS_TABLE) {  // the joined launch needs the LJ table in LDS (forcesJoin)
    if (j.p) {
      const int pair = (int)(((long long)d.n * lanes + PLANE_THREADS - 1) / PLANE_THREADS);
      const int planesFrom = pair + (bondedThreads(d) + PLANE_THREADS - 1) / PLANE_THREADS;
      const int all = planesFrom + j.p->K[0];
      switch (lanes) {
        case 16:
          forcesJoinedKernel<16, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<all, PLANE_THREADS, lds, s>>>(
              d, *j.p, alpha, j.alpha, j.terms, pair, planesFrom, j.search);
          return;
        case 32:
          forcesJoinedKernel<32, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<all, PLANE_THREADS, lds, s>>>(
              d, *j.p, alpha, j.alpha, j.terms, pair, planesFrom, j.search);
          return;
        case 64:
          forcesJoinedKernel<64, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<all, PLANE_THREADS, lds, s>>>(
              d, *j.p, alpha, j.alpha, j.terms, pair, planesFrom, j.search);
          return;
      }
    }
  }
  switch (lanes) {
    case 1: nonbondedKernel<1, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
    case 2: nonbondedKernel<2, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
    case 4: nonbondedKernel<4, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
    case 8: nonbondedKernel<8, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
    case 16: nonbondedKernel<16, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
    case 32: nonbondedKernel<32, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
    default: nonbondedKernel<64, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<blocks, 256, lds, s>>>(d, alpha); break;
  }
}

template <bool ENERGIES, bool LDS_TABLE, bool RECT, bool RATIONAL>
static void launchBatchLanes(const BatchView *v, const TaskTable &t, int lanes, float alpha, size_t lds,
                             hipStream_t s) {
  switch (lanes) {
    case 8:
      nonbondedBatchKernel<8, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<t.blocks, 256, lds, s>>>(v, t.task, alpha);
      break;
    case 16:
      nonbondedBatchKernel<16, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<t.blocks, 256, lds, s>>>(v, t.task, alpha);
      break;
    case 32:
      nonbondedBatchKernel<32, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<t.blocks, 256, lds, s>>>(v, t.task, alpha);
      break;
    default:
      nonbondedBatchKernel<64, ENERGIES, LDS_TABLE, RECT, RATIONAL><<<t.blocks, 256, lds, s>>>(v, t.task, alpha);
      break;
  }
}

// The kernel variant of a launch: energies, LJ table in LDS, rectangular
// box, rational Coulomb force - each a compile-time choice.
struct Variant {
  bool energies, lds, rect, rational;
};

template <bool E, bool L, bool R, bool Q>
static void launchOne(bool batch, const Dev *d, const BatchView *v, const TaskTable *t, float alpha, int lanes,
                      int blocks, size_t lds, hipStream_t s, const JoinedPart &j) {
  if (batch)
    launchBatchLanes<E, L, R, Q>(v, *t, lanes, alpha, lds, s);
  else
    launchLanes<E, L, R, Q>(*d, alpha, lanes, blocks, lds, s, j);
}

static void dispatch(const Variant &k, bool batch, const Dev *d, const BatchView *v, const TaskTable *t, float alpha,
                     int lanes, int blocks, size_t lds, hipStream_t s,
                     const JoinedPart &j = JoinedPart{nullptr, 0.0, 0, 0}) {
  const int code = (k.energies ? 8 : 0) | (k.lds ? 4 : 0) | (k.rect ? 2 : 0) | (k.rational ? 1 : 0);
  switch (code) {
#define RMD_NB_CASE(C)                                                                                     \
  case C:                                                                                                  \
    launchOne<(C & 8) != 0, (C & 4) != 0, (C & 2) != 0, (C & 1) != 0>(batch, d, v, t, alpha, lanes, blocks, \
                                                                       lds, s, j);                        \
    break;
    RMD_NB_CASE(0) RMD_NB_CASE(1) RMD_NB_CASE(2) RMD_NB_CASE(3) RMD_NB_CASE(4) RMD_NB_CASE(5) RMD_NB_CASE(6)
    RMD_NB_CASE(7) RMD_NB_CASE(8) RMD_NB_CASE(9) RMD_NB_CASE(10) RMD_NB_CASE(11) RMD_NB_CASE(12) RMD_NB_CASE(13)
    RMD_NB_CASE(14) RMD_NB_CASE(15)
#undef RMD_NB_CASE
  }
}

int launchNonbonded(const Dev &d, float alpha, int rational, int lanes, int energies, hipStream_t s) {
  const long long threads = (long long)d.n * lanes;
  const int blocks = (int)((threads + 255) / 256);
  const Variant k{energies != 0, tableInLds(d.types), d.rect != 0, rational != 0};
  dispatch(k, false, &d, nullptr, nullptr, alpha, lanes, blocks, tableLdsBytes(d.types), s);
  return blocks;
}

// The joined launch holds the bonded and PME bodies' registers (about 95
// VGPRs against the direct space's 63 at 64 lanes): a system joins when its
// direct space fits in one pass at four waves per SIMD of the 256 CUs, as
// small systems' launch-bound steps do, with the LJ table in LDS, the
// unbinned plane pipeline and at most 64 KB of LDS per workgroup.
bool forcesJoin(const Dev &d, const PmeDev &p, int lanes) {
  constexpr long long ONE_PASS = 256LL * 4 * 4 * 64;  // threads
  return (lanes == 16 || lanes == 32 || lanes == 64) && (long long)d.n * lanes <= ONE_PASS && tableInLds(d.types) &&
         p.planes && !p.binned && std::max(tableLdsBytes(d.types), pmePlaneForwardLds(p)) <= 65536;
}

void launchForcesJoined(const Dev &d, const PmeDev &p, float alpha, int rational, int lanes, int energies,
                        double alphaD, int terms, hipStream_t s, int *pairSlots, int *bondedSlots, int search) {
  const Variant k{energies != 0, true, d.rect != 0, rational != 0};
  const size_t lds = std::max(tableLdsBytes(d.types), pmePlaneForwardLds(p));
  dispatch(k, false, &d, nullptr, nullptr, alpha, lanes, 0, lds, s, JoinedPart{&p, alphaD, terms, search});
  // energy slots as the 256-thread kernels number them
  *pairSlots = (int)(((long long)d.n * lanes + 255) / 256);
  *bondedSlots = (bondedThreads(d) + 255) / 256;
}

// `types` is the largest type count of the batch's replicas: every workgroup
// gets LDS for that table. `rect`: every replica's box is rectangular.
void launchNonbondedBatch(const BatchView *v, const TaskTable &t, int lanes, float alpha, int rational, int energies,
                          int types, int rect, hipStream_t s) {
  const Variant k{energies != 0, tableInLds(types), rect != 0, rational != 0};
  dispatch(k, true, nullptr, v, &t, alpha, lanes, 0, tableLdsBytes(types), s);
}

}  // namespace rmd

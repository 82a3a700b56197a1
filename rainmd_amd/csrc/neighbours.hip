// Cell-list binning by counting sort, and the Verlet-skin neighbour list.
//
// Replaces encodeCellList (webgpu/kernels/cell-list.js:176-226) and
// encodeNeighbours (webgpu/kernels/neighbours.js:46-169). The rebuild is
// decided on the device: the integrator's last block compares the largest
// displacement since the last rebuild with half the skin and sets
// Scalars::rebuild, so these kernels run every step and return at once when
// the list is still valid. That keeps a captured block free of host
// decisions and never steps on a stale list.
//
// The binning arithmetic is FP32 with a fixed operation order and no
// contraction, so the FP32 CPU emulation in the oracle reproduces every cell
// assignment and every pair decision bit for bit.
#include "kernels.h"

// Multiply-adds fuse only within one expression, so each body rounds the same
// way wherever it is inlined (the single-system and the batched kernels).
#pragma clang fp contract(on)

namespace rmd {

constexpr int SORT_CAPACITY = 4096;  // atoms per cell the LDS sort takes

// cellOfPosition (webgpu/kernels/wgsl-common.js:112-123): fractional
// coordinate from the reciprocal rows, wrapped with f - floor(f), scaled,
// truncated and clamped.
__device__ __forceinline__ int cellIndex(float x, float y, float z, const BoxF &b, int c[3]) {
#pragma clang fp contract(off)
  const int *cells = b.cells;
  for (int d = 0; d < 3; ++d) {
    float a0 = x * b.recip[d];
    float a1 = y * b.recip[3 + d];
    float a2 = z * b.recip[6 + d];
    float f = (a0 + a1) + a2;
    float w = f - floorf(f);
    float scaled = w * (float)cells[d];
    int ci = (int)scaled;
    ci = ci > cells[d] - 1 ? cells[d] - 1 : ci;
    ci = ci < 0 ? 0 : ci;
    c[d] = ci;
  }
  return (c[0] * cells[1] + c[1]) * cells[2] + c[2];
}

// Slot k of the cell order gets atom `atom`: its rebuild-time position with
// the atom index (for the search), its current position and charge, its type,
// and the inverse map.
__device__ __forceinline__ void writeSlot(const Dev &d, int k, int atom) {
  const float4 p = d.pos[atom];
  d.cellAtoms[k] = atom;
  d.cellPos[k] = make_float4(p.x, p.y, p.z, __int_as_float(atom));
  d.slotPos[k] = p;
  d.slotType[k] = d.type[atom];
  d.slotOf[atom] = k;
}

__global__ void setRebuildKernel(Scalars *sc) {
  sc->rebuild = 1u;
  sc->maxDisp2Bits = 0u;
  sc->slack = 0.f;
  retireRequirement(sc);
  sc->rebuilds += 1u;
}

__global__ void cellsAssignKernel(Dev d) {
  if (!d.sc->rebuild) return;
  const BoxF &box = *d.box;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const float4 p = d.pos[i];
  int c3[3];
  const int c = cellIndex(p.x, p.y, p.z, box, c3);
  d.cellOfAtom[i] = c;
  atomicAdd(&d.cellCount[c], 1);
  if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicMax(&d.sc->nonFinite, 1u);
}

// Exclusive scan of the per-cell counts in one workgroup; leaves the counts
// zeroed for the next rebuild and the scatter cursors at the cell starts.
__global__ void __launch_bounds__(1024) cellsScanKernel(Dev d) {
  if (!d.sc->rebuild) return;
  __shared__ int partial[1024];
  __shared__ int carry;
  const int total = d.box->cellTotal;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < total; base += 1024) {
    const int c = base + threadIdx.x;
    const int v = c < total ? d.cellCount[c] : 0;
    partial[threadIdx.x] = v;
    __syncthreads();
    for (int offset = 1; offset < 1024; offset <<= 1) {
      int add = threadIdx.x >= offset ? partial[threadIdx.x - offset] : 0;
      __syncthreads();
      partial[threadIdx.x] += add;
      __syncthreads();
    }
    const int exclusive = carry + partial[threadIdx.x] - v;
    if (c < total) {
      d.cellStart[c] = exclusive;
      d.cellCursor[c] = exclusive;
      d.cellCount[c] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += partial[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.cellStart[total] = carry;
}

__global__ void cellsScatterKernel(Dev d) {
  if (!d.sc->rebuild) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int slot = atomicAdd(&d.cellCursor[d.cellOfAtom[i]], 1);
  d.cellAtoms[slot] = i;
}

// The scatter leaves each cell's atoms in arrival order; a bitonic sort in
// LDS puts them in index order so the list, and every sum over it, is the
// same on every run. The sorted cell also gets its positions copied in cell
// order (atom index in .w) for the candidate scan.
__global__ void __launch_bounds__(256) cellsSortKernel(Dev d) {
  if (!d.sc->rebuild) return;
  __shared__ int keys[SORT_CAPACITY];
  const int c = blockIdx.x;
  if (c >= d.box->cellTotal) return;
  const int begin = d.cellStart[c], end = d.cellStart[c + 1];
  const int len = end - begin;
  if (len <= 0) return;
  if (len > SORT_CAPACITY) {
    if (threadIdx.x == 0) atomicMax(&d.sc->cellOverflow, 1u);
    return;
  }
  int size = 1;
  while (size < len) size <<= 1;
  for (int k = threadIdx.x; k < size; k += blockDim.x) keys[k] = k < len ? d.cellAtoms[begin + k] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= size; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < size; t += blockDim.x) {
        const int partner = t ^ j;
        if (partner > t) {
          const bool ascending = (t & k) == 0;
          const int a = keys[t], b = keys[partner];
          if ((a > b) == ascending) {
            keys[t] = b;
            keys[partner] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int k = threadIdx.x; k < len; k += blockDim.x) {
    const int atom = keys[k];
    writeSlot(d, begin + k, atom);
  }
}

// The whole cell list of a small system in one workgroup: assignment with
// LDS-atomic counts, the scan, then each cell's atoms in index order - by
// wave ballots where there are at most 64 cells, else an LDS scatter and a
// sort of each cell's run (insertion by one thread per cell, or where a cell
// is crowded a rank sort: rank = atoms of the same cell with a lower index) -
// the same lists as the four-kernel path, with no global atomics and three
// launches fewer on every step. Used when the atoms and the cells fit the
// LDS arrays below.
constexpr int SMALL_THREADS = 1024;
constexpr int SMALL_MAX_ATOMS = 16384;
constexpr int SMALL_MAX_CELLS = 4096;
constexpr int INSERTION_MAX = 32;  // cells up to this many atoms are sorted by one thread each

__device__ __forceinline__ void cellsBuildSmall(const Dev &d) {
  if (!d.sc->rebuild) return;
  __shared__ int count[SMALL_MAX_CELLS];
  __shared__ int begin[SMALL_MAX_CELLS + 1];
  __shared__ int unsorted[SMALL_MAX_ATOMS];
  __shared__ int partial[SMALL_THREADS];
  const BoxF &box = *d.box;
  const int total = box.cellTotal, n = d.n, t = threadIdx.x;
  for (int c = t; c < total; c += SMALL_THREADS) count[c] = 0;
  __syncthreads();
  for (int i = t; i < n; i += SMALL_THREADS) {
    const float4 p = d.pos[i];
    int c3[3];
    const int c = cellIndex(p.x, p.y, p.z, box, c3);
    d.cellOfAtom[i] = c;
    atomicAdd(&count[c], 1);
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicMax(&d.sc->nonFinite, 1u);
  }
  __syncthreads();
  // exclusive scan: each thread sums a contiguous run of cells, the runs are
  // scanned across the block, then each run is written out
  const int per = (total + SMALL_THREADS - 1) / SMALL_THREADS;
  const int c0 = min(total, t * per), c1 = min(total, c0 + per);
  int run = 0;
  for (int c = c0; c < c1; ++c) run += count[c];
  partial[t] = run;
  __syncthreads();
  for (int offset = 1; offset < SMALL_THREADS; offset <<= 1) {
    const int add = t >= offset ? partial[t - offset] : 0;
    __syncthreads();
    partial[t] += add;
    __syncthreads();
  }
  int at = partial[t] - run;
  bool crowdedRun = false;
  for (int c = c0; c < c1; ++c) {
    begin[c] = at;
    d.cellStart[c] = at;
    crowdedRun |= count[c] > INSERTION_MAX;
    at += count[c];
    count[c] = begin[c];  // now the scatter cursor
  }
  if (t == SMALL_THREADS - 1) {
    begin[total] = partial[t];
    d.cellStart[total] = partial[t];
  }
  const bool crowded = __syncthreads_or(crowdedRun);
  if (total <= 64) {
    // Few (and so crowded) cells: place the atoms in index order directly,
    // SMALL_THREADS at a time. Each wave splits its 64 atoms by cell with one
    // ballot per cell (an atom's place among its wave's atoms of the same
    // cell is the count of lower lanes in its cell's mask); one thread per
    // cell then turns the waves' counts into each wave's start in the cell,
    // in wave order, and advances the cell's cursor. No per-cell rank scan.
    __shared__ int waveStart[SMALL_THREADS / 64][64];
    const int lane = t & 63, wave = t >> 6;
    for (int base = 0; base < n; base += SMALL_THREADS) {
      const int i = base + t;
      const int c = i < n ? d.cellOfAtom[i] : -1;
      unsigned long long same = 0ull;
      for (int k = 0; k < total; ++k) {
        const unsigned long long m = __ballot(c == k);
        if (c == k) same = m;
        if (lane == k) waveStart[wave][k] = __popcll(m);
      }
      __syncthreads();
      if (t < total) {
        int at = count[t];
        for (int w = 0; w < SMALL_THREADS / 64; ++w) {
          const int m = waveStart[w][t];
          waveStart[w][t] = at;
          at += m;
        }
        count[t] = at;
      }
      __syncthreads();
      if (i < n) writeSlot(d, waveStart[wave][c] + __popcll(same & ((1ull << lane) - 1ull)), i);
      __syncthreads();
    }
    return;
  }
  for (int i = t; i < n; i += SMALL_THREADS) unsorted[atomicAdd(&count[d.cellOfAtom[i]], 1)] = i;
  __syncthreads();
  if (!crowded) {
    // every cell holds at most INSERTION_MAX atoms: one thread sorts each
    // cell's run in place, then the slots are written in order (coalesced)
    for (int c = t; c < total; c += SMALL_THREADS) {
      const int b = begin[c], e = begin[c + 1];
      for (int k = b + 1; k < e; ++k) {
        const int v = unsorted[k];
        int m = k - 1;
        while (m >= b && unsorted[m] > v) {
          unsorted[m + 1] = unsorted[m];
          --m;
        }
        unsorted[m + 1] = v;
      }
    }
    __syncthreads();
    for (int k = t; k < n; k += SMALL_THREADS) writeSlot(d, k, unsorted[k]);
    return;
  }
  for (int i = t; i < n; i += SMALL_THREADS) {
    const int c = d.cellOfAtom[i];
    const int b = begin[c], e = begin[c + 1];
    int rank = 0;
    for (int k = b; k < e; ++k) rank += unsorted[k] < i;
    writeSlot(d, b + rank, i);
  }
}

__global__ void __launch_bounds__(SMALL_THREADS) cellsBuildSmallKernel(Dev d) { cellsBuildSmall(d); }

// Batched: one workgroup per replica.
__global__ void __launch_bounds__(SMALL_THREADS) cellsBuildSmallBatchKernel(const BatchView *__restrict__ v,
                                                                           const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  cellsBuildSmall(v[t.x].d);
}

// Per-axis stencil (shared/topology/cells.js:25-29, generalised to cells of
// rl / reach): `reach` cells either side when the axis has at least
// 2 * reach + 1 cells, else every cell of the axis once.
constexpr int REACH_MAX = 3;
constexpr int STENCIL_ROWS = (2 * REACH_MAX + 1) * (2 * REACH_MAX + 1);
constexpr int STENCIL_RANGES = 2 * STENCIL_ROWS;  // a row's z run splits in two where it wraps

__device__ __forceinline__ void axisStencil(int count, int reach, int &base, int &len) {
  if (count >= 2 * reach + 1) {
    base = -reach;
    len = 2 * reach + 1;
  } else {
    base = 0;
    len = count;
  }
}

__device__ __forceinline__ int wrapCell(int v, int count) {
  v %= count;
  return v < 0 ? v + count : v;
}

// Whether candidate pj joins atom's list: not itself, within rl (the FP32
// minimum image of the kernels), not excluded. Exclusions are checked only
// for candidates inside the range of the atom's (sorted) exclusion partners.
template <bool RECT>
__device__ __forceinline__ bool acceptPair(GlobalPtr<int> excl, const BoxF &box, float4 pi, float4 pj, int atom,
                                           int exBegin, int exEnd, int exMin, int exMax) {
  // the distance test first and without a branch, so the candidate's loads
  // are used whole where they are issued
  const int j = __float_as_int(pj.w);
  float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  float s0, s1, s2;
  const float r2 = imageR2<RECT>(dx, dy, dz, box, s0, s1, s2);
  bool accept = (j != atom) & (r2 <= box.rl2);
  if (accept & (j >= exMin) & (j <= exMax))
    for (int e = exBegin; e < exEnd; ++e)
      if (excl[e] == j) accept = false;
  return accept;
}

// One wave per atom. The stencil is a set of z runs, one per (x, y) stencil
// row, each a contiguous range of the cell-ordered positions (two where the
// run wraps). The wave writes the ranges and their prefix sums to LDS and
// streams the concatenated candidates 256 at a time (four per lane, their
// loads issued together), each lane keeping a cursor into the ranges; a
// ballot compacts the accepted ones into the atom's slice in candidate
// order, so the list is the same on every run. Every pair found is counted;
// only `cap` are stored, and the true requirement reaches the host through
// Scalars::required (the reference's overflow contract,
// webgpu/kernels/neighbours.js:128-146).
template <bool RECT>
__device__ __forceinline__ unsigned int neighbours(const Dev &d, int block) {
  __shared__ int rangeBegin[4][STENCIL_RANGES];
  __shared__ int rangeStart[4][STENCIL_RANGES + 1];
  const BoxF box = uniformPairBox(*d.box, false);
  const int cells[3] = {d.box->cells[0], d.box->cells[1], d.box->cells[2]};
  const GlobalPtr<float4> cellPos = globalPtr(d.cellPos);
  const GlobalPtr<int> cellStart = globalPtr(d.cellStart);
  const GlobalPtr<int> excl = globalPtr(d.exclPartners);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = (block * blockDim.x + threadIdx.x) >> 6;  // a slot
  if (row >= d.n) return 0u;
  const int atom = d.cellAtoms[row];
  const float4 pi = d.pos[atom];
  if (lane == 0) d.posRebuild[atom] = pi;
  const int own = d.cellOfAtom[atom];
  const int c3[3] = {own / (cells[1] * cells[2]), (own / cells[2]) % cells[1], own % cells[2]};
  const int exBegin = d.exclStart[atom], exEnd = d.exclStart[atom + 1];
  const int exMin = exBegin < exEnd ? excl[exBegin] : 1;
  const int exMax = exBegin < exEnd ? excl[exEnd - 1] : 0;
  const int reach = d.box->reach;
  int bx, nx, by, ny, bz, nz;
  axisStencil(cells[0], reach, bx, nx);
  axisStencil(cells[1], reach, by, ny);
  axisStencil(cells[2], reach, bz, nz);
  // the z segments, the same for every row: the whole row, or [lo, hi)
  // split where it wraps
  int za0, zb0, za1 = 0, zb1 = 0, segs = 1;
  if (nz >= cells[2]) {
    za0 = 0;
    zb0 = cells[2];
  } else {
    const int lo = c3[2] + bz, hi = lo + nz;
    if (lo < 0) {
      segs = 2;
      za0 = 0, zb0 = hi;
      za1 = lo + cells[2], zb1 = cells[2];
    } else if (hi > cells[2]) {
      segs = 2;
      za0 = 0, zb0 = hi - cells[2];
      za1 = lo, zb1 = cells[2];
    } else {
      za0 = lo, zb0 = hi;
    }
  }
  const int ranges = nx * ny * segs;
  int *rb = rangeBegin[wave], *rs = rangeStart[wave];
  // ranges r = lane and lane + 64: their slots and lengths, then an exclusive
  // prefix over all of them
  int len[2] = {0, 0};
  for (int h = 0; h < 2; ++h) {
    const int r = lane + 64 * h;
    if (r < ranges) {
      const int row = r / segs, seg = r - row * segs;
      const int cx = wrapCell(c3[0] + bx + row / ny, cells[0]);
      const int cy = wrapCell(c3[1] + by + row % ny, cells[1]);
      const int rowBase = (cx * cells[1] + cy) * cells[2];
      const int begin = cellStart[rowBase + (seg ? za1 : za0)];
      const int end = cellStart[rowBase + (seg ? zb1 : zb0)];
      rb[r] = begin;
      len[h] = end - begin;
    }
  }
  int carry = 0;
  for (int h = 0; h < 2; ++h) {
    int inclusive = len[h];
    for (int offset = 1; offset < 64; offset <<= 1) {
      const int v = __shfl_up(inclusive, offset, 64);
      if (lane >= offset) inclusive += v;
    }
    const int r = lane + 64 * h;
    if (r < ranges) rs[r] = carry + inclusive - len[h];
    carry += __shfl(inclusive, 63, 64);
  }
  if (lane == 0) rs[ranges] = carry;
  __builtin_amdgcn_wave_barrier();
  const int total = carry;
  int found = 0;
  const GlobalOut<int> out = globalOut(d.nbrList + (size_t)row * d.cap);
  const int cap = d.cap;
  auto emit = [&](bool accept, int j) {  // j: the candidate's slot
    const unsigned long long mask = __ballot(accept);
    if (accept) {
      const int where = found + __popcll(mask & ((1ull << lane) - 1ull));
      if (where < cap) out[where] = j;
    }
    found += __popcll(mask);
  };
  // Four candidates per lane per pass, base + lane + 64 u. A candidate past
  // the total loads the row's own slot, which the pair test turns down (the
  // atom itself), so every load is in bounds and issued with the others.
  constexpr int CAND = 4;
  int cur = 0;  // the range of this lane's latest candidate
  for (int base = 0; base < total; base += 64 * CAND) {
    int slot[CAND];
#pragma unroll
    for (int u = 0; u < CAND; ++u) {
      const int g = base + lane + 64 * u;
      slot[u] = row;
      if (g < total) {
        while (rs[cur + 1] <= g) ++cur;
        slot[u] = rb[cur] + (g - rs[cur]);
      }
    }
    float4 p[CAND];
#pragma unroll
    for (int u = 0; u < CAND; ++u) p[u] = cellPos[slot[u]];
    // the tests of all four, then the stores (which share the loads'
    // completion counter)
    bool accept[CAND];
#pragma unroll
    for (int u = 0; u < CAND; ++u) accept[u] = acceptPair<RECT>(excl, box, pi, p[u], atom, exBegin, exEnd, exMin, exMax);
#pragma unroll
    for (int u = 0; u < CAND; ++u)
      if (base + 64 * u < total) emit(accept[u], slot[u]);
  }
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) d.nbrCount[row] = found < cap ? found : cap;
  return (unsigned int)found;
}

// One wave per cell. Every atom of a cell has the same stencil, so the wave
// streams the stencil's candidates as above (the same ranges, in the same
// order) and tests each candidate against every atom of the cell, up to 64
// at a time, one per lane, broadcast from its lane in turn. Each atom's
// accepted candidates are compacted into its own slice in candidate order:
// the lists are entry for entry those of one wave per atom, for a fraction
// of the loads and of the per-atom set-up.
template <bool RECT>
__device__ __forceinline__ unsigned int neighboursByCell(const Dev &d, int block) {
  __shared__ int rangeBegin[4][STENCIL_RANGES];
  __shared__ int rangeStart[4][STENCIL_RANGES + 1];
  const BoxF box = uniformPairBox(*d.box, false);
  const int cells[3] = {d.box->cells[0], d.box->cells[1], d.box->cells[2]};
  const GlobalPtr<float4> cellPos = globalPtr(d.cellPos);
  const GlobalPtr<int> cellStart = globalPtr(d.cellStart);
  const GlobalPtr<int> excl = globalPtr(d.exclPartners);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cell = block * 4 + wave;
  if (cell >= d.box->cellTotal) return 0u;
  const int cellBegin = cellStart[cell], cellEnd = cellStart[cell + 1];
  if (cellBegin == cellEnd) return 0u;
  const int c3[3] = {cell / (cells[1] * cells[2]), (cell / cells[2]) % cells[1], cell % cells[2]};
  const int reach = d.box->reach;
  int bx, nx, by, ny, bz, nz;
  axisStencil(cells[0], reach, bx, nx);
  axisStencil(cells[1], reach, by, ny);
  axisStencil(cells[2], reach, bz, nz);
  int za0, zb0, za1 = 0, zb1 = 0, segs = 1;
  if (nz >= cells[2]) {
    za0 = 0;
    zb0 = cells[2];
  } else {
    const int lo = c3[2] + bz, hi = lo + nz;
    if (lo < 0) {
      segs = 2;
      za0 = 0, zb0 = hi;
      za1 = lo + cells[2], zb1 = cells[2];
    } else if (hi > cells[2]) {
      segs = 2;
      za0 = 0, zb0 = hi - cells[2];
      za1 = lo, zb1 = cells[2];
    } else {
      za0 = lo, zb0 = hi;
    }
  }
  // Rectangular boxes: each stencil z run is trimmed at both ends to the
  // cells within reach of the box around this cell's atoms (wrapped into the
  // cell frame), by the periodic gap between the box and each cell's extent,
  // with a margin far above the FP32 rounding of either. A trimmed cell holds
  // no candidate the pair test could accept, so the lists are unchanged.
  // per axis: the box's centre and half width (uniform), the cell width
  float centre[3] = {0.f, 0.f, 0.f}, half[3] = {0.f, 0.f, 0.f};
  const float limit = sqrtf(box.rl2) + 1e-4f, limit2 = limit * limit;
  const float edge[3] = {box.a[0], box.b[1], box.c[2]};
  if (RECT) {
    for (int dd = 0; dd < 3; ++dd) {
      float lo = __builtin_inff(), hi = -__builtin_inff();
      for (int k = cellBegin + lane; k < cellEnd; k += 64) {
        const float4 q = cellPos[k];
        const float x = dd == 0 ? q.x : dd == 1 ? q.y : q.z;
        const float w = x - edge[dd] * floorf(x * box.inv[dd]);
        lo = fminf(lo, w);
        hi = fmaxf(hi, w);
      }
      for (int offset = 32; offset > 0; offset >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, offset, 64));
        hi = fmaxf(hi, __shfl_xor(hi, offset, 64));
      }
      centre[dd] = uniformFloat(0.5f * (lo + hi));
      half[dd] = uniformFloat(0.5f * (hi - lo));
    }
  }
  // the periodic gap along axis dd between the atoms' box and cell index c:
  // the centres' distance folded into [0, L/2], less both half widths
  auto gap = [&](int dd, int c) {
    const float width = edge[dd] / (float)cells[dd];
    float dist = fabsf((c + 0.5f) * width - centre[dd]);
    dist = fminf(dist, edge[dd] - dist);
    return fmaxf(0.f, dist - half[dd] - 0.5f * width);
  };
  const int ranges = nx * ny * segs;
  int *rb = rangeBegin[wave], *rs = rangeStart[wave];
  int len[2] = {0, 0};
  for (int h = 0; h < 2; ++h) {
    const int r = lane + 64 * h;
    if (r < ranges) {
      const int row = r / segs, seg = r - row * segs;
      const int cx = wrapCell(c3[0] + bx + row / ny, cells[0]);
      const int cy = wrapCell(c3[1] + by + row % ny, cells[1]);
      const int rowBase = (cx * cells[1] + cy) * cells[2];
      int zFirst = seg ? za1 : za0, zEnd = seg ? zb1 : zb0;
      if (RECT) {
        const float gx = gap(0, cx), gy = gap(1, cy), gxy = gx * gx + gy * gy;
        if (gxy > limit2) {
          zEnd = zFirst;
        } else {
          while (zFirst < zEnd) {
            const float gz = gap(2, zFirst);
            if (gxy + gz * gz <= limit2) break;
            ++zFirst;
          }
          while (zEnd > zFirst) {
            const float gz = gap(2, zEnd - 1);
            if (gxy + gz * gz <= limit2) break;
            --zEnd;
          }
        }
      }
      const int begin = cellStart[rowBase + zFirst];
      const int end = cellStart[rowBase + zEnd];
      rb[r] = begin;
      len[h] = end - begin;
    }
  }
  int carry = 0;
  for (int h = 0; h < 2; ++h) {
    int inclusive = len[h];
    for (int offset = 1; offset < 64; offset <<= 1) {
      const int v = __shfl_up(inclusive, offset, 64);
      if (lane >= offset) inclusive += v;
    }
    const int r = lane + 64 * h;
    if (r < ranges) rs[r] = carry + inclusive - len[h];
    carry += __shfl(inclusive, 63, 64);
  }
  if (lane == 0) rs[ranges] = carry;
  __builtin_amdgcn_wave_barrier();
  const int total = carry;
  const int cap = d.cap;
  constexpr int CAND = 4;
  unsigned int required = 0;
  for (int first = cellBegin; first < cellEnd; first += 64) {
    // lane l holds the cell's atom first + l (the last one again past the end)
    const int atoms = min(64, cellEnd - first);
    const int own = first + min(lane, atoms - 1);
    const float4 mine = cellPos[own];
    const int myAtom = __float_as_int(mine.w);
    const int myExBegin = d.exclStart[myAtom], myExEnd = d.exclStart[myAtom + 1];
    const int myExMin = myExBegin < myExEnd ? excl[myExBegin] : 1;
    const int myExMax = myExBegin < myExEnd ? excl[myExEnd - 1] : 0;
    if (lane < atoms) d.posRebuild[myAtom] = d.pos[myAtom];
    int myFound = 0;
    int cur = 0;
    for (int base = 0; base < total; base += 64 * CAND) {
      int slot[CAND];
#pragma unroll
      for (int u = 0; u < CAND; ++u) {
        const int g = base + lane + 64 * u;
        slot[u] = own;  // past the total: an in-bounds load, never accepted
        if (g < total) {
          while (rs[cur + 1] <= g) ++cur;
          slot[u] = rb[cur] + (g - rs[cur]);
        }
      }
      float4 p[CAND];
#pragma unroll
      for (int u = 0; u < CAND; ++u) p[u] = cellPos[slot[u]];
      for (int i = 0; i < atoms; ++i) {
        const float4 pi = make_float4(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.x), i)),
                                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.y), i)),
                                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.z), i)), 0.f);
        const int atom = __builtin_amdgcn_readlane(myAtom, i);
        const int exBegin = __builtin_amdgcn_readlane(myExBegin, i);
        const int exEnd = __builtin_amdgcn_readlane(myExEnd, i);
        const int exMin = __builtin_amdgcn_readlane(myExMin, i);
        const int exMax = __builtin_amdgcn_readlane(myExMax, i);
        int found = __builtin_amdgcn_readlane(myFound, i);
        // a lane past the total holds a real atom of the cell: turned down by
        // its position in the stream, not by the pair test
        bool accept[CAND];
#pragma unroll
        for (int u = 0; u < CAND; ++u)
          accept[u] = (base + lane + 64 * u < total) &
                      acceptPair<RECT>(excl, box, pi, p[u], atom, exBegin, exEnd, exMin, exMax);
        const GlobalOut<int> out = globalOut(d.nbrList + (size_t)(first + i) * cap);
#pragma unroll
        for (int u = 0; u < CAND; ++u) {
          const unsigned long long mask = __ballot(accept[u]);
          if (accept[u]) {
            const int where = found + __popcll(mask & ((1ull << lane) - 1ull));
            if (where < cap) out[where] = slot[u];
          }
          found += __popcll(mask);
        }
        if (lane == i) myFound = found;
      }
    }
    if (lane < atoms) {
      d.nbrCount[own] = myFound < cap ? myFound : cap;
      required = max(required, (unsigned int)myFound);
    }
  }
  for (int offset = 32; offset > 0; offset >>= 1) required = max(required, (unsigned int)__shfl_xor((int)required, offset, 64));
  return required;
}

// A block is four atoms or four cells, as the box's plan says. The
// requirement (the longest list found, stored or not) reaches
// Scalars::required once per workgroup, and not at all where the value the
// workgroup read at its start already covers it: every wave updating the one
// word serialised its atomics (about 60 us of a 2854-atom rebuild's 150).
// A workgroup's waves all come here (no early exit), so the barriers hold.
template <bool RECT>
__device__ __forceinline__ void searchBlock(const Dev &d, int block) {
  if (!d.sc->rebuild) return;  // the same for every wave of the workgroup
  const unsigned int seen = __hip_atomic_load(&d.sc->required, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned int need = d.box->cellSearch ? neighboursByCell<RECT>(d, block) : neighbours<RECT>(d, block);
  __shared__ unsigned int part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = need;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int m = max(max(part[0], part[1]), max(part[2], part[3]));
    if (m > seen) atomicMax(&d.sc->required, m);
  }
  __syncthreads();  // part is reused by a batched workgroup's next block
}

template <bool RECT>
__global__ void __launch_bounds__(256) neighboursKernel(Dev d) {
  searchBlock<RECT>(d, blockIdx.x);
}

// The replicas due a rebuild and their search blocks, densely, for the
// batched search: queue[0] the due replicas m, queue[1] their search blocks
// in all, queue[2 .. 2 + m] the prefix of the blocks (its last entry the
// total), queue[3 + replicas ..] the replica of each. One workgroup scans the
// replicas in order, SEARCH_QUEUE_THREADS at a time.
constexpr int SEARCH_QUEUE_THREADS = 1024;

__global__ void __launch_bounds__(SEARCH_QUEUE_THREADS) searchQueueBatchKernel(const BatchView *__restrict__ v,
                                                                                int replicas, int *__restrict__ queue) {
  // (due replicas << 32) + search blocks, scanned together
  __shared__ long long partial[SEARCH_QUEUE_THREADS];
  int *start = queue + 2, *replica = queue + 3 + replicas;
  const int t = threadIdx.x;
  int dueBase = 0, blockBase = 0;
  for (int r0 = 0; r0 < replicas; r0 += SEARCH_QUEUE_THREADS) {
    const int r = r0 + t;
    long long mine = 0;
    if (r < replicas && v[r].d.sc->rebuild) {
      const BoxF &box = *v[r].d.box;
      const int blocks = box.cellSearch ? (box.cellTotal + 3) / 4 : (v[r].d.n + 3) / 4;
      mine = (1LL << 32) + blocks;
    }
    partial[t] = mine;
    __syncthreads();
    for (int offset = 1; offset < SEARCH_QUEUE_THREADS; offset <<= 1) {
      const long long add = t >= offset ? partial[t - offset] : 0;
      __syncthreads();
      partial[t] += add;
      __syncthreads();
    }
    const long long before = partial[t] - mine, sum = partial[SEARCH_QUEUE_THREADS - 1];
    if (mine) {
      const int at = dueBase + (int)(before >> 32);
      start[at] = blockBase + (int)(before & 0xffffffffLL);
      replica[at] = r;
    }
    __syncthreads();  // the next chunk reuses partial
    dueBase += (int)(sum >> 32);
    blockBase += (int)(sum & 0xffffffffLL);
  }
  if (t == 0) {
    queue[0] = dueBase;
    queue[1] = blockBase;
    start[dueBase] = blockBase;
  }
}

// Batched search over the queue: the due replicas' blocks, in replica order,
// are cut into eight equal ranges, one per XCD (workgroup b runs on XCD
// b % 8, and each XCD has its own L2), so a replica's neighbouring cells are
// searched through one L2 and every XCD has the same share whichever
// replicas are due. A workgroup strides through its XCD's range and finds
// each block's replica by bisection of the prefix, staged in LDS.
// (Seven waves per SIMD: the register count the search body needs, without
// spills.)
template <bool RECT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
neighboursBatchKernel(const BatchView *__restrict__ v, const int *__restrict__ queue, int replicas) {
  extern __shared__ int queueLds[];  // the prefix (due + 1 entries), then the replicas (due)
  const int due = queue[0], total = queue[1];
  const int xcd = blockIdx.x % XCD_COUNT, perXcd = gridDim.x / XCD_COUNT;
  const int first = (int)((long long)total * xcd / XCD_COUNT), last = (int)((long long)total * (xcd + 1) / XCD_COUNT);
  int k = first + (int)(blockIdx.x / XCD_COUNT);
  if (k >= last) return;
  int *start = queueLds, *replica = queueLds + due + 1;
  for (int i = threadIdx.x; i <= due; i += blockDim.x) start[i] = queue[2 + i];
  for (int i = threadIdx.x; i < due; i += blockDim.x) replica[i] = queue[3 + replicas + i];
  __syncthreads();
  for (; k < last; k += perXcd) {
    int lo = 0, hi = due - 1;  // the last entry whose prefix is at most k
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= k)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int r = __builtin_amdgcn_readfirstlane(replica[lo]);
    const int block = __builtin_amdgcn_readfirstlane(k - start[lo]);
    searchBlock<RECT>(v[r].d, block);
  }
}

void launchForceRebuildFlag(const Dev &d, hipStream_t s) { setRebuildKernel<<<1, 1, 0, s>>>(d.sc); }

// The slot copy from the positions through the current cell order, for an
// evaluation that did not follow an integrator step (which keeps it current).
__global__ void refreshSlotsKernel(Dev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < d.n) d.slotPos[k] = d.pos[d.cellAtoms[k]];
}

__global__ void __launch_bounds__(256) refreshSlotsBatchKernel(const BatchView *__restrict__ v,
                                                               const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  const Dev &d = v[t.x].d;
  const int k = t.y * blockDim.x + threadIdx.x;
  if (k < d.n) d.slotPos[k] = d.pos[d.cellAtoms[k]];
}

void launchRefreshSlots(const Dev &d, hipStream_t s) { refreshSlotsKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d); }

void launchRefreshSlotsBatch(const BatchView *v, const TaskTable &atoms256, hipStream_t s) {
  refreshSlotsBatchKernel<<<atoms256.blocks, 256, 0, s>>>(v, atoms256.task);
}

void launchRebuild(const Dev &d, hipStream_t s) {
  if (d.n <= SMALL_MAX_ATOMS && d.cellCapacity <= SMALL_MAX_CELLS) {
    cellsBuildSmallKernel<<<1, SMALL_THREADS, 0, s>>>(d);
  } else {
    const int blocks = (d.n + 255) / 256;
    cellsAssignKernel<<<blocks, 256, 0, s>>>(d);
    cellsScanKernel<<<1, 1024, 0, s>>>(d);
    cellsScatterKernel<<<blocks, 256, 0, s>>>(d);
    cellsSortKernel<<<d.cellCapacity, 256, 0, s>>>(d);
  }
  if (d.rect)
    neighboursKernel<true><<<(d.n + 3) / 4, 256, 0, s>>>(d);
  else
    neighboursKernel<false><<<(d.n + 3) / 4, 256, 0, s>>>(d);
}

// Batched rebuild: every replica's cell list in one workgroup of its own (a
// batch takes replicas the single-workgroup build fits), the queue of the due
// replicas' search blocks, then the lists.
void launchRebuildBatch(const BatchView *v, const TaskTable &replicaTable, int replicas, int *queue, int grid,
                        int rect, hipStream_t s) {
  cellsBuildSmallBatchKernel<<<replicaTable.blocks, SMALL_THREADS, 0, s>>>(v, replicaTable.task);
  searchQueueBatchKernel<<<1, SEARCH_QUEUE_THREADS, 0, s>>>(v, replicas, queue);
  const size_t lds = (size_t)(2 * replicas + 1) * sizeof(int);
  if (rect)
    neighboursBatchKernel<true><<<grid, 256, lds, s>>>(v, queue, replicas);
  else
    neighboursBatchKernel<false><<<grid, 256, lds, s>>>(v, queue, replicas);
}

int searchQueueInts(int replicas) { return 3 + 2 * replicas; }

bool fitsSmallCellBuild(int atoms, int cells) { return atoms <= SMALL_MAX_ATOMS && cells <= SMALL_MAX_CELLS; }

// The per-cell search runs in the per-atom search's launch (4 waves per
// workgroup, one workgroup per 4 atoms), so it covers at most that many cells.
static bool cellLaunchCovers(int atoms, int cells) { return cells > 0 && cells <= 4 * ((atoms + 3) / 4); }

bool cellSearchFits(int atoms, int cells) {
  constexpr int MAX_AVERAGE = 32;  // atoms per cell
  return cellLaunchCovers(atoms, cells) && atoms <= MAX_AVERAGE * cells;
}

bool searchByCell(int mode, int atoms, int cells) {
  if (mode == 1) return false;
  if (mode == 2) return cellLaunchCovers(atoms, cells);
  return cellSearchFits(atoms, cells);
}

}  // namespace rmd

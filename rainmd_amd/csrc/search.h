// The cell assignment and the neighbour-list search bodies of the list
// kernels (neighbours.hip), for any workgroup of WAVES waves. FP32 with a
// fixed operation order; multiply-adds fuse only within one expression, so
// every instantiation rounds the same way.
#pragma once
#include "kernels.h"

#pragma clang fp contract(on)

namespace rmd {

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
__device__ __forceinline__ void writeSlotLoaded(const Dev &d, int k, int atom, float4 p, int type) {
  d.cellAtoms[k] = atom;
  d.cellPos[k] = make_float4(p.x, p.y, p.z, __int_as_float(atom));
  d.slotPos[k] = p;
  d.slotType[k] = type;
  d.slotOf[atom] = k;
}

__device__ __forceinline__ void writeSlot(const Dev &d, int k, int atom) {
  writeSlotLoaded(d, k, atom, d.pos[atom], d.type[atom]);
}

// The cell list of a system whose plan has at most FEW_MAX_CELLS cells, by
// one workgroup of THREADS threads: assignment with LDS-atomic counts, the
// exclusive scan by one wave, then the atoms placed in index order THREADS at
// a time. Each wave splits its 64 atoms by cell with one ballot per cell (an
// atom's place among its wave's atoms of the same cell is the count of lower
// lanes in its cell's mask); one thread per cell then turns the waves' counts
// into each wave's start in the cell, in wave order, and advances the cell's
// cursor. Few cells means crowded cells, where a per-cell sort would be long.
constexpr int FEW_MAX_CELLS = 64;

template <int THREADS>
__device__ void cellsBuildFew(const Dev &d) {
  __shared__ int cursor[FEW_MAX_CELLS];
  __shared__ int waveStart[THREADS / 64][FEW_MAX_CELLS];
  const BoxF &box = *d.box;
  const int total = box.cellTotal, n = d.n, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t < FEW_MAX_CELLS) cursor[t] = 0;
  __syncthreads();
  for (int i = t; i < n; i += THREADS) {
    const float4 p = d.pos[i];
    int c3[3];
    const int c = cellIndex(p.x, p.y, p.z, box, c3);
    d.cellOfAtom[i] = c;
    atomicAdd(&cursor[c], 1);
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicMax(&d.sc->nonFinite, 1u);
  }
  __syncthreads();
  if (wave == 0) {
    const int count = lane < total ? cursor[lane] : 0;
    int inclusive = count;
    for (int offset = 1; offset < 64; offset <<= 1) {
      const int v = __shfl_up(inclusive, offset, 64);
      if (lane >= offset) inclusive += v;
    }
    if (lane < total) {
      cursor[lane] = inclusive - count;
      d.cellStart[lane] = inclusive - count;
    }
    if (lane == 63) d.cellStart[total] = inclusive;
  }
  __syncthreads();
  for (int base = 0; base < n; base += THREADS) {
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
      int at = cursor[t];
      for (int w = 0; w < THREADS / 64; ++w) {
        const int m = waveStart[w][t];
        waveStart[w][t] = at;
        at += m;
      }
      cursor[t] = at;
    }
    __syncthreads();
    if (i < n) writeSlot(d, waveStart[wave][c] + __popcll(same & ((1ull << lane) - 1ull)), i);
    __syncthreads();
  }
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
template <bool RECT, int WAVES>
__device__ __forceinline__ unsigned int neighbours(const Dev &d, int block) {
  __shared__ int rangeBegin[WAVES][STENCIL_RANGES];
  __shared__ int rangeStart[WAVES][STENCIL_RANGES + 1];
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
template <bool RECT, int WAVES>
__device__ __forceinline__ unsigned int neighboursByCell(const Dev &d, int block) {
  __shared__ int rangeBegin[WAVES][STENCIL_RANGES];
  __shared__ int rangeStart[WAVES][STENCIL_RANGES + 1];
  const BoxF box = uniformPairBox(*d.box, false);
  const int cells[3] = {d.box->cells[0], d.box->cells[1], d.box->cells[2]};
  const GlobalPtr<float4> cellPos = globalPtr(d.cellPos);
  const GlobalPtr<int> cellStart = globalPtr(d.cellStart);
  const GlobalPtr<int> excl = globalPtr(d.exclPartners);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cell = block * WAVES + wave;
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

// A block is WAVES atoms or WAVES cells (one per wave of the workgroup), as
// the box's plan says. The
// requirement (the longest list found, stored or not) reaches
// Scalars::required once per workgroup, and not at all where the value the
// workgroup read at its start already covers it: every wave updating the one
// word serialised its atomics (about 60 us of a 2854-atom rebuild's 150).
// A workgroup's waves all come here (no early exit), so the barriers hold.
template <bool RECT, int WAVES = 4>
__device__ __forceinline__ void searchBlock(const Dev &d, int block) {
  if (!d.sc->rebuild) return;  // the same for every wave of the workgroup
  const unsigned int seen = __hip_atomic_load(&d.sc->required, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned int need =
      d.box->cellSearch ? neighboursByCell<RECT, WAVES>(d, block) : neighbours<RECT, WAVES>(d, block);
  __shared__ unsigned int part[WAVES];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = need;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int m = 0u;
    for (int w = 0; w < WAVES; ++w) m = max(m, part[w]);
    if (m > seen) atomicMax(&d.sc->required, m);
  }
  __syncthreads();  // part is reused by a batched workgroup's next block
}

}  // namespace rmd

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
#include "search.h"

// Multiply-adds fuse only within one expression, so each body rounds the same
// way wherever it is inlined (the single-system and the batched kernels).
#pragma clang fp contract(on)

namespace rmd {

constexpr int SORT_CAPACITY = 4096;  // atoms per cell the LDS sort takes

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

// Exclusive scan of the per-cell counts in one workgroup, SCAN_TILE cells at
// a time. A tile's counts are loaded coalesced into LDS, every load of a
// thread in flight together (a thread reading its own contiguous run from
// global memory waited one full latency per cell). Each thread then sums its
// run of SCAN_RUN cells in LDS; the runs' sums are scanned across the
// workgroup; each thread writes its run's starts back to LDS; the tile's
// starts go out coalesced. Leaves the counts zeroed for the next rebuild and
// the scatter cursors at the cell starts.
constexpr int SCAN_THREADS = 1024, SCAN_RUN = 8, SCAN_TILE = SCAN_THREADS * SCAN_RUN;

// LDS index of tile entry k, one pad word per 32 so the threads' runs start
// in different banks
__device__ __forceinline__ int scanPad(int k) { return k + (k >> 5); }

__global__ void __launch_bounds__(SCAN_THREADS) cellsScanKernel(Dev d) {
  if (!d.sc->rebuild) return;
  __shared__ int tile[SCAN_TILE + SCAN_TILE / 32];
  __shared__ int partial[SCAN_THREADS];
  const int total = d.box->cellTotal, t = threadIdx.x;
  int carry = 0;
  for (int base = 0; base < total; base += SCAN_TILE) {
    const int here = min(SCAN_TILE, total - base);
    int v[SCAN_RUN];
#pragma unroll
    for (int u = 0; u < SCAN_RUN; ++u) {
      const int k = t + SCAN_THREADS * u;
      v[u] = k < here ? d.cellCount[base + k] : 0;
    }
#pragma unroll
    for (int u = 0; u < SCAN_RUN; ++u) tile[scanPad(t + SCAN_THREADS * u)] = v[u];
    __syncthreads();
    int run = 0;
#pragma unroll
    for (int u = 0; u < SCAN_RUN; ++u) run += tile[scanPad(t * SCAN_RUN + u)];
    partial[t] = run;
    __syncthreads();
    for (int offset = 1; offset < SCAN_THREADS; offset <<= 1) {
      const int add = t >= offset ? partial[t - offset] : 0;
      __syncthreads();
      partial[t] += add;
      __syncthreads();
    }
    const int tileTotal = partial[SCAN_THREADS - 1];
    int at = carry + partial[t] - run;
#pragma unroll
    for (int u = 0; u < SCAN_RUN; ++u) {
      const int k = scanPad(t * SCAN_RUN + u);
      const int count = tile[k];
      tile[k] = at;
      at += count;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SCAN_RUN; ++u) {
      const int k = t + SCAN_THREADS * u;
      if (k < here) {
        const int start = tile[scanPad(k)];
        d.cellStart[base + k] = start;
        d.cellCursor[base + k] = start;
        d.cellCount[base + k] = 0;
      }
    }
    carry += tileTotal;
    __syncthreads();  // the next tile reuses tile and partial
  }
  if (t == 0) d.cellStart[total] = carry;
}

__global__ void cellsScatterKernel(Dev d) {
  if (!d.sc->rebuild) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int slot = atomicAdd(&d.cellCursor[d.cellOfAtom[i]], 1);
  d.cellAtoms[slot] = i;
}

// The scatter leaves each cell's atoms in arrival order; sorting each cell's
// run puts them in index order so the list, and every sum over it, is the
// same on every run. Sorted cells also get their slots written (positions in
// cell order, with the atom index in .w, for the candidate scan). Cells of at
// most SHORT_CELL atoms - every cell of a plan whose cells are a fraction of
// the list cutoff - are sorted by one thread each, by insertion in that
// thread's slice of LDS; longer ones by a bitonic sort over a workgroup
// (cellsSortLongKernel).
constexpr int SHORT_CELL = 32;
constexpr int SHORT_SORT_THREADS = 128;

__global__ void __launch_bounds__(SHORT_SORT_THREADS) cellsSortShortKernel(Dev d) {
  if (!d.sc->rebuild) return;
  __shared__ int keys[SHORT_SORT_THREADS][SHORT_CELL + 1];  // +1: the threads' slices start in different banks
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.box->cellTotal) return;
  const int begin = d.cellStart[c], len = d.cellStart[c + 1] - begin;
  if (len > SHORT_CELL) return;
  int *k = keys[threadIdx.x];
  // the cell's atoms loaded together, then sorted in LDS
  for (int a = 0; a < len; ++a) k[a] = d.cellAtoms[begin + a];
  for (int a = 1; a < len; ++a) {
    const int v = k[a];
    int m = a - 1;
    while (m >= 0 && k[m] > v) {
      k[m + 1] = k[m];
      --m;
    }
    k[m + 1] = v;
  }
  // positions and types SLOT_BATCH at a time, loaded before any slot is
  // written (a slot's stores would otherwise wait on its own loads, one
  // round trip per atom)
  constexpr int SLOT_BATCH = 4;
  for (int a = 0; a < len; a += SLOT_BATCH) {
    float4 p[SLOT_BATCH];
    int type[SLOT_BATCH];
#pragma unroll
    for (int u = 0; u < SLOT_BATCH; ++u) {
      const int atom = k[min(a + u, len - 1)];
      p[u] = d.pos[atom];
      type[u] = d.type[atom];
    }
#pragma unroll
    for (int u = 0; u < SLOT_BATCH; ++u)
      if (a + u < len) writeSlotLoaded(d, begin + a + u, k[a + u], p[u], type[u]);
  }
}

__global__ void __launch_bounds__(256) cellsSortLongKernel(Dev d) {
  if (!d.sc->rebuild) return;
  __shared__ int keys[SORT_CAPACITY];
  for (int c = blockIdx.x; c < d.box->cellTotal; c += gridDim.x) {
    const int begin = d.cellStart[c], end = d.cellStart[c + 1];
    const int len = end - begin;
    if (len <= SHORT_CELL) continue;
    if (len > SORT_CAPACITY) {
      if (threadIdx.x == 0) atomicMax(&d.sc->cellOverflow, 1u);
      continue;
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
    for (int k = threadIdx.x; k < len; k += blockDim.x) writeSlot(d, begin + k, keys[k]);
    __syncthreads();  // keys is reused for the workgroup's next cell
  }
}

// The whole cell list of a small system in one workgroup: assignment with
// LDS-atomic counts, the scan, then each cell's atoms in index order - by
// wave ballots where there are at most FEW_MAX_CELLS cells (cellsBuildFew),
// else an LDS scatter and a sort of each cell's run (insertion by one thread
// per cell, or where a cell is crowded a rank sort: rank = atoms of the same
// cell with a lower index) - the same lists as the four-kernel path, with no
// global atomics and three launches fewer on every step. Used when the atoms
// and the cells fit the LDS arrays below.
constexpr int SMALL_THREADS = 1024;
constexpr int SMALL_MAX_ATOMS = 16384;
constexpr int SMALL_MAX_CELLS = 4096;
constexpr int INSERTION_MAX = 32;  // cells up to this many atoms are sorted by one thread each

__device__ __forceinline__ void cellsBuildSmall(const Dev &d) {
  if (!d.sc->rebuild) return;
  if (d.box->cellTotal <= FEW_MAX_CELLS) {
    cellsBuildFew<SMALL_THREADS>(d);
    return;
  }
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
__global__ void __launch_bounds__(256) refreshSlotsBatchKernel(const BatchView *__restrict__ v,
                                                               const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  const Dev &d = v[t.x].d;
  const int k = t.y * blockDim.x + threadIdx.x;
  if (k < d.n) d.slotPos[k] = d.pos[d.cellAtoms[k]];
}

void launchRefreshSlotsBatch(const BatchView *v, const TaskTable &atoms256, hipStream_t s) {
  refreshSlotsBatchKernel<<<atoms256.blocks, 256, 0, s>>>(v, atoms256.task);
}

bool integratorBuildsCells(const Dev &d, int cellCount) {
  return d.n <= SMALL_MAX_ATOMS && d.cellCapacity <= SMALL_MAX_CELLS && cellCount <= FEW_MAX_CELLS && !splitWater(d);
}

void launchRebuild(const Dev &d, hipStream_t s, bool cellsBuilt, bool searched) {
  if (d.n <= SMALL_MAX_ATOMS && d.cellCapacity <= SMALL_MAX_CELLS) {
    if (!cellsBuilt) cellsBuildSmallKernel<<<1, SMALL_THREADS, 0, s>>>(d);
  } else {
    const int blocks = (d.n + 255) / 256;
    cellsAssignKernel<<<blocks, 256, 0, s>>>(d);
    cellsScanKernel<<<1, SCAN_THREADS, 0, s>>>(d);
    cellsScatterKernel<<<blocks, 256, 0, s>>>(d);
    cellsSortShortKernel<<<(d.cellCapacity + SHORT_SORT_THREADS - 1) / SHORT_SORT_THREADS, SHORT_SORT_THREADS, 0, s>>>(d);
    cellsSortLongKernel<<<std::min(d.cellCapacity, 1024), 256, 0, s>>>(d);
  }
  if (searched) return;
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

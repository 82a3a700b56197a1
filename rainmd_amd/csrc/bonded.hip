// Bonded terms, exception pairs and positional restraints in one launch.
//
// Replaces bondedShader (webgpu/kernels/bonded.js:30-166), exceptionShader
// (webgpu/kernels/nonbonded.js:128-200) and restraintShader
// (webgpu/kernels/restraints.js:29-63). The reference inverts the term
// lists so each thread owns one atom and needs no atomics; here each thread
// owns one TERM and adds its forces to int64 fixed-point accumulators, which
// commute, so the result is identical on every run and there is no
// recomputation per role.
//
// Everything here is evaluated in FP64: a solute's few hundred bonded terms
// (the stretch r - r0 of a stiff bond cancels badly in FP32) and about one
// exception pair per atom; the cost is negligible next to the pair list.
#include "bonded_terms.h"

namespace rmd {

__global__ void __launch_bounds__(256) bondedKernel(Dev d, double alphaD, int terms, int energies) {
  bonded(d, alphaD, terms, energies, blockIdx.x);
}

__global__ void __launch_bounds__(256) bondedBatchKernel(const BatchView *__restrict__ v, const int2 *__restrict__ task,
                                                         double alphaD, int terms, int energies) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  bonded(v[t.x].d, alphaD, terms, energies, t.y);
}

int bondedBlocks(const Dev &d) {
  return (bondedThreads(d) + 255) / 256;
}

void launchBondedBatch(const BatchView *v, const TaskTable &t, double alpha, int terms, int energies, hipStream_t s) {
  if (t.blocks > 0) bondedBatchKernel<<<t.blocks, 256, 0, s>>>(v, t.task, alpha, terms, energies);
}

int launchBonded(const Dev &d, double alpha, int terms, int energies, hipStream_t s, int *slotsUsed) {
  const int blocks = bondedBlocks(d);
  if (blocks > 0) bondedKernel<<<blocks, 256, 0, s>>>(d, alpha, terms, energies);
  if (slotsUsed) *slotsUsed = blocks;
  return blocks;
}

}  // namespace rmd

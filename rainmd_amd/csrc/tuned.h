// Kernel shapes picked by the autotuner (rainmd_amd/autotune.py), which
// writes this file from its measurements (rainmd_amd/tuning.json).
// Regenerate with: python -m rainmd_amd.autotune --write
#pragma once

namespace rmd {

// One row per workload class: single systems (kind 0) by atom count, replica
// batches (kind 1). lanes: direct-space lanes per atom; reach: the widest cell
// stencil reach the plan may take; searchMode: 0 by cell occupancy, 1 one wave
// per atom, 2 one wave per cell; skin: nm when the caller asks for the tuned
// skin (0: the reference's skinFor); pmeBins: 0 by atom count, 1 atom bins,
// 2 every plane scans every atom.
struct TunedShape {
  int kind, minAtoms, lanes, reach, searchMode;
  double skin;
  int pmeBins;
};

constexpr TunedShape TUNED_SHAPES[] = {
    {0, 0, 64, 2, 1, 0.0, 2},
    {0, 2000, 32, 3, 1, 0.4, 0},
    {0, 6000, 16, 3, 2, 0.0, 1},
    {0, 40000, 16, 3, 2, 0.15, 0},
    {1, 0, 16, 3, 0, 0.12, 0},
};

}  // namespace rmd

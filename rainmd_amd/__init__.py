"""rainmd_amd: the MI355X-native molecular-dynamics step behind RainMD's surface.

The product is librainmd_amd.so (HIP kernels for gfx950 plus the host setup,
C ABI in include/rainmd.h). This package is the Python mirror of the
reference's host surface over that ABI.
"""
from .simulation import (  # noqa: F401
    ConstraintFailure, Coordinates, DeviceError, GpuBudgetError, GpuSimulation, NeighbourOverflowError,
    NonFiniteStateError, ParseError, RainmdError, SimulationStopped, System, SystemError_, Topology,
    WORKLOADS, box_vectors_from_lengths_and_angles, device_info, pme_moduli, pme_parameters, read_data,
    workload_system, maxwell_boltzmann, uniform_at, Surrogate, Batch, replica_workload, replica_systems, tuned_shape,
)

"""ctypes declarations for librainmd_amd.so (include/rainmd.h).

The library is loaded from this package directory only; if it is missing the
import fails loudly (build it with ``python -m rainmd_amd.build`` or
``__graft_entry__.build()``). There is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(HERE, "librainmd_amd.so")

RMD_OK = 0
RMD_E_INVALID = -1
RMD_E_PARSE = -2
RMD_E_SYSTEM = -3
RMD_E_DEVICE = -4
RMD_E_NONFINITE = -5
RMD_E_NEIGHBOUR_OVERFLOW = -6
RMD_E_STOPPED = -7
RMD_E_BUDGET = -8
RMD_E_CONSTRAINT = -9

CONSTRAINTS = {"none": 0, "HBonds": 1, "AllBonds": 2}

TERM_BONDED, TERM_DIRECT, TERM_EXCEPTIONS, TERM_RECIPROCAL, TERM_RESTRAINT = 1, 2, 4, 8, 16
TERM_ALL = 31

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class CoordinatesView(C.Structure):
    _fields_ = [
        ("atom_count", C.c_int32),
        ("positions", _dp),
        ("velocities", _dp),
        ("has_box", C.c_int32),
        ("box_lengths", C.c_double * 3),
        ("box_angles", C.c_double * 3),
        ("time", C.c_double),
    ]


class BuildOptions(C.Structure):
    _fields_ = [
        ("constraints", C.c_int32),
        ("rigid_water", C.c_int32),
        ("cutoff", C.c_double),
        ("has_box_vectors", C.c_int32),
        ("box_vectors", C.c_double * 9),
        ("has_box", C.c_int32),
        ("box_lengths", C.c_double * 3),
        ("box_angles", C.c_double * 3),
    ]


class SolvatedOptions(C.Structure):
    _fields_ = [
        ("water_count", C.c_int32),
        ("box_edge", C.c_double),
        ("seed", C.c_uint32),
        ("include_solute", C.c_int32),
        ("solute_clearance", C.c_double),
        ("jitter", C.c_double),
        ("cutoff", C.c_double),
        ("constraints", C.c_int32),
        ("rigid_water", C.c_int32),
    ]


class SystemView(C.Structure):
    _fields_ = [
        ("atom_count", C.c_int32), ("type_count", C.c_int32), ("molecule_count", C.c_int32),
        ("residue_count", C.c_int32),
        ("charges", _dp), ("masses", _dp),
        ("atom_types", _ip), ("atomic_numbers", _ip), ("residue_of_atom", _ip),
        ("lj_a", _dp), ("lj_b", _dp),
        ("bond_count", C.c_int32), ("bond_a", _ip), ("bond_b", _ip), ("bond_k", _dp), ("bond_r0", _dp),
        ("angle_count", C.c_int32), ("angle_a", _ip), ("angle_b", _ip), ("angle_c", _ip),
        ("angle_k", _dp), ("angle_theta0", _dp),
        ("torsion_count", C.c_int32), ("torsion_a", _ip), ("torsion_b", _ip), ("torsion_c", _ip),
        ("torsion_d", _ip), ("torsion_k", _dp), ("torsion_periodicity", _dp), ("torsion_phase", _dp),
        ("constraint_count", C.c_int32), ("constraint_a", _ip), ("constraint_b", _ip),
        ("constraint_distance", _dp),
        ("exception_count", C.c_int32), ("exception_a", _ip), ("exception_b", _ip),
        ("exception_charge_product", _dp), ("exception_lj_a", _dp), ("exception_lj_b", _dp),
        ("exclusion_total", C.c_int32), ("exclusion_start", _ip), ("exclusion_partners", _ip),
        ("molecule_start", _ip), ("molecule_atoms", _ip), ("molecule_of_atom", _ip),
        ("is_periodic", C.c_int32),
        ("box", C.c_double * 9),
        ("cutoff", C.c_double),
        ("degrees_of_freedom", C.c_int32),
        ("constraint_mode", C.c_int32), ("rigid_water", C.c_int32),
        ("positions", _dp), ("velocities", _dp),
        ("water_count", C.c_int32),
    ]


class PmeParameters(C.Structure):
    _fields_ = [("alpha", C.c_double), ("grid", C.c_int32 * 3), ("order", C.c_int32), ("tolerance", C.c_double)]


class Options(C.Structure):
    _fields_ = [
        ("timestep", C.c_double),
        ("temperature", C.c_double),
        ("friction", C.c_double),
        ("seed", C.c_uint32),
        ("tolerance", C.c_double),
        ("max_constraint_iterations", C.c_int32),
        ("rebuild_interval", C.c_int32),
        ("adaptive_rebuild", C.c_int32),
        ("has_pressure", C.c_int32),
        ("pressure", C.c_double),
        ("barostat_interval", C.c_int32),
        ("dispersion_correction", C.c_int32),
        ("restraint_force", C.c_double),
        ("restraint_count", C.c_int32),
        ("restraint_atoms", _ip),
        ("restraint_reference", _dp),
        ("skin", C.c_double),
        ("neighbour_slack", C.c_double),
        ("max_neighbours", C.c_int32),
        ("ewald_tolerance", C.c_double),
        ("alpha", C.c_double),
        ("grid", C.c_int32 * 3),
        ("device", C.c_int32),
        ("lanes_per_atom", C.c_int32),
        ("cell_reach", C.c_int32),
        ("search_mode", C.c_int32),
        ("pme_bins", C.c_int32),
    ]


class EnergyTerms(C.Structure):
    _fields_ = [
        ("step", C.c_int64),
        ("time", C.c_double),
        ("bonded", C.c_double), ("harmonic_bond", C.c_double), ("harmonic_angle", C.c_double),
        ("periodic_torsion", C.c_double),
        ("lennard_jones", C.c_double),
        ("dispersion", C.c_double),
        ("coulomb_direct", C.c_double),
        ("coulomb_reciprocal", C.c_double),
        ("coulomb_self", C.c_double),
        ("coulomb", C.c_double),
        ("position_restraint", C.c_double),
        ("restrained_atoms", C.c_int32),
        ("potential", C.c_double), ("kinetic", C.c_double), ("total", C.c_double),
        ("temperature", C.c_double), ("kinetic_for_temperature", C.c_double),
        ("kinetic_sampled_in_step", C.c_int32),
        ("volume", C.c_double),
        ("box", C.c_double * 9),
        ("dscale_direct", C.c_double), ("dscale_reciprocal", C.c_double),
        ("dscale_molecular", C.c_double),
    ]


class Move(C.Structure):
    _fields_ = [
        ("attempted", C.c_int32), ("accepted", C.c_int32), ("refused", C.c_int32),
        ("weight", C.c_double), ("delta_volume", C.c_double), ("length_scale", C.c_double),
        ("old_potential", C.c_double), ("new_potential", C.c_double), ("volume", C.c_double),
        ("volume_scale", C.c_double), ("acceptance_rate", C.c_double),
    ]


class MinOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("force_tolerance", C.c_double), ("max_step", C.c_double),
                ("max_displacement", C.c_double)]


class MinResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("rejected", C.c_int32), ("stalled", C.c_int32),
                ("converged", C.c_int32), ("potential", C.c_double), ("max_force", C.c_double),
                ("max_step", C.c_double), ("energy_resolution", C.c_double)]


class Diagnostics(C.Structure):
    _fields_ = [
        ("max_shake_iterations", C.c_int32), ("shake_hit_iteration_cap", C.c_int32),
        ("rattle_hit_iteration_cap", C.c_int32),
        ("neighbour_required", C.c_int32), ("neighbour_capacity", C.c_int32),
        ("worst_neighbour_required", C.c_int32), ("neighbour_growths", C.c_int32),
        ("skin", C.c_double), ("worst_displacement", C.c_double),
        ("rebuilds", C.c_int64), ("steps", C.c_int64),
        ("cell_counts", C.c_int32 * 3),
        ("volume", C.c_double),
        ("barostat_attempts", C.c_int32), ("barostat_accepted", C.c_int32), ("barostat_refused", C.c_int32),
        ("volume_scale", C.c_double),
        ("constraint_tolerance", C.c_double),
        ("pme_grid", C.c_int32 * 3),
        ("pme_alpha", C.c_double),
        ("lanes_per_atom", C.c_int32),
        ("water_units", C.c_int32), ("cluster_units", C.c_int32), ("free_units", C.c_int32),
        ("pme_planes", C.c_int32),
        ("cell_reach", C.c_int32), ("search_by_cell", C.c_int32),
        ("steps_since_barostat", C.c_int64),
        ("pme_x_folded", C.c_int32),
    ]


class TunedShape(C.Structure):
    _fields_ = [("lanes_per_atom", C.c_int32), ("cell_reach", C.c_int32), ("search_mode", C.c_int32),
                ("skin", C.c_double), ("pme_bins", C.c_int32)]


class ScreenOptions(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("training_moves", C.c_int32), ("training_steps", C.c_int32),
                ("learning_rate", C.c_float), ("seed", C.c_uint32)]


class ScreenStats(C.Structure):
    _fields_ = [("phase", C.c_int32), ("samples", C.c_int64), ("moves", C.c_int64),
                ("exact_evaluations", C.c_int64), ("screened_out", C.c_int64),
                ("second_stage_rejections", C.c_int64), ("accepted", C.c_int64),
                ("training_loss", C.c_double), ("residual_rms", C.c_double), ("estimate_rms", C.c_double)]


SURROGATE_INPUTS = 16
SURROGATE_PARAMS = 5313


class DeviceInfo(C.Structure):
    _fields_ = [("device_count", C.c_int32), ("name", C.c_char * 128), ("arch", C.c_char * 64),
                ("compute_units", C.c_int32), ("global_memory", C.c_int64)]


# Every function the header declares, with its ctypes signature.
class BatchInfo(C.Structure):
    _fields_ = [("replicas", C.c_int32), ("atoms", C.c_int64), ("blocks_run", C.c_int64),
                ("step_workgroups", C.c_int64), ("nonbonded_workgroups", C.c_int64)]


SIGNATURES = {
    "rmd_abi_version": (C.c_int32, []),
    "rmd_last_global_error": (C.c_char_p, []),
    "rmd_topology_parse_prmtop": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "rmd_topology_free": (None, [C.c_void_p]),
    "rmd_coordinates_parse_inpcrd": (C.c_int, [C.c_char_p, C.c_size_t, C.c_int32, C.POINTER(C.c_void_p)]),
    "rmd_coordinates_free": (None, [C.c_void_p]),
    "rmd_coordinates_view_get": (C.c_int, [C.c_void_p, C.POINTER(CoordinatesView)]),
    "rmd_box_vectors_from_lengths_and_angles": (C.c_int, [_dp, _dp, _dp]),
    "rmd_build_options_default": (None, [C.POINTER(BuildOptions)]),
    "rmd_system_build": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BuildOptions), C.POINTER(C.c_void_p)]),
    "rmd_system_free": (None, [C.c_void_p]),
    "rmd_solvated_options_default": (None, [C.POINTER(SolvatedOptions)]),
    "rmd_system_build_solvated": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                            C.POINTER(SolvatedOptions), C.POINTER(C.c_void_p)]),
    "rmd_system_view_get": (C.c_int, [C.c_void_p, C.POINTER(SystemView)]),
    "rmd_system_rebox": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_void_p)]),
    "rmd_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Options), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_void_p)]),
    "rmd_batch_destroy": (None, [C.c_void_p]),
    "rmd_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "rmd_batch_replica": (C.c_void_p, [C.c_void_p, C.c_int32]),
    "rmd_batch_info_get": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "rmd_batch_initialise": (C.c_int, [C.c_void_p]),
    "rmd_batch_evaluate": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_batch_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    "rmd_batch_request_stop": (C.c_int, [C.c_void_p]),
    "rmd_batch_set_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_batch_set_concurrent_forces": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_batch_profile_nonbonded": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "rmd_batch_screen_configure": (C.c_int, [C.c_void_p, C.POINTER(ScreenOptions)]),
    "rmd_batch_screen_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ScreenStats)]),
    "rmd_system_solute_heavy_atoms": (C.c_int, [C.c_void_p, _ip, C.c_int32, C.POINTER(C.c_int32)]),
    "rmd_pme_parameters_derive": (C.c_int, [_dp, C.c_double, C.c_double, C.c_double, _ip,
                                            C.POINTER(PmeParameters)]),
    "rmd_pme_moduli": (C.c_int, [C.c_int32, C.c_int32, _dp]),
    "rmd_skin_for": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    "rmd_dispersion_energy": (C.c_double, [C.c_void_p, C.c_double]),
    "rmd_options_default": (None, [C.POINTER(Options)]),
    "rmd_create": (C.c_int, [C.c_void_p, C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "rmd_destroy": (None, [C.c_void_p]),
    "rmd_last_error": (C.c_char_p, [C.c_void_p]),
    "rmd_set_state": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int64, C.c_int32]),
    "rmd_set_noise": (C.c_int, [C.c_void_p, _fp]),
    "rmd_initialise": (C.c_int, [C.c_void_p]),
    "rmd_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    "rmd_request_stop": (C.c_int, [C.c_void_p]),
    "rmd_energy": (C.c_int, [C.c_void_p, C.POINTER(EnergyTerms)]),
    "rmd_read_positions": (C.c_int, [C.c_void_p, _fp]),
    "rmd_read_state": (C.c_int, [C.c_void_p, _fp, _fp, _fp, _dp, _dp]),
    "rmd_scale_volume": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "rmd_attempt_volume_change": (C.c_int, [C.c_void_p, C.POINTER(Move)]),
    "rmd_min_options_default": (None, [C.POINTER(MinOptions)]),
    "rmd_minimise": (C.c_int, [C.c_void_p, C.POINTER(MinOptions), C.POINTER(MinResult)]),
    "rmd_diagnostics_get": (C.c_int, [C.c_void_p, C.POINTER(Diagnostics)]),
    "rmd_evaluate": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(EnergyTerms), _dp]),
    "rmd_read_cells": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "rmd_read_neighbours": (C.c_int, [C.c_void_p, _ip, _ip]),
    "rmd_step_once": (C.c_int, [C.c_void_p]),
    "rmd_debug_fft_forward": (C.c_int, [C.c_void_p, _fp, _fp]),
    "rmd_debug_noise": (C.c_int, [C.c_void_p, C.c_int64, _fp, _fp]),
    "rmd_atom_count": (C.c_int32, [C.c_void_p]),
    "rmd_tuned_shape_get": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p]),
    "rmd_profile_pme_stages": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "rmd_set_lanes": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_set_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_set_pme_pipeline": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_set_concurrent_forces": (C.c_int, [C.c_void_p, C.c_int32]),
    "rmd_copy_probe": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "rmd_surrogate_create": (C.c_int, [C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "rmd_surrogate_free": (None, [C.c_void_p]),
    "rmd_surrogate_get_params": (C.c_int, [C.c_void_p, _fp]),
    "rmd_surrogate_set_params": (C.c_int, [C.c_void_p, _fp]),
    "rmd_surrogate_forward": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "rmd_surrogate_forward_noalloc": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "rmd_surrogate_gradient": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, _fp, _fp]),
    "rmd_surrogate_train": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, C.c_int32, C.c_float, _fp]),
    "rmd_screen_options_default": (None, [C.POINTER(ScreenOptions)]),
    "rmd_screen_configure": (C.c_int, [C.c_void_p, C.POINTER(ScreenOptions)]),
    "rmd_screen_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ScreenStats)]),
    "rmd_screen_samples": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, C.POINTER(C.c_int32)]),
    "rmd_get_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "rmd_profile_kernel": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "rmd_device_info_get": (C.c_int, [C.c_int32, C.POINTER(DeviceInfo)]),
}

_lib = None


def load():
    """Load librainmd_amd.so from the package directory (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            f"{LIBRARY_PATH} is missing: the HIP library has not been built. "
            "Run `python -m rainmd_amd.build` (or __graft_entry__.build()) first.")
    lib = C.CDLL(LIBRARY_PATH)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib

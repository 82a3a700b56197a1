"""The Python host surface over the C ABI.

Mirrors the reference's JavaScript host: parsePrmtop / parseInpcrd /
buildSystem (shared/amber/*.js, shared/topology/system.js) and the
GpuSimulation class (webgpu/simulation.js:82-1341), with the same methods,
option names (snake_case), return fields and error classes. Every call goes
through librainmd_amd.so; nothing here computes physics.
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


# ---- errors (the reference's named Error subclasses) ----------------------

class RainmdError(RuntimeError):
    """A plain Error: validation or device failure (execution.js:48-55)."""


class ParseError(RainmdError):
    """A refusal from parsePrmtop / parseInpcrd."""


class SystemError_(RainmdError):
    """A refusal from buildSystem (system.js:80-87)."""


class NonFiniteStateError(RainmdError):
    """shared/finite.js:63-95."""


class NeighbourOverflowError(RainmdError):
    """webgpu/kernels/neighbours.js:172-183."""


class SimulationStopped(RainmdError):
    """webgpu/runtime/yield.js:44-50."""


class GpuBudgetError(RainmdError):
    """webgpu/runtime/allocator.js:141-154."""


class ConstraintFailure(RainmdError):
    """cpu/integrate/langevin.js:160-166."""


class DeviceError(RainmdError):
    """A HIP failure or no device."""


_ERRORS = {
    L.RMD_E_PARSE: ParseError,
    L.RMD_E_SYSTEM: SystemError_,
    L.RMD_E_DEVICE: DeviceError,
    L.RMD_E_NONFINITE: NonFiniteStateError,
    L.RMD_E_NEIGHBOUR_OVERFLOW: NeighbourOverflowError,
    L.RMD_E_STOPPED: SimulationStopped,
    L.RMD_E_BUDGET: GpuBudgetError,
    L.RMD_E_CONSTRAINT: ConstraintFailure,
}


def _check(rc, ctx=None):
    if rc == L.RMD_OK:
        return
    lib = L.load()
    msg = (lib.rmd_last_error(ctx) if ctx else lib.rmd_last_global_error()) or b""
    raise _ERRORS.get(rc, RainmdError)(msg.decode(errors="replace"))


def _error(rc, message):
    return _ERRORS.get(rc, RainmdError)(message)


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def read_data(name):
    with open(os.path.join(DATA, name), "rb") as f:
        return f.read()


# ---- topology, coordinates, system ----------------------------------------

class Topology:
    """parsePrmtop (shared/amber/prmtop.js:111-385)."""

    def __init__(self, text):
        lib = L.load()
        raw = text.encode() if isinstance(text, str) else text
        h = C.c_void_p()
        _check(lib.rmd_topology_parse_prmtop(raw, len(raw), C.byref(h)))
        self._h = h

    @classmethod
    def load(cls, name_or_path):
        path = name_or_path if os.path.exists(name_or_path) else os.path.join(DATA, name_or_path)
        with open(path, "rb") as f:
            return cls(f.read())

    def __del__(self):
        if getattr(self, "_h", None):
            L.load().rmd_topology_free(self._h)
            self._h = None


class Coordinates:
    """parseInpcrd (shared/amber/inpcrd.js:34-132)."""

    def __init__(self, text, expected_atom_count=0):
        lib = L.load()
        raw = text.encode() if isinstance(text, str) else text
        h = C.c_void_p()
        _check(lib.rmd_coordinates_parse_inpcrd(raw, len(raw), int(expected_atom_count), C.byref(h)))
        self._h = h
        v = L.CoordinatesView()
        _check(lib.rmd_coordinates_view_get(h, C.byref(v)))
        self.atom_count = v.atom_count
        self.positions = _arr(v.positions, 3 * v.atom_count, np.float64)
        self.velocities = _arr(v.velocities, 3 * v.atom_count, np.float64) if v.velocities else None
        self.box = {"lengths": list(v.box_lengths), "angles": list(v.box_angles)} if v.has_box else None
        self.time = v.time

    @classmethod
    def load(cls, name_or_path, expected_atom_count=0):
        path = name_or_path if os.path.exists(name_or_path) else os.path.join(DATA, name_or_path)
        with open(path, "rb") as f:
            return cls(f.read(), expected_atom_count)

    def __del__(self):
        if getattr(self, "_h", None):
            L.load().rmd_coordinates_free(self._h)
            self._h = None


def box_vectors_from_lengths_and_angles(lengths, angles):
    out = (C.c_double * 9)()
    _check(L.load().rmd_box_vectors_from_lengths_and_angles((C.c_double * 3)(*lengths), (C.c_double * 3)(*angles), out))
    return np.array(out[:])


class System:
    """buildSystem (shared/topology/system.js:40-236), flattened for the kernels.

    Every array of the built system is exposed as a numpy copy (see
    rmd_system_view in include/rainmd.h).
    """

    def __init__(self, handle):
        self._h = handle
        lib = L.load()
        v = L.SystemView()
        _check(lib.rmd_system_view_get(handle, C.byref(v)))
        n, t = v.atom_count, v.type_count
        self.atom_count = n
        self.type_count = t
        self.molecule_count = v.molecule_count
        self.residue_count = v.residue_count
        self.charges = _arr(v.charges, n, np.float64)
        self.masses = _arr(v.masses, n, np.float64)
        self.atom_types = _arr(v.atom_types, n, np.int32)
        self.atomic_numbers = _arr(v.atomic_numbers, n, np.int32)
        self.residue_of_atom = _arr(v.residue_of_atom, n, np.int32)
        self.lj_a = _arr(v.lj_a, t * t, np.float64)
        self.lj_b = _arr(v.lj_b, t * t, np.float64)
        self.bond_a = _arr(v.bond_a, v.bond_count, np.int32)
        self.bond_b = _arr(v.bond_b, v.bond_count, np.int32)
        self.bond_k = _arr(v.bond_k, v.bond_count, np.float64)
        self.bond_r0 = _arr(v.bond_r0, v.bond_count, np.float64)
        self.angle_a = _arr(v.angle_a, v.angle_count, np.int32)
        self.angle_b = _arr(v.angle_b, v.angle_count, np.int32)
        self.angle_c = _arr(v.angle_c, v.angle_count, np.int32)
        self.angle_k = _arr(v.angle_k, v.angle_count, np.float64)
        self.angle_theta0 = _arr(v.angle_theta0, v.angle_count, np.float64)
        self.torsion_a = _arr(v.torsion_a, v.torsion_count, np.int32)
        self.torsion_b = _arr(v.torsion_b, v.torsion_count, np.int32)
        self.torsion_c = _arr(v.torsion_c, v.torsion_count, np.int32)
        self.torsion_d = _arr(v.torsion_d, v.torsion_count, np.int32)
        self.torsion_k = _arr(v.torsion_k, v.torsion_count, np.float64)
        self.torsion_periodicity = _arr(v.torsion_periodicity, v.torsion_count, np.float64)
        self.torsion_phase = _arr(v.torsion_phase, v.torsion_count, np.float64)
        self.constraint_a = _arr(v.constraint_a, v.constraint_count, np.int32)
        self.constraint_b = _arr(v.constraint_b, v.constraint_count, np.int32)
        self.constraint_distance = _arr(v.constraint_distance, v.constraint_count, np.float64)
        self.exception_a = _arr(v.exception_a, v.exception_count, np.int32)
        self.exception_b = _arr(v.exception_b, v.exception_count, np.int32)
        self.exception_charge_product = _arr(v.exception_charge_product, v.exception_count, np.float64)
        self.exception_lj_a = _arr(v.exception_lj_a, v.exception_count, np.float64)
        self.exception_lj_b = _arr(v.exception_lj_b, v.exception_count, np.float64)
        self.exclusion_start = _arr(v.exclusion_start, n + 1, np.int32)
        self.exclusion_partners = _arr(v.exclusion_partners, v.exclusion_total, np.int32)
        self.molecule_start = _arr(v.molecule_start, v.molecule_count + 1, np.int32)
        self.molecule_atoms = _arr(v.molecule_atoms, n, np.int32)
        self.molecule_of_atom = _arr(v.molecule_of_atom, n, np.int32)
        self.is_periodic = bool(v.is_periodic)
        self.box = np.array(v.box[:])
        self.cutoff = v.cutoff
        self.degrees_of_freedom = v.degrees_of_freedom
        self.positions = _arr(v.positions, 3 * n, np.float64) if v.positions else None
        self.velocities = _arr(v.velocities, 3 * n, np.float64) if v.velocities else None
        self.water_count = v.water_count

    @classmethod
    def build(cls, topology, coordinates=None, constraints="HBonds", rigid_water=True, cutoff=1.0,
              box_vectors=None, box=None):
        lib = L.load()
        o = L.BuildOptions()
        lib.rmd_build_options_default(C.byref(o))
        if constraints not in L.CONSTRAINTS:
            raise SystemError_(
                f"unsupported constraints mode {constraints!r}; this port implements none, HBonds and AllBonds "
                "(HAngles is refused rather than approximated)")
        o.constraints = L.CONSTRAINTS[constraints]
        o.rigid_water = 1 if rigid_water else 0
        o.cutoff = cutoff
        if box_vectors is not None:
            o.has_box_vectors = 1
            o.box_vectors[:] = [float(x) for x in box_vectors]
        elif box is not None:
            o.has_box = 1
            o.box_lengths[:] = box["lengths"]
            o.box_angles[:] = box["angles"]
        h = C.c_void_p()
        _check(lib.rmd_system_build(topology._h, coordinates._h if coordinates is not None else None,
                                    C.byref(o), C.byref(h)))
        return cls(h)

    @classmethod
    def solvated(cls, water_count, box_edge, seed, include_solute=True, cutoff=0.9, constraints="HBonds",
                 rigid_water=True, clearance=0.25, jitter=0.02):
        """The seeded solvated-peptide builder (include/rainmd.h)."""
        lib = L.load()
        o = L.SolvatedOptions()
        lib.rmd_solvated_options_default(C.byref(o))
        o.water_count = water_count
        o.box_edge = box_edge
        o.seed = seed
        o.include_solute = 1 if include_solute else 0
        o.solute_clearance = clearance
        o.jitter = jitter
        o.cutoff = cutoff
        o.constraints = L.CONSTRAINTS[constraints]
        o.rigid_water = 1 if rigid_water else 0
        templ = read_data("ala-wat.prmtop")
        solute = read_data("ala-vac.inpcrd")
        h = C.c_void_p()
        _check(lib.rmd_system_build_solvated(templ, len(templ), solute, len(solute), C.byref(o), C.byref(h)))
        return cls(h)

    def rebox(self, box_vectors):
        """The same system in another (reduced, row-major) box: the live box a
        barostat stage leaves for the next (web/app.js:354-364)."""
        h = C.c_void_p()
        box = (C.c_double * 9)(*[float(x) for x in np.asarray(box_vectors).reshape(9)])
        _check(L.load().rmd_system_rebox(self._h, box, C.byref(h)))
        return System(h)

    def solute_heavy_atoms(self):
        """soluteHeavyAtoms (shared/topology/restraints.js:46-49)."""
        lib = L.load()
        count = C.c_int32()
        _check(lib.rmd_system_solute_heavy_atoms(self._h, None, 0, C.byref(count)))
        out = np.zeros(max(count.value, 1), np.int32)
        _check(lib.rmd_system_solute_heavy_atoms(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), len(out),
                                                 C.byref(count)))
        return out[:count.value].copy()

    def skin_for(self, timestep=0.002, temperature=298.0):
        return L.load().rmd_skin_for(self._h, timestep, temperature)

    def dispersion_energy(self, volume):
        return L.load().rmd_dispersion_energy(self._h, volume)

    def __del__(self):
        if getattr(self, "_h", None):
            L.load().rmd_system_free(self._h)
            self._h = None


def pme_parameters(box, cutoff, tolerance=5e-4, alpha=None, grid=None):
    """pmeParameters (shared/pme/parameters.js:227-252)."""
    out = L.PmeParameters()
    g = (C.c_int32 * 3)(*(grid or (0, 0, 0)))
    _check(L.load().rmd_pme_parameters_derive((C.c_double * 9)(*box), cutoff, tolerance, alpha or 0.0, g,
                                               C.byref(out)))
    return {"alpha": out.alpha, "grid": list(out.grid), "order": out.order, "tolerance": out.tolerance}


def pme_moduli(size, order=5):
    out = (C.c_double * size)()
    _check(L.load().rmd_pme_moduli(size, order, out))
    return np.array(out[:])


def tuned_shape(kind, atoms):
    """The autotuner's compiled-in shapes for a single system ('single') or a
    replica batch ('batch') of `atoms` atoms (rmd_tuned_shape_get)."""
    out = L.TunedShape()
    _check(L.load().rmd_tuned_shape_get({"single": 0, "batch": 1}[kind], int(atoms), C.byref(out)))
    return _struct_dict(out)


def device_info(device=0):
    info = L.DeviceInfo()
    rc = L.load().rmd_device_info_get(device, C.byref(info))
    return {"device_count": info.device_count, "name": info.name.decode(), "arch": info.arch.decode(),
            "compute_units": info.compute_units, "global_memory": info.global_memory, "ok": rc == L.RMD_OK}


def _struct_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


# ---- the simulation -------------------------------------------------------

def _options(options):
    """rmd_options from the reference's constructor options; returns the
    options and the arrays they point into (kept alive by the caller)."""
    lib = L.load()
    o = L.Options()
    lib.rmd_options_default(C.byref(o))
    simple = ["timestep", "temperature", "friction", "seed", "tolerance", "max_constraint_iterations",
              "rebuild_interval", "barostat_interval", "skin", "neighbour_slack", "max_neighbours",
              "ewald_tolerance", "alpha", "device", "lanes_per_atom", "cell_reach", "search_mode", "pme_bins"]
    for key in simple:
        if options.get(key) is not None:
            setattr(o, key, options[key])
    if options.get("adaptive_rebuild") is not None:
        o.adaptive_rebuild = 1 if options["adaptive_rebuild"] else 0
    if options.get("dispersion_correction") is not None:
        o.dispersion_correction = 1 if options["dispersion_correction"] else 0
    if options.get("pressure") is not None:
        o.has_pressure = 1
        o.pressure = options["pressure"]
    if options.get("grid") is not None:
        o.grid[:] = list(options["grid"])
    keep = []
    if options.get("restraint_force"):
        atoms = np.ascontiguousarray(options["restraint_atoms"], dtype=np.int32)
        ref = np.ascontiguousarray(options["restraint_reference"], dtype=np.float64)
        keep += [atoms, ref]
        o.restraint_force = options["restraint_force"]
        o.restraint_count = len(atoms)
        o.restraint_atoms = atoms.ctypes.data_as(L._ip)
        o.restraint_reference = ref.ctypes.data_as(L._dp)
    return o, keep


class GpuSimulation:
    """new GpuSimulation(device, system, options) (webgpu/simulation.js:83-190).

    Options mirror the reference's constructor options (simulation.js:93-188):
    timestep, temperature, friction, seed, tolerance, max_constraint_iterations,
    rebuild_interval, adaptive_rebuild, pressure, barostat_interval,
    dispersion_correction, restraint_force / restraint_atoms /
    restraint_reference, skin, neighbour_slack, max_neighbours,
    ewald_tolerance, alpha, grid; plus device and lanes_per_atom.
    """

    def __init__(self, system, **options):
        lib = L.load()
        self.system = system
        o, self._keep = _options(options)
        h = C.c_void_p()
        _check(lib.rmd_create(system._h, C.byref(o), C.byref(h)))
        self._h = h
        self.atom_count = system.atom_count

    @classmethod
    def _member(cls, handle, system, owner):
        """A batch member's own context (not owned: the batch frees it)."""
        sim = cls.__new__(cls)
        sim.system = system
        sim._keep = [owner]
        sim._h = handle
        sim.atom_count = system.atom_count
        return sim

    # -- state --
    def set_state(self, positions, velocities=None, step=0, supplied_noise=False):
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        v = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1)
        _check(L.load().rmd_set_state(self._h, p.ctypes.data_as(L._dp), None if v is None else v.ctypes.data_as(L._dp),
                                      int(step), 1 if supplied_noise else 0), self._h)

    def set_noise(self, noise):
        n = np.ascontiguousarray(noise, dtype=np.float32).reshape(-1)
        _check(L.load().rmd_set_noise(self._h, n.ctypes.data_as(L._fp)), self._h)

    def initialise(self):
        _check(L.load().rmd_initialise(self._h), self._h)

    def run(self, count, block_size=0):
        step = C.c_int64()
        _check(L.load().rmd_run(self._h, int(count), int(block_size), C.byref(step)), self._h)
        return step.value

    def request_stop(self):
        _check(L.load().rmd_request_stop(self._h), self._h)

    def step_once(self):
        _check(L.load().rmd_step_once(self._h), self._h)

    # -- energies and state readback --
    def energy(self):
        e = L.EnergyTerms()
        _check(L.load().rmd_energy(self._h, C.byref(e)), self._h)
        d = _struct_dict(e)
        d["kinetic_sampled_at"] = "before the closing half kick" if e.kinetic_sampled_in_step else "after the step"
        return d

    def evaluate(self, terms=L.TERM_ALL, forces=True, energies=True):
        """(energies, forces) at the current positions with a fresh list;
        energies=False runs a step's force-only launches and returns None for
        the energies."""
        e = L.EnergyTerms()
        f = np.zeros(3 * self.atom_count) if forces else None
        _check(L.load().rmd_evaluate(self._h, terms, C.byref(e) if energies else None,
                                     f.ctypes.data_as(L._dp) if forces else None), self._h)
        return (_struct_dict(e) if energies else None), (f.reshape(-1, 3) if forces else None)

    def read_positions(self):
        out = np.zeros(3 * self.atom_count, dtype=np.float32)
        _check(L.load().rmd_read_positions(self._h, out.ctypes.data_as(L._fp)), self._h)
        return out

    def read_state(self):
        n3 = 3 * self.atom_count
        p = np.zeros(n3, dtype=np.float32)
        v = np.zeros(n3, dtype=np.float32)
        f = np.zeros(n3, dtype=np.float32)
        box = (C.c_double * 9)()
        vol = C.c_double()
        _check(L.load().rmd_read_state(self._h, p.ctypes.data_as(L._fp), v.ctypes.data_as(L._fp),
                                       f.ctypes.data_as(L._fp), box, C.byref(vol)), self._h)
        return {"positions": p, "velocities": v, "forces": f, "box_vectors": np.array(box[:]), "volume": vol.value}

    # -- barostat --
    def scale_volume(self, length_scale):
        u = C.c_double()
        _check(L.load().rmd_scale_volume(self._h, float(length_scale), C.byref(u)), self._h)
        return u.value

    def attempt_volume_change(self):
        m = L.Move()
        _check(L.load().rmd_attempt_volume_change(self._h, C.byref(m)), self._h)
        if not m.attempted:
            return None
        return _struct_dict(m)

    # -- minimiser --
    def minimise(self, max_iterations=200, force_tolerance=500.0, max_step=0.002, max_displacement=0.05):
        o = L.MinOptions(max_iterations, force_tolerance, max_step, max_displacement)
        r = L.MinResult()
        _check(L.load().rmd_minimise(self._h, C.byref(o), C.byref(r)), self._h)
        return _struct_dict(r)

    def diagnostics(self):
        d = L.Diagnostics()
        _check(L.load().rmd_diagnostics_get(self._h, C.byref(d)), self._h)
        return _struct_dict(d)

    # -- inspection used by the parity tests --
    def read_cells(self):
        diag = self.diagnostics()
        n = self.atom_count
        cells = int(np.prod(diag["cell_counts"]))
        coa = np.zeros(n, dtype=np.int32)
        start = np.zeros(cells + 1, dtype=np.int32)
        atoms = np.zeros(n, dtype=np.int32)
        _check(L.load().rmd_read_cells(self._h, coa.ctypes.data_as(L._ip), start.ctypes.data_as(L._ip),
                                       atoms.ctypes.data_as(L._ip)), self._h)
        return coa, start, atoms

    def read_neighbours(self):
        cap = self.diagnostics()["neighbour_capacity"]
        counts = np.zeros(self.atom_count, dtype=np.int32)
        lists = np.zeros(self.atom_count * cap, dtype=np.int32)
        _check(L.load().rmd_read_neighbours(self._h, counts.ctypes.data_as(L._ip), lists.ctypes.data_as(L._ip)),
               self._h)
        return counts, lists.reshape(self.atom_count, cap)

    def debug_fft_forward(self, real):
        grid = self.diagnostics()["pme_grid"]
        r = np.ascontiguousarray(real, dtype=np.float32).reshape(-1)
        half = grid[2] // 2 + 1
        out = np.zeros(grid[0] * grid[1] * half * 2, dtype=np.float32)
        _check(L.load().rmd_debug_fft_forward(self._h, r.ctypes.data_as(L._fp), out.ctypes.data_as(L._fp)), self._h)
        c = out.reshape(grid[0], grid[1], half, 2)
        return c[..., 0] + 1j * c[..., 1]

    def debug_noise(self, step):
        """The thermostat's device draws at `step` (rmd_debug_noise): the
        uniforms (N x 4) and the normals (N x 3), both float32."""
        n = self.system.atom_count
        u = np.zeros(4 * n, dtype=np.float32)
        g = np.zeros(3 * n, dtype=np.float32)
        _check(L.load().rmd_debug_noise(self._h, int(step), u.ctypes.data_as(L._fp), g.ctypes.data_as(L._fp)),
               self._h)
        return u.reshape(n, 4), g.reshape(n, 3)

    def set_lanes(self, lanes):
        _check(L.load().rmd_set_lanes(self._h, int(lanes)), self._h)

    def set_graphs(self, enabled):
        _check(L.load().rmd_set_graphs(self._h, 1 if enabled else 0), self._h)

    # -- the barostat's learned screen --
    def configure_screen(self, enabled=True, training_moves=2000, training_steps=4000, learning_rate=1e-3,
                         seed=0x5eed):
        o = L.ScreenOptions(1 if enabled else 0, int(training_moves), int(training_steps), float(learning_rate),
                            int(seed))
        _check(L.load().rmd_screen_configure(self._h, C.byref(o)), self._h)

    def screen_stats(self):
        st = L.ScreenStats()
        _check(L.load().rmd_screen_stats_get(self._h, C.byref(st)), self._h)
        return _struct_dict(st)

    def screen_samples(self):
        n = C.c_int32()
        _check(L.load().rmd_screen_samples(self._h, None, None, 0, C.byref(n)), self._h)
        x = np.zeros((max(n.value, 1), L.SURROGATE_INPUTS), dtype=np.float32)
        t = np.zeros(max(n.value, 1), dtype=np.float32)
        _check(L.load().rmd_screen_samples(self._h, x.ctypes.data_as(L._fp), t.ctypes.data_as(L._fp), n.value,
                                           C.byref(n)), self._h)
        return x[:n.value], t[:n.value]

    def copy_probe(self, nbytes, reps=20):
        """Average duration (ms) of a plain device copy of nbytes (counter calibration)."""
        ms = C.c_double()
        _check(L.load().rmd_copy_probe(self._h, int(nbytes), int(reps), C.byref(ms)), self._h)
        return ms.value

    def set_concurrent_forces(self, enabled):
        _check(L.load().rmd_set_concurrent_forces(self._h, 1 if enabled else 0), self._h)

    def set_pme_pipeline(self, mode):
        """'auto', 'planes', 'lines', 'planes-scan' (planes without atom bins)
        or 'planes-folded' (planes-scan with the x lines folded into the plane
        gathers on force-only evaluations)."""
        ids = {"auto": 0, "planes": 1, "lines": 2, "planes-scan": 3, "planes-folded": 4}
        _check(L.load().rmd_set_pme_pipeline(self._h, ids[mode]), self._h)

    def stream(self):
        s = C.c_void_p()
        _check(L.load().rmd_get_stream(self._h, C.byref(s)), self._h)
        return s.value

    def profile_kernel(self, kernel, reps=50):
        """Average launch duration (ms) of a kernel family, HIP events on the
        context's stream: 'nonbonded', 'pme', 'integrate', 'bonded', 'rebuild'."""
        ids = {"nonbonded": 0, "pme": 1, "integrate": 2, "bonded": 3, "rebuild": 4}
        ms = C.c_double()
        _check(L.load().rmd_profile_kernel(self._h, ids[kernel], int(reps), C.byref(ms)), self._h)
        return ms.value

    def profile_pme_stages(self, reps=50):
        """Average durations (ms) of the PME chain's stages: spread (with the
        forward transforms it carries), transforms and convolution, gather
        (with the inverse transforms)."""
        out = np.zeros(3)
        _check(L.load().rmd_profile_pme_stages(self._h, int(reps), out.ctypes.data_as(L._dp)), self._h)
        return dict(zip(("spread", "transform", "gather"), out.tolist()))

    def release(self):
        if getattr(self, "_h", None):
            L.load().rmd_destroy(self._h)
            self._h = None

    def __del__(self):
        self.release()


class Batch:
    """A batch of replicas stepped together (include/rainmd.h, rmd_batch_*):
    many independent systems, each with its own box, PME grid, seed and
    barostat, advanced by batched launches. replica(r) is replica r's own
    GpuSimulation (set_state, minimise, energy, read_state, diagnostics, and
    a run of that replica alone)."""

    def __init__(self, systems, seeds=None, **options):
        lib = L.load()
        self.systems = list(systems)
        self.replicas = len(self.systems)
        o, self._keep = _options(options)
        handles = (C.c_void_p * self.replicas)(*[sys_._h.value if isinstance(sys_._h, C.c_void_p) else sys_._h
                                                 for sys_ in self.systems])
        sd = None if seeds is None else (C.c_uint32 * self.replicas)(*[int(x) for x in seeds])
        h = C.c_void_p()
        _check(lib.rmd_batch_create(handles, self.replicas, C.byref(o), sd, C.byref(h)))
        self._h = h
        self._members = {}

    def _check(self, rc):
        if rc == L.RMD_OK:
            return
        msg = L.load().rmd_batch_last_error(self._h) or b""
        raise _ERRORS.get(rc, RainmdError)(msg.decode(errors="replace"))

    def replica(self, r):
        if r not in self._members:
            handle = L.load().rmd_batch_replica(self._h, int(r))
            if not handle:
                raise RainmdError(f"no replica {r} in a batch of {self.replicas}")
            self._members[r] = GpuSimulation._member(C.c_void_p(handle), self.systems[r], self)
        return self._members[r]

    def initialise(self):
        self._check(L.load().rmd_batch_initialise(self._h))

    def evaluate(self, terms=L.TERM_ALL):
        """Evaluate `terms` for every replica; forces via replica(r).read_state()."""
        self._check(L.load().rmd_batch_evaluate(self._h, int(terms)))

    def run(self, count, block_size=0):
        step = C.c_int64()
        self._check(L.load().rmd_batch_run(self._h, int(count), int(block_size), C.byref(step)))
        return step.value

    def request_stop(self):
        L.load().rmd_batch_request_stop(self._h)

    def set_graphs(self, enabled):
        self._check(L.load().rmd_batch_set_graphs(self._h, 1 if enabled else 0))

    def set_concurrent_forces(self, enabled):
        self._check(L.load().rmd_batch_set_concurrent_forces(self._h, 1 if enabled else 0))

    def profile_nonbonded(self, reps=20):
        ms = C.c_double()
        self._check(L.load().rmd_batch_profile_nonbonded(self._h, int(reps), C.byref(ms)))
        return ms.value

    def configure_screen(self, enabled=True, training_moves=2000, training_steps=4000, learning_rate=1e-3,
                         seed=0x5eed):
        """The learned barostat screen shared by the replicas (rmd_batch_screen_configure)."""
        o = L.ScreenOptions(1 if enabled else 0, int(training_moves), int(training_steps), float(learning_rate),
                            int(seed))
        self._check(L.load().rmd_batch_screen_configure(self._h, C.byref(o)))

    def screen_stats(self):
        st = L.ScreenStats()
        self._check(L.load().rmd_batch_screen_stats_get(self._h, C.byref(st)))
        return _struct_dict(st)

    def info(self):
        i = L.BatchInfo()
        self._check(L.load().rmd_batch_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in L.BatchInfo._fields_}

    def release(self):
        if getattr(self, "_h", None):
            for m in self._members.values():
                m._h = None
            self._members = {}
            L.load().rmd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.release()


# ---- initial velocities ----------------------------------------------------

def _mix(v):
    v = v ^ (v >> np.uint32(16))
    v = v * np.uint32(0x7feb352d)
    v = v ^ (v >> np.uint32(15))
    v = v * np.uint32(0x846ca68b)
    return v ^ (v >> np.uint32(16))


def uniform_at(seed, a, b, c):
    """uniformAt (shared/random.js:59-64), vectorised over numpy uint32 arrays."""
    with np.errstate(over="ignore"):
        a = np.asarray(a, dtype=np.uint32)
        st = _mix(np.uint32(seed) ^ (a * np.uint32(0x9e3779b1)))
        st = _mix(st ^ (np.asarray(b, dtype=np.uint32) * np.uint32(0x85ebca6b)))
        st = _mix(st ^ (np.asarray(c, dtype=np.uint32) * np.uint32(0xc2b2ae35)))
    return ((st >> np.uint32(9)).astype(np.float64) + 0.5) * (1.0 / 8388608.0)


def maxwell_boltzmann(system, temperature, seed=0):
    """setVelocitiesToTemperature (cpu/integrate/langevin.js:211-230): per atom
    gaussianTriple(seed, atom, 0x5eed) scaled by sqrt(kT/m); massless atoms
    get zero."""
    n = system.atom_count
    atoms = np.arange(n, dtype=np.uint32)
    u = [uniform_at(seed, atoms, 0x5eed, k) for k in range(4)]
    r1 = np.sqrt(-2.0 * np.log(u[0]))
    r2 = np.sqrt(-2.0 * np.log(u[2]))
    g = np.stack([r1 * np.cos(2 * np.pi * u[1]), r1 * np.sin(2 * np.pi * u[1]), r2 * np.cos(2 * np.pi * u[3])], 1)
    kT = 0.008314462618 * temperature
    m = system.masses
    sigma = np.where(m > 0, np.sqrt(kT / np.where(m > 0, m, 1.0)), 0.0)
    return (g * sigma[:, None]).reshape(-1)


# ---- the learned screen's network on its own ---------------------------------

class Surrogate:
    """The 16 -> 64 -> 64 -> 1 tanh MLP of the barostat's screen on one device
    (include/rainmd.h, rmd_surrogate_*): BF16 MFMA layer products with FP32
    accumulation, Adam training on the GPU."""

    INPUTS = L.SURROGATE_INPUTS
    PARAMS = L.SURROGATE_PARAMS

    def __init__(self, seed=0, device=0):
        h = C.c_void_p()
        self._h = None
        st = L.load().rmd_surrogate_create(int(device), int(seed), C.byref(h))
        self._ok(st)
        self._h = h

    def release(self):
        if self._h:
            L.load().rmd_surrogate_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def _ok(self, st):
        if st != 0:
            raise _error(st, (L.load().rmd_last_global_error() or b"").decode(errors="replace"))

    @property
    def params(self):
        p = np.zeros(self.PARAMS, dtype=np.float32)
        self._ok(L.load().rmd_surrogate_get_params(self._h, p.ctypes.data_as(L._fp)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.PARAMS,)
        self._ok(L.load().rmd_surrogate_set_params(self._h, p.ctypes.data_as(L._fp)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.INPUTS)
        y = np.zeros(len(x), dtype=np.float32)
        self._ok(L.load().rmd_surrogate_forward(self._h, x.ctypes.data_as(L._fp), len(x), y.ctypes.data_as(L._fp)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.INPUTS)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.PARAMS, dtype=np.float32)
        self._ok(L.load().rmd_surrogate_gradient(self._h, x.ctypes.data_as(L._fp), t.ctypes.data_as(L._fp), len(x),
                                                 loss.ctypes.data_as(L._fp), g.ctypes.data_as(L._fp)))
        return float(loss[0]), g

    def train(self, x, t, steps, learning_rate=1e-3):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.INPUTS)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._ok(L.load().rmd_surrogate_train(self._h, x.ctypes.data_as(L._fp), t.ctypes.data_as(L._fp), len(x),
                                              int(steps), float(learning_rate), loss.ctypes.data_as(L._fp)))
        return float(loss[0])


# ---- the synthetic workloads ------------------------------------------------

# Solvated alanine dipeptide (ala-wat.prmtop's ACE-ALA-NME + TIP3P) on a
# seeded jittered lattice, and the pure water box.
WORKLOADS = {
    "ala-389w": dict(water_count=389, box_edge=2.2816, seed=101, include_solute=True, cutoff=0.9),
    "ala-944w": dict(water_count=944, box_edge=3.0539, seed=202, include_solute=True, cutoff=0.9),
    "water-33400": dict(water_count=33400, box_edge=10.0, seed=303, include_solute=False, cutoff=1.0),
}


def workload_system(name):
    return System.solvated(**WORKLOADS[name])


# The replica workload: `count` replicas of the ACE-ALA-NME + TIP3P composition
# with cube edges drawn uniformly from [2.2, 5.0] nm by the counter hash
# (seed 404, replica r at counter r), each filled with waters to 1.000 g/cm^3
# (ACE-ALA-NME 144.17 g/mol, TIP3P 18.016 g/mol) and built with seed 4000 + r.
REPLICA_SOLUTE_MASS = 144.17
REPLICA_WATER_MASS = 18.016


def replica_workload(count=256, seed=404, low=2.2, high=5.0):
    """[(edge_nm, water_count, seed)] of the replica workload."""
    out = []
    for r in range(count):
        edge = low + (high - low) * float(uniform_at(seed, r, 0, 0))
        mass = 602.214076 * edge ** 3  # g/mol in edge^3 nm^3 at 1 g/cm^3
        waters = int(round((mass - REPLICA_SOLUTE_MASS) / REPLICA_WATER_MASS))
        out.append((edge, waters, 4000 + r))
    return out


def replica_systems(count=256, seed=404, cutoff=0.9):
    return [System.solvated(water_count=w, box_edge=e, seed=s, include_solute=True, cutoff=cutoff)
            for e, w, s in replica_workload(count, seed)]

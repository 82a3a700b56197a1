"""TEST INFRASTRUCTURE ONLY: ctypes binding of the FP64 oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module. It restates the reference's host algorithms (see rmd_oracle.c);
the product never calls it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class OracleSystem(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("types", C.c_int32),
        ("charges", _dp), ("masses", _dp), ("atom_types", _ip), ("lj_a", _dp), ("lj_b", _dp),
        ("nbond", C.c_int32), ("bond_a", _ip), ("bond_b", _ip), ("bond_k", _dp), ("bond_r0", _dp),
        ("nangle", C.c_int32), ("angle_a", _ip), ("angle_b", _ip), ("angle_c", _ip), ("angle_k", _dp),
        ("angle_theta0", _dp),
        ("ntorsion", C.c_int32), ("tor_a", _ip), ("tor_b", _ip), ("tor_c", _ip), ("tor_d", _ip),
        ("tor_k", _dp), ("tor_n", _dp), ("tor_phase", _dp),
        ("nconstraint", C.c_int32), ("con_a", _ip), ("con_b", _ip), ("con_d", _dp),
        ("nexception", C.c_int32), ("exc_a", _ip), ("exc_b", _ip), ("exc_qq", _dp), ("exc_lja", _dp),
        ("exc_ljb", _dp),
        ("nmolecule", C.c_int32), ("mol_start", _ip), ("mol_atoms", _ip),
        ("periodic", C.c_int32), ("box", C.c_double * 9), ("cutoff", C.c_double), ("dof", C.c_int32),
    ]


class OracleOptions(C.Structure):
    _fields_ = [
        ("method", C.c_int32), ("alpha", C.c_double), ("grid", C.c_int32 * 3), ("terms", C.c_int32),
        ("f32_membership", C.c_int32), ("use_cells", C.c_int32),
        ("restraint_k", C.c_double), ("restraint_count", C.c_int32), ("restraint_atoms", _ip),
        ("restraint_reference", _dp),
    ]


class OracleTerms(C.Structure):
    _fields_ = [(name, C.c_double) for name in [
        "restraint", "bond", "angle", "torsion", "lj_direct", "lj_exception", "coulomb_direct",
        "coulomb_exception", "reciprocal", "self", "net", "lennard_jones", "coulomb_direct_total", "coulomb",
        "bonded", "total"]]


class OracleMinOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("force_tolerance", C.c_double), ("max_step", C.c_double),
                ("max_displacement", C.c_double), ("tolerance", C.c_double), ("restore_forces", C.c_int32)]


class OracleMinResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("rejected", C.c_int32), ("stalled", C.c_int32),
                ("converged", C.c_int32), ("potential", C.c_double), ("max_force", C.c_double),
                ("max_step", C.c_double)]


class OracleIntegrator(C.Structure):
    _fields_ = [("timestep", C.c_double), ("temperature", C.c_double), ("friction", C.c_double),
                ("tolerance", C.c_double), ("seed", C.c_uint32), ("step", C.c_int64),
                ("max_iterations", C.c_int32), ("noise", _dp)]


_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIBRARY


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            build()
        lib = C.CDLL(LIBRARY)
        lib.oracle_erfc.restype = C.c_double
        lib.oracle_erfc.argtypes = [C.c_double]
        lib.oracle_erf.restype = C.c_double
        lib.oracle_erf.argtypes = [C.c_double]
        lib.oracle_potential.argtypes = [C.POINTER(OracleSystem), _dp, C.POINTER(OracleOptions),
                                         C.POINTER(OracleTerms), _dp]
        lib.oracle_ewald_reciprocal.restype = C.c_double
        lib.oracle_ewald_reciprocal.argtypes = [C.POINTER(OracleSystem), _dp, C.c_double, C.c_int32, _dp]
        lib.oracle_alpha_from_tolerance.restype = C.c_double
        lib.oracle_alpha_from_tolerance.argtypes = [C.c_double, C.c_double]
        lib.oracle_bspline.argtypes = [C.c_double, C.c_int32, _dp, _dp]
        lib.oracle_moduli.argtypes = [C.c_int32, C.c_int32, _dp]
        lib.oracle_dispersion.restype = C.c_double
        lib.oracle_dispersion.argtypes = [C.POINTER(OracleSystem), C.c_double, C.c_double]
        lib.oracle_self_energy.argtypes = [C.POINTER(OracleSystem), C.c_double, C.c_double, _dp, _dp]
        lib.oracle_mix.restype = C.c_uint32
        lib.oracle_mix.argtypes = [C.c_uint32]
        lib.oracle_uniform.restype = C.c_double
        lib.oracle_uniform.argtypes = [C.c_uint32] * 4
        lib.oracle_gaussian.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _dp]
        lib.oracle_langevin_step.argtypes = [C.POINTER(OracleSystem), _dp, _dp, _dp, C.POINTER(OracleIntegrator)]
        lib.oracle_langevin_finish.argtypes = [C.POINTER(OracleSystem), _dp, _dp, _dp, C.POINTER(OracleIntegrator)]
        lib.oracle_shake.argtypes = [C.POINTER(OracleSystem), _dp, _dp, _dp, C.c_double, C.c_int32]
        lib.oracle_rattle.argtypes = [C.POINTER(OracleSystem), _dp, _dp, C.c_double, C.c_int32]
        lib.oracle_settle.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        lib.oracle_kinetic.restype = C.c_double
        lib.oracle_kinetic.argtypes = [C.POINTER(OracleSystem), _dp]
        lib.oracle_velocities.argtypes = [C.POINTER(OracleSystem), _dp, C.c_double, C.c_uint32]
        lib.oracle_scale_molecules.argtypes = [C.POINTER(OracleSystem), _dp, C.c_double]
        lib.oracle_metropolis_weight.restype = C.c_double
        lib.oracle_metropolis_weight.argtypes = [C.c_double] * 5 + [C.c_int32]
        lib.oracle_skin_for.restype = C.c_double
        lib.oracle_skin_for.argtypes = [C.POINTER(OracleSystem), C.c_double, C.c_double, _dp]
        lib.oracle_plan_cells.restype = C.c_int32
        lib.oracle_plan_cells.argtypes = [C.POINTER(OracleSystem), C.c_double, C.c_double, _dp, _ip, _ip, _dp]
        lib.oracle_minimise.argtypes = [C.POINTER(OracleSystem), _dp, C.POINTER(OracleOptions),
                                        C.POINTER(OracleMinOptions), C.POINTER(OracleMinResult)]
        lib.oracle_pairs_f32_binned.restype = C.c_int32
        lib.oracle_pairs_f32_binned.argtypes = [C.c_int32, _fp, _dp, _fp, _fp, C.c_float, _ip, _ip, C.c_int32,
                                                _ip, _ip]
        lib.oracle_cells_f32.argtypes = [C.c_int32, _fp, _fp, _ip, _ip]
        lib.oracle_r2_f32.restype = C.c_float
        lib.oracle_r2_f32.argtypes = [_fp, _fp, _fp, _fp]
        lib.oracle_pairs_f32.restype = C.c_int32
        lib.oracle_pairs_f32.argtypes = [C.c_int32, _fp, _fp, _fp, C.c_float, _ip, _ip, C.c_int32, _ip, _ip]
        _lib = lib
    return _lib


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Oracle:
    """The oracle bound to one system's arrays (a rainmd_amd.System or any
    object with the same attribute names)."""

    def __init__(self, system, box=None, cutoff=None):
        self.lib = load()
        s = system
        self._keep = []

        def keep(a, dtype):
            arr = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(arr)
            return arr

        o = OracleSystem()
        o.n = s.atom_count
        o.types = s.type_count
        o.charges = _p(keep(s.charges, np.float64), C.c_double)
        o.masses = _p(keep(s.masses, np.float64), C.c_double)
        o.atom_types = _p(keep(s.atom_types, np.int32), C.c_int32)
        o.lj_a = _p(keep(s.lj_a, np.float64), C.c_double)
        o.lj_b = _p(keep(s.lj_b, np.float64), C.c_double)
        o.nbond = len(s.bond_a)
        o.bond_a = _p(keep(s.bond_a, np.int32), C.c_int32)
        o.bond_b = _p(keep(s.bond_b, np.int32), C.c_int32)
        o.bond_k = _p(keep(s.bond_k, np.float64), C.c_double)
        o.bond_r0 = _p(keep(s.bond_r0, np.float64), C.c_double)
        o.nangle = len(s.angle_a)
        o.angle_a = _p(keep(s.angle_a, np.int32), C.c_int32)
        o.angle_b = _p(keep(s.angle_b, np.int32), C.c_int32)
        o.angle_c = _p(keep(s.angle_c, np.int32), C.c_int32)
        o.angle_k = _p(keep(s.angle_k, np.float64), C.c_double)
        o.angle_theta0 = _p(keep(s.angle_theta0, np.float64), C.c_double)
        o.ntorsion = len(s.torsion_a)
        o.tor_a = _p(keep(s.torsion_a, np.int32), C.c_int32)
        o.tor_b = _p(keep(s.torsion_b, np.int32), C.c_int32)
        o.tor_c = _p(keep(s.torsion_c, np.int32), C.c_int32)
        o.tor_d = _p(keep(s.torsion_d, np.int32), C.c_int32)
        o.tor_k = _p(keep(s.torsion_k, np.float64), C.c_double)
        o.tor_n = _p(keep(s.torsion_periodicity, np.float64), C.c_double)
        o.tor_phase = _p(keep(s.torsion_phase, np.float64), C.c_double)
        o.nconstraint = len(s.constraint_a)
        o.con_a = _p(keep(s.constraint_a, np.int32), C.c_int32)
        o.con_b = _p(keep(s.constraint_b, np.int32), C.c_int32)
        o.con_d = _p(keep(s.constraint_distance, np.float64), C.c_double)
        o.nexception = len(s.exception_a)
        o.exc_a = _p(keep(s.exception_a, np.int32), C.c_int32)
        o.exc_b = _p(keep(s.exception_b, np.int32), C.c_int32)
        o.exc_qq = _p(keep(s.exception_charge_product, np.float64), C.c_double)
        o.exc_lja = _p(keep(s.exception_lj_a, np.float64), C.c_double)
        o.exc_ljb = _p(keep(s.exception_lj_b, np.float64), C.c_double)
        o.nmolecule = s.molecule_count
        o.mol_start = _p(keep(s.molecule_start, np.int32), C.c_int32)
        o.mol_atoms = _p(keep(s.molecule_atoms, np.int32), C.c_int32)
        o.periodic = 1 if s.is_periodic else 0
        o.box[:] = list(box if box is not None else s.box)
        o.cutoff = cutoff if cutoff is not None else s.cutoff
        o.dof = s.degrees_of_freedom
        self.s = o
        self.n = o.n

    def set_box(self, box):
        self.s.box[:] = list(box)

    @staticmethod
    def _energy_options(method="PME", alpha=None, grid=None, terms=31, f32_membership=False, use_cells=False):
        o = OracleOptions()
        o.method = 0 if method == "PME" else 1
        o.alpha = alpha or 0.0
        o.grid[:] = list(grid or (0, 0, 0))
        o.terms = terms
        o.f32_membership = 1 if f32_membership else 0
        o.use_cells = 1 if use_cells else 0
        return o

    def skin_for(self, timestep=0.002, temperature=298.0, box=None):
        """skinFor (shared/topology/cells.js:77-94)."""
        b = np.ascontiguousarray(self.s.box if box is None else box, dtype=np.float64).reshape(9)
        return self.lib.oracle_skin_for(C.byref(self.s), timestep, temperature, _p(b, C.c_double))

    def plan_cells(self, skin, slack=2.0, box=None):
        """planCellList (shared/topology/cells.js:111-181): counts, stencil
        entries, neighbour capacity and the estimate it came from."""
        b = np.ascontiguousarray(self.s.box if box is None else box, dtype=np.float64).reshape(9)
        counts = np.zeros(3, dtype=np.int32)
        stencil = C.c_int32()
        estimated = C.c_double()
        cap = self.lib.oracle_plan_cells(C.byref(self.s), skin, slack, _p(b, C.c_double), _p(counts, C.c_int32),
                                         C.byref(stencil), C.byref(estimated))
        return dict(counts=counts.tolist(), stencil=stencil.value, max_neighbours=cap, estimated=estimated.value)

    def minimise(self, positions, max_iterations=200, force_tolerance=500.0, max_step=0.002, max_displacement=0.05,
                 tolerance=1e-10, restore_forces=True, **energy):
        """GpuSimulation.minimise (webgpu/simulation.js:880-979) in FP64;
        returns the final positions and the result fields."""
        x = np.array(positions, dtype=np.float64).reshape(-1)
        eo = self._energy_options(**energy)
        mo = OracleMinOptions(max_iterations, force_tolerance, max_step, max_displacement, tolerance,
                              1 if restore_forces else 0)
        r = OracleMinResult()
        if self.lib.oracle_minimise(C.byref(self.s), _p(x, C.c_double), C.byref(eo), C.byref(mo), C.byref(r)) != 0:
            raise MemoryError("oracle_minimise could not allocate")
        return x, {name: getattr(r, name) for name, _ in OracleMinResult._fields_}

    def potential(self, positions, method="PME", alpha=None, grid=None, terms=31, f32_membership=False,
                  use_cells=False, restraint=None, forces=True):
        """potentialEnergy (cpu/energy.js:25-91): per-term energies and forces."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        o = OracleOptions()
        o.method = 0 if method == "PME" else 1
        o.alpha = alpha or 0.0
        o.grid[:] = list(grid or (0, 0, 0))
        o.terms = terms
        o.f32_membership = 1 if f32_membership else 0
        o.use_cells = 1 if use_cells else 0
        if restraint is not None:
            atoms = np.ascontiguousarray(restraint["atoms"], dtype=np.int32)
            ref = np.ascontiguousarray(restraint["reference"], dtype=np.float64)  # 3 per restrained atom
            o.restraint_k = restraint["force"]
            o.restraint_count = len(atoms)
            o.restraint_atoms = _p(atoms, C.c_int32)
            o.restraint_reference = _p(ref, C.c_double)
        t = OracleTerms()
        f = np.zeros(3 * self.n) if forces else None
        rc = self.lib.oracle_potential(C.byref(self.s), _p(pos, C.c_double), C.byref(o), C.byref(t),
                                       _p(f, C.c_double) if forces else None)
        if rc != 0:
            raise RuntimeError("oracle_potential refused the input")
        terms_out = {name: getattr(t, name) for name, _ in OracleTerms._fields_}
        return terms_out, (f.reshape(-1, 3) if forces else None)

    def ewald_reciprocal(self, positions, alpha, kmax, forces=False):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        f = np.zeros(3 * self.n) if forces else None
        e = self.lib.oracle_ewald_reciprocal(C.byref(self.s), _p(pos, C.c_double), alpha, kmax,
                                             _p(f, C.c_double) if forces else None)
        return e, (f.reshape(-1, 3) if forces else None)

    def dispersion(self, volume, cutoff):
        return self.lib.oracle_dispersion(C.byref(self.s), volume, cutoff)

    def self_energy(self, alpha, volume):
        a, b = C.c_double(), C.c_double()
        self.lib.oracle_self_energy(C.byref(self.s), alpha, volume, C.byref(a), C.byref(b))
        return a.value, b.value

    def kinetic(self, velocities):
        v = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1)
        return self.lib.oracle_kinetic(C.byref(self.s), _p(v, C.c_double))

    def velocities(self, temperature, seed):
        v = np.zeros(3 * self.n)
        self.lib.oracle_velocities(C.byref(self.s), _p(v, C.c_double), temperature, seed)
        return v

    def _integrator(self, timestep, temperature, friction, tolerance, seed, step, max_iterations, noise):
        it = OracleIntegrator()
        it.timestep, it.temperature, it.friction, it.tolerance = timestep, temperature, friction, tolerance
        it.seed, it.step, it.max_iterations = seed, step, max_iterations
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
            self._keep.append(nz)
            it.noise = _p(nz, C.c_double)
        return it

    def langevin_step(self, positions, velocities, forces, timestep=0.002, temperature=298.0, friction=1.0,
                      tolerance=1e-5, seed=0, step=0, max_iterations=150, noise=None):
        """B R A S O R A S (cpu/integrate/langevin.js:30-98); returns new (x, v)."""
        x = np.array(positions, dtype=np.float64).reshape(-1)
        v = np.array(velocities, dtype=np.float64).reshape(-1)
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        it = self._integrator(timestep, temperature, friction, tolerance, seed, step, max_iterations, noise)
        rc = self.lib.oracle_langevin_step(C.byref(self.s), _p(x, C.c_double), _p(v, C.c_double),
                                           _p(f, C.c_double), C.byref(it))
        return x, v, rc

    def langevin_finish(self, positions, velocities, forces, timestep=0.002, tolerance=1e-5, max_iterations=150):
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        v = np.array(velocities, dtype=np.float64).reshape(-1)
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        it = self._integrator(timestep, 0, 0, tolerance, 0, 0, max_iterations, None)
        rc = self.lib.oracle_langevin_finish(C.byref(self.s), _p(x, C.c_double), _p(v, C.c_double),
                                             _p(f, C.c_double), C.byref(it))
        return v, rc

    def shake(self, positions, reference, tolerance, max_iterations=1000):
        x = np.array(positions, dtype=np.float64).reshape(-1)
        r = np.ascontiguousarray(reference, dtype=np.float64).reshape(-1)
        it = self.lib.oracle_shake(C.byref(self.s), _p(x, C.c_double), _p(r, C.c_double), None, tolerance,
                                   max_iterations)
        return x, it

    def rattle(self, positions, velocities, tolerance, max_iterations=1000):
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        v = np.array(velocities, dtype=np.float64).reshape(-1)
        it = self.lib.oracle_rattle(C.byref(self.s), _p(x, C.c_double), _p(v, C.c_double), tolerance,
                                    max_iterations)
        return v, it

    def scale_molecules(self, positions, scale):
        x = np.array(positions, dtype=np.float64).reshape(-1)
        self.lib.oracle_scale_molecules(C.byref(self.s), _p(x, C.c_double), scale)
        return x


def erfc(x):
    return load().oracle_erfc(float(x))


def erf(x):
    return load().oracle_erf(float(x))


def alpha_from_tolerance(cutoff, tolerance):
    return load().oracle_alpha_from_tolerance(cutoff, tolerance)


def bspline(offset, order=5):
    w = np.zeros(order)
    d = np.zeros(order)
    load().oracle_bspline(offset, order, _p(w, C.c_double), _p(d, C.c_double))
    return w, d


def moduli(size, order=5):
    out = np.zeros(size)
    load().oracle_moduli(size, order, _p(out, C.c_double))
    return out


def uniform(seed, a, b, c):
    return load().oracle_uniform(seed & 0xffffffff, a & 0xffffffff, b & 0xffffffff, c & 0xffffffff)


def gaussian(seed, atom, step):
    out = np.zeros(3)
    load().oracle_gaussian(seed, atom, step, _p(out, C.c_double))
    return out


def mix(v):
    return load().oracle_mix(v & 0xffffffff)


def settle(old, new, m_o, m_h, d_oh, d_hh):
    o = np.ascontiguousarray(old, dtype=np.float64).reshape(-1)
    n = np.array(new, dtype=np.float64).reshape(-1)
    load().oracle_settle(_p(o, C.c_double), _p(n, C.c_double), m_o, m_h, d_oh, d_hh)
    return n


def metropolis_weight(delta_energy, delta_volume, volume_ratio, pressure_kj, kT, molecules):
    return load().oracle_metropolis_weight(delta_energy, delta_volume, volume_ratio, pressure_kj, kT, molecules)


def f32_geometry(box):
    """The f32 box rows, inverse widths and reciprocal rows the device uses."""
    box = np.asarray(box, dtype=np.float64)
    ax, bx, by, cx, cy, cz = box[0], box[3], box[4], box[6], box[7], box[8]
    recip = np.array([1 / ax, 0, 0, -bx / (ax * by), 1 / by, 0,
                      (bx * cy - by * cx) / (ax * by * cz), -cy / (by * cz), 1 / cz])
    inv = np.array([1.0 / box[0], 1.0 / box[4], 1.0 / box[8]])
    return box.astype(np.float32), inv.astype(np.float32), recip.astype(np.float32)


def cells_f32(positions_f32, box, counts):
    pos = np.ascontiguousarray(positions_f32, dtype=np.float32).reshape(-1)
    _, _, recip = f32_geometry(box)
    c = np.ascontiguousarray(counts, dtype=np.int32)
    out = np.zeros(len(pos) // 3, dtype=np.int32)
    load().oracle_cells_f32(len(out), _p(pos, C.c_float), _p(recip, C.c_float), _p(c, C.c_int32),
                            _p(out, C.c_int32))
    return out


def pairs_f32(positions_f32, box, list_cutoff, excl_start, excl_partners, cap=None):
    """Brute-force FP32 neighbour sets (counts, lists sorted per atom)."""
    pos = np.ascontiguousarray(positions_f32, dtype=np.float32).reshape(-1)
    n = len(pos) // 3
    box9, inv3, _ = f32_geometry(box)
    rl2 = np.float32(list_cutoff * list_cutoff)
    es = np.ascontiguousarray(excl_start, dtype=np.int32)
    ep = np.ascontiguousarray(excl_partners, dtype=np.int32)
    counts = np.zeros(n, dtype=np.int32)
    widest = load().oracle_pairs_f32(n, _p(pos, C.c_float), _p(box9, C.c_float), _p(inv3, C.c_float), rl2,
                                     _p(es, C.c_int32), _p(ep, C.c_int32), 0, _p(counts, C.c_int32), None)
    cap = cap or max(1, widest)
    lists = np.zeros(n * cap, dtype=np.int32)
    load().oracle_pairs_f32(n, _p(pos, C.c_float), _p(box9, C.c_float), _p(inv3, C.c_float), rl2,
                            _p(es, C.c_int32), _p(ep, C.c_int32), cap, _p(counts, C.c_int32), _p(lists, C.c_int32))
    return counts, lists.reshape(n, cap)


def pairs_f32_binned(positions_f32, box, list_cutoff, excl_start, excl_partners, cap=None):
    """pairs_f32's sets over coarse-cell candidates (the same FP32 pair test),
    for systems too large for the all-pairs loop."""
    pos = np.ascontiguousarray(positions_f32, dtype=np.float32).reshape(-1)
    n = len(pos) // 3
    box64 = np.ascontiguousarray(box, dtype=np.float64).reshape(9)
    box9, inv3, _ = f32_geometry(box64)
    rl2 = np.float32(list_cutoff * list_cutoff)
    es = np.ascontiguousarray(excl_start, dtype=np.int32)
    ep = np.ascontiguousarray(excl_partners, dtype=np.int32)
    counts = np.zeros(n, dtype=np.int32)
    args = (n, _p(pos, C.c_float), _p(box64, C.c_double), _p(box9, C.c_float), _p(inv3, C.c_float), rl2,
            _p(es, C.c_int32), _p(ep, C.c_int32))
    widest = load().oracle_pairs_f32_binned(*args, 0, _p(counts, C.c_int32), None)
    if widest < 0:
        raise MemoryError("oracle_pairs_f32_binned could not allocate")
    cap = cap or max(1, widest)
    lists = np.zeros(n * cap, dtype=np.int32)
    load().oracle_pairs_f32_binned(*args, cap, _p(counts, C.c_int32), _p(lists, C.c_int32))
    return counts, lists.reshape(n, cap)

"""ctypes mirror of include/lima.h (structs only) and the host-side system container.

`LimaSystem` holds one system as numpy arrays in the lima_system layout, which
restates the reference's Box/Compound/CompoundState/Solvent/LJ_Ignores/
CompoundBridgeCompact fields with runtime sizes (Simulation.cuh:35-92,
Bodies.cuh:19-25, 169-472, 644-732).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

MAX_ATOM_TYPES = 32

LIMA_OK = 0
LIMA_E_HIP = 1
LIMA_E_CAPACITY = 2
LIMA_E_CRITICAL_FORCE = 3
LIMA_E_NAN = 4
LIMA_E_ARG = 5
LIMA_E_UNSUPPORTED = 6
STATUS_NAMES = {0: "LIMA_OK", 1: "LIMA_E_HIP", 2: "LIMA_E_CAPACITY", 3: "LIMA_E_CRITICAL_FORCE",
                4: "LIMA_E_NAN", 5: "LIMA_E_ARG", 6: "LIMA_E_UNSUPPORTED"}

THERMOSTAT_NONE = 0
THERMOSTAT_BERENDSEN = 1
THERMOSTAT_REFERENCE = 2   # Engine::hostMaster cadence: T0/T every 40th step, for one step
ACCUM_FP32 = 0
ACCUM_FIXED64 = 1


class ParticleParams(C.Structure):
    _fields_ = [("mass", C.c_float), ("sigma", C.c_float), ("epsilon", C.c_float)]


class ForceField(C.Structure):
    _fields_ = [("particle_parameters", ParticleParams * MAX_ATOM_TYPES)]


class PairBond(C.Structure):
    _fields_ = [("b0", C.c_float), ("kb", C.c_float), ("atom_indexes", C.c_uint32 * 2)]


class AngleBond(C.Structure):
    _fields_ = [("theta_0", C.c_float), ("k_theta", C.c_float), ("atom_indexes", C.c_uint32 * 3)]


class DihedralBond(C.Structure):
    _fields_ = [("phi_0", C.c_float), ("k_phi", C.c_float), ("n", C.c_uint32), ("atom_indexes", C.c_uint32 * 4)]


PAIR_DTYPE = np.dtype([("b0", "<f4"), ("kb", "<f4"), ("idx", "<u4", (2,))])
ANGLE_DTYPE = np.dtype([("theta_0", "<f4"), ("k_theta", "<f4"), ("idx", "<u4", (3,))])
DIHEDRAL_DTYPE = np.dtype([("phi_0", "<f4"), ("k_phi", "<f4"), ("n", "<u4"), ("idx", "<u4", (4,))])
assert PAIR_DTYPE.itemsize == C.sizeof(PairBond)
assert ANGLE_DTYPE.itemsize == C.sizeof(AngleBond)
assert DIHEDRAL_DTYPE.itemsize == C.sizeof(DihedralBond)

_pi = C.POINTER(C.c_int32)
_pf = C.POINTER(C.c_float)


class CSystem(C.Structure):
    _fields_ = [
        ("box_len", C.c_float), ("dt", C.c_double), ("thermostat_target", C.c_float),
        ("forcefield", ForceField),
        ("n_compounds", C.c_int32), ("compound_offset", _pi), ("key_particle_index", _pi),
        ("confining_particle_sphere", _pf), ("chain_positions", _pf), ("chain_prev_positions", _pf),
        ("chain_atom_types", C.POINTER(C.c_uint8)),
        ("bond_offset", _pi), ("bonds", C.POINTER(PairBond)),
        ("angle_offset", _pi), ("angles", C.POINTER(AngleBond)),
        ("dihedral_offset", _pi), ("dihedrals", C.POINTER(DihedralBond)),
        ("ignore_offset", _pi), ("ignore_compound", _pi), ("ignore_local", _pi),
        ("n_bridges", C.c_int32), ("bridge_particle_offset", _pi), ("bridge_particle_compound", _pi),
        ("bridge_particle_local", _pi),
        ("bridge_bond_offset", _pi), ("bridge_bonds", C.POINTER(PairBond)),
        ("bridge_angle_offset", _pi), ("bridge_angles", C.POINTER(AngleBond)),
        ("bridge_dihedral_offset", _pi), ("bridge_dihedrals", C.POINTER(DihedralBond)),
        ("n_solvents", C.c_int32), ("solvent_positions", _pf), ("solvent_prev_positions", _pf),
    ]


class CRunParams(C.Structure):
    _fields_ = [
        ("cutoff", C.c_float), ("nlist_interval", C.c_int32), ("thermostat", C.c_int32),
        ("thermostat_tau", C.c_double), ("accumulation", C.c_int32), ("enable_solvents", C.c_int32),
        ("enable_dihedrals", C.c_int32), ("max_compound_neighbors", C.c_int32),
        ("max_compound_solvents", C.c_int32), ("max_solvent_neighbors", C.c_int32), ("device", C.c_int32),
    ]


class CEnergies(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "bond", "angle", "dihedral", "lj_intra", "lj_inter", "lj_compound_solvent", "lj_solvent_compound",
        "lj_solvent_solvent", "potential", "kinetic", "temperature", "virial", "thermostat_scalar")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CAdjointSpec(C.Structure):
    """lima_adjoint_spec: frames, checkpoint interval, weights and targets (NULL = zeros)."""
    _fields_ = [("n_frames", C.c_int32), ("checkpoint_interval", C.c_int32),
                ("energy_weight", C.POINTER(C.c_double)), ("force_weight", C.POINTER(C.c_double)),
                ("energy_target", C.POINTER(C.c_double)), ("chain_force_target", C.POINTER(C.c_float)),
                ("solvent_force_target", C.POINTER(C.c_float))]


N_PARAMETERS = 68   # sigma[32], epsilon[32], kb, b0, k_theta, theta_0


@dataclass
class RunParams:
    """Runtime restatement of the Constants.cuh switches and capacities."""
    cutoff: float = 0.6
    nlist_interval: int = 10
    thermostat: int = THERMOSTAT_BERENDSEN
    thermostat_tau: float = 1e-4          # ns (0.1 ps)
    accumulation: int = ACCUM_FP32
    enable_solvents: bool = True
    enable_dihedrals: bool = True
    max_compound_neighbors: int = 256
    max_compound_solvents: int = 16384
    max_solvent_neighbors: int = 512
    device: int = 0

    def to_c(self) -> CRunParams:
        return CRunParams(self.cutoff, self.nlist_interval, self.thermostat, self.thermostat_tau,
                          self.accumulation, int(self.enable_solvents), int(self.enable_dihedrals),
                          self.max_compound_neighbors, self.max_compound_solvents,
                          self.max_solvent_neighbors, self.device)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class LimaSystem:
    """One system in lima_system layout (numpy arrays; float32 positions)."""
    box_len: float
    dt: float
    thermostat_target: float
    forcefield: np.ndarray                 # (32, 3) float32: mass, sigma, epsilon
    compound_offset: np.ndarray            # (nc+1,) int32
    key_particle_index: np.ndarray         # (nc,) int32
    confining_particle_sphere: np.ndarray  # (nc,) float32
    chain_positions: np.ndarray            # (n_chain, 3) float32
    chain_prev_positions: np.ndarray       # (n_chain, 3) float32
    chain_atom_types: np.ndarray           # (n_chain,) uint8
    bond_offset: np.ndarray
    bonds: np.ndarray                      # PAIR_DTYPE
    angle_offset: np.ndarray
    angles: np.ndarray                     # ANGLE_DTYPE
    dihedral_offset: np.ndarray
    dihedrals: np.ndarray                  # DIHEDRAL_DTYPE
    ignore_offset: np.ndarray              # (n_chain+1,)
    ignore_compound: np.ndarray
    ignore_local: np.ndarray
    bridge_particle_offset: np.ndarray     # (nb+1,)
    bridge_particle_compound: np.ndarray
    bridge_particle_local: np.ndarray
    bridge_bond_offset: np.ndarray
    bridge_bonds: np.ndarray
    bridge_angle_offset: np.ndarray
    bridge_angles: np.ndarray
    bridge_dihedral_offset: np.ndarray
    bridge_dihedrals: np.ndarray
    solvent_positions: np.ndarray          # (ns, 3) float32
    solvent_prev_positions: np.ndarray     # (ns, 3) float32
    meta: dict = field(default_factory=dict)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_compounds(self) -> int:
        return len(self.compound_offset) - 1

    @property
    def n_chain(self) -> int:
        return int(self.compound_offset[-1])

    @property
    def n_solvents(self) -> int:
        return len(self.solvent_positions)

    @property
    def n_bridges(self) -> int:
        return len(self.bridge_particle_offset) - 1

    def to_c(self) -> CSystem:
        self._keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            self._keep.append(a)
            return a

        s = CSystem()
        s.box_len = self.box_len
        s.dt = self.dt
        s.thermostat_target = self.thermostat_target
        ff = np.ascontiguousarray(self.forcefield, dtype=np.float32)
        for t in range(MAX_ATOM_TYPES):
            s.forcefield.particle_parameters[t] = ParticleParams(*[float(v) for v in ff[t]])
        s.n_compounds = self.n_compounds
        s.compound_offset = _ptr(arr(self.compound_offset, np.int32), C.c_int32)
        s.key_particle_index = _ptr(arr(self.key_particle_index, np.int32), C.c_int32)
        s.confining_particle_sphere = _ptr(arr(self.confining_particle_sphere, np.float32), C.c_float)
        s.chain_positions = _ptr(arr(self.chain_positions, np.float32), C.c_float)
        s.chain_prev_positions = _ptr(arr(self.chain_prev_positions, np.float32), C.c_float)
        s.chain_atom_types = _ptr(arr(self.chain_atom_types, np.uint8), C.c_uint8)
        s.bond_offset = _ptr(arr(self.bond_offset, np.int32), C.c_int32)
        s.bonds = _ptr(arr(self.bonds, PAIR_DTYPE), PairBond)
        s.angle_offset = _ptr(arr(self.angle_offset, np.int32), C.c_int32)
        s.angles = _ptr(arr(self.angles, ANGLE_DTYPE), AngleBond)
        s.dihedral_offset = _ptr(arr(self.dihedral_offset, np.int32), C.c_int32)
        s.dihedrals = _ptr(arr(self.dihedrals, DIHEDRAL_DTYPE), DihedralBond)
        s.ignore_offset = _ptr(arr(self.ignore_offset, np.int32), C.c_int32)
        s.ignore_compound = _ptr(arr(self.ignore_compound, np.int32), C.c_int32)
        s.ignore_local = _ptr(arr(self.ignore_local, np.int32), C.c_int32)
        s.n_bridges = self.n_bridges
        s.bridge_particle_offset = _ptr(arr(self.bridge_particle_offset, np.int32), C.c_int32)
        s.bridge_particle_compound = _ptr(arr(self.bridge_particle_compound, np.int32), C.c_int32)
        s.bridge_particle_local = _ptr(arr(self.bridge_particle_local, np.int32), C.c_int32)
        s.bridge_bond_offset = _ptr(arr(self.bridge_bond_offset, np.int32), C.c_int32)
        s.bridge_bonds = _ptr(arr(self.bridge_bonds, PAIR_DTYPE), PairBond)
        s.bridge_angle_offset = _ptr(arr(self.bridge_angle_offset, np.int32), C.c_int32)
        s.bridge_angles = _ptr(arr(self.bridge_angles, ANGLE_DTYPE), AngleBond)
        s.bridge_dihedral_offset = _ptr(arr(self.bridge_dihedral_offset, np.int32), C.c_int32)
        s.bridge_dihedrals = _ptr(arr(self.bridge_dihedrals, DIHEDRAL_DTYPE), DihedralBond)
        s.n_solvents = self.n_solvents
        s.solvent_positions = _ptr(arr(self.solvent_positions, np.float32), C.c_float)
        s.solvent_prev_positions = _ptr(arr(self.solvent_prev_positions, np.float32), C.c_float)
        return s

    # ---- persistence of fixtures (npz, no pickle) ----
    _ARRAYS = ("forcefield", "compound_offset", "key_particle_index", "confining_particle_sphere",
               "chain_positions", "chain_prev_positions", "chain_atom_types", "bond_offset", "bonds",
               "angle_offset", "angles", "dihedral_offset", "dihedrals", "ignore_offset", "ignore_compound",
               "ignore_local", "bridge_particle_offset", "bridge_particle_compound", "bridge_particle_local",
               "bridge_bond_offset", "bridge_bonds", "bridge_angle_offset", "bridge_angles",
               "bridge_dihedral_offset", "bridge_dihedrals", "solvent_positions", "solvent_prev_positions")

    def save(self, path: str, **extra):
        d = {k: getattr(self, k) for k in self._ARRAYS}
        d["scalars"] = np.array([self.box_len, self.dt, self.thermostat_target], dtype=np.float64)
        d.update(extra)
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path: str) -> "LimaSystem":
        z = np.load(path, allow_pickle=False)
        sc = z["scalars"]
        kw = {k: z[k] for k in cls._ARRAYS}
        return cls(box_len=float(sc[0]), dt=float(sc[1]), thermostat_target=float(sc[2]), **kw)

    def copy(self) -> "LimaSystem":
        kw = {k: np.array(getattr(self, k), copy=True) for k in self._ARRAYS}
        return LimaSystem(box_len=self.box_len, dt=self.dt, thermostat_target=self.thermostat_target,
                          meta=dict(self.meta), **kw)

"""System ingest: CHARMM .prm files plus a GROMACS .gro/.top pair -> LimaSystem.

ctypes binding of include/lima_ingest.h (quantom_amd/liblima_ingest.so, host code).
The three stages mirror the reference's build of a simulation from files:

* `merge_forcefield`   — LIMA_ForcefieldMaker's mergeForcefieldFiles
                         (LIMA_ForcefieldMaker.cpp:116-141)
* `prepare_forcefield` — its prepareForcefieldForSimulation (:70-114)
* `build_system`       — Environment::Environment's molecule and box build
                         (Environment.cpp:16-45: ForceFieldMaker::buildForcefield,
                         CompoundBuilder::buildMolecule, BoxBuilder)

`ingest` runs all three.  Files between the stages have the reference's names and
formats, so a directory prepared by either side can be read by the other.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .abi import (ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE, STATUS_NAMES, CSystem, LimaSystem)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblima_ingest.so")
_lib = None

# The reference's parameter files, in its merge order (LIMA_ForcefieldMaker.cpp:124-126).
CHARMM_FILES = ("par_all36_lipid.prm", "par_all36_na.prm", "par_all36m_prot.prm")


class IngestError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class CIngestParams(C.Structure):
    _fields_ = [("min_residue_id", C.c_int32), ("max_residue_id", C.c_int32), ("ignore_hydrogens", C.c_int32),
                ("max_compound_particles", C.c_int32), ("max_atoms_in_residue", C.c_int32),
                ("max_bridge_particles", C.c_int32), ("max_exclusions", C.c_int32), ("add_solvents", C.c_int32),
                ("box_len", C.c_float), ("dt", C.c_double), ("thermostat_target", C.c_float), ("seed", C.c_uint32)]


class CIngestInfo(C.Structure):
    _fields_ = [("n_compounds", C.c_int32), ("n_bridges", C.c_int32), ("n_chain", C.c_int32),
                ("n_solvents", C.c_int32), ("n_atom_types", C.c_int32), ("n_solvent_candidates", C.c_int32),
                ("n_random_draws", C.c_int64), ("offset", C.c_float * 3)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # LIMA_INGEST_LIB: another build of the same source (the sanitizer test's ASan/UBSan build)
        path = os.environ.get("LIMA_INGEST_LIB", LIB_PATH)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `python -m quantom_amd.build`")
        L = C.CDLL(path)
        L.lima_ff_merge.argtypes = [C.POINTER(C.c_char_p), C.c_int32, C.c_char_p]
        L.lima_ff_prepare.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        L.lima_ingest_default_params.argtypes = [C.POINTER(CIngestParams)]
        L.lima_ingest_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(CIngestParams),
                                         C.POINTER(C.c_void_p)]
        L.lima_ingest_system.argtypes = [C.c_void_p, C.POINTER(CSystem)]
        L.lima_ingest_get_info.argtypes = [C.c_void_p, C.POINTER(CIngestInfo)]
        L.lima_ingest_global_ids.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.lima_ingest_destroy.argtypes = [C.c_void_p]
        L.lima_ingest_last_error.restype = C.c_char_p
        _lib = L
    return _lib


EXPORTS = ("lima_ff_merge", "lima_ff_prepare", "lima_ingest_default_params", "lima_ingest_create",
           "lima_ingest_system", "lima_ingest_get_info", "lima_ingest_global_ids", "lima_ingest_destroy",
           "lima_ingest_last_error")


def _check(status: int):
    if status != 0:
        raise IngestError(status, lib().lima_ingest_last_error().decode())


def default_params() -> dict:
    p = CIngestParams()
    lib().lima_ingest_default_params(C.byref(p))
    return {name: getattr(p, name) for name, _ in CIngestParams._fields_}


def merge_forcefield(prm_paths, out_dir: str) -> None:
    """Writes out_dir/LIMA_ffnonbonded.txt and out_dir/LIMA_ffbonded.txt."""
    paths = (C.c_char_p * len(prm_paths))(*[os.fsencode(p) for p in prm_paths])
    _check(lib().lima_ff_merge(paths, len(prm_paths), os.fsencode(out_dir)))


def prepare_forcefield(conf_gro: str, topol_top: str, ff_dir: str) -> None:
    """Writes ff_dir/LIMA_ffnonbonded_filtered.txt and ff_dir/LIMA_ffbonded_filtered.txt."""
    _check(lib().lima_ff_prepare(os.fsencode(conf_gro), os.fsencode(topol_top), os.fsencode(ff_dir)))


@dataclass
class IngestInfo:
    n_compounds: int
    n_bridges: int
    n_chain: int
    n_solvents: int
    n_atom_types: int
    n_solvent_candidates: int
    n_random_draws: int
    offset: tuple
    global_ids: np.ndarray


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    buf = C.cast(ptr, C.POINTER(C.c_uint8 * (n * np.dtype(dtype).itemsize))).contents
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def _from_c(s: CSystem) -> LimaSystem:
    nc, nb, ns = s.n_compounds, s.n_bridges, s.n_solvents
    coff = _copy(s.compound_offset, nc + 1, np.int32)
    nch = int(coff[-1])
    off = lambda p, n: _copy(p, n + 1, np.int32)
    bo, ao, do_ = off(s.bond_offset, nc), off(s.angle_offset, nc), off(s.dihedral_offset, nc)
    io = off(s.ignore_offset, nch)
    bpo, bbo, bao, bdo = (off(s.bridge_particle_offset, nb), off(s.bridge_bond_offset, nb),
                          off(s.bridge_angle_offset, nb), off(s.bridge_dihedral_offset, nb))
    ff = np.array([[p.mass, p.sigma, p.epsilon] for p in s.forcefield.particle_parameters], np.float32)
    return LimaSystem(
        box_len=float(s.box_len), dt=float(s.dt), thermostat_target=float(s.thermostat_target), forcefield=ff,
        compound_offset=coff, key_particle_index=_copy(s.key_particle_index, nc, np.int32),
        confining_particle_sphere=_copy(s.confining_particle_sphere, nc, np.float32),
        chain_positions=_copy(s.chain_positions, 3 * nch, np.float32).reshape(-1, 3),
        chain_prev_positions=_copy(s.chain_prev_positions, 3 * nch, np.float32).reshape(-1, 3),
        chain_atom_types=_copy(s.chain_atom_types, nch, np.uint8),
        bond_offset=bo, bonds=_copy(s.bonds, int(bo[-1]), PAIR_DTYPE),
        angle_offset=ao, angles=_copy(s.angles, int(ao[-1]), ANGLE_DTYPE),
        dihedral_offset=do_, dihedrals=_copy(s.dihedrals, int(do_[-1]), DIHEDRAL_DTYPE),
        ignore_offset=io, ignore_compound=_copy(s.ignore_compound, int(io[-1]), np.int32),
        ignore_local=_copy(s.ignore_local, int(io[-1]), np.int32),
        bridge_particle_offset=bpo, bridge_particle_compound=_copy(s.bridge_particle_compound, int(bpo[-1]), np.int32),
        bridge_particle_local=_copy(s.bridge_particle_local, int(bpo[-1]), np.int32),
        bridge_bond_offset=bbo, bridge_bonds=_copy(s.bridge_bonds, int(bbo[-1]), PAIR_DTYPE),
        bridge_angle_offset=bao, bridge_angles=_copy(s.bridge_angles, int(bao[-1]), ANGLE_DTYPE),
        bridge_dihedral_offset=bdo, bridge_dihedrals=_copy(s.bridge_dihedrals, int(bdo[-1]), DIHEDRAL_DTYPE),
        solvent_positions=_copy(s.solvent_positions, 3 * ns, np.float32).reshape(-1, 3),
        solvent_prev_positions=_copy(s.solvent_prev_positions, 3 * ns, np.float32).reshape(-1, 3))


def build_system(conf_gro: str, topol_top: str, ff_dir: str, **params) -> tuple[LimaSystem, IngestInfo]:
    """Stage 3; keyword arguments override lima_ingest_default_params (the reference's values)."""
    p = CIngestParams()
    lib().lima_ingest_default_params(C.byref(p))
    for k, v in params.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown ingest parameter {k!r}")
        setattr(p, k, v)
    h = C.c_void_p()
    _check(lib().lima_ingest_create(os.fsencode(conf_gro), os.fsencode(topol_top), os.fsencode(ff_dir),
                                    C.byref(p), C.byref(h)))
    try:
        cs = CSystem()
        _check(lib().lima_ingest_system(h, C.byref(cs)))
        system = _from_c(cs)
        inf = CIngestInfo()
        _check(lib().lima_ingest_get_info(h, C.byref(inf)))
        ids = np.zeros(max(inf.n_chain, 1), np.int32)
        _check(lib().lima_ingest_global_ids(h, ids.ctypes.data_as(C.POINTER(C.c_int32))))
        info = IngestInfo(inf.n_compounds, inf.n_bridges, inf.n_chain, inf.n_solvents, inf.n_atom_types,
                          inf.n_solvent_candidates, int(inf.n_random_draws), tuple(inf.offset), ids[:inf.n_chain])
    finally:
        lib().lima_ingest_destroy(h)
    system.meta["source"] = os.path.abspath(conf_gro)
    return system, info


def ingest(conf_gro: str, topol_top: str, prm_paths, work_dir: str, **params) -> tuple[LimaSystem, IngestInfo]:
    """All three stages; the intermediate files are written to work_dir."""
    os.makedirs(work_dir, exist_ok=True)
    merge_forcefield(prm_paths, work_dir)
    prepare_forcefield(conf_gro, topol_top, work_dir)
    return build_system(conf_gro, topol_top, work_dir, **params)

"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from quantom_amd.abi import CEnergies, CRunParams, CSystem, LimaSystem, RunParams

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs: dict = {}

KIND_CC, KIND_CS, KIND_SC, KIND_SS = 0, 1, 2, 3


def lib(fast: bool = False) -> C.CDLL:
    name = "liblima_oracle_fast.so" if fast else "liblima_oracle.so"
    if name not in _libs:
        # LIMA_ORACLE_DIR: another build of the same source (the sanitizer test's ASan/UBSan build)
        path = os.path.join(os.environ.get("LIMA_ORACLE_DIR", _HERE), name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(CSystem), C.POINTER(CRunParams), C.c_int, C.POINTER(C.c_int)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_step.argtypes = [C.c_void_p, C.c_int]
        L.oracle_compute_forces.argtypes = [C.c_void_p, dp, dp]
        L.oracle_get_forces.argtypes = [C.c_void_p, dp, dp]
        L.oracle_rebuild_neighbors.argtypes = [C.c_void_p]
        L.oracle_get_state.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.oracle_set_state.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.oracle_get_energies.argtypes = [C.c_void_p, C.POINTER(CEnergies)]
        L.oracle_get_lambda.restype = C.c_double
        L.oracle_get_lambda.argtypes = [C.c_void_p]
        L.oracle_set_lambda.argtypes = [C.c_void_p, C.c_double]
        L.oracle_get_cs_forces.argtypes = [C.c_void_p, dp]
        L.oracle_get_ref_temperature.restype = C.c_double
        L.oracle_get_ref_temperature.argtypes = [C.c_void_p]
        L.oracle_get_step.restype = C.c_long
        L.oracle_get_step.argtypes = [C.c_void_p]
        L.oracle_get_neighbor_counts.argtypes = [C.c_void_p, C.c_int, ip]
        L.oracle_get_neighbors.argtypes = [C.c_void_p, C.c_int, ip, C.c_int]
        L.oracle_set_neighbors.argtypes = [C.c_void_p, C.c_int, ip, ip, C.c_int]
        L.oracle_num_threads.restype = C.c_int
        fp = C.POINTER(C.c_float)
        L.oracle_lj_force_f32.argtypes = [fp, fp, C.c_float, C.c_float, fp]
        L.oracle_lj_force_f64.argtypes = [dp, dp, C.c_double, C.c_double, dp]
        L.oracle_pairbond_f32.argtypes = [fp, fp, C.c_float, C.c_float, fp, dp]
        L.oracle_pairbond_f64.argtypes = [dp, dp, C.c_double, C.c_double, dp, dp]
        L.oracle_anglebond_f32.argtypes = [fp, fp, fp, C.c_float, C.c_float, fp, dp]
        L.oracle_anglebond_f64.argtypes = [dp, dp, dp, C.c_double, C.c_double, dp, dp]
        L.oracle_dihedral_f32.argtypes = [fp, fp, fp, fp, C.c_float, C.c_float, C.c_int, fp, dp]
        L.oracle_dihedral_f64.argtypes = [dp, dp, dp, dp, C.c_double, C.c_double, C.c_int, dp, dp]
        L.oracle_integrate_f32.argtypes = [fp, fp, fp, C.c_double, C.c_double, C.c_float, C.c_float]
        L.oracle_hyperpos_f32.argtypes = [fp, fp, C.c_float]
        L.oracle_pbc_f32.argtypes = [fp, C.c_float]
        L.oracle_get_angle_f32.restype = C.c_float
        L.oracle_get_angle_f32.argtypes = [fp, fp]
        L.oracle_within_f32.restype = C.c_int
        L.oracle_within_f32.argtypes = [fp, fp, C.c_float, C.c_float, C.c_float]
        _libs[name] = L
    return _libs[name]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Oracle:
    """One system stepped on the CPU with the reference semantics (FP32 or FP64)."""

    def __init__(self, system: LimaSystem, params: RunParams | None = None, precision: int = 32,
                 fast: bool = False):
        self.sys = system
        self.params = params or RunParams()
        self.L = lib(fast)
        cs = system.to_c()
        cp = self.params.to_c()
        st = C.c_int(0)
        self.h = self.L.oracle_create(C.byref(cs), C.byref(cp), precision, C.byref(st))
        self.status = st.value
        self.nchain = system.n_chain
        self.ns = system.n_solvents
        self.nc = system.n_compounds

    def close(self):
        if getattr(self, "h", None):
            self.L.oracle_destroy(self.h)
            self.h = None

    __del__ = close

    def step(self, n: int = 1) -> int:
        return self.L.oracle_step(self.h, n)

    def compute_forces(self):
        cf = np.zeros((self.nchain, 3))
        sf = np.zeros((self.ns, 3))
        st = self.L.oracle_compute_forces(self.h, _dp(cf), _dp(sf))
        return cf, sf, st

    def forces(self):
        cf = np.zeros((self.nchain, 3))
        sf = np.zeros((self.ns, 3))
        self.L.oracle_get_forces(self.h, _dp(cf), _dp(sf))
        return cf, sf

    def cs_forces(self):
        """computeSolventToCompoundLJForces of the last force pass per chain particle (-1s when not run)."""
        out = np.zeros((self.nchain, 3))
        self.L.oracle_get_cs_forces(self.h, _dp(out))
        return out

    def rebuild_neighbors(self) -> int:
        return self.L.oracle_rebuild_neighbors(self.h)

    def state(self):
        cx = np.zeros((self.nchain, 3)); cxm = np.zeros((self.nchain, 3))
        sx = np.zeros((self.ns, 3)); sxm = np.zeros((self.ns, 3))
        self.L.oracle_get_state(self.h, _dp(cx), _dp(cxm), _dp(sx), _dp(sxm))
        return cx, cxm, sx, sxm

    def set_state(self, cx, cxm, sx, sxm):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (cx, cxm, sx, sxm)]
        self.L.oracle_set_state(self.h, *[_dp(v) for v in a])

    def energies(self) -> dict:
        e = CEnergies()
        self.L.oracle_get_energies(self.h, C.byref(e))
        return e.as_dict()

    @property
    def lam(self) -> float:
        return self.L.oracle_get_lambda(self.h)

    @lam.setter
    def lam(self, v: float):
        self.L.oracle_set_lambda(self.h, v)

    @property
    def ref_temperature(self) -> float:
        """The temperature the reference-cadence thermostat last evaluated."""
        return self.L.oracle_get_ref_temperature(self.h)

    @property
    def step_count(self) -> int:
        return self.L.oracle_get_step(self.h)

    def neighbors(self, kind: int) -> list:
        n = self.nc if kind in (KIND_CC, KIND_CS) else self.ns
        counts = np.zeros(max(n, 1), dtype=np.int32)
        self.L.oracle_get_neighbor_counts(self.h, kind, _ip(counts))
        stride = int(counts.max()) if n else 0
        stride = max(stride, 1)
        ids = np.full((max(n, 1), stride), -1, dtype=np.int32)
        self.L.oracle_get_neighbors(self.h, kind, _ip(ids), stride)
        return [ids[i, :counts[i]].copy() for i in range(n)]

    def set_neighbors(self, kind: int, lists: list):
        n = len(lists)
        counts = np.array([len(l) for l in lists] or [0], dtype=np.int32)
        stride = max(1, int(counts.max()))
        ids = np.zeros((max(n, 1), stride), dtype=np.int32)
        for i, l in enumerate(lists):
            ids[i, :len(l)] = l
        self.L.oracle_set_neighbors(self.h, kind, _ip(counts), _ip(ids), stride)


def num_threads(fast: bool = False) -> int:
    return lib(fast).oracle_num_threads()

"""Python binding of the gfx950 engine over the C ABI (include/lima.h).

`Engine` keeps the reference's step API (Engine.cuh:109-208):
`Engine(simulation, params)` builds the device state and the first neighbour
lists (Engine.cu:9-38); `deviceMaster()` queues one step (Engine.cu:44-47,
400-451) as a captured-graph replay; `hostMaster()` is the reference's per-step
host hook (Engine.cu:49-75), whose work (thermostat, neighbour-list rebuild) now
runs on the device inside the step, so here it waits for the device once per
list period and surfaces errors.  A `Simulation` holds one or more systems
(replicas) stepped together.

The native library must be present: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (CAdjointSpec, CEnergies, CRunParams, CSystem, LimaSystem, N_PARAMETERS, RunParams, STATUS_NAMES,
                  LIMA_OK)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblima_hip.so")
_lib = None

EXPORTS = (
    "lima_version", "lima_device_count", "lima_create", "lima_destroy", "lima_last_error", "lima_step",
    "lima_step_async", "lima_sync", "lima_rebuild_neighbors", "lima_get_state", "lima_set_state",
    "lima_get_forces", "lima_compute_forces", "lima_get_energies", "lima_get_neighbor_counts",
    "lima_get_neighbors", "lima_set_neighbors", "lima_get_timings", "lima_time_kernels", "lima_set_kernel_clock", "lima_get_kernel_clock", "lima_get_config",
    "lima_autotune", "lima_get_parameters", "lima_set_parameters", "lima_adjoint_run", "lima_record_frames",
    "lima_step_observed", "lima_step_observed_noalloc", "lima_step_logged",
)


class LimaStatusError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def load_library() -> C.CDLL:
    """Load liblima_hip.so; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m quantom_amd.build`")
    L = C.CDLL(LIB_PATH)
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.lima_version.restype = C.c_char_p
    L.lima_device_count.argtypes = [C.POINTER(C.c_int)]
    L.lima_create.argtypes = [C.POINTER(CSystem), C.c_int32, C.POINTER(CRunParams), C.POINTER(vp)]
    L.lima_destroy.argtypes = [vp]
    L.lima_last_error.restype = C.c_char_p
    L.lima_last_error.argtypes = [vp]
    for name in ("lima_step", "lima_step_async", "lima_autotune"):
        getattr(L, name).argtypes = [vp, C.c_int32]
    L.lima_sync.argtypes = [vp]
    L.lima_rebuild_neighbors.argtypes = [vp]
    L.lima_get_state.argtypes = [vp, fp, fp, fp, fp]
    L.lima_set_state.argtypes = [vp, fp, fp, fp, fp]
    L.lima_get_forces.argtypes = [vp, fp, fp]
    L.lima_compute_forces.argtypes = [vp, fp, fp]
    L.lima_get_energies.argtypes = [vp, C.POINTER(CEnergies)]
    L.lima_get_neighbor_counts.argtypes = [vp, ip, ip, ip]
    L.lima_get_neighbors.argtypes = [vp, ip, ip, ip]
    L.lima_set_neighbors.argtypes = [vp, ip, ip, ip, ip, ip, ip]
    L.lima_get_timings.argtypes = [vp, C.POINTER(C.c_double)]
    L.lima_get_config.argtypes = [vp, ip, C.c_int32]
    L.lima_time_kernels.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.lima_set_kernel_clock.argtypes = [vp, C.c_int32]
    L.lima_get_kernel_clock.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    dp = C.POINTER(C.c_double)
    L.lima_get_parameters.argtypes = [vp, C.c_int32, dp]
    L.lima_set_parameters.argtypes = [vp, C.c_int32, dp]
    L.lima_adjoint_run.argtypes = [vp, C.POINTER(CAdjointSpec), dp, dp]
    L.lima_record_frames.argtypes = [vp, C.c_int32, fp, fp, dp]
    L.lima_step_observed.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(CEnergies)]
    L.lima_step_observed_noalloc.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(CEnergies)]
    L.lima_step_logged.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, fp, fp]
    _lib = L
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


class Simulation:
    """Host description of the systems stepped together (the reference's Simulation, Simulation.cuh:97-178)."""

    def __init__(self, systems, params: RunParams | None = None):
        self.systems = list(systems) if isinstance(systems, (list, tuple)) else [systems]
        self.params = params or RunParams()
        self.n_chain = sum(s.n_chain for s in self.systems)
        self.n_solvents = sum(s.n_solvents for s in self.systems)
        self.n_compounds = sum(s.n_compounds for s in self.systems)
        self.chain_begin = np.cumsum([0] + [s.n_chain for s in self.systems])
        self.solvent_begin = np.cumsum([0] + [s.n_solvents for s in self.systems])
        self.compound_begin = np.cumsum([0] + [s.n_compounds for s in self.systems])


class Engine:
    """One engine handle (single-threaded use)."""

    def __init__(self, simulation: Simulation | LimaSystem | list, params: RunParams | None = None):
        if not isinstance(simulation, Simulation):
            simulation = Simulation(simulation, params)
        self.simulation = simulation
        self.params = simulation.params
        self.L = load_library()
        arr = (CSystem * len(simulation.systems))()
        self._keep = []
        for k, s in enumerate(simulation.systems):
            arr[k] = s.to_c()
            self._keep.append(s._keep)
        self._cparams = self.params.to_c()
        h = C.c_void_p()
        st = self.L.lima_create(arr, len(simulation.systems), C.byref(self._cparams), C.byref(h))
        if st != LIMA_OK:
            raise LimaStatusError(st, self.L.lima_last_error(None).decode())
        self.h = h
        self.steps = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.lima_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != LIMA_OK:
            raise LimaStatusError(st, self.L.lima_last_error(self.h).decode())

    # ---- the reference's step API ----
    def deviceMaster(self):
        """Queues one step on the device (Engine::deviceMaster, Engine.cu:44-47)."""
        self.step_async(1)
        self._queued = getattr(self, "_queued", 0) + 1

    def hostMaster(self):
        """Per-step host hook (Engine::hostMaster, Engine.cu:49-75).  Its work runs on the device
        inside the step; this waits once per list period and raises device errors."""
        if getattr(self, "_queued", 0) >= max(1, self.params.nlist_interval):
            self.sync()

    def step(self, n: int = 1):
        self._check(self.L.lima_step(self.h, n))
        self.steps += n

    def step_async(self, n: int):
        self._check(self.L.lima_step_async(self.h, n))
        self.steps += n

    def sync(self):
        self._queued = 0
        self._check(self.L.lima_sync(self.h))

    def rebuild_neighbors(self):
        self._check(self.L.lima_rebuild_neighbors(self.h))

    # ---- state ----
    def get_state(self):
        sim = self.simulation
        cx = np.zeros((sim.n_chain, 3), np.float32); cxm = np.zeros_like(cx)
        sx = np.zeros((sim.n_solvents, 3), np.float32); sxm = np.zeros_like(sx)
        self._check(self.L.lima_get_state(self.h, _f(cx), _f(cxm), _f(sx), _f(sxm)))
        return cx, cxm, sx, sxm

    def set_state(self, cx=None, cxm=None, sx=None, sxm=None):
        """Overwrites the given arrays of the state; None keeps that array (NULL in the C call)."""
        sim = self.simulation
        shapes = ((sim.n_chain, 3), (sim.n_chain, 3), (sim.n_solvents, 3), (sim.n_solvents, 3))
        args = []
        for v, shape in zip((cx, cxm, sx, sxm), shapes):
            if v is None:
                args.append(None)
                continue
            v = np.ascontiguousarray(v, dtype=np.float32)
            if v.shape != shape:
                raise ValueError(f"state array of shape {v.shape}, expected {shape}")
            args.append(v)
        self._keep_state = args     # alive for the duration of the call
        self._check(self.L.lima_set_state(self.h, *[_f(v) for v in args]))

    def forces(self):
        sim = self.simulation
        cf = np.zeros((sim.n_chain, 3), np.float32); sf = np.zeros((sim.n_solvents, 3), np.float32)
        self._check(self.L.lima_get_forces(self.h, _f(cf), _f(sf)))
        return cf, sf

    def compute_forces(self):
        sim = self.simulation
        cf = np.zeros((sim.n_chain, 3), np.float32); sf = np.zeros((sim.n_solvents, 3), np.float32)
        self._check(self.L.lima_compute_forces(self.h, _f(cf), _f(sf)))
        return cf, sf

    def energies(self) -> list:
        n = len(self.simulation.systems)
        arr = (CEnergies * n)()
        self._check(self.L.lima_get_energies(self.h, arr))
        return [arr[k].as_dict() for k in range(n)]

    def step_observed(self, n: int, every: int = 1, noalloc: bool = False) -> list:
        """n steps; the first and every `every`-th step after it record each system's energies.
        Returns [record][system] dicts (lima_energies fields).  noalloc: the entry point that
        stages the series in the library's fixed device array (no device allocation)."""
        R = len(self.simulation.systems)
        nrec = (n + every - 1) // every
        arr = (CEnergies * max(nrec * R, 1))()
        fn = self.L.lima_step_observed_noalloc if noalloc else self.L.lima_step_observed
        self._check(fn(self.h, n, every, arr))
        self.steps += n
        return [[arr[k * R + r].as_dict() for r in range(R)] for k in range(nrec)]

    def step_logged(self, n: int, system: int = 0, slots_per_compound: int = 64, train: bool = True):
        """n steps with the reference's per-step logs of one system (logdata.bin, sim_traindata.bin).
        Returns (logdata [n, 10], traindata [n, n_compounds, slots, 3, 3] or None)."""
        s = self.simulation.systems[system]
        log = np.zeros((max(n, 1), 10), np.float32)
        tr = np.zeros((max(n, 1), s.n_compounds, slots_per_compound, 3, 3), np.float32) if train else None
        self._check(self.L.lima_step_logged(self.h, n, system, slots_per_compound, _f(log),
                                            _f(tr) if train else None))
        self.steps += n
        return log[:n], (tr[:n] if train else None)

    # ---- neighbour lists (global ids within the concatenation) ----
    def neighbor_counts(self):
        sim = self.simulation
        cc = np.zeros(max(sim.n_compounds, 1), np.int32); cs = np.zeros_like(cc)
        ss = np.zeros(max(sim.n_solvents, 1), np.int32)
        self._check(self.L.lima_get_neighbor_counts(self.h, _i(cc), _i(cs), _i(ss)))
        return cc[:sim.n_compounds], cs[:sim.n_compounds], ss[:sim.n_solvents]

    def neighbors(self):
        """Returns (compound->compound, compound->solvent, solvent->solvent) lists of global ids."""
        sim, p = self.simulation, self.params
        ccn, csn, ssn = self.neighbor_counts()
        cci = np.zeros((max(sim.n_compounds, 1), p.max_compound_neighbors), np.int32)
        csi = np.zeros((max(sim.n_compounds, 1), p.max_compound_solvents), np.int32)
        ssi = np.zeros((max(sim.n_solvents, 1), p.max_solvent_neighbors), np.int32)
        self._check(self.L.lima_get_neighbors(self.h, _i(cci), _i(csi), _i(ssi)))
        return ([cci[c, :ccn[c]].copy() for c in range(sim.n_compounds)],
                [csi[c, :csn[c]].copy() for c in range(sim.n_compounds)],
                [ssi[s, :ssn[s]].copy() for s in range(sim.n_solvents)])

    def set_neighbors(self, cc=None, cs=None, ss=None):
        sim, p = self.simulation, self.params

        def pack(lists, n, cap):
            if lists is None:
                return None, None
            cnt = np.array([len(l) for l in lists] or [0], np.int32)
            ids = np.zeros((max(n, 1), cap), np.int32)
            for k, l in enumerate(lists):
                ids[k, :min(len(l), cap)] = l[:cap]
            return cnt, ids

        a = pack(cc, sim.n_compounds, p.max_compound_neighbors)
        b = pack(cs, sim.n_compounds, p.max_compound_solvents)
        c = pack(ss, sim.n_solvents, p.max_solvent_neighbors)
        self._check(self.L.lima_set_neighbors(self.h, _i(a[0]), _i(a[1]), _i(b[0]), _i(b[1]), _i(c[0]), _i(c[1])))

    def timings(self) -> dict:
        t = (C.c_double * 5)()
        self._check(self.L.lima_get_timings(self.h, t))
        return {"device_us": t[0], "rebuild_us": t[1], "host_us": t[2], "steps": int(t[3]), "rebuilds": int(t[4])}

    KERNELS = ("solvent_chain", "compound_listed", "compound_integrate", "solvent_integrate", "bridge", "rebuild",
               "solvent_listed", "compound_local")

    def time_kernels(self, reps: int = 20) -> dict:
        """Average microseconds per launch of each step kernel, launched alone (HIP events)."""
        t = (C.c_double * 8)()
        self._check(self.L.lima_time_kernels(self.h, reps, t))
        return dict(zip(self.KERNELS, list(t)))

    def set_kernel_clock(self, on: bool = True):
        """Time the all-pairs kernel inside every step from now on (re-captures the step graph)."""
        self._check(self.L.lima_set_kernel_clock(self.h, 1 if on else 0))

    def kernel_clock(self) -> tuple[float, int]:
        """Total in-step span of the all-pairs kernel since the clock was switched on (us), and the
        launches counted; differences of two reads give the mean over the steps between them."""
        us = C.c_double()
        n = C.c_int64()
        self._check(self.L.lima_get_kernel_clock(self.h, C.byref(us), C.byref(n)))
        return us.value, n.value

    def config(self) -> dict:
        c = (C.c_int32 * 12)()
        self._check(self.L.lima_get_config(self.h, c, 12))
        keys = ("solvents_per_thread", "chunk_particles", "solvent_chain_items", "max_chunks", "use_graph",
                "n_cells", "n_solvent_tiles", "n_replicas", "listed_grid", "listed_lanes", "listed_packed",
                "fused_thermostat")
        return dict(zip(keys, list(c)))

    def autotune(self, trial_steps: int = 20) -> dict:
        self._check(self.L.lima_autotune(self.h, trial_steps))
        return self.config()


    # ---- parameters and reverse-mode gradients ----
    def get_parameters(self, system: int = 0) -> np.ndarray:
        """[sigma(32), epsilon(32), kb, b0, k_theta, theta_0] of one system."""
        v = np.zeros(N_PARAMETERS, np.float64)
        self._check(self.L.lima_get_parameters(self.h, system, _d(v)))
        return v

    def set_parameters(self, values, system: int = 0):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (N_PARAMETERS,):
            raise ValueError(f"expected {N_PARAMETERS} parameters")
        self._check(self.L.lima_set_parameters(self.h, system, _d(v)))

    def record_frames(self, n_frames: int):
        """Forces and potential energies of frames 0..K-1 from the current state (state unchanged)."""
        sim = self.simulation
        cf = np.zeros((n_frames, sim.n_chain, 3), np.float32)
        sf = np.zeros((n_frames, sim.n_solvents, 3), np.float32)
        pot = np.zeros((n_frames, len(sim.systems)), np.float64)
        self._check(self.L.lima_record_frames(self.h, n_frames, _f(cf), _f(sf), _d(pot)))
        return cf, sf, pot

    def adjoint(self, n_frames: int, energy_weight=None, force_weight=None, energy_target=None,
                chain_force_target=None, solvent_force_target=None, checkpoint_interval: int = 0,
                gradient: bool = True):
        """Loss per system and, with `gradient`, dLoss/dparameters [n_systems, 68] (state unchanged)."""
        sim = self.simulation
        R = len(sim.systems)

        def arr(v, shape, dt):
            if v is None:
                return None
            a = np.ascontiguousarray(v, dtype=dt)
            if a.size != int(np.prod(shape)):
                raise ValueError(f"expected {shape} values")
            return a.reshape(shape)

        keep = [arr(energy_weight, (R,), np.float64), arr(force_weight, (R,), np.float64),
                arr(energy_target, (n_frames, R), np.float64),
                arr(chain_force_target, (n_frames, sim.n_chain, 3), np.float32),
                arr(solvent_force_target, (n_frames, sim.n_solvents, 3), np.float32)]
        spec = CAdjointSpec(n_frames, checkpoint_interval, _d(keep[0]), _d(keep[1]), _d(keep[2]), _f(keep[3]),
                            _f(keep[4]))
        loss = np.zeros(R, np.float64)
        grad = np.zeros((R, N_PARAMETERS), np.float64) if gradient else None
        self._check(self.L.lima_adjoint_run(self.h, C.byref(spec), _d(loss), _d(grad)))
        return (loss, grad) if gradient else loss


def version() -> str:
    return load_library().lima_version().decode()


def device_count() -> int:
    n = C.c_int(0)
    load_library().lima_device_count(C.byref(n))
    return n.value

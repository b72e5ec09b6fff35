"""GPU parity on the reference's own demo system: T4 lysozyme (Demo/6lzm) as the ingest
builds it (tests/golden/t4lysozyme.npz: 36 compounds, 35 bridges, 1,293 heavy atoms,
5,000 waters, 7 nm box), stepped with the reference's constants: CUTOFF 0.7 nm
(Constants.cuh:52), dt 0.1 fs (Simulation.cuh:152), lists every 200 steps
(Constants.cuh:71) and the reference thermostat cadence (Engine.cu:49-75).

Bars:
  * neighbour lists: set-equal, bit-exact;
  * forces on identical positions: relative rms error <= 1e-5 against the FP32 and the
    FP64 oracle (as on the synthetic systems);
  * trajectories: the system's start is far from equilibrium (its Rmin-as-sigma
    parameters overlap neighbours; ForcefieldMerger.h:65), so rounding differences grow
    quickly.  The GPU must stay as close to the FP32 oracle as the FP32 oracle stays to
    the FP64 oracle: max |dx| (GPU - oracle32) <= 3 x max |dx| (oracle32 - oracle64) + 1e-6
    after the same number of steps.
"""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN, as_sets, rel_rms

pytestmark = pytest.mark.gpu


def _system():
    from quantom_amd.abi import LimaSystem
    return LimaSystem.load(os.path.join(GOLDEN, "t4lysozyme.npz"))


def _params(**kw):
    from quantom_amd.abi import THERMOSTAT_REFERENCE, RunParams
    base = dict(cutoff=0.7, nlist_interval=200, thermostat=THERMOSTAT_REFERENCE, max_solvent_neighbors=1024,
                max_compound_solvents=4096)
    base.update(kw)
    return RunParams(**base)


def _max_dx(a, b, L):
    d = np.abs(np.concatenate([np.ravel(x) - np.ravel(y) for x, y in zip(a, b)]))
    return float(np.minimum(d, L - d).max())


def test_t4lysozyme_lists_and_forces():
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SS, Oracle
    from quantom_amd.engine import Engine
    print("case: T4 lysozyme demo, lists + forces", flush=True)
    s = _system()
    p = _params()
    eng = Engine([s], p)
    o32 = Oracle(s, p, precision=32)
    o64 = Oracle(s, p, precision=64)
    cc, cs, ss = eng.neighbors()
    assert as_sets(cc) == as_sets(o32.neighbors(KIND_CC))
    assert as_sets(cs) == as_sets(o32.neighbors(KIND_CS))
    assert as_sets(ss) == as_sets(o32.neighbors(KIND_SS))
    cf, sf = eng.compute_forces()
    f32 = o32.compute_forces()
    f64 = o64.compute_forces()
    assert f32[2] == 0 and f64[2] == 0
    e32 = rel_rms(np.concatenate([cf.ravel(), sf.ravel()]), np.concatenate([f32[0].ravel(), f32[1].ravel()]))
    e64 = rel_rms(np.concatenate([cf.ravel(), sf.ravel()]), np.concatenate([f64[0].ravel(), f64[1].ravel()]))
    print(f"force rel rms vs oracle32 {e32:.2e} vs oracle64 {e64:.2e}", flush=True)
    assert e32 <= 1e-5 and e64 <= 1e-5
    eng.close()


@pytest.mark.parametrize("accum", [0, 1])
def test_t4lysozyme_trajectory(accum):
    from oracle.oracle import Oracle
    from quantom_amd.engine import Engine
    print(f"case: T4 lysozyme demo, 1 + 9 steps, accumulation={accum}", flush=True)
    s = _system()
    p = _params(accumulation=accum)
    eng = Engine([s], p)
    o32 = Oracle(s, p, precision=32)
    o64 = Oracle(s, p, precision=64)
    for n in (1, 9):
        eng.step(n)
        assert o32.step(n) == 0 and o64.step(n) == 0
        g, a, b = eng.get_state(), o32.state(), o64.state()
        d_gpu = _max_dx((g[0], g[2]), (a[0], a[2]), s.box_len)
        d_ref = _max_dx((a[0], a[2]), (b[0], b[2]), s.box_len)
        print(f"after {eng.steps} steps: gpu-oracle32 {d_gpu:.2e}, "
              f"oracle32-oracle64 {d_ref:.2e}", flush=True)
        assert d_gpu <= 3 * d_ref + 1e-6
    e = eng.energies()[0]
    eo = o32.energies()
    assert e["temperature"] == pytest.approx(eo["temperature"], rel=1e-3)
    eng.close()


def test_parameter_round_trip_keeps_per_term_bonded_constants():
    """The demo's bonds and angles carry their own constants (ingested per term).  Passing the
    parameter vector back unchanged keeps them: the forces stay bit-equal.  Changing a shared
    bonded entry is refused, since it would replace every term's constant."""
    from quantom_amd.abi import LIMA_E_UNSUPPORTED
    from quantom_amd.engine import Engine, LimaStatusError
    print("case: demo get/set parameters round trip", flush=True)
    s = _system()
    assert len(np.unique(s.bonds["kb"])) > 1 and len(np.unique(s.angles["k_theta"])) > 1
    eng = Engine([s], _params(nlist_interval=0))
    f0 = eng.compute_forces()
    v = eng.get_parameters(0)
    eng.set_parameters(v, 0)
    f1 = eng.compute_forces()
    np.testing.assert_array_equal(f0[0], f1[0])
    np.testing.assert_array_equal(f0[1], f1[1])
    for idx in (64, 66):
        w = v.copy()
        w[idx] *= 1.01
        with pytest.raises(LimaStatusError) as ex:
            eng.set_parameters(w, 0)
        assert ex.value.code == LIMA_E_UNSUPPORTED
    # the LJ entries of the same vector may still change
    w = v.copy()
    w[33] *= 1.01
    eng.set_parameters(w, 0)
    f2 = eng.compute_forces()
    assert np.abs(f2[0] - f0[0]).max() > 0

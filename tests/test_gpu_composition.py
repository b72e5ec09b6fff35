"""The HIP path on the hand-built micro-systems of tests/microsystems.py, through the C ABI.

Same expectations as tests/test_oracle_composition.py (derived by hand from the reference's
call sites, not from the oracle): the epsilon rule of each LJ call site, LJ_Ignores in the
intra- and inter-compound loops, the bridge carry-in stored once per step, the Berendsen
scalar, and the FP64 LJ energy per site.  Both accumulation forms are checked.

Bars: per-particle relative force error <= 5e-5 (FP32 pairs with a hardware reciprocal,
sums of 3-6 pairs with partial cancellation; the rule swaps these tests guard against move
the forces by more than 10%), energies <= 2e-5 relative, lambda <= 2e-7 relative.
"""
import math

import numpy as np
import pytest

from tests import microsystems as M

pytestmark = pytest.mark.gpu

ACCUMS = ["fp32", "fixed64"]


def _engine(system, accum="fp32", **kw):
    from quantom_amd.abi import ACCUM_FIXED64, ACCUM_FP32, THERMOSTAT_NONE, RunParams
    from quantom_amd.engine import Engine
    kw.setdefault("nlist_interval", 0)
    kw.setdefault("thermostat", THERMOSTAT_NONE)
    return Engine([system], RunParams(accumulation=ACCUM_FIXED64 if accum == "fixed64" else ACCUM_FP32, **kw))


def per_particle_rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))


@pytest.mark.parametrize("accum", ACCUMS)
def test_call_site_epsilon_rules(accum):
    print(f"case: micro sites system, {accum}", flush=True)
    s = M.sites_system()
    fc_want, fs_want = M.sites_expected(s)
    e = _engine(s, accum)
    fc, fs = e.compute_forces()
    assert per_particle_rel(fc, fc_want) < 5e-5
    assert per_particle_rel(fs, fs_want) < 5e-5


@pytest.mark.parametrize("accum", ACCUMS)
def test_lj_ignores_in_intra_and_inter_loops(accum):
    print(f"case: micro exclusions system, {accum}", flush=True)
    s = M.exclusions_system()
    want = M.exclusions_expected(s)
    fc, _ = _engine(s, accum).compute_forces()
    assert per_particle_rel(fc, want) < 5e-5


@pytest.mark.parametrize("accum", ACCUMS)
def test_bridge_force_is_stored_once_per_step(accum):
    print(f"case: micro bridge system, {accum}", flush=True)
    s = M.bridge_system()
    want = M.bridge_expected(s)
    e = _engine(s, accum)
    fc, _ = e.compute_forces()
    assert per_particle_rel(fc, want) < 5e-5
    fc2, _ = e.compute_forces()
    np.testing.assert_array_equal(fc, fc2)
    # the forces a step uses equal a fresh engine's on the same state
    e.step(1)
    cx, cxm, sx, sxm = e.get_state()
    fresh = _engine(s, accum)
    fresh.set_state(cx, cxm, sx, sxm)
    f_fresh, _ = fresh.compute_forces()
    e.step(1)
    f_used, _ = e.forces()
    np.testing.assert_array_equal(f_used, f_fresh)


def _berendsen(T, T0, dt, tau):
    lam = math.sqrt(1.0 + dt / tau * (T0 / T - 1.0))
    return min(max(lam, 0.9), 1.1)


@pytest.mark.parametrize("T_init,tau_steps", [(150.0, 10), (600.0, 4), (1200.0, 1)])
def test_berendsen_lambda_closed_form(T_init, tau_steps):
    from quantom_amd.abi import THERMOSTAT_BERENDSEN
    print(f"case: free solvents at {T_init} K, tau {tau_steps} dt", flush=True)
    s = M.free_solvents(T_init)
    dt = s.dt
    tau = tau_steps * dt
    x0 = s.solvent_positions.astype(np.float32)
    x1 = (np.float32(2) * x0 - s.solvent_prev_positions.astype(np.float32)).astype(np.float32)
    disp = (x1 - x0).astype(np.float32)
    m = M.f32(M.MASS_SOL)
    T1 = float((m * (disp.astype(float) ** 2).sum(1)).sum() / dt ** 2 / (3 * len(x0) * 8.3145))
    lam1 = _berendsen(T1, 300.0, dt, tau)
    assert abs(lam1 - 1.0) > 0.02
    e = _engine(s, thermostat=THERMOSTAT_BERENDSEN, thermostat_tau=tau)
    e.step(1)
    _, _, sx1, _ = e.get_state()
    np.testing.assert_array_equal(sx1, x1)
    e.step(1)
    _, _, sx2, _ = e.get_state()
    x2 = (x1 + (disp * np.float32(lam1)).astype(np.float32)).astype(np.float32)
    np.testing.assert_allclose(sx2, x2, rtol=0, atol=4e-7)
    assert abs(e.energies()[0]["thermostat_scalar"] - lam1) < 2e-7 * lam1     # applied on step 2
    d2 = (sx2 - sx1).astype(float)
    T2 = float((m * (d2 ** 2).sum(1)).sum() / dt ** 2 / (3 * len(x0) * 8.3145))
    e.step(1)
    assert abs(e.energies()[0]["thermostat_scalar"] - _berendsen(T2, 300.0, dt, tau)) < 1e-5


@pytest.mark.parametrize("accum", ACCUMS)
def test_lj_energies_per_site(accum):
    print(f"case: micro sites energies, {accum}", flush=True)
    s = M.sites_system()
    want = M.sites_energies(s)
    got = _engine(s, accum).energies()[0]
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=2e-5), k
    # two solvents at the LJ minimum: one pair of energy -eps
    sig, eps = M.f32(M.SIG_SOL), M.f32(M.EPS_SOL)
    s2 = M.build(6.0, [], [(2.0, 2.0, 2.0), (2.0 + 2 ** (1 / 6) * sig, 2.0, 2.0)])
    r32 = float(s2.solvent_positions[1, 0]) - float(s2.solvent_positions[0, 0])
    got2 = _engine(s2, accum).energies()[0]
    assert got2["lj_solvent_solvent"] == pytest.approx(M.lj_energy(r32, sig, eps), rel=2e-5)
    assert got2["lj_solvent_solvent"] == pytest.approx(-eps, rel=1e-4)

"""Pins how the oracle composes its pinned routines into the reference's kernels.

test_oracle_closed_forms.py pins each routine (calcLJForce, the bonded terms, the
integrator).  These tests pin what the kernels do with them, on hand-built systems of two
to five particles whose forces are derived by hand from the reference's call sites in
tests/microsystems.py, never from the oracle's code:

* the epsilon rule of each of the five LJ call sites (Engine.cu:643-646, 668-670, 689-691,
  711-713, 724-727), with types whose arithmetic and geometric means differ by 40-70%;
* LJ_Ignores applied in the intra- and inter-compound loops (Engine.cu:660, 683);
* the bridge force stored as the compound's carry-in and zeroed on load (Engine.cu:1211,
  914; Bodies.cuh:437-438): present once, never accumulated across steps;
* the Berendsen scalar lambda = clamp(sqrt(1 + dt/tau (T0/T - 1)), 0.9, 1.1) on the
  displacement, from a force-free system of known temperature;
* the FP64 LJ energy 4 eps (s^2 - s) per pair: known answers at r = sigma and at the
  minimum, per-site sums, and minus its gradient equal to the force.

Each test also checks that the plausible mistake it guards against (the other mixing rule,
a dropped exclusion, an added carry-in) moves the expected value far beyond its tolerance.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from quantom_amd.abi import THERMOSTAT_BERENDSEN, THERMOSTAT_NONE, RunParams
from tests import microsystems as M

P0 = RunParams(nlist_interval=0, thermostat=THERMOSTAT_NONE)


@pytest.fixture
def mixing(request):
    """Expectations mixed in FP64 for the FP64 oracle, in FP32 (the reference's rounding) otherwise."""
    M.MIX32 = request.param != 64
    yield request.param
    M.MIX32 = True


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def per_particle_rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))


@pytest.mark.parametrize("mixing,tol", [(64, 1e-10), (32, 2e-5)], indirect=["mixing"])
def test_call_site_epsilon_rules(mixing, tol):
    precision = mixing
    s = M.sites_system()
    fc_want, fs_want = M.sites_expected(s)
    o = Oracle(s, P0, precision=precision)
    fc, fs, st = o.compute_forces()
    assert st == 0
    assert per_particle_rel(fc, fc_want) < tol
    assert per_particle_rel(fs, fs_want) < tol
    # solvent <-> compound pairs are not action-reaction symmetric (arithmetic one way,
    # geometric the other): the solvent-side forces are not minus the chain-side ones
    assert abs(np.linalg.norm(fs.sum(0) + fc.sum(0))) > 1e-3 * np.linalg.norm(fs)


def test_call_site_rules_are_distinguishable():
    """Swapping the epsilon rule at any one site moves some force far past the tolerance."""
    s = M.sites_system()
    fc_want, fs_want = M.sites_expected(s)
    orig_a, orig_g = M.eps_arith, M.eps_geo
    try:
        M.eps_arith, M.eps_geo = orig_g, orig_a
        fc_swapped, fs_swapped = M.sites_expected(s)
    finally:
        M.eps_arith, M.eps_geo = orig_a, orig_g
    assert per_particle_rel(fc_swapped, fc_want) > 0.1
    assert per_particle_rel(fs_swapped, fs_want) > 0.1
    assert abs(M.eps_arith(1, 2) - M.eps_geo(1, 2)) / M.eps_geo(1, 2) > 0.6
    assert abs(M.eps_arith(0, 1) - M.eps_geo(1, 0)) / M.eps_geo(1, 0) > 0.4


@pytest.mark.parametrize("mixing,tol", [(64, 1e-10), (32, 2e-5)], indirect=["mixing"])
def test_lj_ignores_in_intra_and_inter_loops(mixing, tol):
    precision = mixing
    s = M.exclusions_system()
    want = M.exclusions_expected(s)
    o = Oracle(s, P0, precision=precision)
    fc, _, st = o.compute_forces()
    assert st == 0
    assert per_particle_rel(fc, want) < tol
    # the excluded pairs are close (0.36 and 0.37 nm): dropping either exclusion is visible
    cp = s.chain_positions.astype(float)
    t = s.chain_atom_types
    f_intra = M.lj(cp[0], cp[1], M.sigma_mix(t[0], t[1]), M.eps_geo(t[0], t[1]))
    f_inter = M.lj(cp[2], cp[3], M.sigma_mix(t[2], t[3]), M.eps_arith(t[2], t[3]))
    assert np.linalg.norm(f_intra) > 0.02 * np.linalg.norm(want[0])
    assert np.linalg.norm(f_inter) > 0.02 * np.linalg.norm(want[2])


@pytest.mark.parametrize("mixing,tol", [(64, 1e-10), (32, 2e-5)], indirect=["mixing"])
def test_bridge_force_is_stored_once_per_step(mixing, tol):
    precision = mixing
    s = M.bridge_system()
    want = M.bridge_expected(s)
    o = Oracle(s, P0, precision=precision)
    fc, _, st = o.compute_forces()
    assert st == 0
    assert per_particle_rel(fc, want) < tol
    # a second force pass on the same state gives the same forces (nothing carried over)
    fc2, _, _ = o.compute_forces()
    np.testing.assert_array_equal(fc, fc2)
    # after a step, the forces used by the next step are those of a fresh pass on the new
    # state: the carry-in holds that step's bridge terms only
    o.step(1)
    fresh = Oracle(s, P0, precision=precision)
    cx, cxm, sx, sxm = o.state()
    fresh.set_state(cx, cxm, sx, sxm)
    f_fresh, _, _ = fresh.compute_forces()
    o.step(1)
    f_used, _ = o.forces()
    np.testing.assert_array_equal(f_used, f_fresh)
    # the bond force is a large part of the bridged particles' forces, so adding the previous
    # step's carry-in (or dropping it) would be far outside the tolerance
    fa, _ = M.pair_bond(s.chain_positions[1], s.chain_positions[2], M.f32(M.BRIDGE_B0), M.f32(M.BRIDGE_KB))
    assert np.linalg.norm(fa) > 0.5 * np.linalg.norm(want[1])


def _berendsen(T, T0, dt, tau):
    lam = math.sqrt(1.0 + dt / tau * (T0 / T - 1.0))
    return min(max(lam, 0.9), 1.1)


def _first_displacement(s):
    """x1 - x0 of a force-free first step with lambda = 1, in the FP32 arithmetic of
    integratePosition (Engine.cu:849-869): x1 = (2 x0 - x_prev), disp = x1 - x0."""
    x0 = s.solvent_positions.astype(np.float32)
    xm = s.solvent_prev_positions.astype(np.float32)
    x1 = (np.float32(2) * x0 - xm).astype(np.float32)
    return x0, x1, (x1 - x0).astype(np.float32)


def _temperature(disp, masses, dt):
    d = disp.astype(float)
    return float((masses * (d * d).sum(1)).sum() / dt ** 2 / (3 * len(d) * 8.3145))


@pytest.mark.parametrize("T_init,tau_steps,clamped", [(150.0, 10, False), (600.0, 4, False), (1200.0, 1, True)])
def test_berendsen_lambda_closed_form(T_init, tau_steps, clamped):
    s = M.free_solvents(T_init)
    dt = s.dt
    tau = tau_steps * dt
    x0, x1, disp = _first_displacement(s)
    T1 = _temperature(disp, np.full(len(x0), M.f32(M.MASS_SOL)), dt)
    assert abs(T1 / T_init - 1) < 1e-3
    lam1 = _berendsen(T1, 300.0, dt, tau)
    assert (lam1 in (0.9, 1.1)) == clamped
    assert abs(lam1 - 1.0) > 0.04                      # far from 1: a thermostat that did nothing fails
    o = Oracle(s, RunParams(nlist_interval=0, thermostat=THERMOSTAT_BERENDSEN, thermostat_tau=tau), precision=32)
    o.step(1)
    _, _, sx, _ = o.state()
    np.testing.assert_array_equal(sx.astype(np.float32), x1)       # lambda starts at 1
    assert abs(o.lam - lam1) < 2e-7 * lam1
    # step 2: displacement scaled by lambda_1; T_2 = lambda_1^2 T_1 sets lambda_2
    o.step(1)
    _, _, sx2, _ = o.state()
    x2 = (x1 + (disp * np.float32(lam1)).astype(np.float32)).astype(np.float32)
    np.testing.assert_allclose(sx2, x2, rtol=0, atol=4e-7)
    T2 = _temperature((sx2 - sx).astype(np.float32), np.full(len(x0), M.f32(M.MASS_SOL)), dt)
    assert abs(T2 / (lam1 ** 2 * T1) - 1) < 2e-3
    assert abs(o.lam - _berendsen(T2, 300.0, dt, tau)) < 1e-5


@pytest.mark.parametrize("mixing", [64], indirect=True)
def test_lj_energy_known_answers(mixing):
    sig, eps = M.f32(M.SIG_SOL), M.f32(M.EPS_SOL)
    for r, known in ((2 ** (1 / 6) * sig, -eps), (sig, 0.0), (1.1, None)):
        s = M.build(6.0, [], [(2.0, 2.0, 2.0), (2.0 + r, 2.0, 2.0)])
        r32 = float(s.solvent_positions[1, 0]) - float(s.solvent_positions[0, 0])   # the stored distance
        want = M.lj_energy(r32, sig, eps)
        if known is not None:
            assert abs(want - known) < 0.02       # one FP32 position ulp away from the known answer
        e = Oracle(s, P0, precision=64).energies()
        assert abs(e["lj_solvent_solvent"] - want) < 1e-9 * eps
        assert e["potential"] == pytest.approx(e["lj_solvent_solvent"], abs=1e-12)
    # per-site sums of the sites system
    s = M.sites_system()
    e = Oracle(s, P0, precision=64).energies()
    for k, v in M.sites_energies(s).items():
        assert e[k] == pytest.approx(v, rel=1e-9, abs=1e-9), k
    # compound<-solvent and solvent<-compound differ (geometric vs arithmetic epsilon)
    assert abs(e["lj_compound_solvent"] - e["lj_solvent_compound"]) > 1e-3 * abs(e["lj_solvent_compound"])


@pytest.mark.parametrize("mixing", [64], indirect=True)
def test_lj_energy_gradient_is_minus_force(mixing):
    """On symmetric sites (solvent-solvent, intra- and inter-compound) the FP64 energy is a
    potential of the FP64 forces: central differences of the total energy give -F."""
    s = M.build(6.0, [([(2.0, 2.0, 2.0), (2.0, 2.0, 2.4)], [1, 2]), ([(2.37, 2.0, 2.1)], [2])],
                [(3.0, 2.5, 2.0), (3.0, 2.9, 2.2)])
    # the solvent<->compound sites are asymmetric, so the check is per site: the chain-chain
    # energies against chain particle 0's chain-chain force, the solvent-solvent energy against
    # solvent 0's solvent-solvent force (both hand-derived)
    o = Oracle(s, P0, precision=64)
    h = 1e-6

    def site_energy(cx, sx, keys):
        o2 = Oracle(s, P0, precision=64)
        o2.set_state(cx, cx, sx, sx)
        e = o2.energies()
        return sum(e[k] for k in keys)

    cx0, _, sx0, _ = o.state()
    # chain particle 0: intra + inter energies vs the chain-chain part of its force
    cp = s.chain_positions.astype(float)
    t = s.chain_atom_types
    f_cc = M.lj(cp[0], cp[1], M.sigma_mix(t[0], t[1]), M.eps_geo(t[0], t[1])) + \
        M.lj(cp[0], cp[2], M.sigma_mix(t[0], t[2]), M.eps_arith(t[0], t[2]))
    for ax in range(3):
        cp_, cm_ = cx0.copy(), cx0.copy()
        cp_[0, ax] += h
        cm_[0, ax] -= h
        g = (site_energy(cp_, sx0, ("lj_intra", "lj_inter")) - site_energy(cm_, sx0, ("lj_intra", "lj_inter"))) / (2 * h)
        assert -g == pytest.approx(f_cc[ax], rel=1e-5, abs=1e-3)
    # solvent 0: solvent-solvent energy vs its solvent-solvent force
    sp = s.solvent_positions.astype(float)
    f_ss = M.lj(sp[0], sp[1], M.f32(M.SIG_SOL), M.f32(M.EPS_SOL))
    for ax in range(3):
        sp_, sm_ = sx0.copy(), sx0.copy()
        sp_[0, ax] += h
        sm_[0, ax] -= h
        g = (site_energy(cx0, sp_, ("lj_solvent_solvent",)) - site_energy(cx0, sm_, ("lj_solvent_solvent",))) / (2 * h)
        assert -g == pytest.approx(f_ss[ax], rel=1e-5, abs=1e-3)

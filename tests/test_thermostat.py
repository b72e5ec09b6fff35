"""The reference's thermostat cadence (LIMA_THERMOSTAT_REFERENCE): Engine::hostMaster
(Engine.cu:49-75), handleBoxtemp (Engine.cu:362-393), getBoxTemperature
(Engine.cu:314-359) and LIMAENG::calcKineticEnergy (Engine.cuh:31-44).

The oracle is pinned with a force-free system (every epsilon zero, no bonded
terms): each particle then moves with its initial velocity v, so the two-step
kinetic energy of the reference gives T = sum m |v|^2 / (3 R N) in closed form,
and after the one rescaled step the next evaluation reads lambda^2 T.  The GPU
path is then compared with the FP32 oracle over 85 steps of the small system.
"""
import numpy as np
import pytest

from quantom_amd.abi import (ANGLE_DTYPE, DIHEDRAL_DTYPE, PAIR_DTYPE, THERMOSTAT_REFERENCE, LimaSystem, RunParams)

F32 = np.float32
GAS_R = 8.3145


def free_system(seed=5, L=4.0, n_comp=2, n_per=4, n_solv=24, T0=300.0):
    """Compounds of n_per unbonded particles plus solvents, all epsilon zero, Maxwell-like velocities."""
    rng = np.random.default_rng(seed)
    ff = np.zeros((32, 3), F32)
    ff[:] = (0.014, 0.32, 0.0)
    ff[0] = (0.018, 0.3166, 0.0)
    cpos = []
    for c in range(n_comp):
        base = rng.uniform(1.0, L - 1.0, 3)
        for i in range(n_per):
            cpos.append(base + [0.0, 0.4 * i, 0.0])
    cpos = np.array(cpos, F32)
    sx = rng.uniform(0.2, L - 0.2, (n_solv, 3)).astype(F32)
    dt = 2e-6
    vc = rng.normal(0.0, 400.0, cpos.shape)          # nm/ns
    vs = rng.normal(0.0, 400.0, sx.shape)
    cprev = (cpos - vc * dt).astype(F32)
    sprev = (sx - vs * dt).astype(F32)
    z = lambda n: np.zeros(n, np.int32)
    nch = n_comp * n_per
    return LimaSystem(
        box_len=L, dt=dt, thermostat_target=T0, forcefield=ff,
        compound_offset=np.arange(n_comp + 1, dtype=np.int32) * n_per, key_particle_index=z(n_comp),
        confining_particle_sphere=np.full(n_comp, 1.3, F32), chain_positions=cpos, chain_prev_positions=cprev,
        chain_atom_types=np.ones(nch, np.uint8), bond_offset=z(n_comp + 1), bonds=np.zeros(0, PAIR_DTYPE),
        angle_offset=z(n_comp + 1), angles=np.zeros(0, ANGLE_DTYPE), dihedral_offset=z(n_comp + 1),
        dihedrals=np.zeros(0, DIHEDRAL_DTYPE), ignore_offset=z(nch + 1), ignore_compound=z(0), ignore_local=z(0),
        bridge_particle_offset=z(1), bridge_particle_compound=z(0), bridge_particle_local=z(0),
        bridge_bond_offset=z(1), bridge_bonds=np.zeros(0, PAIR_DTYPE), bridge_angle_offset=z(1),
        bridge_angles=np.zeros(0, ANGLE_DTYPE), bridge_dihedral_offset=z(1),
        bridge_dihedrals=np.zeros(0, DIHEDRAL_DTYPE), solvent_positions=sx, solvent_prev_positions=sprev)


def closed_form_temperature(s: LimaSystem) -> float:
    """sum m |v|^2 / (3 R N) with v = (x - x_prev) / dt (uniform motion, no forces)."""
    vc = (s.chain_positions.astype(np.float64) - s.chain_prev_positions) / s.dt
    vs = (s.solvent_positions.astype(np.float64) - s.solvent_prev_positions) / s.dt
    ff = s.forcefield.astype(np.float64)
    m2 = (ff[s.chain_atom_types, 0] * (vc ** 2).sum(1)).sum() + (ff[0, 0] * (vs ** 2).sum(1)).sum()
    return m2 / (3 * GAS_R * (s.n_chain + s.n_solvents))


def test_reference_thermostat_cadence_closed_form():
    from oracle.oracle import Oracle
    base = free_system()
    t_exp = closed_form_temperature(base)
    s = free_system(T0=float(F32(1.05 * t_exp)))          # lambda = T0 / T = 1.05, inside the clamp
    o = Oracle(s, RunParams(nlist_interval=10, thermostat=THERMOSTAT_REFERENCE), precision=64)
    lams = []
    for k in range(85):
        assert o.step(1) == 0
        lams.append(o.lam)                                # the scalar the next step applies
        if k + 1 == 40:
            t40 = o.ref_temperature
    # evaluated only with 40 and 80 steps done, applied to exactly one step each
    assert all(lams[k] == 1.0 for k in range(85) if k + 1 not in (40, 80)), lams
    assert abs(t40 - t_exp) / t_exp < 1e-6, (t40, t_exp)
    lam40 = float(F32(F32(s.thermostat_target) / F32(t40)))
    assert abs(lams[39] - lam40) < 1e-7 and abs(lams[39] - 1.05) < 1e-5
    # one step with the displacement scaled by lambda, then free flight: T80 = lambda^2 T40
    t80 = o.ref_temperature
    assert abs(t80 - lam40 ** 2 * t40) / t80 < 1e-6, (t80, lam40 ** 2 * t40)
    assert abs(lams[79] - float(F32(F32(s.thermostat_target) / F32(t80)))) < 1e-7


def test_reference_thermostat_clamps_and_zero_guard():
    from oracle.oracle import Oracle
    t_exp = closed_form_temperature(free_system())
    for ratio, want in ((2.0, 1.1), (0.5, 0.9)):
        s = free_system(T0=float(t_exp * ratio))
        o = Oracle(s, RunParams(nlist_interval=10, thermostat=THERMOSTAT_REFERENCE), precision=64)
        o.step(40)
        assert abs(o.lam - want) < 1e-6
    # at rest: T = 0 is replaced by 1 K (Engine.cu:368), so lambda = clamp(T0 / 1) = 1.1
    s = free_system()
    s.chain_prev_positions = s.chain_positions.copy()
    s.solvent_prev_positions = s.solvent_positions.copy()
    o = Oracle(s, RunParams(nlist_interval=10, thermostat=THERMOSTAT_REFERENCE), precision=64)
    o.step(40)
    assert o.ref_temperature == 1.0 and abs(o.lam - 1.1) < 1e-6


def test_reference_thermostat_uses_positions_two_steps_apart(tiny_system):
    """With forces the one-step and two-step velocities differ; the evaluation with 40 steps
    done must read the positions after steps 40 and 38 (traj slots 39 and 37, Engine.cu:322-323)."""
    from oracle.oracle import Oracle
    s = tiny_system
    o = Oracle(s, RunParams(nlist_interval=10, thermostat=THERMOSTAT_REFERENCE), precision=64)
    o.step(38)
    cx38, _, sx38, _ = o.state()
    o.step(1)
    cx39, _, sx39, _ = o.state()
    o.step(1)
    cx40, _, sx40, _ = o.state()
    L = s.box_len
    ff = s.forcefield.astype(np.float64)
    m = np.concatenate([ff[s.chain_atom_types, 0], np.full(s.n_solvents, ff[0, 0])])

    def temp(a, b, span):
        d = a - b
        d -= L * np.round(d / L)
        v2 = (d ** 2).sum(1) / (span * s.dt) ** 2
        return float((m * v2).sum() / (3 * GAS_R * len(m)))

    x38, x39, x40 = (np.concatenate(p) for p in ((cx38, sx38), (cx39, sx39), (cx40, sx40)))
    t_two, t_one = temp(x40, x38, 2), temp(x40, x39, 1)
    assert abs(t_two - t_one) / t_two > 1e-4                # the two readings are distinguishable here
    assert abs(o.ref_temperature - t_two) / t_two < 1e-6, (o.ref_temperature, t_two, t_one)


@pytest.mark.gpu
def test_reference_thermostat_matches_oracle(small_system):
    from oracle.oracle import Oracle
    from quantom_amd.engine import Engine
    print("case: 85 steps, reference thermostat cadence, small system", flush=True)
    s = small_system.copy()
    s.thermostat_target = 310.0                            # the reference's target (Engine.cu:363)
    p = RunParams(nlist_interval=10, thermostat=THERMOSTAT_REFERENCE)
    eng = Engine([s], p)
    recs = eng.step_observed(85, 1)
    lam_gpu = [r[0]["thermostat_scalar"] for r in recs]    # the scalar applied on each step
    o = Oracle(s, p, precision=32)
    lam_ora = [1.0]
    for _ in range(84):
        o.step(1)
        lam_ora.append(o.lam)
    o.step(1)
    assert all(lam_gpu[k] == 1.0 for k in range(85) if k not in (40, 80)), lam_gpu
    for k in (40, 80):
        assert lam_gpu[k] != 1.0
        assert abs(lam_gpu[k] - lam_ora[k]) <= 1e-3 * lam_ora[k], (k, lam_gpu[k], lam_ora[k])
    cx, _, sx, _ = eng.get_state()
    ocx, _, osx, _ = o.state()
    d = np.concatenate([np.abs(cx - ocx).ravel(), np.abs(sx - osx).ravel()])
    d = np.minimum(d, s.box_len - d)
    print("max |dx|", d.max(), "lambda", lam_gpu[40], lam_ora[40], lam_gpu[80], lam_ora[80])
    assert d.max() <= 1e-3
    eng.close()


def test_boxbuilder_initial_velocities(small_system):
    """BoxBuilder's start (BoxBuilder.cu:248-276): solvents at v_rms = sqrt(3 R T / M) with
    T = 313 K, R = 8.3144 and the solvent mass; compounds at rest.  The box temperature the
    engine's one-step kinetic energy then reads is N_solvent / N of that, in closed form."""
    import math

    from oracle.oracle import Oracle
    from quantom_amd import systems
    s = systems.reference_initial_velocities(small_system)
    v_rms = math.sqrt(3 * 8.3144 * 313.0 / (float(np.float32(18.01528)) * 1e-3))
    v = np.linalg.norm((s.solvent_positions.astype(np.float64) - s.solvent_prev_positions) / s.dt, axis=1)
    assert np.allclose(v, v_rms, rtol=1e-3)                 # FP32 positions: |dx| = 1.3e-3 nm, ulp 5e-7
    assert np.array_equal(s.chain_prev_positions, s.chain_positions)
    assert np.array_equal(s.solvent_positions, small_system.solvent_positions)
    m_sol = float(s.forcefield[0, 0])
    t_expect = s.n_solvents * m_sol * v_rms ** 2 / (3 * GAS_R * (s.n_chain + s.n_solvents))
    o = Oracle(s, RunParams(nlist_interval=10, thermostat=THERMOSTAT_REFERENCE), precision=64)
    t = o.energies()["temperature"]
    assert abs(t - t_expect) / t_expect < 2e-3, (t, t_expect)

"""Per-step energy series (lima_step_observed) and the reference's energy.bin /
temperature.bin layout (Environment.cpp:146-148, 256-266; Analyzer.cu:137-207)."""
import numpy as np
import pytest

from quantom_amd.abi import RunParams


def _record(pot_b, pot_lj, kin, temp):
    return {"bond": pot_b, "angle": 0.0, "dihedral": 0.0, "potential": pot_b + pot_lj, "kinetic": kin,
            "temperature": temp}


class _Sys:
    n_chain, n_solvents = 30, 70


class _Eng:
    class simulation:
        systems = [_Sys()]


def test_reference_dump_layout(tmp_path):
    from quantom_amd.observe import write_reference_dumps
    recs = [[_record(100.0, -50.0, 20.0, 300.0)], [_record(110.0, -40.0, 30.0, 310.0)]]
    paths = write_reference_dumps(recs, _Eng(), str(tmp_path))
    e = np.fromfile(paths["energy"], dtype="<f4")
    t = np.fromfile(paths["temperature"], dtype="<f4")
    assert e.size == 2 * 3 and t.size == 2          # one Float3 / one float per record
    np.testing.assert_allclose(e.reshape(2, 3), [[1.0, 0.2, 1.2], [1.1, 0.3, 1.4]], rtol=1e-6)
    np.testing.assert_array_equal(t, np.float32([300.0, 310.0]))
    paths = write_reference_dumps(recs, _Eng(), str(tmp_path), lj=True)
    e = np.fromfile(paths["energy"], dtype="<f4").reshape(2, 3)
    np.testing.assert_allclose(e[:, 0], [0.5, 0.7], rtol=1e-6)


@pytest.mark.gpu
def test_observed_series_matches_energies_and_plain_steps(small_system):
    from quantom_amd.engine import Engine
    p = RunParams(nlist_interval=10)
    a = Engine([small_system], p)
    recs = a.step_observed(20, 5)
    assert len(recs) == 4 and len(recs[0]) == 1
    b = Engine([small_system], p)
    for k in range(4):
        e = b.energies()[0]
        r = recs[k][0]
        # same positions up to the rounding of the energy kernels' force sums
        for key in ("bond", "angle", "dihedral", "lj_intra", "lj_inter", "lj_compound_solvent",
                    "lj_solvent_compound", "lj_solvent_solvent"):
            assert abs(r[key] - e[key]) <= 1e-5 * max(abs(e[key]), 1.0), (k, key, r[key], e[key])
        assert 150.0 < r["temperature"] < 600.0
        b.step(5)
    for x, y in zip(a.get_state(), b.get_state()):
        d = np.abs(x - y)
        d = np.minimum(d, small_system.box_len - d)
        assert d.max() < 1e-5
    a.close()
    b.close()


@pytest.mark.gpu
def test_observed_every_step_and_zero(small_system):
    from quantom_amd.engine import Engine
    e = Engine([small_system], RunParams(nlist_interval=10))
    assert e.step_observed(0, 3) == []
    recs = e.step_observed(7, 1)
    assert len(recs) == 7
    lam = [r[0]["thermostat_scalar"] for r in recs]
    assert all(0.9 <= v <= 1.1 for v in lam)
    recs = e.step_observed(7, 3)                     # records at steps 0, 3, 6
    assert len(recs) == 3
    e.close()


def test_reference_log_layout(tmp_path):
    from quantom_amd.observe import write_reference_logs
    log = np.arange(3 * 10, dtype=np.float32).reshape(3, 10)
    tr = np.arange(3 * 2 * 4 * 9, dtype=np.float32).reshape(3, 2, 4, 3, 3)
    paths = write_reference_logs(log, tr, str(tmp_path))
    assert np.array_equal(np.fromfile(paths["logdata"], "<f4"), log.ravel())          # 10 floats per step
    t = np.fromfile(paths["traindata"], "<f4").reshape(3, 2, 4, 3, 3)                # [step][compound][slot][3][xyz]
    assert np.array_equal(t, tr)
    paths = write_reference_logs(log, tr, str(tmp_path / "crit"), critical_error=True)
    assert "traindata" not in paths                                                   # Environment.cpp:137


@pytest.mark.gpu
def test_step_logged_matches_state_forces_and_oracle(small_system):
    """The logs of 30 steps against the same steps run one at a time on a second engine
    (positions and forces bit-equal), and against the FP32 oracle on the initial state
    (compound <- solvent LJ force, Engine.cu:971; the first dihedral's energy, Engine.cu:612)."""
    import ctypes as C

    from oracle.oracle import Oracle, lib
    from quantom_amd.engine import Engine
    print("case: step_logged, small system, 30 steps, 128 slots", flush=True)
    p = RunParams(nlist_interval=10)
    s = small_system
    a = Engine([s], p)
    log, tr = a.step_logged(30, 0, 128)
    b = Engine([s], p)
    states, forces = [b.get_state()], []
    for _ in range(30):
        b.step(1)
        states.append(b.get_state())
        forces.append(b.forces()[0])
    off = s.compound_offset
    n0 = off[1] - off[0]
    for k in range(30):
        cx, cxm = states[k + 1][0], states[k + 1][1]
        for c in range(s.n_compounds):
            n = off[c + 1] - off[c]
            assert np.array_equal(tr[k, c, :n, 0], cx[off[c]:off[c + 1]]), (k, c)
            assert np.all(tr[k, c, n:, 0] == -1.0)
            assert np.all(tr[k, c, :, 2] == 0.0)
        # particle 0 of compound 0: speed, one-step kinetic energy, |F|
        d = cx[0].astype(np.float64) - cxm[0]
        d -= s.box_len * np.round(d / s.box_len)
        v = np.linalg.norm(d) / s.dt
        assert abs(log[k, 0] - v) <= 1e-5 * v
        m = float(s.forcefield[s.chain_atom_types[0], 0])
        assert abs(log[k, 1] - 0.5 * m * (0.5 * v) ** 2) <= 1e-4 * log[k, 1]
        f = np.linalg.norm(forces[k][0].astype(np.float64))
        assert abs(log[k, 3] - f) <= 1e-6 * f
        assert np.all(log[k, 4:] == 0.0)
    # the first step against the oracle on the same initial state and lists
    o = Oracle(s, p, precision=32)
    o.compute_forces()
    want = o.cs_forces()
    got = tr[0, :, :, 1].reshape(-1, 3)
    rows = np.concatenate([np.arange(off[c], off[c + 1]) - off[c] + c * 128 for c in range(s.n_compounds)])
    got = got[rows]
    assert np.sqrt(((got - want) ** 2).sum() / (want ** 2).sum()) <= 1e-5
    # thread 0's potE slot: the energy of compound 0's first dihedral at the initial positions
    x = states[0][0][:n0].astype(np.float32)
    dh = s.dihedrals[s.dihedral_offset[0]]
    idx = [int(i) for i in dh["idx"]]
    fp = lambda v: np.ascontiguousarray(v, np.float32).ctypes.data_as(C.POINTER(C.c_float))
    out = np.zeros(12, np.float32)
    pe = C.c_double(0.0)
    lib().oracle_dihedral_f32(fp(x[idx[0]]), fp(x[idx[1]]), fp(x[idx[2]]), fp(x[idx[3]]), C.c_float(dh["phi_0"]),
                              C.c_float(dh["k_phi"]), int(dh["n"]), out.ctypes.data_as(C.POINTER(C.c_float)),
                              C.byref(pe))
    assert abs(log[0, 2] - pe.value) <= 1e-5 * abs(pe.value) + 1e-3, (log[0, 2], pe.value)   # cosf ulps x k_phi
    a.close()
    b.close()

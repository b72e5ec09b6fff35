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

"""lima_step_observed_noalloc (the energy series staged in the library's fixed device array,
no device allocation) against lima_step_observed on twin engines: two calls in a row, the
second smaller, then a smaller engine on the same device (the array keeps the last call's
contents), and a request past the array's size refused before any step runs."""
import numpy as np
import pytest

from quantom_amd.abi import RunParams

FIELDS = ("bond", "angle", "dihedral", "lj_intra", "lj_inter", "lj_compound_solvent", "lj_solvent_compound",
          "lj_solvent_solvent", "potential", "kinetic", "temperature", "virial")


def _same_records(got, want):
    assert len(got) == len(want)
    for k, (rg, rw) in enumerate(zip(got, want)):
        assert len(rg) == len(rw)
        for r, (g, w) in enumerate(zip(rg, rw)):
            # the FP64 energy sums are block partials added with device atomics, in no fixed order
            for key in FIELDS:
                assert abs(g[key] - w[key]) <= 1e-11 * max(abs(w[key]), 1.0), (k, r, key, g[key], w[key])
            assert g["thermostat_scalar"] == w["thermostat_scalar"], (k, r)


def _same_state(a, b):
    for x, y in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
def test_noalloc_series_equals_the_allocating_entry_point(small_system, tiny_system):
    from quantom_amd.engine import Engine
    print("case: noalloc observed series, batch of 2 (small + tiny), 23 steps every 3 then 7 every 4", flush=True)
    p = RunParams(nlist_interval=10)
    a = Engine([small_system, tiny_system], p)
    b = Engine([small_system, tiny_system], p)
    _same_records(a.step_observed(23, 3, noalloc=True), b.step_observed(23, 3))
    _same_state(a, b)
    _same_records(a.step_observed(7, 4, noalloc=True), b.step_observed(7, 4))
    _same_state(a, b)
    a.close()
    b.close()

    print("case: noalloc observed series, tiny system alone, 5 steps every 5", flush=True)
    c = Engine([tiny_system], p)
    d = Engine([tiny_system], p)
    _same_records(c.step_observed(5, 5, noalloc=True), d.step_observed(5, 5))
    _same_state(c, d)
    c.close()
    d.close()


@pytest.mark.gpu
def test_noalloc_refuses_a_series_past_its_staging(tiny_system):
    from quantom_amd.abi import LIMA_E_ARG
    from quantom_amd.engine import Engine, LimaStatusError
    print("case: noalloc observed series past the 4 MiB staging is refused", flush=True)
    e = Engine([tiny_system, tiny_system], RunParams(nlist_interval=10))
    before = e.get_state()
    with pytest.raises(LimaStatusError) as ex:
        e.step_observed(30000, 1, noalloc=True)        # 30,000 records x 2 systems x 84 bytes > 4 MiB
    assert ex.value.code == LIMA_E_ARG and "no-alloc staging" in str(ex.value)
    for x, y in zip(before, e.get_state()):
        np.testing.assert_array_equal(x, y)
    assert len(e.step_observed(3, 1, noalloc=True)) == 3
    e.close()

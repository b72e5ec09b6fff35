"""GPU parity: the HIP path through the C ABI against the CPU oracle on identical inputs.

Bars (stated per test):
  * neighbour lists: set-equal, bit-exact (same FP32 distance arithmetic);
  * forces: relative rms error (||F_gpu - F_oracle|| / ||F_oracle||) <= 1e-5
    against both the FP32 and the FP64 oracle.  The FP32 oracle is itself
    ~2e-6 from FP64 on the small system, so 1e-5 is five times the FP32
    rounding scale (summation order and a 1-ulp reciprocal differ);
  * positions after one step: |dx| <= 2e-6 nm;
  * energies: relative 1e-6 per term.
"""
import numpy as np
import pytest

from tests.conftest import as_sets, rel_rms

pytestmark = pytest.mark.gpu


def _engine(system, **kw):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    return Engine([system], RunParams(**kw))


def _oracle(system, precision=32, **kw):
    from oracle.oracle import Oracle
    from quantom_amd.abi import RunParams
    return Oracle(system, RunParams(**kw), precision=precision)


def _copy_lists_to_gpu(eng, ora):
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SS
    eng.set_neighbors(cc=ora.neighbors(KIND_CC), cs=ora.neighbors(KIND_CS), ss=ora.neighbors(KIND_SS))


def test_library_loads_and_sees_a_device():
    from quantom_amd import engine
    print(engine.version(), flush=True)
    assert engine.device_count() >= 1


def test_initial_neighbor_lists_bit_exact(small_system):
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SS
    print("case: initial nlist, small system", flush=True)
    eng = _engine(small_system)
    ora = _oracle(small_system)
    cc, cs, ss = eng.neighbors()
    assert as_sets(cc) == as_sets(ora.neighbors(KIND_CC))
    assert as_sets(cs) == as_sets(ora.neighbors(KIND_CS))
    assert as_sets(ss) == as_sets(ora.neighbors(KIND_SS))
    assert sum(len(x) for x in ss) > 0


def test_rebuild_after_motion_bit_exact(small_system):
    """Lists rebuilt on a moved state, starting from the same previous lists (hysteresis)."""
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SS
    print("case: rebuild after 9 oracle steps", flush=True)
    eng = _engine(small_system, nlist_interval=0)
    ora = _oracle(small_system, nlist_interval=0)
    ora.step(9)
    cx, cxm, sx, sxm = ora.state()
    eng.set_state(cx, cxm, sx, sxm)
    ora.rebuild_neighbors()
    eng.rebuild_neighbors()
    cc, cs, ss = eng.neighbors()
    assert as_sets(cc) == as_sets(ora.neighbors(KIND_CC))
    assert as_sets(cs) == as_sets(ora.neighbors(KIND_CS))
    assert as_sets(ss) == as_sets(ora.neighbors(KIND_SS))


@pytest.mark.parametrize("accum", [0, 1])
def test_forces_match_oracle(small_system, accum):
    print(f"case: forces, accumulation={accum}", flush=True)
    eng = _engine(small_system, accumulation=accum)
    ora = _oracle(small_system)
    _copy_lists_to_gpu(eng, ora)
    cf, sf = eng.compute_forces()
    ocf, osf, st = ora.compute_forces()
    assert st == 0
    assert np.all(np.isfinite(cf)) and np.all(np.isfinite(sf))
    e_c, e_s = rel_rms(cf, ocf), rel_rms(sf, osf)
    # the same pair sets in FP64
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SC, KIND_SS
    o64 = _oracle(small_system, precision=64)
    for k in (KIND_CC, KIND_CS, KIND_SC, KIND_SS):
        o64.set_neighbors(k, ora.neighbors(k))
    dcf, dsf, _ = o64.compute_forces()
    d_c, d_s = rel_rms(cf, dcf), rel_rms(sf, dsf)
    print(f"rel rms vs FP32 oracle: chain {e_c:.3e} solvent {e_s:.3e}; vs FP64: chain {d_c:.3e} solvent {d_s:.3e}")
    assert e_c <= 1e-5 and e_s <= 1e-5
    assert d_c <= 1e-5 and d_s <= 1e-5


def test_one_step_positions(small_system):
    print("case: one step positions", flush=True)
    eng = _engine(small_system, nlist_interval=0)
    ora = _oracle(small_system, nlist_interval=0)
    _copy_lists_to_gpu(eng, ora)
    eng.step(1)
    ora.step(1)
    cx, cxm, sx, sxm = eng.get_state()
    ocx, ocxm, osx, osxm = ora.state()
    assert np.max(np.abs(cx - ocx)) <= 2e-6
    assert np.max(np.abs(sx - osx)) <= 2e-6
    assert np.array_equal(cxm, ocxm.astype(np.float32))
    assert np.array_equal(sxm, osxm.astype(np.float32))


def test_twenty_steps_with_rebuilds_and_thermostat(small_system):
    print("case: 20 steps, rebuild every 10, Berendsen", flush=True)
    eng = _engine(small_system, nlist_interval=10)
    ora = _oracle(small_system, nlist_interval=10)
    eng.step(20)
    ora.step(20)
    cx, _, sx, _ = eng.get_state()
    ocx, _, osx, _ = ora.state()
    d = np.concatenate([np.abs(cx - ocx).ravel(), np.abs(sx - osx).ravel()])
    d = np.minimum(d, small_system.box_len - d)   # positions compared modulo the box
    print("max |dx|", d.max())
    assert d.max() <= 1e-4
    # the thermostat scalar applied on the last step (FP64 KE reduction on both sides)
    eg = eng.energies()[0]
    o_prev = ora.lam  # lambda for the next step; the applied one is compared through temperatures
    assert abs(eg["temperature"] - ora.energies()["temperature"]) / ora.energies()["temperature"] < 1e-4
    assert 0.9 <= eg["thermostat_scalar"] <= 1.1 and 0.9 <= o_prev <= 1.1


def test_energies_match_oracle(small_system):
    print("case: energies", flush=True)
    eng = _engine(small_system)
    ora = _oracle(small_system)
    _copy_lists_to_gpu(eng, ora)
    eg = eng.energies()[0]
    eo = ora.energies()
    for k in ("bond", "angle", "dihedral", "lj_intra", "lj_inter", "lj_compound_solvent", "lj_solvent_compound",
              "lj_solvent_solvent", "kinetic", "virial"):
        scale = max(abs(eo[k]), 1e-3 * abs(eo["potential"]), 1e-9)
        assert abs(eg[k] - eo[k]) / scale <= 1e-5, (k, eg[k], eo[k])


def test_batched_replicas_equal_individual_runs(small_system, tiny_system):
    """Two replicas stepped in one engine give the same states as each stepped alone."""
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: batch of 2 vs single", flush=True)
    b = Engine([small_system, tiny_system], RunParams())
    b.step(15)
    cx, cxm, sx, sxm = b.get_state()
    sim = b.simulation
    for k, sysk in enumerate((small_system, tiny_system)):
        e = Engine([sysk], RunParams())
        e.step(15)
        x, xm, s, sm = e.get_state()
        c0, c1 = sim.chain_begin[k], sim.chain_begin[k + 1]
        s0, s1 = sim.solvent_begin[k], sim.solvent_begin[k + 1]
        assert np.array_equal(x, cx[c0:c1])
        assert np.array_equal(s, sx[s0:s1])


def test_fixed_point_is_deterministic(small_system):
    print("case: fixed-point determinism", flush=True)
    runs = []
    for _ in range(2):
        e = _engine(small_system, accumulation=1)
        e.step(12)
        runs.append(e.get_state())
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_autotune_leaves_state_and_fixed_point_trajectory_unchanged(small_system, monkeypatch, tmp_path):
    """lima_autotune times whole steps from the current state and restores it: positions and
    neighbour lists are unchanged, and a fixed-point trajectory after tuning is bit-equal to
    one stepped with the default launch shape (int64 sums do not depend on the shape)."""
    monkeypatch.setenv("LIMA_AUTOTUNE_CACHE", str(tmp_path / "tune.txt"))
    print("case: autotune restores the state", flush=True)
    a = _engine(small_system, accumulation=1, nlist_interval=10)
    a.step(7)                                       # mid-cycle, off the graph alignment
    before = a.get_state()
    lists_before = [as_sets(l) for l in a.neighbors()]
    cfg = a.autotune(10)
    for x, y in zip(before, a.get_state()):
        assert np.array_equal(x, y)
    assert [as_sets(l) for l in a.neighbors()] == lists_before
    assert cfg["solvents_per_thread"] in (1, 2, 4) and cfg["chunk_particles"] in (256, 512, 1024, 2048)
    b = _engine(small_system, accumulation=1, nlist_interval=10)
    b.step(7)
    a.step(25)
    b.step(25)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)


def test_kernel_clock_times_the_all_pairs_kernel_inside_steps(small_system):
    """lima_set_kernel_clock / lima_get_kernel_clock: one measured launch per list period of 10
    steps, a span close to the kernel launched alone, and the trajectory unchanged by the timing."""
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: in-step kernel clock, small system", flush=True)
    a = Engine([small_system], RunParams(nlist_interval=10))
    b = Engine([small_system], RunParams(nlist_interval=10))
    a.set_kernel_clock(True)
    a.step(30)
    b.step(30)
    total, n = a.kernel_clock()
    us = total / max(n, 1)
    alone = a.time_kernels(10)["solvent_chain"]
    print(f"in-step {us:.2f} us over {n} launches, alone {alone:.2f} us", flush=True)
    assert n == 3 and us > 0
    assert 0.3 * alone < us < 10 * alone
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    a.set_kernel_clock(False)
    a.step(5)
    a.close()
    b.close()


def test_solvent_lists_at_the_bound_and_across_faces():
    """The solvent-solvent list build drops far candidates with an approximate minimum-image
    distance before the exact predicate; pairs placed within a few ulps of the bound (C + C,
    Engine.cu:198, 264), inside the box and across each periodic face, must come out exactly as
    the oracle lists them."""
    from oracle.oracle import KIND_SS, Oracle
    from quantom_amd import systems
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: solvent pairs at the list bound, 5 nm box", flush=True)
    s = systems.build_system(0x51A7 + 91, 1, 2, 5.0, 4, 300, compound_size=16)
    rng = np.random.default_rng(11)
    pts, pairs = [], []
    for k in range(-6, 7):                       # distance 1.2 nm x (1 + k 1.2e-7)
        for wrap in (False, True):
            for axis in range(3):
                u = rng.normal(size=3)
                u[axis] = abs(u[axis]) + 2.0
                u /= np.linalg.norm(u)
                a = rng.uniform(1.5, 3.5, size=3)
                if wrap:
                    a[axis] = rng.uniform(0.01, 0.2)   # the partner lies across the face
                b = np.mod(a - 1.2 * (1.0 + k * 1.2e-7) * u, s.box_len)
                pairs.append((len(pts), len(pts) + 1, k))
                pts += [a, b]
    base = len(s.solvent_positions)
    s.solvent_positions = np.concatenate([s.solvent_positions, np.array(pts, np.float32)]).astype(np.float32)
    s.solvent_prev_positions = s.solvent_positions.copy()
    p = RunParams()
    want = Oracle(s, p, precision=32).neighbors(KIND_SS)
    member = [(base + j) in set(want[base + i]) for i, j, _ in pairs]
    assert any(member) and not all(member)       # the placed pairs straddle the bound
    eng = Engine([s], p)
    got = eng.neighbors()[2]
    diff = [(i, sorted(set(g) ^ set(w))) for i, (g, w) in enumerate(zip(got, want)) if set(g) != set(w)]
    for i, d in diff[:8]:
        print(f"solvent {i}: differs in {d}", flush=True)
    assert not diff
    eng.close()


@pytest.mark.parametrize("box,n_solvents", [(4.7, 3700), (3.5, 3700)])
def test_solvent_lists_beyond_one_lds_stage(box, n_solvents):
    """Home cells whose 27-cell neighbourhood holds more candidates than one LDS stage (3,456) take
    the candidates in chunks, each home solvent every chunk in turn: lists as the oracle's.  4.7 nm:
    3 cells per side (candidates staged shifted into the home cell's image); 3.5 nm: 2 cells per
    side (minimum-image prefilter).  Either way every cell's neighbourhood is the whole box."""
    from oracle.oracle import KIND_SS, Oracle
    from quantom_amd import systems
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print(f"case: solvent lists, {n_solvents} solvents in a {box} nm box", flush=True)
    s = systems.build_system(0x51A7 + 93, 1, 1, box, 4, n_solvents, compound_size=16, auto_density=True)
    assert s.n_solvents == n_solvents
    p = RunParams(max_solvent_neighbors=2048)
    want = Oracle(s, p, precision=32).neighbors(KIND_SS)
    eng = Engine([s], p)
    got = eng.neighbors()[2]
    diff = [(i, sorted(set(g) ^ set(w))) for i, (g, w) in enumerate(zip(got, want)) if set(g) != set(w)]
    for i, d in diff[:8]:
        print(f"solvent {i}: differs in {d}", flush=True)
    assert not diff
    assert all(len(g) == len(set(g)) for g in got)
    print(f"mean list length {np.mean([len(g) for g in got]):.1f}", flush=True)
    eng.close()


def test_compound_solvent_lists_at_the_bound():
    """Solvents placed within a few ulps of a compound's list radius: C + r + C from its key
    particle, r the confining sphere (Engine.cu:146-151, 188-200), in random directions and
    wrapped into the box.  The compound->solvent lists must equal the oracle's."""
    from oracle.oracle import KIND_CS, KIND_SS, Oracle
    from quantom_amd import systems
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: solvents at the compound list radius, 5 nm box", flush=True)
    s = systems.build_system(0x51A7 + 91, 1, 2, 5.0, 4, 300, compound_size=16)
    p = RunParams()
    rng = np.random.default_rng(5)
    pts, tags = [], []
    for c in range(len(s.key_particle_index)):
        key = s.chain_positions[s.compound_offset[c] + s.key_particle_index[c]].astype(np.float64)
        radius = p.cutoff + float(s.confining_particle_sphere[c]) + p.cutoff
        for k in range(-6, 7):                   # radius x (1 + k 1.2e-7)
            for _ in range(4):
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                pts.append(np.mod(key + radius * (1.0 + k * 1.2e-7) * u, s.box_len))
                tags.append(c)
    base = len(s.solvent_positions)
    s.solvent_positions = np.concatenate([s.solvent_positions, np.array(pts, np.float32)]).astype(np.float32)
    s.solvent_prev_positions = s.solvent_positions.copy()
    ora = Oracle(s, p, precision=32)
    want_cs, want_ss = ora.neighbors(KIND_CS), ora.neighbors(KIND_SS)
    member = [(base + i) in set(want_cs[c]) for i, c in enumerate(tags)]
    assert any(member) and not all(member)       # the placed solvents straddle the radius
    eng = Engine([s], p)
    _, got_cs, got_ss = eng.neighbors()
    for name, got, want in (("compound->solvent", got_cs, want_cs), ("solvent->solvent", got_ss, want_ss)):
        diff = [(i, sorted(set(g) ^ set(w))) for i, (g, w) in enumerate(zip(got, want)) if set(g) != set(w)]
        for i, d in diff[:8]:
            print(f"{name} row {i}: differs in {d}", flush=True)
        assert not diff, name
    eng.close()


def test_compound_compound_lists_at_the_bound():
    """Two compounds whose key particles sit within a few ulps of the list radius that an empty
    list adds at, C + ((r_lo + r_lo) + C) (Engine.cu:146-151, 170-186), one offset per system
    of a 13-system batch.  Each system's compound lists must equal the oracle's."""
    import copy
    from oracle.oracle import KIND_CC, Oracle
    from quantom_amd import systems
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: compound pairs at the list radius, 13 systems of 5 nm", flush=True)
    p = RunParams()
    C = np.float32(p.cutoff)
    base = systems.build_system(0x51A7 + 91, 1, 2, 5.0, 4, 300, compound_size=16)
    r0 = np.float32(base.confining_particle_sphere[0])
    radius = float(C + ((r0 + r0) + C))
    o1, o2 = int(base.compound_offset[1]), int(base.compound_offset[2])
    key0 = base.chain_positions[base.compound_offset[0] + base.key_particle_index[0]].astype(np.float64)
    key1 = base.chain_positions[o1 + base.key_particle_index[1]].astype(np.float64)
    u = np.ones(3) / np.sqrt(3.0)                # every component under half the box
    syss, want = [], []
    for k in range(-6, 7):                       # radius x (1 + k 1.2e-7)
        s = copy.deepcopy(base)
        shift = key0 + radius * (1.0 + k * 1.2e-7) * u - key1
        s.chain_positions[o1:o2] = (s.chain_positions[o1:o2] + shift).astype(np.float32)
        s.chain_prev_positions[o1:o2] = (s.chain_prev_positions[o1:o2] + shift).astype(np.float32)
        syss.append(s)
        want.append(Oracle(s, p, precision=32).neighbors(KIND_CC))
    linked = [1 in set(w[0]) for w in want]
    assert any(linked) and not all(linked)       # the placed pairs straddle the radius
    eng = Engine(syss, p)
    got = eng.neighbors()[0]
    for r, w in enumerate(want):                 # system r holds compounds 2r and 2r + 1
        for c in range(2):
            assert set(int(x) - 2 * r for x in got[2 * r + c]) == set(w[c]), (r, c)
    eng.close()


def test_four_lanes_per_solvent_listed_form(small_system, monkeypatch, tmp_path):
    """The FP32 listed solvent kernel with four lanes per solvent (chosen by the autotuner for
    small systems): forces within the oracle bar, and after 20 steps the same images and a
    trajectory as close to the oracle as the one-lane form's."""
    from oracle.oracle import Oracle
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: four lanes per solvent, small system", flush=True)
    s = small_system
    key = f"acc0_sol1_{(s.n_chain + 64) // 128}x{(s.n_solvents + 256) // 512}:1"
    path = tmp_path / "tune.txt"
    path.write_text(f"{key} 2 1024 4096 4\n")
    monkeypatch.setenv("LIMA_AUTOTUNE_CACHE", str(path))
    p = RunParams(nlist_interval=10)
    four = Engine([s], p)
    assert four.autotune(10)["listed_lanes"] == 4
    one = Engine([s], p)
    ora = Oracle(s, p, precision=32)
    f4 = np.concatenate([x.ravel() for x in four.compute_forces()])
    f1 = np.concatenate([x.ravel() for x in one.compute_forces()])
    fo = np.concatenate([x.ravel() for x in ora.compute_forces()[:2]])
    print(f"force rel rms: four lanes {rel_rms(f4, fo):.2e}, one lane {rel_rms(f1, fo):.2e}", flush=True)
    assert rel_rms(f4, fo) <= 1e-5
    four.step(20)
    one.step(20)
    assert ora.step(20) == 0
    g4, g1, go = four.get_state(), one.get_state(), ora.state()
    def max_dx(x, y):   # periodic: a coordinate wrapped on one side only differs by about L
        d = np.abs(np.concatenate([x[0].ravel() - y[0].ravel(), x[2].ravel() - y[2].ravel()]))
        return float(np.minimum(d, s.box_len - d).max())
    d4, d1 = max_dx(g4, go), max_dx(g1, go)
    print(f"after 20 steps max |dx| vs oracle: four lanes {d4:.2e}, one lane {d1:.2e}", flush=True)
    assert d4 <= 3 * d1 + 1e-5
    four.close()
    one.close()


def tune_key(systems, accumulation):
    """The autotune cache key of a batch: the histogram of (chain / 128, solvents / 512) shapes."""
    from collections import Counter
    hist = Counter(((s.n_chain + 64) // 128, (s.n_solvents + 256) // 512) for s in systems)
    return f"acc{accumulation}_sol1" + "".join(f"_{c}x{v}:{n}" for (c, v), n in sorted(hist.items()))


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("mode", ["berendsen", "reference"])
def test_thermostat_in_the_integrators_bit_equal(small_system, tiny_system, mode, form, monkeypatch, tmp_path):
    """The thermostat run by each system's last integrator block (an autotuner choice: both
    integrators count, or the compound one alone, run after the solvent one) gives the same
    temperatures, scalars and trajectories, bit for bit, as its own kernel after the
    integrators: two systems, 100 steps (the reference cadence rescales at steps 40 and 80)."""
    from quantom_amd.abi import THERMOSTAT_BERENDSEN, THERMOSTAT_REFERENCE, RunParams
    from quantom_amd.engine import Engine
    print(f"case: thermostat in the integrators, form {form}, {mode}", flush=True)
    syss = [small_system, tiny_system]
    p = RunParams(nlist_interval=10,
                  thermostat=THERMOSTAT_BERENDSEN if mode == "berendsen" else THERMOSTAT_REFERENCE)
    engines = []
    for fused in (form, 0):   # the same launch shape otherwise
        path = tmp_path / f"tune{fused}.txt"
        path.write_text(f"{tune_key(syss, 0)} 2 1024 4096 1 1 {fused}\n")
        monkeypatch.setenv("LIMA_AUTOTUNE_CACHE", str(path))
        e = Engine(syss, p)
        assert e.autotune(10)["fused_thermostat"] == fused
        engines.append(e)
    fused, own = engines
    for _ in range(5):
        fused.step(20)
        own.step(20)
        for x, y in zip(fused.get_state(), own.get_state()):
            assert np.array_equal(x, y)
        for a, b in zip(fused.energies(), own.energies()):
            # the energy record's kinetic term is an FP64 sum of atomic adds (order varies)
            assert a["temperature"] == pytest.approx(b["temperature"], rel=1e-12)
            assert a["thermostat_scalar"] == b["thermostat_scalar"]
    fused.close()
    own.close()


def test_listed_compound_forms(tiny_system, monkeypatch, tmp_path):
    """Compounds of at most 64 particles: the packed listed kernel (the default), the plain one and,
    in FP32, the plain one with each item's sources split over both wavefronts (autotuner choices).
    Every form gives forces within the oracle bar; packed and plain give bit-equal sums (one
    sequential sum per particle and item); int64 accumulation never splits, so a requested split
    form steps bit-equal to the packed default."""
    from oracle.oracle import Oracle
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: listed compound forms, tiny system", flush=True)
    s = tiny_system
    fo = np.concatenate([x.ravel() for x in Oracle(s, RunParams(nlist_interval=10), precision=32).compute_forces()[:2]])
    forces = {}
    for acc in (0, 1):
        key = f"acc{acc}_sol1_{(s.n_chain + 64) // 128}x{(s.n_solvents + 256) // 512}:1"
        for form in (1, 0, 2):
            # the library reads a cache file once per path: one file per form
            path = tmp_path / f"tune_{acc}_{form}.txt"
            path.write_text(f"{key} 2 1024 4096 1 {form}\n")
            monkeypatch.setenv("LIMA_AUTOTUNE_CACHE", str(path))
            e = Engine([s], RunParams(nlist_interval=10, accumulation=acc))
            cfg = e.autotune(10)
            assert cfg["listed_packed"] == (form if not (acc == 1 and form == 2) else 0), (acc, form, cfg)
            f = np.concatenate([x.ravel() for x in e.compute_forces()])
            e.step(15)
            forces[(acc, form)] = (f, e.get_state())
            e.close()
            print(f"accumulation {acc} form {form}: rel rms vs oracle {rel_rms(f, fo):.2e}", flush=True)
            assert rel_rms(f, fo) <= 1e-5
    for acc in (0, 1):
        np.testing.assert_array_equal(forces[(acc, 1)][0], forces[(acc, 0)][0])
    for x, y in zip(forces[(1, 1)][1], forces[(1, 2)][1]):
        np.testing.assert_array_equal(x, y)

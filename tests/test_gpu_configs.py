"""GPU parity and run-length tests at the benchmark configurations' own sizes.

  production  12,800 chain + ~24,000 solvent, 10 nm, FP32 and int64 accumulation
  large       51,200 chain + ~159,000 solvent, 18 nm, int64 accumulation, lists every 20 steps
  replicas    256 uneven systems (chain 128-1,024, solvent 500-6,000, 3-6 nm boxes), one batch
  adjoint     one 2,000-particle replica (512 chain + 1,488 solvent, 4.2 nm, 16 types)

Bars (stated per test):
  * neighbour lists: set-equal to the oracle's (bit-exact FP32 predicate);
  * forces on identical lists: relative rms <= 2e-5 against the FP32 oracle (the FP32 rounding
    scale of these sums is ~1e-5) and <= 1e-4 against the FP64 oracle (the accuracy target);
  * a replica stepped in the 256-batch equals the same replica stepped alone, bit for bit;
  * int64 accumulation gives bit-equal trajectories under two launch shapes;
  * configured run lengths (10,000 / 5,000 / 2,000 steps): LIMA_OK, finite state, temperature
    within 15 K of the target, every compound within 0.9 of half a box of its particle 0 (the
    particle it is re-imaged against) and within 0.5 nm of its list sphere;
  * adjoint gradients within 2e-3 relative (+1e-4 of the largest) of FP64 central differences.
"""
import numpy as np
import pytest

from tests.conftest import as_sets, rel_rms

pytestmark = pytest.mark.gpu


def _kinds():
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SC, KIND_SS
    return (KIND_CC, KIND_CS, KIND_SC, KIND_SS)


def _oracle_reference(system, params):
    """FP32 oracle lists and forces, and FP64 forces on the same lists."""
    from oracle.oracle import Oracle
    o32 = Oracle(system, params, precision=32)
    lists = {k: o32.neighbors(k) for k in _kinds()}
    c32, s32, st = o32.compute_forces()
    assert st == 0
    o32.close()
    o64 = Oracle(system, params, precision=64)
    for k, v in lists.items():
        o64.set_neighbors(k, v)
    c64, s64, st = o64.compute_forces()
    assert st == 0
    o64.close()
    return dict(lists=lists, f32=np.concatenate([c32, s32]), f64=np.concatenate([c64, s64]))


def _set_lists(eng, lists):
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SS
    eng.set_neighbors(cc=lists[KIND_CC], cs=lists[KIND_CS], ss=lists[KIND_SS])


def _assert_lists_equal(eng, lists):
    from oracle.oracle import KIND_CC, KIND_CS, KIND_SS
    cc, cs, ss = eng.neighbors()
    assert as_sets(cc) == as_sets(lists[KIND_CC])
    assert as_sets(cs) == as_sets(lists[KIND_CS])
    assert as_sets(ss) == as_sets(lists[KIND_SS])


def _forces_within_bars(eng, ref, label):
    cf, sf = eng.compute_forces()
    f = np.concatenate([cf, sf])
    assert np.all(np.isfinite(f))
    e32, e64 = rel_rms(f, ref["f32"]), rel_rms(f, ref["f64"])
    print(f"{label}: rel rms vs FP32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e}", flush=True)
    assert e32 <= 2e-5, e32
    assert e64 <= 1e-4, e64


def _compound_geometry(system, cx, chain_begin=0):
    """(largest |x - x_particle0| per dimension / half box, largest distance beyond the list sphere)."""
    L = system.box_len
    w0, ws = 0.0, -np.inf
    for k in range(system.n_compounds):
        a, b = chain_begin + system.compound_offset[k], chain_begin + system.compound_offset[k + 1]
        p = cx[a:b].astype(np.float64)
        d0 = p - p[0]
        d0 -= L * np.round(d0 / L)
        w0 = max(w0, float(np.abs(d0).max()) / (0.5 * L))
        dk = p - p[system.key_particle_index[k]]
        dk -= L * np.round(dk / L)
        ws = max(ws, float(np.sqrt((dk ** 2).sum(1)).max()) - float(system.confining_particle_sphere[k]))
    return w0, ws


def _run_configured_length(eng, systems_, steps, chunk, label):
    sim = eng.simulation
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        eng.step(n)                # raises LimaStatusError on any device error word
        done += n
        cx, cxm, sx, sxm = eng.get_state()
        assert all(np.all(np.isfinite(a)) for a in (cx, cxm, sx, sxm))
        ens = eng.energies()
        worst0, worst_s, worst_t = 0.0, -np.inf, 0.0
        for r, s in enumerate(systems_):
            w0, ws = _compound_geometry(s, cx, int(sim.chain_begin[r]))
            worst0, worst_s = max(worst0, w0), max(worst_s, ws)
            worst_t = max(worst_t, abs(ens[r]["temperature"] - s.thermostat_target))
        print(f"{label} step {done}: |T - T0| <= {worst_t:.1f} K, far from particle 0 {worst0:.3f} of half a box, "
              f"beyond the list sphere {worst_s:.3f} nm", flush=True)
        assert worst_t <= 15.0
        assert worst0 <= 0.9
        assert worst_s <= 0.5


# ------------------------------------------------------------------------------- production
@pytest.fixture(scope="module")
def production():
    from quantom_amd import systems
    from quantom_amd.abi import RunParams
    print("case: production system generation and oracle reference", flush=True)
    s = systems.production()
    return dict(system=s, ref=_oracle_reference(s, RunParams(nlist_interval=0)))


def test_production_initial_lists_bit_exact(production):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: production initial lists", flush=True)
    eng = Engine([production["system"]], RunParams(nlist_interval=0))
    _assert_lists_equal(eng, production["ref"]["lists"])
    eng.close()


@pytest.mark.parametrize("accum", [0, 1])
def test_production_forces(production, accum):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print(f"case: production forces, accumulation={accum}", flush=True)
    eng = Engine([production["system"]], RunParams(nlist_interval=0, accumulation=accum))
    _set_lists(eng, production["ref"]["lists"])
    _forces_within_bars(eng, production["ref"], f"production accumulation={accum}")
    eng.close()


def test_production_rebuild_after_motion_bit_exact(production):
    """Nine steps on the GPU, then the GPU's state into the oracle: both rebuild from the same
    previous lists (the hysteresis depends on them) and must agree."""
    from oracle.oracle import Oracle
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: production rebuild after 9 steps", flush=True)
    s = production["system"]
    eng = Engine([s], RunParams(nlist_interval=0))
    eng.step(9)
    ora = Oracle(s, RunParams(nlist_interval=0), precision=32)
    ora.set_state(*eng.get_state())
    assert ora.rebuild_neighbors() == 0
    eng.rebuild_neighbors()
    _assert_lists_equal(eng, {k: ora.neighbors(k) for k in _kinds()})
    ora.close()
    eng.close()


def test_production_runs_its_configured_10000_steps(production):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    print("case: production, 10,000 steps, FP32, Berendsen", flush=True)
    s = production["system"]
    eng = Engine([s], RunParams(nlist_interval=10))
    _run_configured_length(eng, [s], 10000, 2500, "production")
    eng.close()


# ------------------------------------------------------------------------------------ large
def _large_params(**kw):
    from quantom_amd.abi import RunParams
    p = RunParams(**{"nlist_interval": 20, "accumulation": 1, **kw})
    p.max_compound_solvents = 32768
    return p


@pytest.fixture(scope="module")
def large():
    from quantom_amd import systems
    print("case: large system generation and oracle reference", flush=True)
    s = systems.large()
    return dict(system=s, ref=_oracle_reference(s, _large_params(nlist_interval=0)))


def test_large_lists_and_fixed_point_forces(large):
    from quantom_amd.engine import Engine
    print("case: large initial lists and int64 forces", flush=True)
    eng = Engine([large["system"]], _large_params(nlist_interval=0))
    _assert_lists_equal(eng, large["ref"]["lists"])
    _forces_within_bars(eng, large["ref"], "large accumulation=1")
    eng.close()


def test_large_fixed_point_trajectory_is_shape_independent(large, monkeypatch, tmp_path):
    """int64 force sums are exact and order-free: the default launch shape and a forced other
    one (1 solvent per lane, chunks of 512 chain particles, listed grid 2048) step bit-equal."""
    from quantom_amd.engine import Engine
    print("case: large, two launch shapes, 40 steps", flush=True)
    s = large["system"]
    key = f"acc1_sol1_{(s.n_chain + 64) // 128}x{(s.n_solvents + 256) // 512}:1"
    path = tmp_path / "tune.txt"
    path.write_text(f"{key} 1 512 2048 1\n")
    monkeypatch.setenv("LIMA_AUTOTUNE_CACHE", str(path))
    a = Engine([s], _large_params())
    b = Engine([s], _large_params())
    cfg = b.autotune(10)
    assert (cfg["solvents_per_thread"], cfg["chunk_particles"]) == (1, 512)
    assert (cfg["solvents_per_thread"], cfg["chunk_particles"]) != (a.config()["solvents_per_thread"],
                                                                    a.config()["chunk_particles"])
    a.step(40)
    b.step(40)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


def test_large_runs_its_configured_5000_steps(large):
    from quantom_amd.engine import Engine
    print("case: large, 5,000 steps, int64 accumulation, Berendsen", flush=True)
    s = large["system"]
    eng = Engine([s], _large_params())
    _run_configured_length(eng, [s], 5000, 1250, "large")
    eng.close()


# --------------------------------------------------------------------------------- replicas
@pytest.fixture(scope="module")
def replica_set():
    from quantom_amd import systems
    print("case: 256-replica set generation", flush=True)
    syss = systems.replicas(256)
    L = [s.box_len for s in syss]
    # a sample: the smallest and largest boxes, the most particles, and three seeded picks
    n = [s.n_chain + s.n_solvents for s in syss]
    rng = np.random.default_rng(11)
    sample = sorted({int(np.argmin(L)), int(np.argmax(L)), int(np.argmax(n)), *map(int, rng.choice(256, 3, replace=False))})
    return dict(systems=syss, sample=sample)


def test_replica_batch_equals_each_replica_alone(replica_set):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    syss, sample = replica_set["systems"], replica_set["sample"]
    print(f"case: 256-replica batch vs single engines for replicas {sample}, 25 steps", flush=True)
    batch = Engine(syss, RunParams())
    bf = batch.compute_forces()
    batch.step(25)
    cx, cxm, sx, sxm = batch.get_state()
    sim = batch.simulation
    for r in sample:
        c0, c1 = int(sim.chain_begin[r]), int(sim.chain_begin[r + 1])
        s0, s1 = int(sim.solvent_begin[r]), int(sim.solvent_begin[r + 1])
        e = Engine([syss[r]], RunParams())
        f = e.compute_forces()
        assert np.array_equal(f[0], bf[0][c0:c1]) and np.array_equal(f[1], bf[1][s0:s1]), r
        e.step(25)
        x, xm, s, sm = e.get_state()
        for got, want in ((x, cx[c0:c1]), (xm, cxm[c0:c1]), (s, sx[s0:s1]), (sm, sxm[s0:s1])):
            assert np.array_equal(got, want), r
        e.close()
    batch.close()


def test_replica_sample_lists_and_forces_vs_oracle(replica_set):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    syss, sample = replica_set["systems"], replica_set["sample"]
    for r in sample:
        print(f"case: replica {r} (box {syss[r].box_len:.2f} nm, chain {syss[r].n_chain}, "
              f"solvent {syss[r].n_solvents}) lists and forces", flush=True)
        ref = _oracle_reference(syss[r], RunParams(nlist_interval=0))
        e = Engine([syss[r]], RunParams(nlist_interval=0))
        _assert_lists_equal(e, ref["lists"])
        _forces_within_bars(e, ref, f"replica {r}")
        e.close()


def test_replica_batch_runs_its_configured_2000_steps(replica_set):
    from quantom_amd.abi import RunParams
    from quantom_amd.engine import Engine
    syss = replica_set["systems"]
    print("case: 256-replica batch, 2,000 steps, FP32, Berendsen per replica", flush=True)
    eng = Engine(syss, RunParams())
    _run_configured_length(eng, syss, 2000, 500, "replicas")
    eng.close()


# ---------------------------------------------------------------------------------- adjoint
def test_adjoint_replica_gradient_matches_fd_of_fp64_oracle():
    """One replica of the adjoint workload, 20 frames (checkpoints every 10, so the backward pass
    recomputes a segment), a loss over per-frame potential energy and forces against targets at
    perturbed parameters; every used sigma/epsilon and the four bonded constants against central
    differences of the FP64 oracle's loss."""
    from quantom_amd import systems
    from quantom_amd.engine import Engine
    from tests.test_adjoint import PARAMS, base_params, oracle_frames, oracle_loss, used_types, with_params
    K = 20
    s = systems.adjoint_replicas(1)[0]
    print(f"case: adjoint replica ({s.n_chain} chain + {s.n_solvents} solvent), {K} frames", flush=True)
    v0 = base_params(s)
    vt = v0.copy()
    rng = np.random.default_rng(0x51A7 + 4000)
    types = used_types(s)
    for t in types:
        vt[t] *= 1.0 + 0.03 * rng.uniform(-1, 1)
        vt[32 + t] *= 1.0 + 0.10 * rng.uniform(-1, 1)
    fct, fst, et = oracle_frames(with_params(s, vt.astype(np.float32).astype(np.float64)), K)
    bfc, bfs, bpot = oracle_frames(s, K)
    beta = 1.0 / float((fct ** 2).sum() + (fst ** 2).sum())
    alpha = beta * float(((bfc - fct) ** 2).sum() + ((bfs - fst) ** 2).sum()) / float(((bpot - et) ** 2).sum())
    e = Engine([s], PARAMS)
    loss, grad = e.adjoint(K, [alpha], [beta], et[:, None], fct.astype(np.float32), fst.astype(np.float32),
                           checkpoint_interval=10)
    g = grad[0]
    l0 = oracle_loss(s, K, alpha, beta, et, fct.astype(np.float32), fst.astype(np.float32))
    print(f"loss: adjoint {loss[0]:.9e} oracle {l0:.9e}", flush=True)
    assert abs(loss[0] - l0) <= 1e-4 * abs(l0)
    idx = [i for t in types for i in (t, 32 + t)] + [64, 65, 66, 67]
    fd = {}
    for i in idx:
        h = (1e-4 if i != 67 else 1e-3) * max(abs(v0[i]), 1e-3)
        vp, vm = v0.copy(), v0.copy()
        vp[i] = np.float32(v0[i] + h)
        vm[i] = np.float32(v0[i] - h)
        lp = oracle_loss(with_params(s, vp), K, alpha, beta, et, fct.astype(np.float32), fst.astype(np.float32))
        lm = oracle_loss(with_params(s, vm), K, alpha, beta, et, fct.astype(np.float32), fst.astype(np.float32))
        fd[i] = (lp - lm) / (float(vp[i]) - float(vm[i]))
    scale = max(abs(x) for x in fd.values())
    worst = 0.0
    for i in idx:
        rel = abs(g[i] - fd[i]) / max(abs(fd[i]), 1e-30)
        worst = max(worst, abs(g[i] - fd[i]) / (2e-3 * abs(fd[i]) + 1e-4 * scale))
        print(f"param {i:2d}  adjoint {g[i]: .6e}  fd {fd[i]: .6e}  rel {rel:.2e}")
    print(f"worst error / bar: {worst:.3f}", flush=True)
    for i in idx:
        assert abs(g[i] - fd[i]) <= 2e-3 * abs(fd[i]) + 1e-4 * scale, f"parameter {i}: adjoint {g[i]} fd {fd[i]}"
    e.close()


def test_adjoint_replica_200_frames_matches_converged_fd():
    """The adjoint workload's full horizon: 200 frames with checkpoints every 20.  FP64 central
    differences of the oracle's loss are converged there (steps of 1e-3 to 1e-6 relative agree
    to ~4e-3); the adjoint must agree with the 1e-5 step to 1.5e-2 relative for two types' sigma
    and epsilon and the four bonded constants."""
    from quantom_amd import systems
    from quantom_amd.engine import Engine
    from tests.test_adjoint import PARAMS, base_params, oracle_frames, oracle_loss, used_types, with_params
    K = 200
    s = systems.adjoint_replicas(1)[0]
    print(f"case: adjoint replica, {K} frames, checkpoints every 20", flush=True)
    v0 = base_params(s)
    vt = v0.copy()
    rng = np.random.default_rng(0x51A7 + 4000)
    types = used_types(s)
    for t in types:
        vt[t] *= 1.0 + 0.03 * rng.uniform(-1, 1)
        vt[32 + t] *= 1.0 + 0.10 * rng.uniform(-1, 1)
    fct, fst, et = oracle_frames(with_params(s, vt.astype(np.float32).astype(np.float64)), K)
    bfc, bfs, bpot = oracle_frames(s, K)
    beta = 1.0 / float((fct ** 2).sum() + (fst ** 2).sum())
    alpha = beta * float(((bfc - fct) ** 2).sum() + ((bfs - fst) ** 2).sum()) / float(((bpot - et) ** 2).sum())
    fct32, fst32 = fct.astype(np.float32), fst.astype(np.float32)
    e = Engine([s], PARAMS)
    loss, grad = e.adjoint(K, [alpha], [beta], et[:, None], fct32, fst32, checkpoint_interval=20)
    e.close()
    l0 = oracle_loss(s, K, alpha, beta, et, fct32, fst32)
    print(f"loss: adjoint {loss[0]:.9e} oracle {l0:.9e}", flush=True)
    assert abs(loss[0] - l0) <= 1e-3 * abs(l0)
    for i in (types[1], types[5], 32 + types[1], 32 + types[5], 64, 65, 66, 67):
        h = 1e-5 * max(abs(v0[i]), 1e-3)
        vp, vm = v0.copy(), v0.copy()
        vp[i] = np.float32(v0[i] + h)
        vm[i] = np.float32(v0[i] - h)
        fd = (oracle_loss(with_params(s, vp), K, alpha, beta, et, fct32, fst32) -
              oracle_loss(with_params(s, vm), K, alpha, beta, et, fct32, fst32)) / (float(vp[i]) - float(vm[i]))
        rel = abs(grad[0][i] - fd) / abs(fd)
        print(f"param {i:2d}  adjoint {grad[0][i]: .6e}  fd {fd: .6e}  rel {rel:.2e}", flush=True)
        assert rel <= 1.5e-2, f"parameter {i}: adjoint {grad[0][i]} fd {fd}"
